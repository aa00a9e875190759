// Tuning file loader.  See tuning.hpp.
#include "tuning.hpp"

#include <dlfcn.h>

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>

namespace ufs {

namespace {

std::string default_path() {
    Dl_info info;
    if (dladdr((void*)&default_path, &info) && info.dli_fname) {
        std::string p = info.dli_fname;
        const size_t slash = p.rfind('/');
        if (slash != std::string::npos) return p.substr(0, slash + 1) + "ufsecp_tuning.txt";
    }
    return "ufsecp_tuning.txt";
}

void load(Tuning& t) {
    const char* env = std::getenv("UFSECP_TUNING_FILE");
    const std::string path = env && *env ? std::string(env) : default_path();
    std::FILE* f = std::fopen(path.c_str(), "r");
    if (!f) return;
    char line[256];
    while (std::fgets(line, sizeof(line), f)) {
        char key[64];
        int a = 0, b = 0;
        const int got = std::sscanf(line, "%63s %d %d", key, &a, &b);
        if (got < 2 || key[0] == '#') continue;
        if (!std::strcmp(key, "gcomb_w") && a >= 4 && a <= 20) t.gcomb_w = a;
        else if (!std::strcmp(key, "verify_blocks_per_cu") && a >= 0 && a <= 8) t.verify_blocks_per_cu = a;
        else if (!std::strcmp(key, "server_groups") && a >= 1 && a <= 4) t.server_groups = a;
        else if (!std::strcmp(key, "msm_split_from") && a >= 0 && a <= 40) t.msm_split_from = a;
        else if (!std::strcmp(key, "msm_window") && got == 3 && a >= 0 && a < 40 && b >= 2 && b <= 20)
            t.msm_window[a] = b;
    }
    std::fclose(f);
    t.loaded = true;
    std::snprintf(t.source, sizeof(t.source), "%s", path.c_str());
}

}  // namespace

Tuning load_tuning() {
    Tuning t;
    load(t);
    return t;
}

}  // namespace ufs
