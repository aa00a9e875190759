// Provider for the reference engine's GPU column-verify hook.
//
// The reference CPU engine (secp256k1::ecdsa_batch_verify_opaque_columns /
// schnorr_batch_verify_bip340_columns, src/cpu/src/batch_verify.cpp:851-960)
// consults a process-wide hook before its CPU path
// (src/cpu/include/secp256k1/batch_verify.hpp:149-168).  The reference fills
// it from its GPU host layer at static-init time
// (src/gpu/src/gpu_engine_hook.cpp:60-107).  This file is the MI355X
// provider: the same hook contract (1 handled / all valid, 0 handled / some
// invalid, -1 declined -> CPU fallback; every operational error is a decline),
// one lazily created device context for the process, serialised by a mutex.
//
// Installation: at library load, if the engine's installer is visible in the
// process (the engine linked into the executable or loaded RTLD_GLOBAL), the
// hook installs itself -- the reference's self-installing behaviour.
// UFSECP_GPU_ENGINE_HOOK=0 disables that; ufsecp_gpu_ext_engine_hook_install()
// retries later (an engine loaded after this library).
#include <dlfcn.h>

#include <atomic>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <mutex>

#include "backend.hpp"
#include "ufsecp/ufsecp_gpu_ext.h"

namespace {

using HookFn = int (*)(int, const uint8_t*, const uint8_t*, const uint8_t*, size_t, uint8_t*) noexcept;
using InstallFn = HookFn (*)(HookFn) noexcept;

// secp256k1::install_gpu_columns_verify_hook(GpuColumnsVerifyHook) as the
// Itanium C++ ABI names it (batch_verify.hpp:168; the hook type is a noexcept
// function pointer).
constexpr const char* kInstallSymbol = "_ZN9secp256k131install_gpu_columns_verify_hookEPDoFiiPKhS1_S1_mPhE";

std::mutex g_mtx;
// Kept until the process ends (never destroyed: tearing a device context down
// from a static destructor can run after the HIP runtime has unloaded).
ufs::RocmBackend* g_backend = nullptr;
bool g_probed = false;
std::atomic<uint64_t> g_calls{0}, g_handled{0}, g_declined{0}, g_items{0};

// Called with g_mtx held.
ufs::RocmBackend* backend() {
    if (!g_probed) {
        g_probed = true;
        if (ufs::RocmBackend::device_count() > 0) {
            auto b = std::make_unique<ufs::RocmBackend>();
            if (b->init(0) == ufs::GpuError::Ok && b->is_ready()) g_backend = b.release();
        }
    }
    return g_backend;
}

struct AutoInstall {
    AutoInstall() {
        const char* e = std::getenv("UFSECP_GPU_ENGINE_HOOK");
        if (e && std::strcmp(e, "0") == 0) return;
        (void)ufsecp_gpu_ext_engine_hook_install();
    }
};
AutoInstall g_auto_install;

}  // namespace

extern "C" {

int ufsecp_gpu_ext_engine_columns_hook(int kind, const uint8_t* digests32, const uint8_t* keys,
                                       const uint8_t* sigs64, size_t count, uint8_t* out_results) noexcept {
    g_calls.fetch_add(1, std::memory_order_relaxed);
    try {
        if ((kind != 0 && kind != 1) || !digests32 || !keys || !sigs64 || !out_results || count > (1u << 26)) {
            g_declined.fetch_add(1, std::memory_order_relaxed);
            return -1;
        }
        std::lock_guard<std::mutex> lk(g_mtx);
        ufs::RocmBackend* b = backend();
        if (!b) {
            g_declined.fetch_add(1, std::memory_order_relaxed);
            return -1;
        }
        if (count == 0) {
            g_handled.fetch_add(1, std::memory_order_relaxed);
            return 1;
        }
        const ufs::GpuError e = kind == 0 ? b->ecdsa_verify_opaque_columns(digests32, keys, sigs64, count, out_results)
                                          : b->schnorr_verify_batch(digests32, keys, sigs64, count, out_results, false);
        if (e != ufs::GpuError::Ok) {
            g_declined.fetch_add(1, std::memory_order_relaxed);
            return -1;
        }
        g_handled.fetch_add(1, std::memory_order_relaxed);
        g_items.fetch_add(count, std::memory_order_relaxed);
        for (size_t i = 0; i < count; ++i)
            if (out_results[i] == 0) return 0;
        return 1;
    } catch (...) {
        g_declined.fetch_add(1, std::memory_order_relaxed);
        return -1;
    }
}

int ufsecp_gpu_ext_engine_hook_install(void) {
    void* sym = dlsym(RTLD_DEFAULT, kInstallSymbol);
    if (!sym) return 0;
    InstallFn install = reinterpret_cast<InstallFn>(sym);
    (void)install(&ufsecp_gpu_ext_engine_columns_hook);
    return 1;
}

void ufsecp_gpu_ext_engine_hook_stats(ufsecp_gpu_engine_hook_stats_t* out) {
    if (!out) return;
    out->calls = g_calls.load();
    out->handled = g_handled.load();
    out->declined = g_declined.load();
    out->items = g_items.load();
}

}  // extern "C"
