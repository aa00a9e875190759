// Launch parameters chosen by the project's autotuner
// (ultrafastsecp256k1_amd/autotune.py), read at every context creation from a
// small text file of "key value" lines:
//   gcomb_w <4..20>                 width of the fixed-base G comb table
//   verify_blocks_per_cu <1..8>     persistent verify grid (0 = occupancy)
//   msm_window <log2 n> <c>         Pippenger window for n in [2^k, 2^(k+1))
//   msm_split_from <log2 n>         MSMs of at least 2^k terms run the high
//                                   windows' sums on a second stream (40 = never)
//   server_groups <1..4>            launch groups in flight in the server
// The file is UFSECP_TUNING_FILE if set, else ufsecp_tuning.txt beside the
// library.  Missing keys keep the built-in defaults; an environment variable
// (UFSECP_GCOMB_W) still overrides the comb width.
#pragma once

namespace ufs {

struct Tuning {
    int gcomb_w = 16;
    int verify_blocks_per_cu = 0;
    int server_groups = 2;
    int msm_window[40] = {};   // 0 = size formula
    int msm_split_from = 40;
    bool loaded = false;
    char source[512] = {};
};

Tuning load_tuning();

}  // namespace ufs
