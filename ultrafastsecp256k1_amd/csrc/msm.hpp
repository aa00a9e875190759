// Pippenger multi-scalar multiplication on gfx950.
//
// The reference's ufsecp_gpu_msm (src/gpu/src/gpu_backend_cuda.cu:1493-1572)
// runs one full GLV scalar multiplication per term followed by a block tree
// and a single-thread serial sum.  This engine is a bucket method with a
// device-side sort by bucket instead:
//   1. parse + decompress the bases into P and phi(P) = (beta x, y), strict
//      SEC1 points checked per term, on a second stream beside step 2
//      (first failing term decides the error, as in the reference CPU
//      ufsecp_multi_scalar_mul, impl/ufsecp_taproot.cpp:957-991);
//   2. scalars < n checked and split by GLV, k = +-k1 +- lambda k2 with
//      k1, k2 < 2^128, into the terms (k1, +-P) and (k2, +-phi(P));
//      c-bit digits of the 128-bit halves in W = ceil(128/c) windows: signed,
//      |d| <= 2^(c-1), below the top window, unsigned (bits + carry, up to
//      2^c) in the top one; one 8-byte entry (bucket id, term index | sign)
//      per (half-term, window) written in window-major order (a zero digit
//      is a skip entry), then a stable radix sort on the c low bucket-id
//      bits, which differ between the buckets of one window, so every
//      bucket's entries end up contiguous;
//   3. bucket accumulation load-balanced by fixed-size chunks of the sorted
//      entries (mixed additions); a bucket that spans a few chunks is finished
//      by one lane joining its chunk partials (usually the last step), longer
//      runs go through chunked reduction levels, so no lane adds more than 32
//      terms even when one bucket holds half of all terms;
//   4. per-window weighted sum sum_b (b+1) B_b by segments (running sums plus
//      one small scalar multiplication per segment) and a block tree;
//   5. Horner combination of the windows, affine normalisation, compression.
// The result is an exact group element, so it is bit-identical to any
// correct CPU multi-scalar multiplication.
#pragma once
#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

#include "backend.hpp"

namespace ufs {

struct MsmPlan {
    int c;          // window width
    int windows;    // ceil(128 / c) over the GLV halves
    int groups;     // windows + 1 bucket groups: the top window spans two
    int buckets;    // 2^(c-1) per group
    int seg_len;    // buckets per reduction segment
};

int msm_default_window(size_t n);
// c <= 0: the tuned window for n's size band, else the size formula.
MsmPlan msm_plan(size_t n, int c, const Tuning* tun = nullptr);

// Device launchers (msm.hip)
// bases (33-byte compressed, or 64-byte x || y when affine) -> P, phi(P)
hipError_t launch_msm_points(const uint8_t* points, bool affine, uint32_t n, void* pts, uint32_t* err,
                             unsigned cus, hipStream_t st);
// window sums of the bucket groups w0 .. w0 + gn - 1 into seg[0 .. gn * T)
hipError_t launch_msm_segments(const void* bsum, void* seg, const MsmPlan& p, uint32_t w0, uint32_t gn,
                               hipStream_t st);
hipError_t launch_msm_reduce(const void* in, void* out, uint32_t groups, uint32_t len, uint32_t out_len,
                             hipStream_t st);
hipError_t launch_msm_final(const void* rows, uint32_t nwin, bool top, uint32_t c, uint32_t shift, const void* extra,
                            const uint32_t* err, uint8_t* out33, int32_t* status, void* part, hipStream_t st);
hipError_t launch_u32_fill(uint32_t* p, uint32_t v, size_t n, hipStream_t st);

// Device buffers of one MSM, as parts of one byte range: offsets are
// 256-byte aligned; total is the range's size.
struct MsmLayout {
    static constexpr int kParts = 13;
    size_t off[kParts], size[kParts];
    size_t total, max_entries, chunks0, slots0, sort_bytes;
    uint32_t seg_chunk;   // entries per lane of the first accumulation stage
};

class MsmEngine {
public:
    explicit MsmEngine(RocmBackend* be) : be_(be) {}
    ~MsmEngine();
    // Context set-up (the backend's device current): the completion event and
    // the device address of the no-allocation scratch.
    GpuError init();
    // Host buffers; synchronous.  c <= 0 picks the default window.
    GpuError msm_host(const uint8_t* scalars32, const uint8_t* points33, size_t n, uint8_t* out33, int c);
    // Device buffers; asynchronous on st.  *d_status receives the ABI code.
    // Points are 33-byte compressed, or 64-byte affine x || y when `affine`.
    GpuError msm_device(const uint8_t* d_scalars32, const uint8_t* d_points, bool affine, size_t n,
                        uint8_t* d_out33, int32_t* d_status, int c, hipStream_t st);
    // msm_device without any allocation: buffers carved from the library's
    // static per-device scratch (32 MiB), n <= 16384, 33-byte points.
    // Calls through this engine are ordered; calls on one GPU from other
    // contexts share the same scratch and must not overlap.
    GpuError msm_device_noalloc(const uint8_t* d_scalars32, const uint8_t* d_points33, size_t n, uint8_t* d_out33,
                                int32_t* d_status, int c, hipStream_t st);
    // BIP-340 batch check by a random linear combination, one MSM of 2n + 1
    // terms per chunk (host buffers, synchronous): *all_valid = 1 when every
    // signature verifies, 0 otherwise (a wrong 1 has probability about 2^-128
    // per call).  Defined in batch_rlc.cpp.
    GpuError schnorr_batch_check(const uint8_t* msgs, const uint8_t* pubx, const uint8_t* sigs, size_t n,
                                 int* all_valid);

private:
    GpuError enqueue(const uint8_t* d_scalars32, const uint8_t* d_points, bool affine, size_t n, uint8_t* d_out33,
                     int32_t* d_status, int c, hipStream_t st);
    GpuError enqueue_work(const uint8_t* d_scalars32, const uint8_t* d_points, bool affine, size_t n,
                          uint8_t* d_out33, int32_t* d_status, int c, hipStream_t st);
    // seg_chunk 0: the width for this size (msm_seg_chunk)
    GpuError msm_layout(size_t n, const MsmPlan& p, hipStream_t st, MsmLayout& L, uint32_t seg_chunk = 0);
    // split: windows [W/2, W) go first and their sums and Horner share run
    // on the engine's second stream beside the low windows' bucket phase
    GpuError run_msm(const MsmLayout& L, const MsmPlan& p, void* const* ptr, const uint8_t* d_scalars32,
                     const uint8_t* d_points, bool affine, size_t n, uint8_t* d_out33, int32_t* d_status,
                     bool split, hipStream_t st);
    // Sort, accumulate and reduce the buckets of windows [wa, wb) into bsum.
    GpuError bucket_phase(const MsmLayout& L, const MsmPlan& p, void* const* ptr, size_t n, int wa, int wb,
                          hipEvent_t points, hipStream_t st);
    // Window sums of windows [wa, wb) (and the top window's second bucket
    // group when wb is the last window); *rows receives one point per group.
    GpuError window_phase(const MsmPlan& p, void* const* ptr, int wa, int wb, hipStream_t st, const void** rows);
    hipError_t ensure_aux();
    // Every call shares the scratch buffers below.  order(st) makes st wait
    // for the previous call when that ran on another stream; grow() waits on
    // the host for the previous call before a buffer it may still read is
    // reallocated; finish(st) marks the end of this call's work on st.
    hipError_t order(hipStream_t st);
    hipError_t grow(DeviceBuffer& b, size_t bytes);
    hipError_t finish(hipStream_t st);
    RocmBackend* be_;
    hipEvent_t done_ = nullptr;
    // second stream (the bases' decompression, the split form's high
    // windows) with its fork and join events; ready_ marks the bases done
    hipStream_t aux_ = nullptr;
    hipEvent_t fork_ = nullptr, join_ = nullptr, ready_ = nullptr;
    hipStream_t last_stream_ = nullptr;
    bool pending_ = false;
    DeviceBuffer in_s_, in_p_, pts_, entries_, bsum_, seg_, red_a_, red_b_, err_,
        out_, slot_keys_[2], slot_part_[2];
    DeviceBuffer rlc_in_, rlc_terms_, rlc_part_, sorted_, sort_tmp_;
};

}  // namespace ufs
