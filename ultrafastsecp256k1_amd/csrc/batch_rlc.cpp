// BIP-340 batch verification by a random linear combination: the term
// kernel (kernels_rlc.hip) and one Pippenger MSM per chunk.  The reference
// computes the same check on the CPU (schnorr_batch_verify_impl,
// src/cpu/src/batch_verify.cpp:324-424); this is its device form, for
// callers that need only "every signature valid or not" for a large batch.
#include <sys/random.h>

#include <cstring>

#include "msm.hpp"

namespace ufs {

namespace {

// Items per MSM: 2^24 signatures are 2^25 + 1 terms.
constexpr size_t kRlcChunk = size_t(1) << 24;

bool fill_seed(RlcSeed& seed) {
    uint8_t b[32];
    size_t got = 0;
    while (got < sizeof(b)) {
        const ssize_t r = getrandom(b + got, sizeof(b) - got, 0);
        if (r <= 0) return false;
        got += (size_t)r;
    }
    for (int k = 0; k < 8; ++k)
        seed.w[k] = ((uint32_t)b[4 * k] << 24) | ((uint32_t)b[4 * k + 1] << 16) | ((uint32_t)b[4 * k + 2] << 8) |
                    (uint32_t)b[4 * k + 3];
    return true;
}

}  // namespace

GpuError MsmEngine::schnorr_batch_check(const uint8_t* msgs, const uint8_t* pubx, const uint8_t* sigs, size_t n,
                                        int* all_valid) {
    if (!be_->is_ready()) return be_->fail(GpuError::Device, "context not initialised");
    *all_valid = 0;
    if (n == 0) {
        be_->clear_error();
        return GpuError::Ok;
    }
    hipStream_t st = be_->stream();
    hipError_t e = hipSetDevice(be_->device());
    if (e != hipSuccess) return be_->fail(GpuError::Device, "hipSetDevice", e);
    const size_t cmax = n < kRlcChunk ? n : kRlcChunk;
    const size_t terms = 2 * cmax + 1;
    // inputs: msgs | pubx | sigs; terms: scalars | points; scratch: partial sums,
    // then 64 bytes of MSM output (status at +48) and the s-range flag (+56)
    if ((e = order(st)) != hipSuccess) return be_->fail(GpuError::Backend, "MSM stream order", e);
    if ((e = grow(rlc_in_, cmax * 128)) != hipSuccess || (e = grow(rlc_terms_, terms * 65)) != hipSuccess ||
        (e = grow(rlc_part_, (size_t)rlc_partial_words((uint32_t)cmax) * 4 + 64)) != hipSuccess)
        return be_->fail(GpuError::Memory, "hipMalloc(batch check buffers)", e);
    uint8_t* d_in = (uint8_t*)rlc_in_.ptr;
    uint8_t* d_scl = (uint8_t*)rlc_terms_.ptr;
    uint32_t* d_part = (uint32_t*)rlc_part_.ptr;
    uint8_t* d_tail = (uint8_t*)rlc_part_.ptr + (size_t)rlc_partial_words((uint32_t)cmax) * 4;
    for (size_t off = 0; off < n; off += cmax) {
        const size_t c = n - off < cmax ? n - off : cmax;
        RlcSeed seed;
        if (!fill_seed(seed)) return be_->fail(GpuError::Internal, "getrandom failed");
        e = hipMemcpyAsync(d_in, msgs + off * 32, c * 32, hipMemcpyHostToDevice, st);
        if (e == hipSuccess) e = hipMemcpyAsync(d_in + c * 32, pubx + off * 32, c * 32, hipMemcpyHostToDevice, st);
        if (e == hipSuccess) e = hipMemcpyAsync(d_in + c * 64, sigs + off * 64, c * 64, hipMemcpyHostToDevice, st);
        if (e == hipSuccess) e = launch_u32_fill((uint32_t*)(d_tail + 56), 0, 1, st);
        if (e != hipSuccess) return be_->fail(GpuError::Memory, "copy batch check inputs", e);
        // the terms of this chunk are written packed: scalars [0, 2c], points [0, 2c]
        uint8_t* scl = d_scl;
        uint8_t* pts = d_scl + (2 * c + 1) * 32;
        e = launch_rlc_terms(d_in, d_in + c * 32, d_in + c * 64, (uint32_t)c, seed, be_->bip340_midstate(), scl, pts,
                             d_part, (uint32_t*)(d_tail + 56), st);
        if (e != hipSuccess) return be_->fail(GpuError::Launch, "batch check term launch", e);
        const GpuError r = enqueue(scl, pts, false, 2 * c + 1, d_tail, (int32_t*)(d_tail + 48), 0, st);
        if (r != GpuError::Ok) return r;
        uint8_t host[64];
        e = hipMemcpyAsync(host, d_tail, 64, hipMemcpyDeviceToHost, st);
        if (e == hipSuccess) e = hipStreamSynchronize(st);
        if (e != hipSuccess) return be_->fail(GpuError::Backend, "batch check synchronize", e);
        int32_t status;
        uint32_t srange;
        std::memcpy(&status, host + 48, 4);
        std::memcpy(&srange, host + 56, 4);
        if (status == (int32_t)GpuError::Internal) return be_->fail(GpuError::Internal, "MSM bucket id out of range");
        // infinity (ERR_ARITH) is the all-valid outcome; a finite sum or a
        // point that does not decompress means some signature is invalid
        if (srange != 0 || status != (int32_t)GpuError::Arith) {
            be_->clear_error();
            return GpuError::Ok;
        }
    }
    *all_valid = 1;
    be_->clear_error();
    return GpuError::Ok;
}

}  // namespace ufs
