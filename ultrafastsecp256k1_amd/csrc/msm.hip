// Pippenger MSM kernels and host engine.  See msm.hpp for the pipeline.
#include <cstring>

#include <hipcub/hipcub.hpp>

#include "msm.hpp"
#include "secp_point.h"

namespace ufs {

constexpr int kMsmBlock = 256;
constexpr int kReduceBlock = 256;

UFS_DEV void msm_load_scalar(u32 k[8], const u8* p) {
    const uint4* q = reinterpret_cast<const uint4*>(p);
    const uint4 a = q[0], b = q[1];
    const u32 w[8] = {a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w};
#pragma unroll
    for (int i = 0; i < 8; ++i) k[i] = __builtin_bswap32(w[7 - i]);
}

UFS_DEV u32 msm_bits(const u32 k[8], int pos, int w) {
    const int wi = pos >> 5, sh = pos & 31;
    u32 lo = 0, hi = 0;
#pragma unroll
    for (int i = 0; i < 8; ++i) {
        lo = (wi == i) ? k[i] : lo;
        hi = (wi + 1 == i) ? k[i] : hi;
    }
    const u64 v = ((u64)hi << 32) | lo;
    return (u32)(v >> sh) & ((1u << w) - 1u);
}

// Signed digit of window w given the running carry: |d| <= 2^(c-1).
UFS_DEV int msm_digit(const u32 k[8], int w, int c, u32& carry) {
    const u32 raw = msm_bits(k, w * c, c) + carry;
    const u32 half = 1u << (c - 1);
    if (raw > half) {
        carry = 1;
        return (int)raw - (1 << c);
    }
    carry = 0;
    return (int)raw;
}

__global__ void __launch_bounds__(kMsmBlock)
k_msm_prepare(const u8* __restrict__ scalars32, const u8* __restrict__ points33, u32 n, aff* __restrict__ pts,
              sc* __restrict__ scl, u32* __restrict__ counts, int c, int W, u32* __restrict__ err) {
    const u32 i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    u32 k[8];
    msm_load_scalar(k, scalars32 + 32ull * i);
    const bool sok = u256_lt_n(k);
    const u8* pp = points33 + 33ull * i;
    const u32 prefix = pp[0];
    u32 x[8];
#pragma unroll
    for (int l = 0; l < 8; ++l) {
        const u8* q = pp + 1 + 28 - 4 * l;
        x[l] = ((u32)q[0] << 24) | ((u32)q[1] << 16) | ((u32)q[2] << 8) | (u32)q[3];
    }
    aff P;
    bool pok = lift_x(P, x, prefix & 1u);
    pok &= (prefix == 2u) | (prefix == 3u);
    if (!sok) {
        atomicMin(err, 2u * i);
    } else if (!pok) {
        atomicMin(err, 2u * i + 1u);
    }
    pts[i] = P;
    sc s;
#pragma unroll
    for (int l = 0; l < 8; ++l) s.v[l] = k[l];
    scl[i] = s;
    const u32 B = 1u << (c - 1);
    u32 carry = 0;
    for (int w = 0; w < W; ++w) {
        const int d = msm_digit(k, w, c, carry);
        if (d != 0) atomicAdd(&counts[w * B + (u32)(d < 0 ? -d : d) - 1u], 1u);
    }
}

__global__ void __launch_bounds__(kMsmBlock)
k_msm_scatter(const sc* __restrict__ scl, u32 n, u32* __restrict__ cursor, u32* __restrict__ entries, int c,
              int W) {
    const u32 i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const sc s = scl[i];
    const u32 B = 1u << (c - 1);
    u32 carry = 0;
    for (int w = 0; w < W; ++w) {
        const int d = msm_digit(s.v, w, c, carry);
        if (d != 0) {
            const u32 b = w * B + (u32)(d < 0 ? -d : d) - 1u;
            const u32 pos = atomicAdd(&cursor[b], 1u);
            entries[pos] = i | (d < 0 ? 0x80000000u : 0u);
        }
    }
}

__global__ void __launch_bounds__(kMsmBlock)
k_msm_buckets(const aff* __restrict__ pts, const u32* __restrict__ offsets, const u32* __restrict__ entries,
              jac* __restrict__ bsum, u32 nb) {
    const u32 b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= nb) return;
    const u32 beg = offsets[b], end = offsets[b + 1];
    jac acc;
    jac_set_inf(acc);
    aff nxt;
    u32 ent = 0;
    if (beg < end) {
        ent = entries[beg];
        nxt = pts[ent & 0x7fffffffu];
    }
    for (u32 e = beg; e < end; ++e) {
        aff P = nxt;
        const bool neg = (ent >> 31) != 0;
        if (e + 1 < end) {
            ent = entries[e + 1];
            nxt = pts[ent & 0x7fffffffu];
        }
        if (neg) fe_neg(P.y, P.y);
        jac_madd(acc, acc, P);
    }
    bsum[b] = acc;
}

// r = m * a for a small non-negative integer m (double-and-add).
UFS_DEV void jac_mul_small(jac& r, const jac& a, u32 m) {
    jac acc;
    jac_set_inf(acc);
    if (m != 0) {
        const int top = 31 - __builtin_clz(m);
        acc = a;
        for (int b = top - 1; b >= 0; --b) {
            jac_dbl(acc, acc);
            if ((m >> b) & 1u) jac_add(acc, acc, a);
        }
    }
    r = acc;
}

// Lane (w, s): P = sum_{j<L} (s*L + j + 1) * B[w][s*L + j].
__global__ void __launch_bounds__(kMsmBlock)
k_msm_segments(const jac* __restrict__ bsum, jac* __restrict__ seg, u32 W, u32 B, u32 L) {
    const u32 T = B / L;
    const u32 id = blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= W * T) return;
    const u32 w = id / T, s = id % T;
    const jac* bb = bsum + (size_t)w * B + (size_t)s * L;
    jac R, S;
    jac_set_inf(R);
    jac_set_inf(S);
    for (int j = (int)L - 1; j >= 0; --j) {
        jac_add(R, R, bb[j]);
        jac_add(S, S, R);
    }
    jac M;
    jac_mul_small(M, R, s * L);
    jac_add(S, S, M);
    seg[id] = S;
}

// Block tree sum of `len` points per group into out[g * out_len + blockIdx.x].
__global__ void __launch_bounds__(kReduceBlock)
k_msm_reduce(const jac* __restrict__ in, jac* __restrict__ out, u32 len, u32 out_len) {
    __shared__ jac sh[kReduceBlock];
    const u32 g = blockIdx.y, q = blockIdx.x, t = threadIdx.x;
    const u32 idx = q * kReduceBlock + t;
    jac v;
    if (idx < len) {
        v = in[(size_t)g * len + idx];
    } else {
        jac_set_inf(v);
    }
    sh[t] = v;
    __syncthreads();
    for (u32 stride = kReduceBlock / 2; stride > 0; stride >>= 1) {
        if (t < stride) {
            jac a = sh[t];
            jac_add(a, a, sh[t + stride]);
            sh[t] = a;
        }
        __syncthreads();
    }
    if (t == 0) out[(size_t)g * out_len + q] = sh[0];
}

__global__ void k_msm_final(const jac* __restrict__ wsum, u32 W, u32 c, const u32* __restrict__ err,
                            u8* __restrict__ out33, int32_t* __restrict__ status) {
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    const u32 e = *err;
    if (e != 0xFFFFFFFFu) {
        for (int i = 0; i < 33; ++i) out33[i] = 0;
        *status = (e & 1u) ? 3 /* BAD_PUBKEY */ : 5 /* BAD_INPUT */;
        return;
    }
    jac acc = wsum[W - 1];
    for (int w = (int)W - 2; w >= 0; --w) {
        for (u32 d = 0; d < c; ++d) jac_dbl(acc, acc);
        jac_add(acc, acc, wsum[w]);
    }
    if (acc.inf) {
        for (int i = 0; i < 33; ++i) out33[i] = 0;
        *status = 7;   // ERR_ARITH: result is the point at infinity
        return;
    }
    aff A;
    jac_to_aff(A, acc);
    out33[0] = 0x02 | (u8)(A.y.v[0] & 1u);
    u256_to_be(out33 + 1, A.x.v);
    *status = 0;
}

__global__ void k_u32_fill(u32* p, u32 v, size_t n) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// ---------------------------------------------------------------------------
// Launchers
// ---------------------------------------------------------------------------
static inline unsigned blocks_for(size_t n, unsigned b) { return (unsigned)((n + b - 1) / b); }

hipError_t launch_u32_fill(uint32_t* p, uint32_t v, size_t n, hipStream_t st) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_u32_fill, dim3(blocks_for(n, 256)), dim3(256), 0, st, p, v, n);
    return hipGetLastError();
}

hipError_t launch_msm_prepare(const uint8_t* scalars32, const uint8_t* points33, uint32_t n, void* pts, void* scl,
                              uint32_t* counts, const MsmPlan& p, uint32_t* err, hipStream_t st) {
    hipLaunchKernelGGL(k_msm_prepare, dim3(blocks_for(n, kMsmBlock)), dim3(kMsmBlock), 0, st, scalars32, points33,
                       n, (aff*)pts, (sc*)scl, counts, p.c, p.windows, err);
    return hipGetLastError();
}

hipError_t launch_msm_scatter(const void* scl, uint32_t n, uint32_t* cursor, uint32_t* entries, const MsmPlan& p,
                              hipStream_t st) {
    hipLaunchKernelGGL(k_msm_scatter, dim3(blocks_for(n, kMsmBlock)), dim3(kMsmBlock), 0, st, (const sc*)scl, n,
                       cursor, entries, p.c, p.windows);
    return hipGetLastError();
}

hipError_t launch_msm_buckets(const void* pts, const uint32_t* offsets, const uint32_t* entries, void* bsum,
                              const MsmPlan& p, hipStream_t st) {
    const uint32_t nb = (uint32_t)p.windows * (uint32_t)p.buckets;
    hipLaunchKernelGGL(k_msm_buckets, dim3(blocks_for(nb, kMsmBlock)), dim3(kMsmBlock), 0, st, (const aff*)pts,
                       offsets, entries, (jac*)bsum, nb);
    return hipGetLastError();
}

hipError_t launch_msm_segments(const void* bsum, void* seg, const MsmPlan& p, hipStream_t st) {
    const uint32_t T = (uint32_t)(p.buckets / p.seg_len);
    const uint32_t lanes = (uint32_t)p.windows * T;
    hipLaunchKernelGGL(k_msm_segments, dim3(blocks_for(lanes, kMsmBlock)), dim3(kMsmBlock), 0, st,
                       (const jac*)bsum, (jac*)seg, (uint32_t)p.windows, (uint32_t)p.buckets, (uint32_t)p.seg_len);
    return hipGetLastError();
}

hipError_t launch_msm_reduce(const void* in, void* out, uint32_t groups, uint32_t len, uint32_t out_len,
                             hipStream_t st) {
    hipLaunchKernelGGL(k_msm_reduce, dim3(out_len, groups), dim3(kReduceBlock), 0, st, (const jac*)in, (jac*)out,
                       len, out_len);
    return hipGetLastError();
}

hipError_t launch_msm_final(const void* wsum, const MsmPlan& p, const uint32_t* err, uint8_t* out33,
                            int32_t* status, hipStream_t st) {
    hipLaunchKernelGGL(k_msm_final, dim3(1), dim3(64), 0, st, (const jac*)wsum, (uint32_t)p.windows, (uint32_t)p.c,
                       err, out33, status);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Host engine
// ---------------------------------------------------------------------------
int msm_default_window(size_t n) {
    int lg = 0;
    while (((size_t)1 << (lg + 1)) <= n) ++lg;
    int c = lg - 4;
    if (c < 4) c = 4;
    if (c > 16) c = 16;
    return c;
}

MsmPlan msm_plan(size_t n, int c) {
    if (c <= 0) c = msm_default_window(n);
    MsmPlan p;
    p.c = c;
    p.windows = (257 + c - 1) / c;
    p.buckets = 1 << (c - 1);
    p.seg_len = p.buckets < 8 ? p.buckets : 8;
    return p;
}

MsmEngine::~MsmEngine() {
    DeviceBuffer* all[] = {&in_s_, &in_p_, &pts_, &scl_, &counts_, &offsets_, &entries_,
                           &bsum_, &seg_, &red_a_, &red_b_, &err_, &scan_tmp_, &out_};
    for (DeviceBuffer* b : all) b->release();
}

GpuError MsmEngine::enqueue(const uint8_t* d_scalars32, const uint8_t* d_points33, size_t n, uint8_t* d_out33,
                            int32_t* d_status, int c, hipStream_t st) {
    if (c != 0 && (c < 2 || c > 20)) return be_->fail(GpuError::BadInput, "MSM window width out of range");
    const MsmPlan p = msm_plan(n, c);
    const size_t nbk = (size_t)p.windows * p.buckets;
    const size_t nseg = (size_t)p.windows * (p.buckets / p.seg_len);
    const size_t jsz = sizeof(jac);
    hipError_t e;
    if ((e = pts_.ensure(n * sizeof(aff))) != hipSuccess || (e = scl_.ensure(n * sizeof(sc))) != hipSuccess ||
        (e = counts_.ensure((nbk + 1) * 4)) != hipSuccess || (e = offsets_.ensure((nbk + 1) * 4)) != hipSuccess ||
        (e = entries_.ensure(n * p.windows * 4)) != hipSuccess || (e = bsum_.ensure(nbk * jsz)) != hipSuccess ||
        (e = seg_.ensure(nseg * jsz)) != hipSuccess || (e = red_a_.ensure(nseg * jsz)) != hipSuccess ||
        (e = red_b_.ensure(nseg * jsz)) != hipSuccess || (e = err_.ensure(4)) != hipSuccess)
        return be_->fail(GpuError::Memory, "hipMalloc(MSM buffers)", e);
    uint32_t* counts = (uint32_t*)counts_.ptr;
    uint32_t* offsets = (uint32_t*)offsets_.ptr;
    uint32_t* err = (uint32_t*)err_.ptr;
    size_t tmp_bytes = 0;
    e = hipcub::DeviceScan::ExclusiveSum(nullptr, tmp_bytes, counts, offsets, (int)(nbk + 1), st);
    if (e != hipSuccess) return be_->fail(GpuError::Launch, "scan size query", e);
    if ((e = scan_tmp_.ensure(tmp_bytes + 16)) != hipSuccess)
        return be_->fail(GpuError::Memory, "hipMalloc(scan temp)", e);

    if ((e = launch_u32_fill(counts, 0u, nbk + 1, st)) != hipSuccess ||
        (e = launch_u32_fill(err, 0xFFFFFFFFu, 1, st)) != hipSuccess ||
        (e = launch_msm_prepare(d_scalars32, d_points33, (uint32_t)n, pts_.ptr, scl_.ptr, counts, p, err, st)) !=
            hipSuccess)
        return be_->fail(GpuError::Launch, "MSM prepare", e);
    e = hipcub::DeviceScan::ExclusiveSum(scan_tmp_.ptr, tmp_bytes, counts, offsets, (int)(nbk + 1), st);
    if (e != hipSuccess) return be_->fail(GpuError::Launch, "MSM bucket scan", e);
    e = hipMemcpyAsync(counts, offsets, (nbk + 1) * 4, hipMemcpyDeviceToDevice, st);
    if (e != hipSuccess) return be_->fail(GpuError::Memory, "MSM cursor copy", e);
    if ((e = launch_msm_scatter(scl_.ptr, (uint32_t)n, counts, (uint32_t*)entries_.ptr, p, st)) != hipSuccess ||
        (e = launch_msm_buckets(pts_.ptr, offsets, (const uint32_t*)entries_.ptr, bsum_.ptr, p, st)) !=
            hipSuccess ||
        (e = launch_msm_segments(bsum_.ptr, seg_.ptr, p, st)) != hipSuccess)
        return be_->fail(GpuError::Launch, "MSM bucket phase", e);
    // tree-reduce segments per window until one point per window remains
    uint32_t len = (uint32_t)(p.buckets / p.seg_len);
    const void* cur = seg_.ptr;
    void* bufs[2] = {red_a_.ptr, red_b_.ptr};
    int which = 0;
    while (len > 1) {
        const uint32_t out_len = (len + kReduceBlock - 1) / kReduceBlock;
        e = launch_msm_reduce(cur, bufs[which], (uint32_t)p.windows, len, out_len, st);
        if (e != hipSuccess) return be_->fail(GpuError::Launch, "MSM reduce", e);
        cur = bufs[which];
        which ^= 1;
        len = out_len;
    }
    e = launch_msm_final(cur, p, err, d_out33, d_status, st);
    if (e != hipSuccess) return be_->fail(GpuError::Launch, "MSM final", e);
    return GpuError::Ok;
}

GpuError MsmEngine::msm_device(const uint8_t* d_scalars32, const uint8_t* d_points33, size_t n, uint8_t* d_out33,
                               int32_t* d_status, int c, hipStream_t st) {
    if (!be_->is_ready()) return be_->fail(GpuError::Device, "context not initialised");
    GpuError r = enqueue(d_scalars32, d_points33, n, d_out33, d_status, c, st ? st : be_->stream());
    if (r == GpuError::Ok) be_->clear_error();
    return r;
}

GpuError MsmEngine::msm_host(const uint8_t* scalars32, const uint8_t* points33, size_t n, uint8_t* out33, int c) {
    if (!be_->is_ready()) return be_->fail(GpuError::Device, "context not initialised");
    hipStream_t st = be_->stream();
    hipError_t e;
    if ((e = in_s_.ensure(n * 32)) != hipSuccess || (e = in_p_.ensure(n * 33)) != hipSuccess ||
        (e = out_.ensure(64)) != hipSuccess)
        return be_->fail(GpuError::Memory, "hipMalloc(MSM inputs)", e);
    e = hipMemcpyAsync(in_s_.ptr, scalars32, n * 32, hipMemcpyHostToDevice, st);
    if (e == hipSuccess) e = hipMemcpyAsync(in_p_.ptr, points33, n * 33, hipMemcpyHostToDevice, st);
    if (e != hipSuccess) return be_->fail(GpuError::Memory, "copy MSM inputs", e);
    uint8_t* d_out = (uint8_t*)out_.ptr;
    int32_t* d_status = (int32_t*)(d_out + 48);
    GpuError r = enqueue((const uint8_t*)in_s_.ptr, (const uint8_t*)in_p_.ptr, n, d_out, d_status, c, st);
    if (r != GpuError::Ok) return r;
    uint8_t host[64];
    e = hipMemcpyAsync(host, d_out, 64, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess) return be_->fail(GpuError::Backend, "MSM synchronize", e);
    int32_t status;
    std::memcpy(&status, host + 48, 4);
    if (status != 0) {
        const char* m = status == 7 ? "MSM result is infinity"
                        : status == 3 ? "invalid point in array" : "scalar >= n";
        return be_->fail((GpuError)status, m);
    }
    std::memcpy(out33, host, 33);
    be_->clear_error();
    return GpuError::Ok;
}

}  // namespace ufs
