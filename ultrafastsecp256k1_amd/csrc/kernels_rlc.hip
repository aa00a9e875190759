// Term preparation for BIP-340 batch verification by a random linear
// combination (reference: schnorr_batch_verify_impl,
// src/cpu/src/batch_verify.cpp:324-424).  With weights a_i,
//   sum_i a_i (s_i G - R_i - e_i P_i) = O
// holds when every signature is valid, and fails with probability about
// 1/n for a batch holding an invalid one.  The check becomes one
// multi-scalar multiplication over 2N + 1 terms:
//   term i       scalar -(a_i e_i) mod n,  point 02 || P_i.x   (lift, even y)
//   term N + i   scalar -a_i mod n,        point 02 || r_i     (R_i, even y)
//   term 2N      scalar sum a_i s_i,       point G
// whose result is the point at infinity exactly when the combination
// vanishes.  The MSM's own strict decompression rejects r >= p, P.x >= p and
// x without a square root; s = 0 or s >= n is flagged here.
//
// Weights: a_i = SHA256(seed || LE32(i)) mod n (0 -> 1), the reference's
// batch_weight (batch_verify.cpp:271-289) with a seed drawn from the host's
// CSPRNG for every call (the reference hashes the batch and XORs 32 CSPRNG
// bytes; an unpredictable seed alone is what the soundness argument needs).
#include "kernels.h"
#include "secp_ecmult.h"
#include "secp_hash.h"

namespace ufs {

constexpr int kRlcBlock = 256;

UFS_DEV void store_be32(u8* p, const u32 v[8]) { u256_to_be(p, v); }

__global__ void __launch_bounds__(kRlcBlock)
k_rlc_terms(const u8* __restrict__ msgs, const u8* __restrict__ pubx, const u8* __restrict__ sigs, u32 n,
            RlcSeed seed, const u32* __restrict__ mid_g, u8* __restrict__ scalars, u8* __restrict__ points,
            u32* __restrict__ partial, u32* __restrict__ err) {
    __shared__ u32 acc[kRlcBlock][9];
    const u32 i = blockIdx.x * kRlcBlock + threadIdx.x;
    sc as;   // a_i * s_i mod n, 0 for lanes past the end
    sc_set_u32(as, 0);
    if (i < n) {
        u32 m_w[8], px_w[8], r_w[8], s_raw[8];
        for (int k = 0; k < 8; ++k) {
            m_w[k] = ((u32)msgs[32ull * i + 4 * k] << 24) | ((u32)msgs[32ull * i + 4 * k + 1] << 16) |
                     ((u32)msgs[32ull * i + 4 * k + 2] << 8) | (u32)msgs[32ull * i + 4 * k + 3];
            px_w[k] = ((u32)pubx[32ull * i + 4 * k] << 24) | ((u32)pubx[32ull * i + 4 * k + 1] << 16) |
                      ((u32)pubx[32ull * i + 4 * k + 2] << 8) | (u32)pubx[32ull * i + 4 * k + 3];
            r_w[k] = ((u32)sigs[64ull * i + 4 * k] << 24) | ((u32)sigs[64ull * i + 4 * k + 1] << 16) |
                     ((u32)sigs[64ull * i + 4 * k + 2] << 8) | (u32)sigs[64ull * i + 4 * k + 3];
        }
        u256_from_be(s_raw, sigs + 64ull * i + 32);
        if (u256_is_zero(s_raw) || !u256_lt_n(s_raw)) atomicOr(err, 1u);
        u32 mid[8];
#pragma unroll
        for (int k = 0; k < 8; ++k) mid[k] = mid_g[k];
        // e = SHA256(tag || tag || r || P.x || m) mod n
        u32 ew[8];
        bip340_challenge(ew, mid, r_w, px_w, m_w);
        sc e;
        u256_from_be_words(e.v, ew);
        sc_reduce_once(e);
        // a = SHA256(seed || LE32(i)) mod n, one block of 36 bytes
        u32 blk[16];
#pragma unroll
        for (int k = 0; k < 8; ++k) blk[k] = seed.w[k];
        blk[8] = ((i & 0xFFu) << 24) | (((i >> 8) & 0xFFu) << 16) | (((i >> 16) & 0xFFu) << 8) | (i >> 24);
        blk[9] = 0x80000000u;
#pragma unroll
        for (int k = 10; k < 15; ++k) blk[k] = 0;
        blk[15] = 36u * 8u;
        u32 st[8] = {UFS_SHA256_IV};
        sha256_compress(st, blk);
        sc a;
        u256_from_be_words(a.v, st);
        sc_reduce_once(a);
        if (u256_is_zero(a.v)) sc_set_u32(a, 1);
        sc s, ae, t;
#pragma unroll
        for (int k = 0; k < 8; ++k) s.v[k] = s_raw[k];
        sc_reduce_once(s);
        sc_mul(as, a, s);
        sc_mul(ae, a, e);
        sc_neg(t, ae);
        store_be32(scalars + 32ull * i, t.v);
        sc_neg(t, a);
        store_be32(scalars + 32ull * (n + i), t.v);
        u8* pp = points + 33ull * i;
        u8* pr = points + 33ull * (n + i);
        pp[0] = 2;
        pr[0] = 2;
        for (int k = 0; k < 32; ++k) {
            pp[1 + k] = pubx[32ull * i + k];
            pr[1 + k] = sigs[64ull * i + k];
        }
    }
    // block sum of a_i s_i (mod n)
#pragma unroll
    for (int k = 0; k < 8; ++k) acc[threadIdx.x][k] = as.v[k];
    __syncthreads();
    for (int w = kRlcBlock / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
            sc x, y;
#pragma unroll
            for (int k = 0; k < 8; ++k) {
                x.v[k] = acc[threadIdx.x][k];
                y.v[k] = acc[threadIdx.x + w][k];
            }
            sc_add(x, x, y);
#pragma unroll
            for (int k = 0; k < 8; ++k) acc[threadIdx.x][k] = x.v[k];
        }
        __syncthreads();
    }
    if (threadIdx.x < 8) partial[8ull * blockIdx.x + threadIdx.x] = acc[0][threadIdx.x];
}

// Sums the per-block partials and writes term 2N: (sum a_i s_i, G).
__global__ void __launch_bounds__(kRlcBlock)
k_rlc_finish(const u32* __restrict__ partial, u32 nblocks, u8* __restrict__ g_scalar, u8* __restrict__ g_point) {
    __shared__ u32 acc[kRlcBlock][9];
    sc x;
    sc_set_u32(x, 0);
    for (u32 b = threadIdx.x; b < nblocks; b += kRlcBlock) {
        sc y;
#pragma unroll
        for (int k = 0; k < 8; ++k) y.v[k] = partial[8ull * b + k];
        sc_add(x, x, y);
    }
#pragma unroll
    for (int k = 0; k < 8; ++k) acc[threadIdx.x][k] = x.v[k];
    __syncthreads();
    for (int w = kRlcBlock / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
            sc a, b;
#pragma unroll
            for (int k = 0; k < 8; ++k) {
                a.v[k] = acc[threadIdx.x][k];
                b.v[k] = acc[threadIdx.x + w][k];
            }
            sc_add(a, a, b);
#pragma unroll
            for (int k = 0; k < 8; ++k) acc[threadIdx.x][k] = a.v[k];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        u32 v[8];
#pragma unroll
        for (int k = 0; k < 8; ++k) v[k] = acc[0][k];
        store_be32(g_scalar, v);
        aff g;
        set_generator(g);
        g_point[0] = (u8)(2u | (g.y.v[0] & 1u));
        u256_to_be(g_point + 1, g.x.v);
    }
}

hipError_t launch_rlc_terms(const uint8_t* msgs, const uint8_t* pubx, const uint8_t* sigs, uint32_t n,
                            const RlcSeed& seed, const uint32_t* mid, uint8_t* scalars, uint8_t* points,
                            uint32_t* partial, uint32_t* err, hipStream_t st) {
    if (n == 0) return hipSuccess;
    const unsigned blocks = (n + kRlcBlock - 1) / kRlcBlock;
    hipLaunchKernelGGL(k_rlc_terms, dim3(blocks), dim3(kRlcBlock), 0, st, msgs, pubx, sigs, n, seed, mid, scalars,
                       points, partial, err);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_rlc_finish, dim3(1), dim3(kRlcBlock), 0, st, partial, blocks, scalars + 64ull * n,
                       points + 66ull * n);
    return hipGetLastError();
}

uint32_t rlc_partial_words(uint32_t n) { return 8u * ((n + kRlcBlock - 1) / kRlcBlock); }

}  // namespace ufs
