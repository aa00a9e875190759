// Public-key decompression through a device-resident cache, for the serving
// loop.  Decompression (a square root, about 9% of a verification) is split
// out of the verify kernel so that cache hits and misses do not share waves:
// a lookup pass resolves hits and compacts the misses, a dense pass
// decompresses only the misses, and the verify kernel reads finished points.
//
// Strict parsing as in the verify kernel (src/cpu/src/ecdsa.cpp:891-927 for
// compressed keys: prefix 02/03 and x < p; schnorr.cpp:275-300 for x-only
// keys: x < p, even y).  An invalid key is written as a point whose y is all
// one bits, which the verify kernel turns into a 0 verdict.
#include "kernels.h"
#include "secp_ecmult.h"
#include "secp_hash.h"

namespace ufs {

// x of item i and the requested y parity; false if the key is malformed.
UFS_DEV bool parse_key(const BatchDesc& d, u32 i, u32 x[8], u32& parity) {
    if (d.kind == 0) {
        const u8* pk = d.pubs + 33ull * i;
        const u32 prefix = pk[0];
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            const u8* q = pk + 1 + 28 - 4 * k;
            x[k] = ((u32)q[0] << 24) | ((u32)q[1] << 16) | ((u32)q[2] << 8) | (u32)q[3];
        }
        parity = prefix & 1u;
        return ((prefix == 2u) | (prefix == 3u)) && u256_lt_p(x);
    }
    const uint4* q = reinterpret_cast<const uint4*>(d.pubs + 32ull * i);
    const uint4 a = q[0], b = q[1];
    const u32 w[8] = {a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w};
#pragma unroll
    for (int k = 0; k < 8; ++k) x[k] = bswap32(w[7 - k]);
    parity = 0;
    return u256_lt_p(x);
}

UFS_DEV u32 key_hash(const u32 x[8], u32 mask) {
    u32 h = x[0] ^ (x[3] * 0x9E3779B9u) ^ (x[6] * 0x85EBCA6Bu);
    h ^= h >> 15;
    return h & mask;
}

UFS_DEV void write_invalid(aff* p) {
    aff bad;
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        bad.x.v[k] = 0;
        bad.y.v[k] = 0xFFFFFFFFu;
    }
    store_aff(p, bad);
}

// y with the requested parity from either root r of y^2 = x^3 + 7.
UFS_DEV void apply_parity(fe& y, const fe& r, u32 parity) {
    fe a = r, b;
    fe_normalize(a);
    fe_neg(b, a);
    fe_normalize(b);
    const bool take_a = (a.v[0] & 1u) == parity;
#pragma unroll
    for (int k = 0; k < 8; ++k) y.v[k] = take_a ? a.v[k] : b.v[k];
}

__global__ void __launch_bounds__(256)
k_key_resolve(const BatchDesc* __restrict__ descs, const TileDesc* __restrict__ tiles, u32 ntiles,
              const void* const* __restrict__ batch_pts, KeyCacheDesc cache, u32* __restrict__ miss,
              u32* __restrict__ miss_count) {
    const u32 t = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const u32 lane = threadIdx.x & 63u;
    if (t >= ntiles) return;
    const TileDesc tl = tiles[t];
    const BatchDesc d = descs[tl.batch];
    const u32 i = tl.first + lane;
    if (i >= d.count) return;
    aff* pts = reinterpret_cast<aff*>(const_cast<void*>(batch_pts[tl.batch]));
    u32 x[8], parity;
    if (!parse_key(d, i, x, parity)) {
        write_invalid(pts + i);
        return;
    }
    bool hit = false;
    if (cache.entries) {
        // Two ways per set: way 0 holds keys that were found at least once
        // after their insertion, way 1 takes every new key.  A key found in
        // way 1 is copied into way 0, so a stream of keys seen once only
        // replaces way 1 and leaves the reused keys in place.
        aff* set = reinterpret_cast<aff*>(cache.entries) + 2u * key_hash(x, cache.mask >> 1);
#pragma unroll 1
        for (u32 way = 0; way < 2 && !hit; ++way) {
            aff c;
            load_aff(c, set + way);
            bool same = true;
#pragma unroll
            for (int k = 0; k < 8; ++k) same &= c.x.v[k] == x[k];
            if (!same) continue;
            // trust the entry only if (x, y) is on the curve: a torn entry
            // (another lane's copy or insert in flight) fails the test and is
            // taken as a miss
            fe x3, y2, seven;
            fe_sqr(x3, c.x);
            fe_mul(x3, x3, c.x);
            fe_set_u32(seven, 7);
            fe_add(x3, x3, seven);
            fe_sqr(y2, c.y);
            if (fe_equal(x3, y2)) {
                aff P;
                P.x = c.x;
                apply_parity(P.y, c.y, parity);
                store_aff(pts + i, P);
                if (way == 1) store_aff(set, c);
                hit = true;
            }
        }
    }
    if (!hit) {
        const u32 slot = atomicAdd(miss_count, 1u);
        uint2* m = reinterpret_cast<uint2*>(miss);
        m[slot] = make_uint2(tl.batch, i);
    }
}

__global__ void __launch_bounds__(256)
k_key_decompress(const BatchDesc* __restrict__ descs, const void* const* __restrict__ batch_pts, KeyCacheDesc cache,
                 const u32* __restrict__ miss, const u32* __restrict__ miss_count) {
    const u32 n = *miss_count;
    const uint2* m = reinterpret_cast<const uint2*>(miss);
    for (u32 j = blockIdx.x * blockDim.x + threadIdx.x; j < n; j += gridDim.x * blockDim.x) {
        const uint2 e = m[j];
        const BatchDesc d = descs[e.x];
        const u32 i = e.y;
        aff* pts = reinterpret_cast<aff*>(const_cast<void*>(batch_pts[e.x]));
        u32 x[8], parity;
        aff P;
        // parse again (the resolve pass only kept the position)
        (void)parse_key(d, i, x, parity);
        if (!lift_x(P, x, 0u)) {
            write_invalid(pts + i);
            continue;
        }
        if (cache.entries) {
            aff* c = reinterpret_cast<aff*>(cache.entries) + 2u * key_hash(x, cache.mask >> 1) + 1u;
            store_aff(c, P);   // new keys go to way 1; y of even parity
        }
        if (parity) fe_neg(P.y, P.y);
        fe_normalize(P.y);
        store_aff(pts + i, P);
    }
}

hipError_t launch_key_resolve(const BatchDesc* descs, const TileDesc* tiles, uint32_t ntiles,
                              const void* const* batch_pts, KeyCacheDesc cache, uint32_t* miss, uint32_t* miss_count,
                              hipStream_t st) {
    if (ntiles == 0) return hipSuccess;
    const unsigned blocks = (ntiles + 3) / 4;   // four 64-lane tiles per block
    hipLaunchKernelGGL(k_key_resolve, dim3(blocks), dim3(256), 0, st, descs, tiles, ntiles, batch_pts, cache, miss,
                       miss_count);
    return hipGetLastError();
}

hipError_t launch_key_decompress(const BatchDesc* descs, const void* const* batch_pts, KeyCacheDesc cache,
                                 const uint32_t* miss, const uint32_t* miss_count, unsigned grid, hipStream_t st) {
    if (grid == 0) return hipSuccess;
    hipLaunchKernelGGL(k_key_decompress, dim3(grid), dim3(256), 0, st, descs, batch_pts, cache, miss, miss_count);
    return hipGetLastError();
}

}  // namespace ufs
