// Batch ECDSA / BIP-340 verification kernels for gfx950 (one lane per
// signature) plus context set-up kernels (fixed-base G comb table, BIP-340
// challenge midstate) and arithmetic self-test kernels.
//
// Per-item semantics follow the reference CPU path, which is the parity
// contract:
//   ECDSA  : src/cpu/src/impl/ufsecp_ecdsa.cpp:244-275 (strict r,s in [1,n),
//            low-S required), ecdsa.cpp:891-927 (strict compressed pubkey),
//            ecdsa.cpp:764-850 (x(R) == r, including the r + n case).
//   Schnorr: src/cpu/src/schnorr.cpp:355-371 (r < p, 0 < s < n),
//            :275-300 (x-only key x < p, lift_x), :745-806 (R = sG - eP,
//            R finite, even y, x(R) == r).
#include "kernels.h"
#include "secp_ecmult.h"
#include "secp_hash.h"
#include "secp_keycache.h"
#include "secp_wide26.h"

#ifndef UFS_VERIFY_MIN_BLOCKS
#define UFS_VERIFY_MIN_BLOCKS 3
#endif

namespace ufs {

static_assert(kSlotEntries * sizeof(aff) == kSlotBytes, "lane slot size");

UFS_DEV void load_be32_aligned(u32 r[8], const u8* p) {
    const uint4* q = reinterpret_cast<const uint4*>(p);
    const uint4 a = q[0], b = q[1];
    const u32 w[8] = {a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w};
#pragma unroll
    for (int i = 0; i < 8; ++i) r[i] = bswap32(w[7 - i]);
}

UFS_DEV void load_be_words(u32 w[8], const u8* p) {
    const uint4* q = reinterpret_cast<const uint4*>(p);
    const uint4 a = q[0], b = q[1];
    w[0] = bswap32(a.x); w[1] = bswap32(a.y); w[2] = bswap32(a.z); w[3] = bswap32(a.w);
    w[4] = bswap32(b.x); w[5] = bswap32(b.y); w[6] = bswap32(b.z); w[7] = bswap32(b.w);
}

UFS_DEV void load_be32_bytes(u32 r[8], const u8* p) {
#pragma unroll
    for (int i = 0; i < 8; ++i) {
        const u8* q = p + 28 - 4 * i;
        r[i] = ((u32)q[0] << 24) | ((u32)q[1] << 16) | ((u32)q[2] << 8) | (u32)q[3];
    }
}

// x(R) == r (mod n) test without inversion: r*Z^2 == X, or (r+n)*Z^2 == X
// when r + n < p.
UFS_DEV bool ecdsa_xcheck(const jac& R, const u32 r[8]) {
    fe rz, z2, t;
#pragma unroll
    for (int i = 0; i < 8; ++i) rz.v[i] = r[i];
    fe_sqr(z2, R.z);
    fe_mul(t, rz, z2);
    bool ok = fe_equal(t, R.x);
    const u32 PMN[8] = {UFS_PMN_LIMBS};
    const u32 N[8] = {UFS_N_LIMBS};
    if (u256_lt(r, PMN)) {
        u32 c = 0;
#pragma unroll
        for (int i = 0; i < 8; ++i) rz.v[i] = addc(r[i], N[i], c, c);
        fe_mul(t, rz, z2);
        ok |= fe_equal(t, R.x);
    }
    return ok;
}

// Where a batch's public keys come from: the serving loop's key cache
// (`cache`, variant 3), or none (decompress here).  In variants 0..2 `cache`
// is a compile-time null and its code is not emitted.
struct KeySrc {
    aff* cache;
    u32 mask;
    u32* hits;
};

// Key of item x (big-endian words already converted), y of parity `odd`.
// False for x >= p or x^3 + 7 not a square.  Through the cache: a hit uses
// the cached point (counted); a miss is decompressed and inserted.
UFS_DEV bool item_key(aff& P, const KeySrc& ks, const u32 x[8], u32 odd) {
    if (ks.cache) {
        if (u256_lt_p(x) && key_cache_find(P, ks.cache, ks.mask, x, odd)) {
            atomicAdd(ks.hits, 1u);
            return true;
        }
        if (!lift_x(P, x, 0u)) return false;
        key_cache_insert(ks.cache, ks.mask, P);
        if (odd) {
            fe_neg(P.y, P.y);
            fe_normalize(P.y);
        }
        return true;
    }
    return lift_x(P, x, odd);
}

// 256-bit little-endian integer from 32 bytes of any alignment.
UFS_DEV void load_le32_bytes(u32 r[8], const u8* p) {
#pragma unroll
    for (int i = 0; i < 8; ++i) {
        const u8* q = p + 4 * i;
        r[i] = (u32)q[0] | ((u32)q[1] << 8) | ((u32)q[2] << 16) | ((u32)q[3] << 24);
    }
}

// One ECDSA verification of item i of batch d; returns the verdict.
// `slot` is the lane's private table area (kSlotEntries affine entries);
// `ks` says where the public key comes from.
// opaque: the batch is of kKindEcdsaOpaque -- r and s are little-endian
// 256-bit integers, strict in [1,n) (src/cpu/src/impl/ufsecp_ecdsa.cpp:105-117,
// 128-140), and high-S is accepted: the reference normalises s to n - s,
// which leaves x(R) unchanged.  Items are d.stride bytes apart when the
// stride is set (rows), else packed.
UFS_DEV bool ecdsa_verify_item_w(const BatchDesc& d, u32 i, const gcomb& G, aff* slot, const KeySrc& ks,
                                 bool opaque, const sc& w_in);

// s of item i as the inversion sees it: the signature's s when it is in range
// (strict [1, n), and low-S unless opaque), else 1 -- the item's verdict is
// false then, and 1 keeps a shared product of s values invertible.
UFS_DEV void ecdsa_item_s(const BatchDesc& d, u32 i, bool opaque, sc& s) {
    u32 raw[8];
    if (opaque) {
        const u64 st = d.stride;
        load_le32_bytes(raw, d.sigs + (st ? st : 64ull) * i + 32);
    } else {
        load_be32_aligned(raw, d.sigs + 64ull * i + 32);
    }
#pragma unroll
    for (int k = 0; k < 8; ++k) s.v[k] = raw[k];
    bool ok = !u256_is_zero(raw) && u256_lt_n(raw);
    if (!opaque) ok &= !sc_is_high(s);
    if (!ok) sc_set_u32(s, 1);
}

UFS_DEV bool ecdsa_verify_item(const BatchDesc& d, u32 i, const gcomb& G, aff* slot, const KeySrc& ks,
                               bool opaque) {
    sc s, w;
    ecdsa_item_s(d, i, opaque, s);
    sc_inv(w, s);
    return ecdsa_verify_item_w(d, i, G, slot, ks, opaque, w);
}

// ECDSA verification of item i of batch d given w = 1/s mod n (computed by
// the caller, possibly shared with another item).
UFS_DEV bool ecdsa_verify_item_w(const BatchDesc& d, u32 i, const gcomb& G, aff* slot, const KeySrc& ks,
                                 bool opaque, const sc& w_in) {
    u32 z_raw[8], r_raw[8], s_raw[8], x_raw[8];
    const u8* pk;
    if (opaque) {
        const u64 st = d.stride;
        const u8* m = d.msgs + (st ? st : 32ull) * i;
        const u8* sg = d.sigs + (st ? st : 64ull) * i;
        pk = d.pubs + (st ? st : 33ull) * i;
        load_be32_bytes(z_raw, m);
        load_le32_bytes(r_raw, sg);
        load_le32_bytes(s_raw, sg + 32);
    } else {
        load_be32_aligned(z_raw, d.msgs + 32ull * i);
        load_be32_aligned(r_raw, d.sigs + 64ull * i);
        load_be32_aligned(s_raw, d.sigs + 64ull * i + 32);
        pk = d.pubs + 33ull * i;
    }
    const u32 prefix = pk[0];
    load_be32_bytes(x_raw, pk + 1);
    bool ok = !u256_is_zero(r_raw) && u256_lt_n(r_raw) && !u256_is_zero(s_raw) && u256_lt_n(s_raw);
    sc s, r, z;
#pragma unroll
    for (int k = 0; k < 8; ++k) { s.v[k] = s_raw[k]; r.v[k] = r_raw[k]; z.v[k] = z_raw[k]; }
    if (!opaque) ok &= !sc_is_high(s);
    ok &= (prefix == 2u) | (prefix == 3u);
    aff P;
    ok &= item_key(P, ks, x_raw, prefix & 1u);
    if (!ok) {
        // keep the lane on the common path with harmless values
        set_generator(P);
        sc_set_u32(r, 1);
    }
    sc_reduce_once(z);
    sc u1, u2;
    const sc w = w_in;
    sc_mul(u1, z, w);
    sc_mul(u2, r, w);
    sc k1, k2;
    bool n1, n2;
    glv_split(k1, k2, n1, n2, u2);
    build_ptable(slot, P);
    jac R;
    ecmult_dual(R, slot, k1, n1, k2, n2, G, u1);
    const bool v = !R.inf && ecdsa_xcheck(R, r.v);
    return ok && v;
}

UFS_DEV void store_sc(fe* p, const sc& a) {
    fe t;
#pragma unroll
    for (int k = 0; k < 8; ++k) t.v[k] = a.v[k];
    store_fe(p, t);
}

UFS_DEV void load_sc(sc& a, const fe* p) {
    fe t;
    load_fe(t, p);
#pragma unroll
    for (int k = 0; k < 8; ++k) a.v[k] = t.v[k];
}

// ECDSA verification of item i1 and, when `two`, item i2 of batch d, the two
// sharing the inversion of s (Montgomery's trick: 1/s1 = s2 / (s1 s2),
// 1/s2 = s1 / (s1 s2)).  Both inverses wait in the slot's last entry while
// the items run one after the other in a rolled loop (one inline copy of the
// scalar multiplication).  `opaque` as for ecdsa_verify_item.
UFS_DEV void ecdsa_verify_items(const BatchDesc& d, u32 i1, u32 i2, bool two, const gcomb& G, aff* slot,
                                const KeySrc& ks, bool opaque, bool& v1, bool& v2) {
    fe* stash = &slot[kSlotEntries - 1].x;
    {
        sc s1, s2, ss, inv, w;
        ecdsa_item_s(d, i1, opaque, s1);
        if (two) ecdsa_item_s(d, i2, opaque, s2); else sc_set_u32(s2, 1);
        sc_mul(ss, s1, s2);
        sc_inv(inv, ss);
        sc_mul(w, inv, s2);
        store_sc(stash, w);
        sc_mul(w, inv, s1);
        store_sc(stash + 1, w);
    }
    v2 = false;
    const int m = two ? 2 : 1;
#pragma unroll 1
    for (int k = 0; k < m; ++k) {
        sc w;
        load_sc(w, stash + k);
        const bool v = ecdsa_verify_item_w(d, k ? i2 : i1, G, slot, ks, opaque, w);
        if (k == 0) v1 = v; else v2 = v;
    }
}

// BIP-340 verification of item i of batch d up to the y-parity test: R =
// s G - e P in Jacobian form and whether every check but y(R) even passed.
UFS_DEV bool schnorr_verify_head(const BatchDesc& d, u32 i, const u32 mid[8], const gcomb& G, aff* slot,
                                 const KeySrc& ks, jac& R) {
    u32 m_w[8], px_w[8], r_w[8], s_w[8];
    load_be_words(m_w, d.msgs + 32ull * i);
    load_be_words(px_w, d.pubs + 32ull * i);
    load_be_words(r_w, d.sigs + 64ull * i);
    load_be_words(s_w, d.sigs + 64ull * i + 32);
    u32 px[8], rr[8], ss[8];
    u256_from_be_words(px, px_w);
    u256_from_be_words(rr, r_w);
    u256_from_be_words(ss, s_w);
    bool ok = u256_lt_p(rr) && !u256_is_zero(ss) && u256_lt_n(ss);
    aff P;
    ok &= item_key(P, ks, px, 0u);
    sc s;
#pragma unroll
    for (int k = 0; k < 8; ++k) s.v[k] = ss[k];
    if (!ok) {
        set_generator(P);
        sc_set_u32(s, 1);
    }
    u32 ew[8];
    bip340_challenge(ew, mid, r_w, px_w, m_w);
    sc e, b;
    u256_from_be_words(e.v, ew);
    sc_reduce_once(e);
    sc_neg(b, e);   // R = s*G + (n - e)*P
    sc k1, k2;
    bool n1, n2;
    glv_split(k1, k2, n1, n2, b);
    build_ptable(slot, P);
    ecmult_dual(R, slot, k1, n1, k2, n2, G, s);
    if (R.inf) return false;
    // x(R) == r without inversion (r < p): r * Z^2 == X
    fe z2, t, rf;
#pragma unroll
    for (int k = 0; k < 8; ++k) rf.v[k] = rr[k];
    fe_sqr(z2, R.z);
    fe_mul(t, rf, z2);
    return ok && fe_equal(t, R.x);
}

// y(R) even, with zi = 1/Z: y = Y zi^3.
UFS_DEV bool y_is_even(const fe& Y, const fe& zi) {
    fe zi3, y;
    fe_sqr(zi3, zi);
    fe_mul(zi3, zi3, zi);
    fe_mul(y, Y, zi3);
    fe_normalize(y);
    return (y.v[0] & 1u) == 0;
}

// BIP-340 verification of item i1 and, when `two`, item i2 of batch d, the
// two sharing the final inversion (Montgomery's trick: 1/Z1 = Z2 / (Z1 Z2),
// 1/Z2 = Z1 / (Z1 Z2)); a point at infinity, or the missing second item,
// takes Z = 1 in the product (a point at infinity's verdict is false
// already).  The head runs in a rolled loop: one inline copy of the
// scalar multiplication serves both items.
UFS_DEV void schnorr_verify_items(const BatchDesc& d, u32 i1, u32 i2, bool two, const u32 mid[8], const gcomb& G,
                                  aff* slot, const KeySrc& ks, bool& v1, bool& v2) {
    // the first item's Y and Z wait in the slot's last entry while the
    // second item runs (no long-lived registers across the multiplication)
    fe* stash = &slot[kSlotEntries - 1].x;
    fe y1, z1, y2, z2;
    fe_set_u32(z2, 1);
    v2 = false;
    const int m = two ? 2 : 1;
#pragma unroll 1
    for (int k = 0; k < m; ++k) {
        jac R;
        const bool v = schnorr_verify_head(d, k ? i2 : i1, mid, G, slot, ks, R);
        fe z = R.z;
        if (R.inf) fe_set_u32(z, 1);
        if (k == 0) {
            v1 = v;
            store_fe(stash, R.y);
            store_fe(stash + 1, z);
        } else {
            v2 = v;
            y2 = R.y;
            z2 = z;
        }
    }
    load_fe(y1, stash);
    load_fe(z1, stash + 1);
    fe zz, inv, zi;
    fe_mul(zz, z1, z2);
    fe_inv(inv, zz);
    fe_mul(zi, inv, z2);
    v1 = v1 && y_is_even(y1, zi);
    if (two) {
        fe_mul(zi, inv, z1);
        v2 = v2 && y_is_even(y2, zi);
    }
}

// ---------------------------------------------------------------------------
// Persistent verification kernel.  A wave verifies 64-item tiles; each lane
// owns one table slot for the kernel's lifetime.  Tiles come from one batch
// (single mode: wave w takes tiles w, w + W, w + 2W, ...) or from a tile list
// over uneven batches (list mode: waves pull tiles from a device counter
// until it passes the tile count, so the grid drains before the kernel
// returns).  A tile never straddles batches, so a wave runs one signature
// kind per tile.  KIND: 0 = ECDSA, 1 = BIP-340, 2 = per batch, 3 = per
// batch with the keys through the serving loop's key cache.
// ---------------------------------------------------------------------------
template <int KIND>
__global__ void __launch_bounds__(kVerifyBlock, UFS_VERIFY_MIN_BLOCKS)
k_verify(VerifyLaunch a, gcomb G, const u32* __restrict__ mid_g) {
    const u32 gtid = blockIdx.x * blockDim.x + threadIdx.x;
    const u32 lane = threadIdx.x & 63u;
    aff* slot = reinterpret_cast<aff*>(a.slots) + (size_t)gtid * kSlotEntries;
    u32 mid[8];
#pragma unroll
    for (int j = 0; j < 8; ++j) mid[j] = mid_g[j];
    const bool list = a.tiles != nullptr;
    // single mode with a work counter: waves take the next tile (two
    // consecutive tiles when pairing) as they finish instead of a fixed
    // stride, so waves that run slower take fewer tiles (2^20 items: about 6%
    // faster than the stride)
    const bool dyn = !list && a.counter != nullptr;
    const u32 ntiles = list ? *a.ntiles_dev : a.single_tiles;
    const u32 nwaves = (gridDim.x * blockDim.x) >> 6;
    u32 t = gtid >> 6;
    // Single ECDSA (either form) or BIP-340 batch: the lane's item of a second tile joins
    // the one of tile t (shared inversion).  With the counter, units
    // 0 .. np - 1 are pairs of consecutive tiles and the units after them
    // single tiles: as many full rounds of one pair per wave as the batch
    // fills, then for the r < 2W tiles left one unit per wave (r - W pairs
    // and the rest singles when r > W, else r singles), so the last round
    // ends within one unit rather than a pair for some waves and nothing for
    // others.  By stride: t with t + W, the loop stepping by 2W.
    const bool pair_kind = (KIND == 0 || KIND == 1 || KIND == 4) && !list;
    u32 np = 0;
    if (pair_kind && dyn) {
        const u32 r = ntiles % (2 * nwaves);
        np = nwaves * (ntiles / (2 * nwaves)) + (r > nwaves ? r - nwaves : 0u);
    }
    for (;;) {
        u32 t2 = 0;
        bool paired = false;
        if (list || dyn) {
            u32 x = 0;
            if (lane == 0) x = atomicAdd(a.counter, 1u);
            const u32 u = __shfl(x, 0, 64);
            paired = u < np;
            t = paired ? 2 * u : (list ? u : np + u);
            t2 = t + 1;
        } else {
            t2 = t + nwaves;
            paired = pair_kind;
        }
        if (t >= ntiles) break;
        BatchDesc d;
        u32 first;
        KeySrc ks{nullptr, 0u, nullptr};
        if (KIND == 3) {
            ks.cache = reinterpret_cast<aff*>(a.key_cache);
            ks.mask = a.key_cache_mask;
            ks.hits = a.key_hits;
        }
        if (list) {
            const TileDesc tl = a.tiles[t];
            d = a.descs[tl.batch];
            first = tl.first;
        } else {
            d = a.single;
            first = t * kTileItems;
        }
        // without a counter the tiles go by a fixed stride: t, t + W (pairs:
        // the loop then steps by 2W)
        if (!list && !dyn) t += pair_kind ? 2 * nwaves : nwaves;
        const u32 i = first + lane;
        if (i >= d.count) continue;
        const u32 i2 = t2 * kTileItems + lane;
        const bool two = paired && t2 < ntiles && i2 < d.count;
        const u32 kind = KIND == 4 ? kKindEcdsaOpaque : KIND >= 2 ? d.kind : (u32)KIND;
        bool v, v2 = false;
        if (kind == kKindSchnorr) {
            schnorr_verify_items(d, i, i2, two, mid, G, slot, ks, v, v2);
        } else if (KIND == 0 || KIND == 4) {
            ecdsa_verify_items(d, i, i2, two, G, slot, ks, KIND == 4, v, v2);
        } else {
            v = ecdsa_verify_item(d, i, G, slot, ks, kind == kKindEcdsaOpaque);
        }
        if (a.collect) {
            if (v) d.out[i] = 0;
            if (two && v2) d.out[i2] = 0;
        } else {
            d.out[i] = v ? 1 : 0;
            if (two) d.out[i2] = v2 ? 1 : 0;
        }
    }
}

// entry(i, j) = (2j+1) * 2^(w*i) * G for i < nwin; entry(nwin, 0) = 2^(w*nwin) G.
__global__ void __launch_bounds__(64) k_build_gcomb(aff* tab, int w, int nwin) {
    const int half = 1 << (w - 1);
    const int id = blockIdx.x * blockDim.x + threadIdx.x;
    const int total = nwin * half + 1;
    if (id >= total) return;
    const int i = id / half, j = id % half;
    aff g;
    set_generator(g);
    jac B;
    jac_from_aff(B, g);
    for (int d = 0; d < w * i; ++d) jac_dbl(B, B);
    const u32 m = (i == nwin) ? 1u : (u32)(2 * j + 1);
    jac acc = B;
    int top = 31 - __builtin_clz(m);
    for (int b = top - 1; b >= 0; --b) {
        jac_dbl(acc, acc);
        if ((m >> b) & 1u) jac_add(acc, acc, B);
    }
    aff a;
    jac_to_aff(a, acc);
    tab[id] = a;
}

// Midstate of SHA256 after absorbing SHA256(tag) || SHA256(tag) for
// tag = "BIP0340/challenge" (BIP-340 tagged hash).
__global__ void k_bip340_midstate(u32* out) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    const u32 iv[8] = {UFS_SHA256_IV};
    u32 st[8], blk[16];
#pragma unroll
    for (int i = 0; i < 8; ++i) st[i] = iv[i];
    // "BIP0340/challenge" = 17 bytes, single padded block
    blk[0] = 0x42495030U;  // BIP0
    blk[1] = 0x3334302FU;  // 340/
    blk[2] = 0x6368616CU;  // chal
    blk[3] = 0x6C656E67U;  // leng
    blk[4] = 0x65800000U;  // e + pad
    for (int i = 5; i < 15; ++i) blk[i] = 0;
    blk[15] = 17 * 8;
    sha256_compress(st, blk);
    u32 blk2[16];
    for (int i = 0; i < 8; ++i) { blk2[i] = st[i]; blk2[8 + i] = st[i]; }
    u32 mid[8];
    for (int i = 0; i < 8; ++i) mid[i] = iv[i];
    sha256_compress(mid, blk2);
    for (int i = 0; i < 8; ++i) out[i] = mid[i];
}

// Arithmetic self-test kernel: op codes documented in kernels.h.
__global__ void __launch_bounds__(64)
k_selftest(int op, const u32* __restrict__ a, const u32* __restrict__ b, u32* __restrict__ out,
           u32 count, gcomb G, aff* __restrict__ slots) {
    if (op == kOpJac26DblWide || op == kOpJac26AddLanes) {
        // every lane of the wave runs each item of this block in turn
        for (u32 j = 0; j < blockDim.x; ++j) {
            const u32 idx = blockIdx.x * blockDim.x + j;
            if (idx >= count) break;
            // A (and B) in Jacobian form with Z = idx + 3 (idx + 5)
            jac26 ab[2];
            for (int o = 0; o < 2; ++o) {
                const u32* src = o == 0 ? a : b;
                fe ax, ay, z, z2, z3;
#pragma unroll
                for (int l = 0; l < 8; ++l) {
                    ax.v[l] = src[16 * idx + l];
                    ay.v[l] = src[16 * idx + 8 + l];
                }
                fe_set_u32(z, idx + 3 + 2 * o);
                fe_sqr(z2, z);
                fe_mul(z3, z2, z);
                jac Pj;
                fe_mul(Pj.x, ax, z2);
                fe_mul(Pj.y, ay, z3);
                Pj.z = z;
                Pj.inf = 0;
                jac26_from_jac(ab[o], Pj);
            }
            jac26 a26 = ab[0];
            if (idx % 11 == 10) jac26_set_inf(a26);
            if (op == kOpJac26DblWide) {
                jac26_dbl_wide_n(a26, 1 + idx % 7);
            } else {
                jac26 b26 = ab[1];
                if (idx % 13 == 12) jac26_set_inf(b26);
                jac26_add_lanes(a26, a26, b26);
            }
            jac R;
            jac26_to_jac(R, a26);
            if (threadIdx.x == 0) {
                u32* o = out + 32 * idx;
                for (int l = 0; l < 32; ++l) o[l] = 0;
                o[16] = R.inf;
                if (!R.inf) {
                    aff A;
                    jac_to_aff(A, R);
                    for (int l = 0; l < 8; ++l) {
                        o[l] = A.x.v[l];
                        o[8 + l] = A.y.v[l];
                    }
                }
            }
        }
        return;
    }
    const u32 i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    fe x, y, r;
    sc sx, sy, sr;
#pragma unroll
    for (int l = 0; l < 8; ++l) {
        x.v[l] = a[16 * i + l];
        y.v[l] = b[16 * i + l];
        sx.v[l] = x.v[l];
        sy.v[l] = y.v[l];
    }
    u32* o = out + 32 * i;
    for (int l = 0; l < 32; ++l) o[l] = 0;
    switch (op) {
    case kOpFeMul: fe_mul(r, x, y); fe_normalize(r); for (int l = 0; l < 8; ++l) o[l] = r.v[l]; break;
    case kOpFeSqr: fe_sqr(r, x); fe_normalize(r); for (int l = 0; l < 8; ++l) o[l] = r.v[l]; break;
    case kOpFeAdd: fe_add(r, x, y); fe_normalize(r); for (int l = 0; l < 8; ++l) o[l] = r.v[l]; break;
    case kOpFeSub: fe_sub(r, x, y); fe_normalize(r); for (int l = 0; l < 8; ++l) o[l] = r.v[l]; break;
    case kOpFeInv: fe_inv(r, x); fe_normalize(r); for (int l = 0; l < 8; ++l) o[l] = r.v[l]; break;
    case kOpFeInvFermat: fe_inv_fermat(r, x); fe_normalize(r); for (int l = 0; l < 8; ++l) o[l] = r.v[l]; break;
    case kOpFeSqrt: { bool ok = fe_sqrt(r, x); fe_normalize(r); for (int l = 0; l < 8; ++l) o[l] = r.v[l]; o[8] = ok; } break;
    case kOpFeMulChain:   // throughput probe: 4096 dependent multiplications
        r = x;
        for (int k = 0; k < 4096; ++k) fe_mul(r, r, y);
        fe_normalize(r);
        for (int l = 0; l < 8; ++l) o[l] = r.v[l];
        break;
    case kOpFeSqrChain:   // throughput probe: 4096 dependent squarings
        r = x;
        for (int k = 0; k < 4096; ++k) fe_sqr(r, r);
        fe_normalize(r);
        for (int l = 0; l < 8; ++l) o[l] = r.v[l];
        break;
    case kOpScMul: sc_mul(sr, sx, sy); for (int l = 0; l < 8; ++l) o[l] = sr.v[l]; break;
    case kOpScInv: sc_inv(sr, sx); for (int l = 0; l < 8; ++l) o[l] = sr.v[l]; break;
    case kOpScInvFermat: sc_inv_fermat(sr, sx); for (int l = 0; l < 8; ++l) o[l] = sr.v[l]; break;
    case kOpGlv: {
        sc k1, k2; bool n1, n2;
        glv_split(k1, k2, n1, n2, sx);
        for (int l = 0; l < 8; ++l) { o[l] = k1.v[l]; o[8 + l] = k2.v[l]; }
        o[16] = n1; o[17] = n2;
    } break;
    case kOpGenMul: {   // a*G -> affine (x, y), o[16] = infinity flag
        aff* slot = slots + (size_t)i * kSlotEntries;
        aff g;
        set_generator(g);
        build_ptable(slot, g);
        sc zero;
        sc_set_u32(zero, 0);
        jac R;
        ecmult_dual(R, slot, zero, false, zero, false, G, sx);
        o[16] = R.inf;
        if (!R.inf) { aff A; jac_to_aff(A, R); for (int l = 0; l < 8; ++l) { o[l] = A.x.v[l]; o[8 + l] = A.y.v[l]; } }
    } break;
    case kOpVarMul: {   // b*P where P = lift_x(a, even y) -> affine
        aff P;
        u32 xr[8];
        for (int l = 0; l < 8; ++l) xr[l] = a[16 * i + l];
        bool ok = lift_x(P, xr, 0);
        sc k1, k2; bool n1, n2;
        glv_split(k1, k2, n1, n2, sy);
        aff* slot = slots + (size_t)i * kSlotEntries;
        build_ptable(slot, P);
        sc zero;
        sc_set_u32(zero, 0);
        jac R;
        ecmult_dual(R, slot, k1, n1, k2, n2, G, zero);
        o[16] = R.inf; o[17] = ok;
        if (!R.inf) { aff A; jac_to_aff(A, R); for (int l = 0; l < 8; ++l) { o[l] = A.x.v[l]; o[8 + l] = A.y.v[l]; } }
    } break;
    case kOpFe26Mul:
    case kOpFe26Sqr:
    case kOpFe26Norm: {
        fe26 x26, y26, r26;
        for (int l = 0; l < 10; ++l) {
            x26.v[l] = a[16 * i + l];
            y26.v[l] = b[16 * i + l];
        }
        if (op == kOpFe26Norm) {
            fe26_norm(r26, x26);
            for (int l = 0; l < 10; ++l) o[l] = r26.v[l];
            break;
        }
        if (op == kOpFe26Mul) fe26_mul(r26, x26, y26); else fe26_sqr(r26, x26);
        for (int l = 0; l < 10; ++l) o[8 + l] = r26.v[l];
        fe26_to_fe(r, r26);
        for (int l = 0; l < 8; ++l) o[l] = r.v[l];
    } break;
    case kOpFe26FromTo: {   // 8 x 32 -> 10 x 26 -> 8 x 32
        fe26 x26;
        fe26_from_fe(x26, x);
        fe26_to_fe(r, x26);
        for (int l = 0; l < 8; ++l) o[l] = r.v[l];
        for (int l = 0; l < 10; ++l) o[8 + l] = x26.v[l];
    } break;
    case kOpJac26Madd:
    case kOpJac26Dbl:
    case kOpJac26Add: {
        aff A, B;
        for (int l = 0; l < 8; ++l) {
            A.x.v[l] = a[16 * i + l];
            A.y.v[l] = a[16 * i + 8 + l];
            B.x.v[l] = b[16 * i + l];
            B.y.v[l] = b[16 * i + 8 + l];
        }
        // A in Jacobian form with Z = i + 3: (x Z^2, y Z^3, Z)
        fe z, z2, z3;
        fe_set_u32(z, i + 3);
        fe_sqr(z2, z);
        fe_mul(z3, z2, z);
        jac Aj;
        fe_mul(Aj.x, A.x, z2);
        fe_mul(Aj.y, A.y, z3);
        Aj.z = z;
        Aj.inf = 0;
        jac26 a26, r26;
        fe26_from_fe(a26.x, Aj.x);
        fe26_from_fe(a26.y, Aj.y);
        fe26_from_fe(a26.z, Aj.z);
        a26.inf = 0;
        if (i % 11 == 10) jac26_set_inf(a26);
        bool nd = false;
        if (op == kOpJac26Madd) {
            jac26_madd(r26, a26, B, i & 3u, nd);
        } else if (op == kOpJac26Add) {
            // B in Jacobian form with Z = i + 5; B = infinity when i % 13 == 12
            fe zb, zb2, zb3;
            fe_set_u32(zb, i + 5);
            fe_sqr(zb2, zb);
            fe_mul(zb3, zb2, zb);
            jac Bj;
            fe_mul(Bj.x, B.x, zb2);
            fe_mul(Bj.y, B.y, zb3);
            Bj.z = zb;
            Bj.inf = 0;
            jac26 b26;
            jac26_from_jac(b26, Bj);
            if (i % 13 == 12) jac26_set_inf(b26);
            jac26_add(r26, a26, b26);
        } else {
            jac26_dbl(r26, a26);
        }
        jac R;
        jac26_to_jac(R, r26);
        o[16] = R.inf;
        o[17] = nd;
        if (!R.inf) { aff Ra; jac_to_aff(Ra, R); for (int l = 0; l < 8; ++l) { o[l] = Ra.x.v[l]; o[8 + l] = Ra.y.v[l]; } }
    } break;
    default: break;
    }
}

}  // namespace ufs

// ---------------------------------------------------------------------------
// Host-side launchers (plain C++ linkage inside the library)
// ---------------------------------------------------------------------------
namespace ufs {

static inline unsigned grid_for(u32 count, unsigned block) { return (count + block - 1) / block; }

hipError_t launch_verify(int kind, const VerifyLaunch& a, const GcombDesc& g, const uint32_t* mid, unsigned grid,
                         hipStream_t st) {
    if (grid == 0) return hipSuccess;
    gcomb G{reinterpret_cast<const aff*>(g.table), g.w, g.nwin};
    switch (kind) {
    case 0: hipLaunchKernelGGL(k_verify<0>, dim3(grid), dim3(kVerifyBlock), 0, st, a, G, mid); break;
    case 1: hipLaunchKernelGGL(k_verify<1>, dim3(grid), dim3(kVerifyBlock), 0, st, a, G, mid); break;
    case 3: hipLaunchKernelGGL(k_verify<3>, dim3(grid), dim3(kVerifyBlock), 0, st, a, G, mid); break;
    case 4: hipLaunchKernelGGL(k_verify<4>, dim3(grid), dim3(kVerifyBlock), 0, st, a, G, mid); break;
    default: hipLaunchKernelGGL(k_verify<2>, dim3(grid), dim3(kVerifyBlock), 0, st, a, G, mid); break;
    }
    return hipGetLastError();
}

int verify_blocks_per_cu(int kind) {
    int blocks = 0;
    hipError_t e;
    switch (kind) {
    case 0: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, k_verify<0>, kVerifyBlock, 0); break;
    case 1: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, k_verify<1>, kVerifyBlock, 0); break;
    case 3: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, k_verify<3>, kVerifyBlock, 0); break;
    case 4: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, k_verify<4>, kVerifyBlock, 0); break;
    default: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, k_verify<2>, kVerifyBlock, 0); break;
    }
    if (e != hipSuccess || blocks <= 0) return 1;
    return blocks;
}

hipError_t launch_build_gcomb(void* table, int w, int nwin, hipStream_t st) {
    const int total = nwin * (1 << (w - 1)) + 1;
    hipLaunchKernelGGL(k_build_gcomb, dim3((total + 63) / 64), dim3(64), 0, st,
                       reinterpret_cast<aff*>(table), w, nwin);
    return hipGetLastError();
}

hipError_t launch_bip340_midstate(u32* out, hipStream_t st) {
    hipLaunchKernelGGL(k_bip340_midstate, dim3(1), dim3(64), 0, st, out);
    return hipGetLastError();
}

hipError_t launch_selftest(int op, const u32* a, const u32* b, u32* out, u32 count,
                           const GcombDesc& g, void* slots, hipStream_t st) {
    if (count == 0) return hipSuccess;
    gcomb G{reinterpret_cast<const aff*>(g.table), g.w, g.nwin};
    hipLaunchKernelGGL(k_selftest, dim3(grid_for(count, 64)), dim3(64), 0, st, op, a, b, out, count, G,
                       reinterpret_cast<aff*>(slots));
    return hipGetLastError();
}

}  // namespace ufs
