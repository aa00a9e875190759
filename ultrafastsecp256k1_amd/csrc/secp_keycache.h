// Device-resident public-key decompression cache of the serving loop.
//
// 2^k entries of 64 bytes, (x, y with even parity) in the 8 x 32 storage form,
// grouped in 2^(k-1) two-way sets: the set of x is key_hash(x, mask >> 1)
// (mask = entries - 1).  New keys go to way 1 of their set; a key found in
// way 1 is copied to way 0, which only such copies overwrite, so keys seen
// once churn way 1 and leave reused keys in way 0.
//
// Lanes read and write entries concurrently, without locks.  An entry is
// trusted only if its x equals the looked-up x and (x, y) is on the curve, so
// an entry torn by a copy or an insert in flight (words of two different
// keys) is taken as a miss and never gives a wrong point.
#pragma once
#include "secp_ecmult.h"

namespace ufs {

UFS_DEV u32 key_hash(const u32 x[8], u32 mask) {
    u32 h = x[0] ^ (x[3] * 0x9E3779B9u) ^ (x[6] * 0x85EBCA6Bu);
    h ^= h >> 15;
    return h & mask;
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

// Point with x = x (x < p, checked by the caller) and y of parity `parity`
// if the cache holds x; a hit in way 1 is copied to way 0.
UFS_DEV bool key_cache_find(aff& P, aff* entries, u32 mask, const u32 x[8], u32 parity) {
    aff* set = entries + 2u * key_hash(x, mask >> 1);
    bool hit = false;
#pragma unroll 1
    for (u32 way = 0; way < 2 && !hit; ++way) {
        aff c;
        load_aff(c, set + way);
        bool same = true;
#pragma unroll
        for (int k = 0; k < 8; ++k) same &= c.x.v[k] == x[k];
        if (!same) continue;
        fe x3, y2, seven;
        fe_sqr(x3, c.x);
        fe_mul(x3, x3, c.x);
        fe_set_u32(seven, 7);
        fe_add(x3, x3, seven);
        fe_sqr(y2, c.y);
        if (fe_equal(x3, y2)) {
            P.x = c.x;
            apply_parity(P.y, c.y, parity);
            if (way == 1) store_aff(set, c);
            hit = true;
        }
    }
    return hit;
}

// Insert a decompressed key (y of even parity) into way 1 of its set.
UFS_DEV void key_cache_insert(aff* entries, u32 mask, const aff& even) {
    store_aff(entries + 2u * key_hash(even.x.v, mask >> 1) + 1u, even);
}

}  // namespace ufs
