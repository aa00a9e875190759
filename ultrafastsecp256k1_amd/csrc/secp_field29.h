// secp256k1 field arithmetic in 9 x 29-bit limbs for exponentiation chains
// (the square root of key decompression), gfx950.
//
// A chain of squarings and multiplications never adds or subtracts, so every
// multiplication input is a previous product (or an exact conversion) and
// needs no headroom for lazy reduction.  That allows limbs of 29 bits: a
// column of up to 9 limb products stays below 9 * 2^58 < 2^62 in a 64-bit
// v_mad_u64_u32 chain, and a square costs 45 limb products against 55 with
// the 10 x 26-bit limbs of secp_field26.h (81 against 100 for a product),
// with 17 product columns instead of 19 to carry and fold.
//
// Representation: value = sum l[i] * 2^(29 i), limbs 0..7 of 29 bits, limb 8
// of 24 bits (2^256 at the top).  fe29_mul / fe29_sqr accept limbs up to
// 2^29 - 1 + 2^12 (limb 8 up to 2^24 - 1) and return limbs 0..8 within the
// same bounds (only limb 2 can exceed 2^29 - 1, by less than 2^12); the
// worst-case column sums are re-derived by tests/test_field29_bounds.py.
//
// Reduction: 2^261 = Q1 * 2^29 + Q0 (mod p) with Q1 = 2^8, Q0 = 0x7A20
// (2^261 = 32 * 2^256 and 2^256 = 2^32 + 977), so product limb 9 + j folds as
// Q0 into column j and Q1 into column j + 1; limb 17 reaches 2^261 once more
// through its Q1 part and adds Q1 Q0 to column 0 and Q1 Q1 to column 1.  The
// carry t out of bit 256 (limb 8 keeps 24 bits) folds as t (2^32 + 977).
#pragma once
#include "secp_field26.h"

namespace ufs {

struct fe29 { u32 v[9]; };

constexpr u32 kM29 = (1u << 29) - 1;
constexpr u32 kM24 = (1u << 24) - 1;
constexpr u32 kQ0 = 0x7A20u;
constexpr u32 kQ1 = 0x100u;

// 8 x 32 (any value below 2^256) -> 9 x 29, exact limbs.
UFS_DEV void fe29_from_fe(fe29& r, const fe& a) {
#pragma unroll
    for (int i = 0; i < 8; ++i) {
        const int bit = 29 * i, w = bit >> 5, s = bit & 31;
        const u32 x = (s + 29 <= 32) ? (a.v[w] >> s) : __builtin_amdgcn_alignbit(a.v[w + 1], a.v[w], s);
        r.v[i] = x & kM29;
    }
    r.v[8] = a.v[7] >> 8;
}

constexpr int col29_fold_count(int K) { return K > 8 ? 0 : (K == 0 ? 2 : (K == 1 ? 3 : 2)); }

// Folded high limbs h (h[0..7] = 29-bit limbs 9..16 of the product, h[8] =
// the carry out of column 16) as operands of low column K.
template <int K>
UFS_DEV void col29_folds(u32* x, u32* y, const u32 h[9]) {
    if constexpr (K <= 8) {
        x[0] = h[K];
        y[0] = kQ0;
        if constexpr (K >= 1) {
            x[1] = h[K - 1];
            y[1] = kQ1;
        }
        if constexpr (K == 0) {
            x[1] = h[8];
            y[1] = kQ1 * kQ0;
        }
        if constexpr (K == 1) {
            x[2] = h[8];
            y[2] = kQ1 * kQ1;
        }
    }
}

template <int K>
UFS_DEV void mul29_col(u64& acc, const fe29& a, const fe29& b, const u32 h[9]) {
    constexpr int lo = K > 8 ? K - 8 : 0, hi = K > 8 ? 8 : K;
    constexpr int np = hi - lo + 1, N = np + col29_fold_count(K);
    u32 x[N], y[N];
#pragma unroll
    for (int i = 0; i < np; ++i) {
        x[i] = a.v[lo + i];
        y[i] = b.v[K - lo - i];
    }
    col29_folds<K>(x + np, y + np, h);
    Macc<N>::run(acc, x, y);
}

template <int K>
UFS_DEV void sqr29_col(u64& acc, const fe29& a, const u32 d[9], const u32 h[9]) {
    constexpr int lo = K > 8 ? K - 8 : 0;
    constexpr int top = (K - 1) >> 1;
    constexpr int np = (K >= 1 && top >= lo) ? top - lo + 1 : 0;
    constexpr int sq = (K % 2 == 0) ? 1 : 0;
    constexpr int N = np + sq + col29_fold_count(K);
    u32 x[N], y[N];
#pragma unroll
    for (int i = 0; i < np; ++i) {
        x[i] = a.v[lo + i];
        y[i] = d[K - lo - i];
    }
    if constexpr (sq) {
        x[np] = a.v[K / 2];
        y[np] = a.v[K / 2];
    }
    col29_folds<K>(x + np + sq, y + np + sq, h);
    Macc<N>::run(acc, x, y);
}

// Columns 9..16 -> h, then columns 0..8 with the folds -> r, then the carry
// out of bit 256.
template <typename Col>
UFS_DEV void fe29_columns(fe29& r, Col col) {
    u32 h[9];
    u64 acc = 0;
    col(std::integral_constant<int, 9>{}, acc, h);
    h[0] = (u32)acc & kM29; acc >>= 29;
    col(std::integral_constant<int, 10>{}, acc, h);
    h[1] = (u32)acc & kM29; acc >>= 29;
    col(std::integral_constant<int, 11>{}, acc, h);
    h[2] = (u32)acc & kM29; acc >>= 29;
    col(std::integral_constant<int, 12>{}, acc, h);
    h[3] = (u32)acc & kM29; acc >>= 29;
    col(std::integral_constant<int, 13>{}, acc, h);
    h[4] = (u32)acc & kM29; acc >>= 29;
    col(std::integral_constant<int, 14>{}, acc, h);
    h[5] = (u32)acc & kM29; acc >>= 29;
    col(std::integral_constant<int, 15>{}, acc, h);
    h[6] = (u32)acc & kM29; acc >>= 29;
    col(std::integral_constant<int, 16>{}, acc, h);
    h[7] = (u32)acc & kM29; acc >>= 29;
    h[8] = (u32)acc;
    u32 l[9];
    acc = 0;
    col(std::integral_constant<int, 0>{}, acc, h);
    l[0] = (u32)acc & kM29; acc >>= 29;
    col(std::integral_constant<int, 1>{}, acc, h);
    l[1] = (u32)acc & kM29; acc >>= 29;
    col(std::integral_constant<int, 2>{}, acc, h);
    l[2] = (u32)acc & kM29; acc >>= 29;
    col(std::integral_constant<int, 3>{}, acc, h);
    l[3] = (u32)acc & kM29; acc >>= 29;
    col(std::integral_constant<int, 4>{}, acc, h);
    l[4] = (u32)acc & kM29; acc >>= 29;
    col(std::integral_constant<int, 5>{}, acc, h);
    l[5] = (u32)acc & kM29; acc >>= 29;
    col(std::integral_constant<int, 6>{}, acc, h);
    l[6] = (u32)acc & kM29; acc >>= 29;
    col(std::integral_constant<int, 7>{}, acc, h);
    l[7] = (u32)acc & kM29; acc >>= 29;
    col(std::integral_constant<int, 8>{}, acc, h);
    l[8] = (u32)acc & kM24; acc >>= 24;
    const u64 t = acc;
    const u64 x0 = t * 977u + l[0];
    r.v[0] = (u32)x0 & kM29;
    const u64 x1 = (t << 3) + (x0 >> 29) + l[1];
    r.v[1] = (u32)x1 & kM29;
    r.v[2] = l[2] + (u32)(x1 >> 29);
#pragma unroll
    for (int i = 3; i < 9; ++i) r.v[i] = l[i];
}

UFS_DEV void fe29_mul(fe29& r, const fe29& a, const fe29& b) {
    fe29_columns(r, [&](auto k, u64& acc, const u32* h) { mul29_col<decltype(k)::value>(acc, a, b, h); });
}

UFS_DEV void fe29_sqr(fe29& r, const fe29& a) {
    u32 d[9];
#pragma unroll
    for (int i = 0; i < 9; ++i) d[i] = a.v[i] << 1;
    fe29_columns(r, [&](auto k, u64& acc, const u32* h) { sqr29_col<decltype(k)::value>(acc, a, d, h); });
}

// One round of: fold the bits at or above 2^256 (x (2^32 + 977)), then one
// carry pass over limbs 0..7.
UFS_DEV void fe29_fold_carry(u32 l[9]) {
    const u32 x = l[8] >> 24;
    l[8] &= kM24;
    l[0] += x * 977u;
    l[1] += x << 3;
#pragma unroll
    for (int i = 0; i < 8; ++i) {
        l[i + 1] += l[i] >> 29;
        l[i] &= kM29;
    }
}

// fe29_mul / fe29_sqr output (or exact limbs) -> 8 x 32 value below 2^256,
// congruent.  Two fold-and-carry rounds: after the first the value is below
// 2^256 + 2^34, after the second below 2^256.
UFS_DEV void fe29_to_fe(fe& r, const fe29& a) {
    u32 l[9];
#pragma unroll
    for (int i = 0; i < 9; ++i) l[i] = a.v[i];
    fe29_fold_carry(l);
    fe29_fold_carry(l);
    u64 acc = 0;
    int bits = 0, w = 0;
#pragma unroll
    for (int i = 0; i < 9; ++i) {
        acc |= (u64)l[i] << bits;
        bits += 29;
        while (bits >= 32) {
            r.v[w++] = (u32)acc;
            acc >>= 32;
            bits -= 32;
        }
    }
}

// r = a^((p+1)/4) in 8 x 32 form; returns true iff r^2 == a (a is a
// quadratic residue).  The exponent chain (253 squarings, 13
// multiplications) runs on 9 x 29-bit limbs, all inline: the chain's values
// are products, so their magnitude stays at the multiply's output bound.
// (Out-of-line squarings would keep the live values across calls and spill.)
UFS_DEV bool fe_sqrt(fe& r, const fe& a8) {
    fe29 a, x2, x3, x11, x22, x44, x88, t;
    fe29_from_fe(a, a8);
    auto sqr_n = [](fe29& v, int n) {
        for (int i = 0; i < n; ++i) fe29_sqr(v, v);
    };
    fe29_sqr(t, a);
    fe29_mul(x2, t, a);
    fe29_sqr(t, x2);
    fe29_mul(x3, t, a);
    t = x3;
    sqr_n(t, 3);
    fe29_mul(t, t, x3);            // x6
    sqr_n(t, 3);
    fe29_mul(t, t, x3);            // x9
    sqr_n(t, 2);
    fe29_mul(x11, t, x2);
    t = x11;
    sqr_n(t, 11);
    fe29_mul(x22, t, x11);
    t = x22;
    sqr_n(t, 22);
    fe29_mul(x44, t, x22);
    t = x44;
    sqr_n(t, 44);
    fe29_mul(x88, t, x44);
    t = x88;
    sqr_n(t, 88);
    fe29_mul(t, t, x88);           // x176
    sqr_n(t, 44);
    fe29_mul(t, t, x44);           // x220
    sqr_n(t, 3);
    fe29_mul(t, t, x3);            // x223
    sqr_n(t, 23);
    fe29_mul(t, t, x22);
    sqr_n(t, 6);
    fe29_mul(t, t, x2);
    sqr_n(t, 2);
    fe29_to_fe(r, t);
    fe c;
    fe_sqr(c, r);
    return fe_equal(c, a8);
}

}  // namespace ufs
