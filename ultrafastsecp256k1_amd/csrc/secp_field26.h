// secp256k1 field arithmetic in 10 x 26-bit limbs for the scalar-multiplication
// loops (gfx950).
//
// Why a second representation: on gfx950 a 64-bit multiply-add
// (v_mad_u64_u32) and every carry-writing add (v_add_co_u32 / v_addc_co_u32)
// issue at about half the rate of a plain 32-bit add.  With 26-bit limbs a
// column of 10 limb products accumulates in 64 bits without carry-outs, and
// additions / subtractions are limb-wise 32-bit adds with no carry chain
// (lazy reduction), so the doubling and mixed-addition formulas run on
// full-rate adds and carry-free multiply-adds.  Values enter from and leave to
// the 8 x 32 form (secp_field.h) at the boundaries of the loops.
//
// Representation: value = sum l[i] * 2^(26 i), limbs 0..8 nominally 26 bits,
// limb 9 nominally 22 bits (2^256 at the top).  A value of "magnitude m" has
// l[i] <= m * (2^26 - 1) (i < 9) and l[9] <= m * (2^22 - 1).  Limbs are not
// canonical; only fe26_to_fe and the zero tests normalise.
//
// Bounds (worst-case limb-bound propagation over the doubling and mixed
// addition of secp_point26.h, which reach a fixed point of magnitudes
// X <= 6.1, Y <= 10.1, Z <= 5.1; re-derived by tests/test_field26_bounds.py):
//   fe26_mul / fe26_sqr : inputs of magnitude <= 20 (the 64-bit column sums
//                         stay below 2^64); the loops use at most 16.  Output:
//                         limbs <= 2^26 - 1 except l[2] <= 2^26 - 1 + 2^22,
//                         l[9] <= 2^22 - 1 (magnitude about 1.06).
//   fe26_sub<K>         : b limb-wise at most K * p's limb; adds K * p.
//   fe26_norm           : input limbs below 2^31; output limbs 0..8 exact
//                         26-bit, l[9] <= 2^22 + 64, value below 2p.
#pragma once
#include <type_traits>

#include "secp_field.h"

namespace ufs {

struct fe26 { u32 v[10]; };

constexpr u32 kM26 = (1u << 26) - 1;
constexpr u32 kM22 = (1u << 22) - 1;
// p = 2^256 - 2^32 - 977 in 26-bit limbs
constexpr u32 kP26[10] = {0x3FFFC2Fu, 0x3FFFFBFu, kM26, kM26, kM26, kM26, kM26, kM26, kM26, kM22};
// 2^260 = R1 * 2^26 + R0 (mod p)
constexpr u32 kR0 = 0x3D10u;
constexpr u32 kR1 = 0x400u;

// 8 x 32 (any value below 2^256) -> 10 x 26, exact limbs.
UFS_DEV void fe26_from_fe(fe26& r, const fe& a) {
#pragma unroll
    for (int i = 0; i < 9; ++i) {
        const int bit = 26 * i, w = bit >> 5, s = bit & 31;
        const u32 x = (s + 26 <= 32) ? (a.v[w] >> s) : __builtin_amdgcn_alignbit(a.v[w + 1], a.v[w], s);
        r.v[i] = x & kM26;
    }
    r.v[9] = a.v[7] >> 10;
}

UFS_DEV void fe26_add(fe26& r, const fe26& a, const fe26& b) {
#pragma unroll
    for (int i = 0; i < 10; ++i) r.v[i] = a.v[i] + b.v[i];
}

// r = a + K p - b: needs b's limbs at most K times p's limbs.
template <u32 K>
UFS_DEV void fe26_sub(fe26& r, const fe26& a, const fe26& b) {
#pragma unroll
    for (int i = 0; i < 10; ++i) r.v[i] = a.v[i] + (K * kP26[i] - b.v[i]);
}

template <u32 K>
UFS_DEV void fe26_neg(fe26& r, const fe26& b) {
#pragma unroll
    for (int i = 0; i < 10; ++i) r.v[i] = K * kP26[i] - b.v[i];
}

UFS_DEV void fe26_mul3(fe26& r, const fe26& a) {
#pragma unroll
    for (int i = 0; i < 10; ++i) r.v[i] = (a.v[i] << 1) + a.v[i];
}

template <int S>
UFS_DEV void fe26_shl(fe26& r, const fe26& a) {
#pragma unroll
    for (int i = 0; i < 10; ++i) r.v[i] = a.v[i] << S;
}

// acc += sum_i x[i] * y[i] as one chain of v_mad_u64_u32 on the 64-bit
// accumulator (carry-outs unused: the column sums stay below 2^64).  One asm
// statement per column keeps the chain exactly as written: left to the
// compiler, the column would be summed separately and added to the incoming
// carry with a 64-bit add, another half-rate instruction per column.
// (Generated for N = 1..12.)
template <int N> struct Macc;
template <> struct Macc<1> {
    UFS_DEV static void run(u64& acc, const u32* x, const u32* y) {
        u64 dead;
        asm("v_mad_u64_u32 %0, %1, %2, %3, %0"
            : "+v"(acc), "=&s"(dead)
            : "v"(x[0]), "v"(y[0]));
    }
};
template <> struct Macc<2> {
    UFS_DEV static void run(u64& acc, const u32* x, const u32* y) {
        u64 dead;
        asm("v_mad_u64_u32 %0, %1, %2, %3, %0\n\tv_mad_u64_u32 %0, %1, %4, %5, %0"
            : "+v"(acc), "=&s"(dead)
            : "v"(x[0]), "v"(y[0]), "v"(x[1]), "v"(y[1]));
    }
};
template <> struct Macc<3> {
    UFS_DEV static void run(u64& acc, const u32* x, const u32* y) {
        u64 dead;
        asm("v_mad_u64_u32 %0, %1, %2, %3, %0\n\tv_mad_u64_u32 %0, %1, %4, %5, %0\n\tv_mad_u64_u32 %0, %1, %6, %7, %0"
            : "+v"(acc), "=&s"(dead)
            : "v"(x[0]), "v"(y[0]), "v"(x[1]), "v"(y[1]), "v"(x[2]), "v"(y[2]));
    }
};
template <> struct Macc<4> {
    UFS_DEV static void run(u64& acc, const u32* x, const u32* y) {
        u64 dead;
        asm("v_mad_u64_u32 %0, %1, %2, %3, %0\n\tv_mad_u64_u32 %0, %1, %4, %5, %0\n\tv_mad_u64_u32 %0, %1, %6, %7, %0\n\tv_mad_u64_u32 %0, %1, %8, %9, %0"
            : "+v"(acc), "=&s"(dead)
            : "v"(x[0]), "v"(y[0]), "v"(x[1]), "v"(y[1]), "v"(x[2]), "v"(y[2]), "v"(x[3]), "v"(y[3]));
    }
};
template <> struct Macc<5> {
    UFS_DEV static void run(u64& acc, const u32* x, const u32* y) {
        u64 dead;
        asm("v_mad_u64_u32 %0, %1, %2, %3, %0\n\tv_mad_u64_u32 %0, %1, %4, %5, %0\n\tv_mad_u64_u32 %0, %1, %6, %7, %0\n\tv_mad_u64_u32 %0, %1, %8, %9, %0\n\tv_mad_u64_u32 %0, %1, %10, %11, %0"
            : "+v"(acc), "=&s"(dead)
            : "v"(x[0]), "v"(y[0]), "v"(x[1]), "v"(y[1]), "v"(x[2]), "v"(y[2]), "v"(x[3]), "v"(y[3]), "v"(x[4]), "v"(y[4]));
    }
};
template <> struct Macc<6> {
    UFS_DEV static void run(u64& acc, const u32* x, const u32* y) {
        u64 dead;
        asm("v_mad_u64_u32 %0, %1, %2, %3, %0\n\tv_mad_u64_u32 %0, %1, %4, %5, %0\n\tv_mad_u64_u32 %0, %1, %6, %7, %0\n\tv_mad_u64_u32 %0, %1, %8, %9, %0\n\tv_mad_u64_u32 %0, %1, %10, %11, %0\n\tv_mad_u64_u32 %0, %1, %12, %13, %0"
            : "+v"(acc), "=&s"(dead)
            : "v"(x[0]), "v"(y[0]), "v"(x[1]), "v"(y[1]), "v"(x[2]), "v"(y[2]), "v"(x[3]), "v"(y[3]), "v"(x[4]), "v"(y[4]), "v"(x[5]), "v"(y[5]));
    }
};
template <> struct Macc<7> {
    UFS_DEV static void run(u64& acc, const u32* x, const u32* y) {
        u64 dead;
        asm("v_mad_u64_u32 %0, %1, %2, %3, %0\n\tv_mad_u64_u32 %0, %1, %4, %5, %0\n\tv_mad_u64_u32 %0, %1, %6, %7, %0\n\tv_mad_u64_u32 %0, %1, %8, %9, %0\n\tv_mad_u64_u32 %0, %1, %10, %11, %0\n\tv_mad_u64_u32 %0, %1, %12, %13, %0\n\tv_mad_u64_u32 %0, %1, %14, %15, %0"
            : "+v"(acc), "=&s"(dead)
            : "v"(x[0]), "v"(y[0]), "v"(x[1]), "v"(y[1]), "v"(x[2]), "v"(y[2]), "v"(x[3]), "v"(y[3]), "v"(x[4]), "v"(y[4]), "v"(x[5]), "v"(y[5]), "v"(x[6]), "v"(y[6]));
    }
};
template <> struct Macc<8> {
    UFS_DEV static void run(u64& acc, const u32* x, const u32* y) {
        u64 dead;
        asm("v_mad_u64_u32 %0, %1, %2, %3, %0\n\tv_mad_u64_u32 %0, %1, %4, %5, %0\n\tv_mad_u64_u32 %0, %1, %6, %7, %0\n\tv_mad_u64_u32 %0, %1, %8, %9, %0\n\tv_mad_u64_u32 %0, %1, %10, %11, %0\n\tv_mad_u64_u32 %0, %1, %12, %13, %0\n\tv_mad_u64_u32 %0, %1, %14, %15, %0\n\tv_mad_u64_u32 %0, %1, %16, %17, %0"
            : "+v"(acc), "=&s"(dead)
            : "v"(x[0]), "v"(y[0]), "v"(x[1]), "v"(y[1]), "v"(x[2]), "v"(y[2]), "v"(x[3]), "v"(y[3]), "v"(x[4]), "v"(y[4]), "v"(x[5]), "v"(y[5]), "v"(x[6]), "v"(y[6]), "v"(x[7]), "v"(y[7]));
    }
};
template <> struct Macc<9> {
    UFS_DEV static void run(u64& acc, const u32* x, const u32* y) {
        u64 dead;
        asm("v_mad_u64_u32 %0, %1, %2, %3, %0\n\tv_mad_u64_u32 %0, %1, %4, %5, %0\n\tv_mad_u64_u32 %0, %1, %6, %7, %0\n\tv_mad_u64_u32 %0, %1, %8, %9, %0\n\tv_mad_u64_u32 %0, %1, %10, %11, %0\n\tv_mad_u64_u32 %0, %1, %12, %13, %0\n\tv_mad_u64_u32 %0, %1, %14, %15, %0\n\tv_mad_u64_u32 %0, %1, %16, %17, %0\n\tv_mad_u64_u32 %0, %1, %18, %19, %0"
            : "+v"(acc), "=&s"(dead)
            : "v"(x[0]), "v"(y[0]), "v"(x[1]), "v"(y[1]), "v"(x[2]), "v"(y[2]), "v"(x[3]), "v"(y[3]), "v"(x[4]), "v"(y[4]), "v"(x[5]), "v"(y[5]), "v"(x[6]), "v"(y[6]), "v"(x[7]), "v"(y[7]), "v"(x[8]), "v"(y[8]));
    }
};
template <> struct Macc<10> {
    UFS_DEV static void run(u64& acc, const u32* x, const u32* y) {
        u64 dead;
        asm("v_mad_u64_u32 %0, %1, %2, %3, %0\n\tv_mad_u64_u32 %0, %1, %4, %5, %0\n\tv_mad_u64_u32 %0, %1, %6, %7, %0\n\tv_mad_u64_u32 %0, %1, %8, %9, %0\n\tv_mad_u64_u32 %0, %1, %10, %11, %0\n\tv_mad_u64_u32 %0, %1, %12, %13, %0\n\tv_mad_u64_u32 %0, %1, %14, %15, %0\n\tv_mad_u64_u32 %0, %1, %16, %17, %0\n\tv_mad_u64_u32 %0, %1, %18, %19, %0\n\tv_mad_u64_u32 %0, %1, %20, %21, %0"
            : "+v"(acc), "=&s"(dead)
            : "v"(x[0]), "v"(y[0]), "v"(x[1]), "v"(y[1]), "v"(x[2]), "v"(y[2]), "v"(x[3]), "v"(y[3]), "v"(x[4]), "v"(y[4]), "v"(x[5]), "v"(y[5]), "v"(x[6]), "v"(y[6]), "v"(x[7]), "v"(y[7]), "v"(x[8]), "v"(y[8]), "v"(x[9]), "v"(y[9]));
    }
};
template <> struct Macc<11> {
    UFS_DEV static void run(u64& acc, const u32* x, const u32* y) {
        u64 dead;
        asm("v_mad_u64_u32 %0, %1, %2, %3, %0\n\tv_mad_u64_u32 %0, %1, %4, %5, %0\n\tv_mad_u64_u32 %0, %1, %6, %7, %0\n\tv_mad_u64_u32 %0, %1, %8, %9, %0\n\tv_mad_u64_u32 %0, %1, %10, %11, %0\n\tv_mad_u64_u32 %0, %1, %12, %13, %0\n\tv_mad_u64_u32 %0, %1, %14, %15, %0\n\tv_mad_u64_u32 %0, %1, %16, %17, %0\n\tv_mad_u64_u32 %0, %1, %18, %19, %0\n\tv_mad_u64_u32 %0, %1, %20, %21, %0\n\tv_mad_u64_u32 %0, %1, %22, %23, %0"
            : "+v"(acc), "=&s"(dead)
            : "v"(x[0]), "v"(y[0]), "v"(x[1]), "v"(y[1]), "v"(x[2]), "v"(y[2]), "v"(x[3]), "v"(y[3]), "v"(x[4]), "v"(y[4]), "v"(x[5]), "v"(y[5]), "v"(x[6]), "v"(y[6]), "v"(x[7]), "v"(y[7]), "v"(x[8]), "v"(y[8]), "v"(x[9]), "v"(y[9]), "v"(x[10]), "v"(y[10]));
    }
};
template <> struct Macc<12> {
    UFS_DEV static void run(u64& acc, const u32* x, const u32* y) {
        u64 dead;
        asm("v_mad_u64_u32 %0, %1, %2, %3, %0\n\tv_mad_u64_u32 %0, %1, %4, %5, %0\n\tv_mad_u64_u32 %0, %1, %6, %7, %0\n\tv_mad_u64_u32 %0, %1, %8, %9, %0\n\tv_mad_u64_u32 %0, %1, %10, %11, %0\n\tv_mad_u64_u32 %0, %1, %12, %13, %0\n\tv_mad_u64_u32 %0, %1, %14, %15, %0\n\tv_mad_u64_u32 %0, %1, %16, %17, %0\n\tv_mad_u64_u32 %0, %1, %18, %19, %0\n\tv_mad_u64_u32 %0, %1, %20, %21, %0\n\tv_mad_u64_u32 %0, %1, %22, %23, %0\n\tv_mad_u64_u32 %0, %1, %24, %25, %0"
            : "+v"(acc), "=&s"(dead)
            : "v"(x[0]), "v"(y[0]), "v"(x[1]), "v"(y[1]), "v"(x[2]), "v"(y[2]), "v"(x[3]), "v"(y[3]), "v"(x[4]), "v"(y[4]), "v"(x[5]), "v"(y[5]), "v"(x[6]), "v"(y[6]), "v"(x[7]), "v"(y[7]), "v"(x[8]), "v"(y[8]), "v"(x[9]), "v"(y[9]), "v"(x[10]), "v"(y[10]), "v"(x[11]), "v"(y[11]));
    }
};

// Operands of column K (0..18) of a*b, plus for K <= 9 the folded high limbs
// h (h[0..8] = 26-bit limbs 10..18 of the product, h[9] = the carry out of
// column 18):
//   h_j * 2^(26 j) = h_j * 2^(26 (j-10)) * 2^260 and 2^260 = R1 2^26 + R0,
// so h_j adds h_j R0 to column j-10 and h_j R1 to column j-9; for j = 19 the
// R1 part lands on 2^260 once more and adds h_19 R1 R0 to column 0 and
// h_19 R1 R1 to column 1.
constexpr int col_fold_count(int K) { return K > 9 ? 0 : (K == 0 ? 2 : (K == 1 ? 3 : 2)); }

template <int K>
UFS_DEV void col_folds(u32* x, u32* y, const u32 h[10]) {
    if constexpr (K <= 9) {
        x[0] = h[K];
        y[0] = kR0;
        if constexpr (K >= 1) {
            x[1] = h[K - 1];
            y[1] = kR1;
        }
        if constexpr (K == 0) {
            x[1] = h[9];
            y[1] = kR1 * kR0;
        }
        if constexpr (K == 1) {
            x[2] = h[9];
            y[2] = kR1 * kR1;
        }
    }
}

template <int K>
UFS_DEV void mul_col(u64& acc, const fe26& a, const fe26& b, const u32 h[10]) {
    constexpr int lo = K > 9 ? K - 9 : 0, hi = K > 9 ? 9 : K;
    constexpr int np = hi - lo + 1, N = np + col_fold_count(K);
    u32 x[N], y[N];
#pragma unroll
    for (int i = 0; i < np; ++i) {
        x[i] = a.v[lo + i];
        y[i] = b.v[K - lo - i];
    }
    col_folds<K>(x + np, y + np, h);
    Macc<N>::run(acc, x, y);
}

// Square column K: cross products a_i (2 a_j), i < j, and a_{K/2}^2.
template <int K>
UFS_DEV void sqr_col(u64& acc, const fe26& a, const u32 d[10], const u32 h[10]) {
    constexpr int lo = K > 9 ? K - 9 : 0;
    constexpr int top = (K - 1) >> 1;                     // largest i with i < K - i
    constexpr int np = (K >= 1 && top >= lo) ? top - lo + 1 : 0;
    constexpr int sq = (K % 2 == 0) ? 1 : 0;
    constexpr int N = np + sq + col_fold_count(K);
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
    col_folds<K>(x + np + sq, y + np + sq, h);
    Macc<N>::run(acc, x, y);
}

// Columns 10..18 -> h (26-bit limbs, h[9] the final carry), then columns
// 0..9 with the folds -> r.  The carry t out of 2^256 (limb 9 keeps 22 bits)
// is below 2^42 and folds as t (2^32 + 977) into limbs 0..2.
template <typename Col>
UFS_DEV void fe26_columns(fe26& r, Col col) {
    u32 h[10];
    u64 acc = 0;
    col(std::integral_constant<int, 10>{}, acc, h);
    h[0] = (u32)acc & kM26; acc >>= 26;
    col(std::integral_constant<int, 11>{}, acc, h);
    h[1] = (u32)acc & kM26; acc >>= 26;
    col(std::integral_constant<int, 12>{}, acc, h);
    h[2] = (u32)acc & kM26; acc >>= 26;
    col(std::integral_constant<int, 13>{}, acc, h);
    h[3] = (u32)acc & kM26; acc >>= 26;
    col(std::integral_constant<int, 14>{}, acc, h);
    h[4] = (u32)acc & kM26; acc >>= 26;
    col(std::integral_constant<int, 15>{}, acc, h);
    h[5] = (u32)acc & kM26; acc >>= 26;
    col(std::integral_constant<int, 16>{}, acc, h);
    h[6] = (u32)acc & kM26; acc >>= 26;
    col(std::integral_constant<int, 17>{}, acc, h);
    h[7] = (u32)acc & kM26; acc >>= 26;
    col(std::integral_constant<int, 18>{}, acc, h);
    h[8] = (u32)acc & kM26; acc >>= 26;
    h[9] = (u32)acc;
    u32 l[10];
    acc = 0;
    col(std::integral_constant<int, 0>{}, acc, h);
    l[0] = (u32)acc & kM26; acc >>= 26;
    col(std::integral_constant<int, 1>{}, acc, h);
    l[1] = (u32)acc & kM26; acc >>= 26;
    col(std::integral_constant<int, 2>{}, acc, h);
    l[2] = (u32)acc & kM26; acc >>= 26;
    col(std::integral_constant<int, 3>{}, acc, h);
    l[3] = (u32)acc & kM26; acc >>= 26;
    col(std::integral_constant<int, 4>{}, acc, h);
    l[4] = (u32)acc & kM26; acc >>= 26;
    col(std::integral_constant<int, 5>{}, acc, h);
    l[5] = (u32)acc & kM26; acc >>= 26;
    col(std::integral_constant<int, 6>{}, acc, h);
    l[6] = (u32)acc & kM26; acc >>= 26;
    col(std::integral_constant<int, 7>{}, acc, h);
    l[7] = (u32)acc & kM26; acc >>= 26;
    col(std::integral_constant<int, 8>{}, acc, h);
    l[8] = (u32)acc & kM26; acc >>= 26;
    col(std::integral_constant<int, 9>{}, acc, h);
    l[9] = (u32)acc & kM22; acc >>= 22;
    const u64 t = acc;
    const u64 x0 = t * 977u + l[0];
    r.v[0] = (u32)x0 & kM26;
    const u64 x1 = (t << 6) + (x0 >> 26) + l[1];
    r.v[1] = (u32)x1 & kM26;
    r.v[2] = l[2] + (u32)(x1 >> 26);
#pragma unroll
    for (int i = 3; i < 10; ++i) r.v[i] = l[i];
}

UFS_DEV void fe26_mul(fe26& r, const fe26& a, const fe26& b) {
    fe26_columns(r, [&](auto k, u64& acc, const u32* h) { mul_col<decltype(k)::value>(acc, a, b, h); });
}

UFS_DEV void fe26_sqr(fe26& r, const fe26& a) {
    u32 d[10];
#pragma unroll
    for (int i = 0; i < 10; ++i) d[i] = a.v[i] << 1;
    fe26_columns(r, [&](auto k, u64& acc, const u32* h) { sqr_col<decltype(k)::value>(acc, a, d, h); });
}

// Weak normalisation: bits at or above 2^256 folded first (x (2^32 + 977)),
// then one carry pass.  Input limbs below 2^31; output limbs 0..8 exact
// 26-bit, l[9] <= 2^22 + 64, value below 2p.
UFS_DEV void fe26_norm(fe26& r, const fe26& a) {
    const u32 x = a.v[9] >> 22;
    u32 l[10];
#pragma unroll
    for (int i = 0; i < 10; ++i) l[i] = a.v[i];
    l[9] &= kM22;
    l[0] += x * 977u;
    l[1] += x << 6;
#pragma unroll
    for (int i = 0; i < 9; ++i) {
        l[i + 1] += l[i] >> 26;
        l[i] &= kM26;
    }
#pragma unroll
    for (int i = 0; i < 10; ++i) r.v[i] = l[i];
}

// For a fe26_norm output (value below 2p): is the value 0 or p?
UFS_DEV bool fe26_normed_is_zero(const fe26& a) {
    u32 z = 0, q = 0;
#pragma unroll
    for (int i = 0; i < 10; ++i) {
        z |= a.v[i];
        q |= a.v[i] ^ kP26[i];
    }
    return z == 0 || q == 0;
}

// Necessary condition for a fe26_norm output to be 0 or p (limb 2 of both).
UFS_DEV bool fe26_normed_maybe_zero(const fe26& a) { return a.v[2] == 0 || a.v[2] == kM26; }

// Any fe26 with limbs below 2^31 -> 8 x 32 value below 2^256, congruent.
UFS_DEV void fe26_to_fe(fe& r, const fe26& a) {
    fe26 n;
    fe26_norm(n, a);
    // value < 2^256 + 2^241: one more fold if limb 9 overflowed, full carry
    const u32 x = n.v[9] >> 22;
    n.v[9] &= kM22;
    n.v[0] += x * 977u;
    n.v[1] += x << 6;
#pragma unroll
    for (int i = 0; i < 9; ++i) {
        n.v[i + 1] += n.v[i] >> 26;
        n.v[i] &= kM26;
    }
    // the value is now below 2^256 (n.v[9] < 2^22): pack 26-bit limbs
    u64 acc = 0;
    int bits = 0, w = 0;
#pragma unroll
    for (int i = 0; i < 10; ++i) {
        acc |= (u64)n.v[i] << bits;
        bits += 26;
        if (bits >= 32) {
            r.v[w++] = (u32)acc;
            acc >>= 32;
            bits -= 32;
        }
    }
}

}  // namespace ufs
