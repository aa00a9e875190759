// secp256k1 field (mod p) and scalar (mod n) arithmetic for gfx950.
//
// Representation: 8 x 32-bit little-endian limbs.  Field elements are kept
// weakly reduced: any value in [0, 2^256) congruent to the element; only
// comparisons, parity and serialisation normalise to [0, p).
//
// Products use column-wise (product-scanning) accumulation into a 96-bit
// accumulator (64-bit lo + 32-bit hi).  Each 32x32->64 multiply-accumulate is
// one v_mad_u64_u32 whose carry-out feeds one v_addc_co_u32 into the hi word.
// Reduction mod p folds the high half with 2^256 = 2^32 + 977 (mod p);
// reduction mod n folds with 2^256 = NC (mod n), NC = 2^256 - n (129 bits).
//
// Semantics follow the reference's field/scalar layer
// (src/cuda/include/secp256k1.cuh:1066-1547 field ops, :563-843 scalar ops,
// :2614-2782 inverse/sqrt) but the limb layout, reduction and chains are this
// project's own.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace ufs {

typedef uint8_t u8;
typedef uint32_t u32;
typedef uint64_t u64;

#define UFS_DEV __device__ __forceinline__

struct fe { u32 v[8]; };   // field element mod p (weakly reduced)
struct sc { u32 v[8]; };   // scalar mod n (fully reduced unless stated)

// ---------------------------------------------------------------------------
// Curve constants (little-endian 32-bit limbs).  All are closed-form public
// values: p = 2^256 - 2^32 - 977, n = group order, beta^3 = 1 mod p,
// lambda^3 = 1 mod n with lambda*G = (beta*Gx, Gy).
// ---------------------------------------------------------------------------
#define UFS_P_LIMBS   0xFFFFFC2FU, 0xFFFFFFFEU, 0xFFFFFFFFU, 0xFFFFFFFFU, 0xFFFFFFFFU, 0xFFFFFFFFU, 0xFFFFFFFFU, 0xFFFFFFFFU
#define UFS_N_LIMBS   0xD0364141U, 0xBFD25E8CU, 0xAF48A03BU, 0xBAAEDCE6U, 0xFFFFFFFEU, 0xFFFFFFFFU, 0xFFFFFFFFU, 0xFFFFFFFFU
#define UFS_NH_LIMBS  0x681B20A0U, 0xDFE92F46U, 0x57A4501DU, 0x5D576E73U, 0xFFFFFFFFU, 0xFFFFFFFFU, 0xFFFFFFFFU, 0x7FFFFFFFU
#define UFS_GX_LIMBS  0x16F81798U, 0x59F2815BU, 0x2DCE28D9U, 0x029BFCDBU, 0xCE870B07U, 0x55A06295U, 0xF9DCBBACU, 0x79BE667EU
#define UFS_GY_LIMBS  0xFB10D4B8U, 0x9C47D08FU, 0xA6855419U, 0xFD17B448U, 0x0E1108A8U, 0x5DA4FBFCU, 0x26A3C465U, 0x483ADA77U
#define UFS_BETA_LIMBS 0x719501EEU, 0xC1396C28U, 0x12F58995U, 0x9CF04975U, 0xAC3434E9U, 0x6E64479EU, 0x657C0710U, 0x7AE96A2BU
#define UFS_MLAMBDA_LIMBS 0xB51283CFU, 0xE0CFC810U, 0x8EC739C2U, 0xA880B9FCU, 0x77ED9BA4U, 0x5AD9E3FDU, 0x3FA3CF1FU, 0xAC9C52B3U
#define UFS_MB1_LIMBS 0x0ABFE4C3U, 0x6F547FA9U, 0x010E8828U, 0xE4437ED6U, 0U, 0U, 0U, 0U
#define UFS_MB2_LIMBS 0x3DB1562CU, 0xD765CDA8U, 0x0774346DU, 0x8A280AC5U, 0xFFFFFFFEU, 0xFFFFFFFFU, 0xFFFFFFFFU, 0xFFFFFFFFU
#define UFS_G1_LIMBS  0x45DBB031U, 0xE893209AU, 0x71E8CA7FU, 0x3DAA8A14U, 0x9284EB15U, 0xE86C90E4U, 0xA7D46BCDU, 0x3086D221U
#define UFS_G2_LIMBS  0x8AC47F71U, 0x1571B4AEU, 0x9DF506C6U, 0x221208ACU, 0x0ABFE4C4U, 0x6F547FA9U, 0x010E8828U, 0xE4437ED6U
#define UFS_PMN_LIMBS 0x2FC9BAEEU, 0x402DA172U, 0x50B75FC4U, 0x45512319U, 0x00000001U, 0U, 0U, 0U
// 2^256 - n
#define UFS_NC0 0x2FC9BEBFU
#define UFS_NC1 0x402DA173U
#define UFS_NC2 0x50B75FC4U
#define UFS_NC3 0x45512319U

// ---------------------------------------------------------------------------
// 32-bit carry helpers
// ---------------------------------------------------------------------------
UFS_DEV u32 addc(u32 a, u32 b, u32 cin, u32& cout) {
    u32 c;
    u32 r = __builtin_addc(a, b, cin, &c);
    cout = c;
    return r;
}
UFS_DEV u32 subc(u32 a, u32 b, u32 bin, u32& bout) {
    u32 c;
    u32 r = __builtin_subc(a, b, bin, &c);
    bout = c;
    return r;
}

// acc(96) += a*b, with acc = hi:lo64.  One v_mad_u64_u32 (carry-out to an SGPR
// pair) plus one v_addc_co_u32 into the hi word.
UFS_DEV void mac96(u64& lo, u32& hi, u32 a, u32 b) {
    u64 cc, dead;
    asm("v_mad_u64_u32 %0, %1, %2, %3, %0" : "+v"(lo), "=s"(cc) : "v"(a), "v"(b));
    asm("v_addc_co_u32 %0, %1, %0, 0, %2" : "+v"(hi), "=s"(dead) : "s"(cc));
}

// Column multiply-accumulate blocks (generated): (hi:lo) = lo + sum a_k*b_k, hi = carries.
UFS_DEV void col1(u64& lo, u32& hi, u32 a0, u32 b0) {
    u64 sa, sb;
    asm("v_mad_u64_u32 %0, %2, %4, %5, %0\n\ts_nop 0\n\tv_addc_co_u32 %1, %2, 0, 0, %2"
        : "+v"(lo), "=&v"(hi), "=&s"(sa), "=&s"(sb)
        : "v"(a0), "v"(b0));
}
UFS_DEV void col2(u64& lo, u32& hi, u32 a0, u32 b0, u32 a1, u32 b1) {
    u64 sa, sb;
    asm("v_mad_u64_u32 %0, %2, %4, %5, %0\n\tv_mad_u64_u32 %0, %3, %6, %7, %0\n\tv_addc_co_u32 %1, %2, 0, 0, %2\n\tv_addc_co_u32 %1, %3, %1, 0, %3"
        : "+v"(lo), "=&v"(hi), "=&s"(sa), "=&s"(sb)
        : "v"(a0), "v"(b0), "v"(a1), "v"(b1));
}
UFS_DEV void col3(u64& lo, u32& hi, u32 a0, u32 b0, u32 a1, u32 b1, u32 a2, u32 b2) {
    u64 sa, sb;
    asm("v_mad_u64_u32 %0, %2, %4, %5, %0\n\tv_mad_u64_u32 %0, %3, %6, %7, %0\n\tv_addc_co_u32 %1, %2, 0, 0, %2\n\tv_mad_u64_u32 %0, %2, %8, %9, %0\n\tv_addc_co_u32 %1, %3, %1, 0, %3\n\tv_addc_co_u32 %1, %2, %1, 0, %2"
        : "+v"(lo), "=&v"(hi), "=&s"(sa), "=&s"(sb)
        : "v"(a0), "v"(b0), "v"(a1), "v"(b1), "v"(a2), "v"(b2));
}
UFS_DEV void col4(u64& lo, u32& hi, u32 a0, u32 b0, u32 a1, u32 b1, u32 a2, u32 b2, u32 a3, u32 b3) {
    u64 sa, sb;
    asm("v_mad_u64_u32 %0, %2, %4, %5, %0\n\tv_mad_u64_u32 %0, %3, %6, %7, %0\n\tv_addc_co_u32 %1, %2, 0, 0, %2\n\tv_mad_u64_u32 %0, %2, %8, %9, %0\n\tv_addc_co_u32 %1, %3, %1, 0, %3\n\tv_mad_u64_u32 %0, %3, %10, %11, %0\n\tv_addc_co_u32 %1, %2, %1, 0, %2\n\tv_addc_co_u32 %1, %3, %1, 0, %3"
        : "+v"(lo), "=&v"(hi), "=&s"(sa), "=&s"(sb)
        : "v"(a0), "v"(b0), "v"(a1), "v"(b1), "v"(a2), "v"(b2), "v"(a3), "v"(b3));
}
UFS_DEV void col5(u64& lo, u32& hi, u32 a0, u32 b0, u32 a1, u32 b1, u32 a2, u32 b2, u32 a3, u32 b3, u32 a4, u32 b4) {
    u64 sa, sb;
    asm("v_mad_u64_u32 %0, %2, %4, %5, %0\n\tv_mad_u64_u32 %0, %3, %6, %7, %0\n\tv_addc_co_u32 %1, %2, 0, 0, %2\n\tv_mad_u64_u32 %0, %2, %8, %9, %0\n\tv_addc_co_u32 %1, %3, %1, 0, %3\n\tv_mad_u64_u32 %0, %3, %10, %11, %0\n\tv_addc_co_u32 %1, %2, %1, 0, %2\n\tv_mad_u64_u32 %0, %2, %12, %13, %0\n\tv_addc_co_u32 %1, %3, %1, 0, %3\n\tv_addc_co_u32 %1, %2, %1, 0, %2"
        : "+v"(lo), "=&v"(hi), "=&s"(sa), "=&s"(sb)
        : "v"(a0), "v"(b0), "v"(a1), "v"(b1), "v"(a2), "v"(b2), "v"(a3), "v"(b3), "v"(a4), "v"(b4));
}
UFS_DEV void col6(u64& lo, u32& hi, u32 a0, u32 b0, u32 a1, u32 b1, u32 a2, u32 b2, u32 a3, u32 b3, u32 a4, u32 b4, u32 a5, u32 b5) {
    u64 sa, sb;
    asm("v_mad_u64_u32 %0, %2, %4, %5, %0\n\tv_mad_u64_u32 %0, %3, %6, %7, %0\n\tv_addc_co_u32 %1, %2, 0, 0, %2\n\tv_mad_u64_u32 %0, %2, %8, %9, %0\n\tv_addc_co_u32 %1, %3, %1, 0, %3\n\tv_mad_u64_u32 %0, %3, %10, %11, %0\n\tv_addc_co_u32 %1, %2, %1, 0, %2\n\tv_mad_u64_u32 %0, %2, %12, %13, %0\n\tv_addc_co_u32 %1, %3, %1, 0, %3\n\tv_mad_u64_u32 %0, %3, %14, %15, %0\n\tv_addc_co_u32 %1, %2, %1, 0, %2\n\tv_addc_co_u32 %1, %3, %1, 0, %3"
        : "+v"(lo), "=&v"(hi), "=&s"(sa), "=&s"(sb)
        : "v"(a0), "v"(b0), "v"(a1), "v"(b1), "v"(a2), "v"(b2), "v"(a3), "v"(b3), "v"(a4), "v"(b4), "v"(a5), "v"(b5));
}
UFS_DEV void col7(u64& lo, u32& hi, u32 a0, u32 b0, u32 a1, u32 b1, u32 a2, u32 b2, u32 a3, u32 b3, u32 a4, u32 b4, u32 a5, u32 b5, u32 a6, u32 b6) {
    u64 sa, sb;
    asm("v_mad_u64_u32 %0, %2, %4, %5, %0\n\tv_mad_u64_u32 %0, %3, %6, %7, %0\n\tv_addc_co_u32 %1, %2, 0, 0, %2\n\tv_mad_u64_u32 %0, %2, %8, %9, %0\n\tv_addc_co_u32 %1, %3, %1, 0, %3\n\tv_mad_u64_u32 %0, %3, %10, %11, %0\n\tv_addc_co_u32 %1, %2, %1, 0, %2\n\tv_mad_u64_u32 %0, %2, %12, %13, %0\n\tv_addc_co_u32 %1, %3, %1, 0, %3\n\tv_mad_u64_u32 %0, %3, %14, %15, %0\n\tv_addc_co_u32 %1, %2, %1, 0, %2\n\tv_mad_u64_u32 %0, %2, %16, %17, %0\n\tv_addc_co_u32 %1, %3, %1, 0, %3\n\tv_addc_co_u32 %1, %2, %1, 0, %2"
        : "+v"(lo), "=&v"(hi), "=&s"(sa), "=&s"(sb)
        : "v"(a0), "v"(b0), "v"(a1), "v"(b1), "v"(a2), "v"(b2), "v"(a3), "v"(b3), "v"(a4), "v"(b4), "v"(a5), "v"(b5), "v"(a6), "v"(b6));
}
UFS_DEV void col8(u64& lo, u32& hi, u32 a0, u32 b0, u32 a1, u32 b1, u32 a2, u32 b2, u32 a3, u32 b3, u32 a4, u32 b4, u32 a5, u32 b5, u32 a6, u32 b6, u32 a7, u32 b7) {
    u64 sa, sb;
    asm("v_mad_u64_u32 %0, %2, %4, %5, %0\n\tv_mad_u64_u32 %0, %3, %6, %7, %0\n\tv_addc_co_u32 %1, %2, 0, 0, %2\n\tv_mad_u64_u32 %0, %2, %8, %9, %0\n\tv_addc_co_u32 %1, %3, %1, 0, %3\n\tv_mad_u64_u32 %0, %3, %10, %11, %0\n\tv_addc_co_u32 %1, %2, %1, 0, %2\n\tv_mad_u64_u32 %0, %2, %12, %13, %0\n\tv_addc_co_u32 %1, %3, %1, 0, %3\n\tv_mad_u64_u32 %0, %3, %14, %15, %0\n\tv_addc_co_u32 %1, %2, %1, 0, %2\n\tv_mad_u64_u32 %0, %2, %16, %17, %0\n\tv_addc_co_u32 %1, %3, %1, 0, %3\n\tv_mad_u64_u32 %0, %3, %18, %19, %0\n\tv_addc_co_u32 %1, %2, %1, 0, %2\n\tv_addc_co_u32 %1, %3, %1, 0, %3"
        : "+v"(lo), "=&v"(hi), "=&s"(sa), "=&s"(sb)
        : "v"(a0), "v"(b0), "v"(a1), "v"(b1), "v"(a2), "v"(b2), "v"(a3), "v"(b3), "v"(a4), "v"(b4), "v"(a5), "v"(b5), "v"(a6), "v"(b6), "v"(a7), "v"(b7));
}


// Full 256x256 -> 512-bit product, t[0..15], by columns (product scanning).
// Each column is one asm block of v_mad_u64_u32 / v_addc_co_u32 pairs whose
// carry flags alternate between two SGPR pairs, so every carry read is one
// instruction away from its write (no hazard wait states).  A single chain per
// column measured faster than two interleaved column chains: at 3 waves per
// SIMD the 10-cycle dependent-issue latency is already hidden and the extra
// carry merging costs more than the added ILP.
UFS_DEV void mul_512(u32 t[16], const u32 a[8], const u32 b[8]) {
    u64 lo = 0;
    u32 hi;
    col1(lo, hi, a[0], b[0]);
    t[0] = (u32)lo; lo = (lo >> 32) | ((u64)hi << 32);
    col2(lo, hi, a[0], b[1], a[1], b[0]);
    t[1] = (u32)lo; lo = (lo >> 32) | ((u64)hi << 32);
    col3(lo, hi, a[0], b[2], a[1], b[1], a[2], b[0]);
    t[2] = (u32)lo; lo = (lo >> 32) | ((u64)hi << 32);
    col4(lo, hi, a[0], b[3], a[1], b[2], a[2], b[1], a[3], b[0]);
    t[3] = (u32)lo; lo = (lo >> 32) | ((u64)hi << 32);
    col5(lo, hi, a[0], b[4], a[1], b[3], a[2], b[2], a[3], b[1], a[4], b[0]);
    t[4] = (u32)lo; lo = (lo >> 32) | ((u64)hi << 32);
    col6(lo, hi, a[0], b[5], a[1], b[4], a[2], b[3], a[3], b[2], a[4], b[1], a[5], b[0]);
    t[5] = (u32)lo; lo = (lo >> 32) | ((u64)hi << 32);
    col7(lo, hi, a[0], b[6], a[1], b[5], a[2], b[4], a[3], b[3], a[4], b[2], a[5], b[1], a[6], b[0]);
    t[6] = (u32)lo; lo = (lo >> 32) | ((u64)hi << 32);
    col8(lo, hi, a[0], b[7], a[1], b[6], a[2], b[5], a[3], b[4], a[4], b[3], a[5], b[2], a[6], b[1], a[7], b[0]);
    t[7] = (u32)lo; lo = (lo >> 32) | ((u64)hi << 32);
    col7(lo, hi, a[1], b[7], a[2], b[6], a[3], b[5], a[4], b[4], a[5], b[3], a[6], b[2], a[7], b[1]);
    t[8] = (u32)lo; lo = (lo >> 32) | ((u64)hi << 32);
    col6(lo, hi, a[2], b[7], a[3], b[6], a[4], b[5], a[5], b[4], a[6], b[3], a[7], b[2]);
    t[9] = (u32)lo; lo = (lo >> 32) | ((u64)hi << 32);
    col5(lo, hi, a[3], b[7], a[4], b[6], a[5], b[5], a[6], b[4], a[7], b[3]);
    t[10] = (u32)lo; lo = (lo >> 32) | ((u64)hi << 32);
    col4(lo, hi, a[4], b[7], a[5], b[6], a[6], b[5], a[7], b[4]);
    t[11] = (u32)lo; lo = (lo >> 32) | ((u64)hi << 32);
    col3(lo, hi, a[5], b[7], a[6], b[6], a[7], b[5]);
    t[12] = (u32)lo; lo = (lo >> 32) | ((u64)hi << 32);
    col2(lo, hi, a[6], b[7], a[7], b[6]);
    t[13] = (u32)lo; lo = (lo >> 32) | ((u64)hi << 32);
    col1(lo, hi, a[7], b[7]);
    t[14] = (u32)lo;
    t[15] = (u32)(lo >> 32);   // hi is 0: the full product is < 2^512
}

// 256-bit square -> 512 bits: the 28 off-diagonal products once (column
// blocks as in mul_512), doubled with funnel shifts, plus the 8 squares.
UFS_DEV void sqr_512(u32 t[16], const u32 a[8]) {
    u32 o[16];
    u64 lo = 0;
    u32 hi;
    o[0] = 0;
    col1(lo, hi, a[0], a[1]);
    o[1] = (u32)lo; lo = (lo >> 32) | ((u64)hi << 32);
    col1(lo, hi, a[0], a[2]);
    o[2] = (u32)lo; lo = (lo >> 32) | ((u64)hi << 32);
    col2(lo, hi, a[0], a[3], a[1], a[2]);
    o[3] = (u32)lo; lo = (lo >> 32) | ((u64)hi << 32);
    col2(lo, hi, a[0], a[4], a[1], a[3]);
    o[4] = (u32)lo; lo = (lo >> 32) | ((u64)hi << 32);
    col3(lo, hi, a[0], a[5], a[1], a[4], a[2], a[3]);
    o[5] = (u32)lo; lo = (lo >> 32) | ((u64)hi << 32);
    col3(lo, hi, a[0], a[6], a[1], a[5], a[2], a[4]);
    o[6] = (u32)lo; lo = (lo >> 32) | ((u64)hi << 32);
    col4(lo, hi, a[0], a[7], a[1], a[6], a[2], a[5], a[3], a[4]);
    o[7] = (u32)lo; lo = (lo >> 32) | ((u64)hi << 32);
    col3(lo, hi, a[1], a[7], a[2], a[6], a[3], a[5]);
    o[8] = (u32)lo; lo = (lo >> 32) | ((u64)hi << 32);
    col3(lo, hi, a[2], a[7], a[3], a[6], a[4], a[5]);
    o[9] = (u32)lo; lo = (lo >> 32) | ((u64)hi << 32);
    col2(lo, hi, a[3], a[7], a[4], a[6]);
    o[10] = (u32)lo; lo = (lo >> 32) | ((u64)hi << 32);
    col2(lo, hi, a[4], a[7], a[5], a[6]);
    o[11] = (u32)lo; lo = (lo >> 32) | ((u64)hi << 32);
    col1(lo, hi, a[5], a[7]);
    o[12] = (u32)lo; lo = (lo >> 32) | ((u64)hi << 32);
    col1(lo, hi, a[6], a[7]);
    o[13] = (u32)lo; lo = (lo >> 32) | ((u64)hi << 32);
    o[14] = (u32)lo;
    o[15] = (u32)(lo >> 32);
    // t = 2*o + sum a_i^2 * 2^(64 i)
    u32 d[16];
#pragma unroll
    for (int i = 0; i < 8; ++i) {
        const u64 s = (u64)a[i] * a[i];
        d[2 * i] = (u32)s;
        d[2 * i + 1] = (u32)(s >> 32);
    }
    u32 c = 0;
#pragma unroll
    for (int k = 0; k < 16; ++k) {
        const u32 dbl = (k == 0) ? (o[0] << 1) : __builtin_amdgcn_alignbit(o[k], o[k - 1], 31);
        t[k] = addc(dbl, d[k], c, c);
    }
}

// ---------------------------------------------------------------------------
// Field arithmetic mod p
// ---------------------------------------------------------------------------
UFS_DEV void fe_set_u32(fe& r, u32 x) {
    r.v[0] = x;
#pragma unroll
    for (int i = 1; i < 8; ++i) r.v[i] = 0;
}

UFS_DEV void fe_copy(fe& r, const fe& a) {
#pragma unroll
    for (int i = 0; i < 8; ++i) r.v[i] = a.v[i];
}

// r += c * (2^32 + 977) mod 2^256 for a small c (c < 2^35), wrapping once more
// if needed.  Keeps r in [0, 2^256) and congruent mod p.
UFS_DEV void fe_fold(fe& r, u64 c) {
    u64 t = (u64)r.v[0] + c * 977u;
    r.v[0] = (u32)t;
    t = (t >> 32) + (u64)r.v[1] + c;
    r.v[1] = (u32)t;
    t >>= 32;
    u32 cy = (u32)t;
#pragma unroll
    for (int i = 2; i < 8; ++i) r.v[i] = addc(r.v[i], 0, cy, cy);
    // A second wrap leaves r < c*K < 2^68, so one more fold touches limbs 0..2.
    u32 c2 = cy;
    t = (u64)r.v[0] + (u64)c2 * 977u;
    r.v[0] = (u32)t;
    t = (t >> 32) + (u64)r.v[1] + c2;
    r.v[1] = (u32)t;
    r.v[2] += (u32)(t >> 32);
}

// 512-bit t -> weakly reduced field element.
UFS_DEV void fe_reduce512(fe& r, const u32 t[16]) {
    // r = t_lo + 977*t_hi + (t_hi << 32)
    u64 c = 0;
#pragma unroll
    for (int i = 0; i < 8; ++i) {
        u64 acc = (u64)t[8 + i] * 977u + (u64)t[i] + c;
        if (i > 0) acc += t[8 + i - 1];
        r.v[i] = (u32)acc;
        c = acc >> 32;
    }
    c += t[15];          // (t_hi[7] << 32) lands at 2^256
    fe_fold(r, c);
}

// Inline bodies, used inside the out-of-line point operations.
UFS_DEV void fe_mul_i(fe& r, const fe& a, const fe& b) {
    u32 t[16];
    mul_512(t, a.v, b.v);
    fe_reduce512(r, t);
}

UFS_DEV void fe_sqr_i(fe& r, const fe& a) {
    u32 t[16];
    sqr_512(t, a.v);
    fe_reduce512(r, t);
}

// Out-of-line multiplication and squaring for everything else (arguments
// and results travel in VGPRs): one copy of each body keeps the code small
// enough for the instruction cache.
__device__ __noinline__ fe fe_mul_v(fe a, fe b) {
    u32 t[16];
    fe r;
    mul_512(t, a.v, b.v);
    fe_reduce512(r, t);
    return r;
}

__device__ __noinline__ fe fe_sqr_v(fe a) {
    u32 t[16];
    fe r;
    sqr_512(t, a.v);
    fe_reduce512(r, t);
    return r;
}

UFS_DEV void fe_mul(fe& r, const fe& a, const fe& b) { r = fe_mul_v(a, b); }
UFS_DEV void fe_sqr(fe& r, const fe& a) { r = fe_sqr_v(a); }

UFS_DEV void fe_add(fe& r, const fe& a, const fe& b) {
    u32 c = 0;
#pragma unroll
    for (int i = 0; i < 8; ++i) r.v[i] = addc(a.v[i], b.v[i], c, c);
    fe_fold(r, c);
}

// r -= c*(2^32+977) mod 2^256 for c in {0,1}; re-borrows once if needed.
UFS_DEV void fe_unfold1(fe& r, u32 c) {
    u32 b = 0;
    r.v[0] = subc(r.v[0], c * 977u, 0, b);
    r.v[1] = subc(r.v[1], c, b, b);
#pragma unroll
    for (int i = 2; i < 8; ++i) r.v[i] = subc(r.v[i], 0, b, b);
    // A second borrow leaves r >= 2^256 - K; subtracting K again borrows at
    // most into limb 2.
    u32 c2 = b;
    r.v[0] = subc(r.v[0], c2 * 977u, 0, b);
    r.v[1] = subc(r.v[1], c2, b, b);
    r.v[2] = subc(r.v[2], 0, b, b);
}

UFS_DEV void fe_sub(fe& r, const fe& a, const fe& b) {
    u32 bw = 0;
#pragma unroll
    for (int i = 0; i < 8; ++i) r.v[i] = subc(a.v[i], b.v[i], bw, bw);
    fe_unfold1(r, bw);
}

UFS_DEV void fe_neg(fe& r, const fe& a) {
    fe z;
    fe_set_u32(z, 0);
    fe_sub(r, z, a);
}

// r = a * k for small k (k < 2^16)
UFS_DEV void fe_mul_small(fe& r, const fe& a, u32 k) {
    u64 c = 0;
#pragma unroll
    for (int i = 0; i < 8; ++i) {
        u64 acc = (u64)a.v[i] * k + c;
        r.v[i] = (u32)acc;
        c = acc >> 32;
    }
    fe_fold(r, c);
}

// Canonical form in [0, p).
UFS_DEV void fe_normalize(fe& a) {
    // t = a + (2^256 - p); carry-out means a >= p and a - p = t mod 2^256.
    fe t;
    u32 c = 0;
    t.v[0] = addc(a.v[0], 977u, 0, c);
    t.v[1] = addc(a.v[1], 1u, c, c);
#pragma unroll
    for (int i = 2; i < 8; ++i) t.v[i] = addc(a.v[i], 0, c, c);
#pragma unroll
    for (int i = 0; i < 8; ++i) a.v[i] = c ? t.v[i] : a.v[i];
}

UFS_DEV bool fe_is_zero(const fe& a) {
    // a in [0, 2^256): a == 0 mod p iff a == 0 or a == p
    u32 z = 0, q = 0;
#pragma unroll
    for (int i = 0; i < 8; ++i) z |= a.v[i];
    q = (a.v[0] ^ 0xFFFFFC2FU) | (a.v[1] ^ 0xFFFFFFFEU);
#pragma unroll
    for (int i = 2; i < 8; ++i) q |= ~a.v[i];
    return (z == 0) | (q == 0);
}

UFS_DEV bool fe_equal(const fe& a, const fe& b) {
    fe d;
    fe_sub(d, a, b);
    return fe_is_zero(d);
}

UFS_DEV bool fe_is_odd(const fe& a) {
    fe t = a;
    fe_normalize(t);
    return (t.v[0] & 1u) != 0;
}

// Raw 256-bit strict check: a < p.
UFS_DEV bool u256_lt_p(const u32 a[8]) {
    // a < p  <=>  a + (2^256 - p) does not carry
    u32 c = 0;
    (void)addc(a[0], 977u, 0, c);
    (void)addc(a[1], 1u, c, c);
#pragma unroll
    for (int i = 2; i < 8; ++i) (void)addc(a[i], 0, c, c);
    return c == 0;
}

UFS_DEV void fe_sqr_n(fe& r, const fe& a, int n) {
    fe t = a;
    for (int i = 0; i < n; ++i) fe_sqr(t, t);
    r = t;
}

// Shared prefix of the p-2 and (p+1)/4 exponent chains: x223 = a^(2^223-1),
// plus x2 = a^3 and x22 = a^(2^22-1).
UFS_DEV void fe_pow_x223(fe& x223, fe& x22, fe& x2, const fe& a) {
    fe x3, x6, x9, x11, x44, x88, x176, x220, t;
    fe_sqr(t, a);        fe_mul(x2, t, a);
    fe_sqr(t, x2);       fe_mul(x3, t, a);
    fe_sqr_n(t, x3, 3);  fe_mul(x6, t, x3);
    fe_sqr_n(t, x6, 3);  fe_mul(x9, t, x3);
    fe_sqr_n(t, x9, 2);  fe_mul(x11, t, x2);
    fe_sqr_n(t, x11, 11); fe_mul(x22, t, x11);
    fe_sqr_n(t, x22, 22); fe_mul(x44, t, x22);
    fe_sqr_n(t, x44, 44); fe_mul(x88, t, x44);
    fe_sqr_n(t, x88, 88); fe_mul(x176, t, x88);
    fe_sqr_n(t, x176, 44); fe_mul(x220, t, x44);
    fe_sqr_n(t, x220, 3); fe_mul(x223, t, x3);
}

// a^(p-2): 255 squarings + 15 multiplications.  The kernels invert with
// divsteps (fe_inv); this form is kept as an independent cross-check.
UFS_DEV void fe_inv_fermat(fe& r, const fe& a) {
    fe x223, x22, x2, t;
    fe_pow_x223(x223, x22, x2, a);
    fe_sqr_n(t, x223, 23); fe_mul(t, t, x22);
    fe_sqr_n(t, t, 5);     fe_mul(t, t, a);
    fe_sqr_n(t, t, 3);     fe_mul(t, t, x2);
    fe_sqr_n(t, t, 2);     fe_mul(r, t, a);
}

// Big-endian 32 bytes -> raw limbs (no reduction).
UFS_DEV void u256_from_be(u32 r[8], const u8* b) {
#pragma unroll
    for (int i = 0; i < 8; ++i) {
        const u8* q = b + 28 - 4 * i;
        r[i] = ((u32)q[0] << 24) | ((u32)q[1] << 16) | ((u32)q[2] << 8) | (u32)q[3];
    }
}

// Raw limbs from eight big-endian 32-bit words (word 0 is most significant).
UFS_DEV void u256_from_be_words(u32 r[8], const u32 w[8]) {
#pragma unroll
    for (int i = 0; i < 8; ++i) r[i] = w[7 - i];
}

UFS_DEV void u256_to_be(u8* b, const u32 a[8]) {
#pragma unroll
    for (int i = 0; i < 8; ++i) {
        u8* q = b + 28 - 4 * i;
        q[0] = (u8)(a[i] >> 24);
        q[1] = (u8)(a[i] >> 16);
        q[2] = (u8)(a[i] >> 8);
        q[3] = (u8)a[i];
    }
}

// ---------------------------------------------------------------------------
// Scalar arithmetic mod n
// ---------------------------------------------------------------------------
UFS_DEV bool u256_is_zero(const u32 a[8]) {
    u32 z = 0;
#pragma unroll
    for (int i = 0; i < 8; ++i) z |= a[i];
    return z == 0;
}

// a < b (unsigned 256-bit)
UFS_DEV bool u256_lt(const u32 a[8], const u32 b[8]) {
    u32 bw = 0;
#pragma unroll
    for (int i = 0; i < 8; ++i) (void)subc(a[i], b[i], bw, bw);
    return bw != 0;
}

UFS_DEV void sc_set_u32(sc& r, u32 x) {
    r.v[0] = x;
#pragma unroll
    for (int i = 1; i < 8; ++i) r.v[i] = 0;
}

UFS_DEV bool u256_lt_n(const u32 a[8]) {
    const u32 N[8] = {UFS_N_LIMBS};
    return u256_lt(a, N);
}

// a > n/2 (a assumed < n)
UFS_DEV bool sc_is_high(const sc& a) {
    const u32 H[8] = {UFS_NH_LIMBS};
    return u256_lt(H, a.v);
}

// r = a - n if a >= n (one conditional subtraction; a < 2^256).
UFS_DEV void sc_reduce_once(sc& a) {
    const u32 N[8] = {UFS_N_LIMBS};
    u32 t[8], bw = 0;
#pragma unroll
    for (int i = 0; i < 8; ++i) t[i] = subc(a.v[i], N[i], bw, bw);
#pragma unroll
    for (int i = 0; i < 8; ++i) a.v[i] = bw ? a.v[i] : t[i];
}

UFS_DEV void sc_add(sc& r, const sc& a, const sc& b) {
    const u32 N[8] = {UFS_N_LIMBS};
    u32 c = 0, t[8], bw = 0;
#pragma unroll
    for (int i = 0; i < 8; ++i) r.v[i] = addc(a.v[i], b.v[i], c, c);
#pragma unroll
    for (int i = 0; i < 8; ++i) t[i] = subc(r.v[i], N[i], bw, bw);
    // subtract n when the sum carried out or is >= n
    const bool take = c | !bw;
#pragma unroll
    for (int i = 0; i < 8; ++i) r.v[i] = take ? t[i] : r.v[i];
}

// r = n - a (a in [0, n)); 0 maps to 0.
UFS_DEV void sc_neg(sc& r, const sc& a) {
    const u32 N[8] = {UFS_N_LIMBS};
    u32 bw = 0;
    const bool z = u256_is_zero(a.v);
#pragma unroll
    for (int i = 0; i < 8; ++i) r.v[i] = subc(N[i], a.v[i], bw, bw);
#pragma unroll
    for (int i = 0; i < 8; ++i) r.v[i] = z ? 0u : r.v[i];
}

// 512-bit t -> t mod n (fully reduced).
UFS_DEV void sc_reduce512(sc& r, const u32 t[16]) {
    const u32 C[5] = {UFS_NC0, UFS_NC1, UFS_NC2, UFS_NC3, 1u};
    // u = t_lo + t_hi * C   (13 limbs)
    u32 u[13];
    {
        u64 lo = 0;
        u32 hi = 0;
#pragma unroll
        for (int k = 0; k < 13; ++k) {
            if (k < 8) {
                u64 s = lo + t[k];
                hi += (s < lo);
                lo = s;
            }
#pragma unroll
            for (int i = 0; i < 8; ++i) {
                const int j = k - i;
                if (j < 0 || j > 4) continue;
                if (j == 4) {
                    u64 s = lo + t[8 + i];
                    hi += (s < lo);
                    lo = s;
                } else {
                    mac96(lo, hi, t[8 + i], C[j]);
                }
            }
            u[k] = (u32)lo;
            lo = (lo >> 32) | ((u64)hi << 32);
            hi = 0;
        }
        // lo is zero here: t_hi*C < 2^385, t_lo < 2^256, sum < 2^386 < 2^416.
    }
    // v = u_lo + u_hi * C, u_hi = u[8..12] (< 2^130)   (9 limbs)
    u32 v[9];
    {
        u64 lo = 0;
        u32 hi = 0;
#pragma unroll
        for (int k = 0; k < 9; ++k) {
            if (k < 8) {
                u64 s = lo + u[k];
                hi += (s < lo);
                lo = s;
            }
#pragma unroll
            for (int i = 0; i < 5; ++i) {
                const int j = k - i;
                if (j < 0 || j > 4) continue;
                if (j == 4) {
                    u64 s = lo + u[8 + i];
                    hi += (s < lo);
                    lo = s;
                } else {
                    mac96(lo, hi, u[8 + i], C[j]);
                }
            }
            v[k] = (u32)lo;
            lo = (lo >> 32) | ((u64)hi << 32);
            hi = 0;
        }
    }
    // w = v_lo + v[8] * C, v[8] < 2^5
    {
        u64 acc = (u64)v[0] + (u64)v[8] * C[0];
        r.v[0] = (u32)acc; acc >>= 32;
        acc += (u64)v[1] + (u64)v[8] * C[1];
        r.v[1] = (u32)acc; acc >>= 32;
        acc += (u64)v[2] + (u64)v[8] * C[2];
        r.v[2] = (u32)acc; acc >>= 32;
        acc += (u64)v[3] + (u64)v[8] * C[3];
        r.v[3] = (u32)acc; acc >>= 32;
        acc += (u64)v[4] + (u64)v[8];
        r.v[4] = (u32)acc; acc >>= 32;
        u32 c = (u32)acc;
        r.v[5] = addc(v[5], 0, c, c);
        r.v[6] = addc(v[6], 0, c, c);
        r.v[7] = addc(v[7], 0, c, c);
        // c == 1: value wrapped past 2^256; add C once more (result is small)
        u32 k = c, cc = 0;
        r.v[0] = addc(r.v[0], k ? C[0] : 0u, 0, cc);
        r.v[1] = addc(r.v[1], k ? C[1] : 0u, cc, cc);
        r.v[2] = addc(r.v[2], k ? C[2] : 0u, cc, cc);
        r.v[3] = addc(r.v[3], k ? C[3] : 0u, cc, cc);
        r.v[4] = addc(r.v[4], k, cc, cc);
        r.v[5] = addc(r.v[5], 0, cc, cc);
        r.v[6] = addc(r.v[6], 0, cc, cc);
        r.v[7] = addc(r.v[7], 0, cc, cc);
    }
    sc_reduce_once(r);
}

__device__ __noinline__ sc sc_mul_v(sc a, sc b) {
    u32 t[16];
    sc r;
    mul_512(t, a.v, b.v);
    sc_reduce512(r, t);
    return r;
}

__device__ __noinline__ sc sc_sqr_v(sc a) {
    u32 t[16];
    sc r;
    sqr_512(t, a.v);
    sc_reduce512(r, t);
    return r;
}

UFS_DEV void sc_mul(sc& r, const sc& a, const sc& b) { r = sc_mul_v(a, b); }
UFS_DEV void sc_sqr(sc& r, const sc& a) { r = sc_sqr_v(a); }

// Sliding-window schedule for the low 129 bits of n-2 after the leading run
// of 127 ones: (squarings, odd power) pairs.
__constant__ const u8 kScInvSteps[26][2] = {
    {5, 11}, {3, 5}, {4, 5}, {4, 7}, {5, 13}, {2, 3}, {5, 7}, {6, 13}, {5, 11}, {4, 13},
    {3, 1}, {6, 5}, {10, 7}, {4, 7}, {5, 15}, {4, 15}, {5, 9}, {6, 11}, {4, 13}, {5, 3},
    {6, 13}, {10, 13}, {4, 9}, {9, 9}, {4, 15}, {1, 1}};

UFS_DEV void sc_sqr_n(sc& r, const sc& a, int n) {
    sc t = a;
    for (int i = 0; i < n; ++i) sc_sqr(t, t);
    r = t;
}

// r = a^(n-2) mod n (Fermat inverse).  The kernels invert with divsteps
// (sc_inv); this form is kept as an independent cross-check.
UFS_DEV void sc_inv_fermat(sc& r, const sc& a) {
    sc pw[8];   // a^1, a^3, ..., a^15
    sc a2, t;
    sc_sqr(a2, a);
    pw[0] = a;
#pragma unroll
    for (int i = 1; i < 8; ++i) sc_mul(pw[i], pw[i - 1], a2);
    // x127 = a^(2^127 - 1)
    // pw[3] = a^7 = a^(2^3 - 1)
    sc x3 = pw[3], x6, x7, x14, x28, x56, x112, x126;
    sc_sqr_n(t, x3, 3);   sc_mul(x6, t, x3);
    sc_sqr(t, x6);        sc_mul(x7, t, a);
    sc_sqr_n(t, x7, 7);   sc_mul(x14, t, x7);
    sc_sqr_n(t, x14, 14); sc_mul(x28, t, x14);
    sc_sqr_n(t, x28, 28); sc_mul(x56, t, x28);
    sc_sqr_n(t, x56, 56); sc_mul(x112, t, x56);
    sc_sqr_n(t, x112, 14); sc_mul(x126, t, x14);
    sc_sqr(t, x126);      sc_mul(t, t, a);     // x127
    for (int s = 0; s < 26; ++s) {
        const int nsq = kScInvSteps[s][0];
        const int idx = kScInvSteps[s][1] >> 1;
        sc_sqr_n(t, t, nsq);
        sc m = pw[0];
#pragma unroll
        for (int j = 1; j < 8; ++j)
            if (idx == j) m = pw[j];
        sc_mul(t, t, m);
    }
    r = t;
}

}  // namespace ufs
