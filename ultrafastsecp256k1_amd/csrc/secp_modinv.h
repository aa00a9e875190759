// Modular inversion mod p and mod n by Bernstein-Yang divsteps
// (D. J. Bernstein, B.-Y. Yang, "Fast constant-time gcd computation and
// modular inversion", TCHES 2019), half-delta variant.
//
// State: f = M, g = x, d = 0, e = 1 with f = d*x and g = e*x (mod M).  A
// batch of 30 divsteps is decided by the low 32 bits of f and g alone and
// yields a 2x2 matrix T (entries |.| <= 2^30) with
//     (f, g) <- T (f, g) / 2^30        (exact division)
//     (d, e) <- T (d, e) / 2^30 mod M  (made exact by adding multiples of M).
// After 590 divsteps g = 0 and f = +-1 for every 256-bit input (the proven
// bound for this variant), so x^-1 = f * d mod M.  The kernels run the
// batches until every lane of the wave has reached g = 0 (at most 20
// batches = 600 divsteps); lanes that finish early are unaffected by further
// batches, since once g = 0 every step is the identity on f and d.
//
// Big integers are held in signed radix 2^30: nine int32 limbs, limbs 0..7
// in [0, 2^30), limb 8 signed.  |f|, |g| <= M throughout; |d|, |e| grow by at
// most M/2 per batch (|u| + |v| <= 2^30 and the correction |md| <= 2^29), so
// they stay below 11M < 2^260 and every limb product fits an int64 sum.
//
// Cost per inversion: about 20 x (30 x 24 + 170) VALU instructions, against
// 255 squarings + 15 multiplications for a^(p-2) and about 300 mod-n
// multiplications for a^(n-2).
#pragma once
#include "secp_field.h"

namespace ufs {

typedef int32_t i32;
typedef int64_t i64;

struct s30 { i32 v[9]; };

constexpr i32 kM30 = 0x3FFFFFFF;

// Modulus limbs (radix 2^30) and M^-1 mod 2^30.
template <int MOD> struct ModConst;
template <> struct ModConst<0> {   // p = 2^256 - 2^32 - 977
    static constexpr i32 m[9] = {0x3FFFFC2F, 0x3FFFFFFB, 0x3FFFFFFF, 0x3FFFFFFF, 0x3FFFFFFF,
                                 0x3FFFFFFF, 0x3FFFFFFF, 0x3FFFFFFF, 0x0000FFFF};
    static constexpr u32 minv = 0x2DDACACFu;
};
template <> struct ModConst<1> {   // n, the group order
    static constexpr i32 m[9] = {0x10364141, 0x3F497A33, 0x348A03BB, 0x2BB739AB, 0x3FFFFEBA,
                                 0x3FFFFFFF, 0x3FFFFFFF, 0x3FFFFFFF, 0x0000FFFF};
    static constexpr u32 minv = 0x2A774EC1u;
};

// 8 x 32-bit limbs (value < 2^256) -> radix 2^30.
UFS_DEV void to_s30(s30& r, const u32 a[8]) {
#pragma unroll
    for (int i = 0; i < 9; ++i) {
        const int k = (30 * i) / 32, s = (30 * i) % 32;
        const u64 w = (u64)a[k] | (k + 1 < 8 ? (u64)a[k + 1] << 32 : 0ull);
        r.v[i] = (i32)((u32)(w >> s) & (u32)kM30);
    }
}

// 30 half-delta divsteps on the low words of f and g; d2 = 2*delta (odd).
// Output: 2^30 f' = u f + v g, 2^30 g' = q f + r g.
UFS_DEV void divsteps30(i32& d2, u32 f, u32 g, i32& u_o, i32& v_o, i32& q_o, i32& r_o) {
    i32 u = 1, v = 0, q = 0, r = 1;
    for (int k = 0; k < 30; ++k) {
        const bool odd = (g & 1u) != 0;
        const bool swap = odd && (d2 > 0);
        // swap case: (f, g, u, v, q, r, delta) -> (g, -f, q, r, -u, -v, -delta)
        const u32 f1 = swap ? g : f;
        const u32 g1 = swap ? 0u - f : g;
        const i32 u1 = swap ? q : u;
        const i32 v1 = swap ? r : v;
        const i32 q1 = swap ? -u : q;
        const i32 r1 = swap ? -v : r;
        d2 = swap ? -d2 : d2;
        // g odd: g += f (both cases now read g' = (g + f) / 2)
        const u32 g2 = odd ? g1 + f1 : g1;
        q = odd ? q1 + u1 : q1;
        r = odd ? r1 + v1 : r1;
        f = f1;
        g = g2 >> 1;
        u = (i32)((u32)u1 << 1);
        v = (i32)((u32)v1 << 1);
        d2 += 2;
    }
    u_o = u; v_o = v; q_o = q; r_o = r;
}

UFS_DEV void update_fg(s30& f, s30& g, i32 u, i32 v, i32 q, i32 r) {
    i64 cf = (i64)u * f.v[0] + (i64)v * g.v[0];
    i64 cg = (i64)q * f.v[0] + (i64)r * g.v[0];
    cf >>= 30;   // the low 30 bits are zero by construction
    cg >>= 30;
#pragma unroll
    for (int i = 1; i < 9; ++i) {
        cf += (i64)u * f.v[i] + (i64)v * g.v[i];
        cg += (i64)q * f.v[i] + (i64)r * g.v[i];
        f.v[i - 1] = (i32)cf & kM30;
        g.v[i - 1] = (i32)cg & kM30;
        cf >>= 30;
        cg >>= 30;
    }
    f.v[8] = (i32)cf;
    g.v[8] = (i32)cg;
}

template <int MOD>
UFS_DEV void update_de(s30& d, s30& e, i32 u, i32 v, i32 q, i32 r) {
    // md, me in [-2^29, 2^29) with u d + v e + md M = 0 (mod 2^30), same for e.
    const u32 td = (u32)u * (u32)d.v[0] + (u32)v * (u32)e.v[0];
    const u32 te = (u32)q * (u32)d.v[0] + (u32)r * (u32)e.v[0];
    const i32 md = (i32)((0u - td * ModConst<MOD>::minv) << 2) >> 2;
    const i32 me = (i32)((0u - te * ModConst<MOD>::minv) << 2) >> 2;
    i64 cd = (i64)u * d.v[0] + (i64)v * e.v[0] + (i64)md * ModConst<MOD>::m[0];
    i64 ce = (i64)q * d.v[0] + (i64)r * e.v[0] + (i64)me * ModConst<MOD>::m[0];
    cd >>= 30;
    ce >>= 30;
#pragma unroll
    for (int i = 1; i < 9; ++i) {
        cd += (i64)u * d.v[i] + (i64)v * e.v[i] + (i64)md * ModConst<MOD>::m[i];
        ce += (i64)q * d.v[i] + (i64)r * e.v[i] + (i64)me * ModConst<MOD>::m[i];
        d.v[i - 1] = (i32)cd & kM30;
        e.v[i - 1] = (i32)ce & kM30;
        cd >>= 30;
        ce >>= 30;
    }
    d.v[8] = (i32)cd;
    e.v[8] = (i32)ce;
}

// 9-word little-endian integer (< 2^288).
struct u288 { u32 w[9]; };

// sign * d (|d| < 11 M) -> integer in [0, 27 M) congruent to sign * d mod M.
template <int MOD>
UFS_DEV void s30_to_u288(u288& t, const s30& d, bool negate) {
    i64 c = 0;
    u32 o[9];
#pragma unroll
    for (int i = 0; i < 9; ++i) {
        const i32 di = negate ? -d.v[i] : d.v[i];
        c += (i64)di + 16 * (i64)ModConst<MOD>::m[i];
        if (i < 8) {
            o[i] = (u32)c & (u32)kM30;
            c >>= 30;
        } else {
            o[i] = (u32)c;   // total is positive and < 2^261
        }
    }
#pragma unroll
    for (int j = 0; j < 9; ++j) {
        const int i = (32 * j) / 30, s = (32 * j) % 30;   // s <= 16
        const u64 w = (u64)o[i] | (i + 1 < 9 ? (u64)o[i + 1] << 30 : 0ull);
        t.w[j] = (u32)(w >> s);
    }
}

// x^-1 mod M for x in [0, M) (0 maps to 0), as an integer below 27 M that is
// congruent to the inverse (the caller reduces it).  Out of line with
// by-value arguments, so the state stays in VGPRs.
template <int MOD>
__device__ __noinline__ u288 modinv_core(fe x) {
    s30 f, g, d, e;
#pragma unroll
    for (int i = 0; i < 9; ++i) {
        f.v[i] = ModConst<MOD>::m[i];
        d.v[i] = 0;
        e.v[i] = 0;
    }
    e.v[0] = 1;
    to_s30(g, x.v);
    i32 d2 = 1;
    for (int b = 0; b < 20; ++b) {
        i32 u, v, q, r;
        divsteps30(d2, (u32)f.v[0] | ((u32)f.v[1] << 30), (u32)g.v[0] | ((u32)g.v[1] << 30), u, v, q, r);
        update_fg(f, g, u, v, q, r);
        update_de<MOD>(d, e, u, v, q, r);
        i32 nz = 0;
#pragma unroll
        for (int i = 0; i < 9; ++i) nz |= g.v[i];
        if (!__any(nz != 0)) break;   // every lane of the wave is done
    }
    // f = +-1 now; limb 8 is -1 exactly when f = -1.
    u288 t;
    s30_to_u288<MOD>(t, d, f.v[8] < 0);
    return t;
}

UFS_DEV void u288_to_wide(u32 t[16], const u288& a) {
#pragma unroll
    for (int i = 0; i < 9; ++i) t[i] = a.w[i];
#pragma unroll
    for (int i = 9; i < 16; ++i) t[i] = 0;
}

// Field inverse (a weakly reduced; 0 -> 0), weakly reduced result.
UFS_DEV void fe_inv(fe& r, const fe& a) {
    fe x = a;
    fe_normalize(x);
    u32 t[16];
    u288_to_wide(t, modinv_core<0>(x));
    fe_reduce512(r, t);
}

// Scalar inverse (a < n; 0 -> 0), fully reduced result.
UFS_DEV void sc_inv(sc& r, const sc& a) {
    fe x;
#pragma unroll
    for (int i = 0; i < 8; ++i) x.v[i] = a.v[i];
    u32 t[16];
    u288_to_wide(t, modinv_core<1>(x));
    sc_reduce512(r, t);
}

}  // namespace ufs
