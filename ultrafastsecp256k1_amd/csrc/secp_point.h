// secp256k1 group law on gfx950: Jacobian doubling, mixed (Jacobian+affine)
// addition and general Jacobian addition, all exception-complete so that the
// verdicts stay exact for adversarial public inputs (P == +-Q, infinity).
//
// The doubling and the mixed addition of the scalar-multiplication loops are
// out-of-line functions whose field arithmetic is inlined: the point travels
// in a 28-word vector (x, y, z, inf) that the calling convention keeps in
// v0-v27, so a chain of calls passes the accumulator through without moves,
// and register allocation inside each body is unconstrained by call
// boundaries.  (Aggregates above 16 words would be passed through scratch
// memory, hence the vector type.)  The addend of the mixed addition is read
// from device memory inside the call.
//
// Formula choices (a = 0 curve):
//   dbl   : dbl-2009-l             2M + 5S
//   madd  : madd-2007-bl           7M + 4S
//   add   : add-2007-bl            11M + 5S
// Reference counterparts: src/cuda/include/secp256k1.cuh:1551-1782 (dbl/madd),
// :2410-2491 (jacobian_add).
#pragma once
#include "secp_field29.h"
#include "secp_modinv.h"

namespace ufs {

struct aff { fe x, y; };
struct jac { fe x, y, z; u32 inf; };

typedef u32 jvec __attribute__((ext_vector_type(28)));

UFS_DEV jvec jac_pack(const jac& a) {
    jvec v;
#pragma unroll
    for (int i = 0; i < 8; ++i) {
        v[i] = a.x.v[i];
        v[8 + i] = a.y.v[i];
        v[16 + i] = a.z.v[i];
    }
    v[24] = a.inf;
    v[25] = v[26] = v[27] = 0;
    return v;
}

UFS_DEV void jac_unpack(jac& r, const jvec& v) {
#pragma unroll
    for (int i = 0; i < 8; ++i) {
        r.x.v[i] = v[i];
        r.y.v[i] = v[8 + i];
        r.z.v[i] = v[16 + i];
    }
    r.inf = v[24];
}

template <bool INL> UFS_DEV void fmul(fe& r, const fe& a, const fe& b) {
    if constexpr (INL) fe_mul_i(r, a, b); else fe_mul(r, a, b);
}
template <bool INL> UFS_DEV void fsqr(fe& r, const fe& a) {
    if constexpr (INL) fe_sqr_i(r, a); else fe_sqr(r, a);
}

UFS_DEV void jac_set_inf(jac& r) {
    fe_set_u32(r.x, 1);
    fe_set_u32(r.y, 1);
    fe_set_u32(r.z, 0);
    r.inf = 1;
}

UFS_DEV void jac_from_aff(jac& r, const aff& a) {
    r.x = a.x;
    r.y = a.y;
    fe_set_u32(r.z, 1);
    r.inf = 0;
}

UFS_DEV void jac_select(jac& r, bool c, const jac& a) {
#pragma unroll
    for (int i = 0; i < 8; ++i) {
        r.x.v[i] = c ? a.x.v[i] : r.x.v[i];
        r.y.v[i] = c ? a.y.v[i] : r.y.v[i];
        r.z.v[i] = c ? a.z.v[i] : r.z.v[i];
    }
    r.inf = c ? a.inf : r.inf;
}

// r = 2a.  secp256k1 has no points of order 2, so 2a = inf iff a = inf.
template <bool INL>
UFS_DEV void jac_dbl_body(jac& r, const jac& a) {
    fe A, B, C, D, E, F, t;
    fsqr<INL>(A, a.x);         // A = X^2
    fsqr<INL>(B, a.y);         // B = Y^2
    fsqr<INL>(C, B);           // C = B^2
    fe_add(t, a.x, B);
    fsqr<INL>(t, t);
    fe_sub(t, t, A);
    fe_sub(t, t, C);
    fe_add(D, t, t);           // D = 2((X+B)^2 - A - C)
    fe_add(E, A, A);
    fe_add(E, E, A);           // E = 3A
    fsqr<INL>(F, E);           // F = E^2
    fe Z3;
    fmul<INL>(Z3, a.y, a.z);
    fe_add(r.z, Z3, Z3);       // Z3 = 2YZ
    fe_add(t, D, D);
    fe_sub(r.x, F, t);         // X3 = F - 2D
    fe_sub(t, D, r.x);
    fmul<INL>(t, E, t);
    fe_mul_small(C, C, 8);
    fe_sub(r.y, t, C);         // Y3 = E(D - X3) - 8C
    r.inf = a.inf;
}

__device__ __noinline__ jvec jac_dbl_v(jvec av) {
    jac a, r;
    jac_unpack(a, av);
    jac_dbl_body<true>(r, a);
    return jac_pack(r);
}

UFS_DEV void jac_dbl(jac& r, const jac& a) { jac_unpack(r, jac_dbl_v(jac_pack(a))); }

// Degenerate mixed addition (a finite, H == 0): a == b gives 2a, a == -b
// gives infinity.  Rare; out of line.
__device__ __noinline__ jvec jac_madd_degenerate(jvec av, u32 same) {
    if (same) return jac_dbl_v(av);
    jac r;
    jac_set_inf(r);
    return jac_pack(r);
}

// r = a + b with b affine (never infinity).  Handles a == inf, a == b and
// a == -b.  The degenerate branches are rare and taken per lane.
UFS_DEV void jac_madd(jac& r, const jac& a, const aff& b) {
    fe z1z1, u2, s2, h, hh, i4, j, rr, v, t;
    fe_sqr(z1z1, a.z);
    fe_mul(u2, b.x, z1z1);
    fe_mul(s2, a.z, z1z1);
    fe_mul(s2, b.y, s2);
    fe_sub(h, u2, a.x);        // H = U2 - X1
    fe_sub(rr, s2, a.y);       // S2 - Y1
    const bool hz = fe_is_zero(h);
    if (hz & !a.inf) {
        jac_unpack(r, jac_madd_degenerate(jac_pack(a), fe_is_zero(rr)));
        return;
    }
    fe_sqr(hh, h);             // HH = H^2
    fe_add(i4, hh, hh);
    fe_add(i4, i4, i4);        // I = 4HH
    fe_mul(j, h, i4);          // J = H*I
    fe_add(rr, rr, rr);        // r = 2(S2 - Y1)
    fe_mul(v, a.x, i4);        // V = X1*I
    jac out;
    fe_sqr(t, rr);
    fe_sub(t, t, j);
    fe_sub(t, t, v);
    fe_sub(out.x, t, v);       // X3 = r^2 - J - 2V
    fe_sub(t, v, out.x);
    fe_mul(t, rr, t);
    fe_mul(j, a.y, j);
    fe_add(j, j, j);
    fe_sub(out.y, t, j);       // Y3 = r(V - X3) - 2 Y1 J
    fe_add(t, a.z, h);
    fe_sqr(t, t);
    fe_sub(t, t, z1z1);
    fe_sub(out.z, t, hh);      // Z3 = (Z1 + H)^2 - Z1Z1 - HH
    out.inf = 0;
    if (a.inf) jac_from_aff(out, b);
    r = out;
}

// Flags of the memory-operand mixed addition.
constexpr u32 kAddNeg = 1;    // use -b
constexpr u32 kAddBeta = 2;   // use phi(b) = (beta * x, y)

UFS_DEV void load_aff_g(aff& b, const aff* bp) {
    const uint4* q = reinterpret_cast<const uint4*>(bp);
    const uint4 q0 = q[0], q1 = q[1], q2 = q[2], q3 = q[3];
    b.x.v[0] = q0.x; b.x.v[1] = q0.y; b.x.v[2] = q0.z; b.x.v[3] = q0.w;
    b.x.v[4] = q1.x; b.x.v[5] = q1.y; b.x.v[6] = q1.z; b.x.v[7] = q1.w;
    b.y.v[0] = q2.x; b.y.v[1] = q2.y; b.y.v[2] = q2.z; b.y.v[3] = q2.w;
    b.y.v[4] = q3.x; b.y.v[5] = q3.y; b.y.v[6] = q3.z; b.y.v[7] = q3.w;
}

// out = a + b for b affine (never infinity), flags as above, field
// arithmetic inline.  The degenerate cases are reported instead of handled:
// a == b (finite) leaves out = a and sets `need_dbl` (the caller doubles);
// a == -b gives infinity.
UFS_DEV void jac_madd_reg_body(jac& out, const jac& a, aff b, u32 flags, bool& need_dbl) {
    fe z1z1, u2, s2, h, hh, i4, j, rr, v, t;
    fe_sqr_i(z1z1, a.z);
    if (flags & kAddBeta) {
        const fe BETA = {{UFS_BETA_LIMBS}};
        fe_mul_i(b.x, b.x, BETA);
    }
    {
        fe ny;
        fe_neg(ny, b.y);
        const bool ng = (flags & kAddNeg) != 0;
#pragma unroll
        for (int l = 0; l < 8; ++l) b.y.v[l] = ng ? ny.v[l] : b.y.v[l];
    }
    fe_mul_i(u2, b.x, z1z1);
    fe_mul_i(s2, a.z, z1z1);
    fe_mul_i(s2, b.y, s2);
    fe_sub(h, u2, a.x);        // H = U2 - X1
    fe_sub(rr, s2, a.y);       // S2 - Y1
    const bool degen = fe_is_zero(h) & !a.inf;
    const bool same = fe_is_zero(rr);
    fe_sqr_i(hh, h);           // HH = H^2
    fe_add(i4, hh, hh);
    fe_add(i4, i4, i4);        // I = 4HH
    fe_mul_i(j, h, i4);        // J = H*I
    fe_add(rr, rr, rr);        // r = 2(S2 - Y1)
    fe_mul_i(v, a.x, i4);      // V = X1*I
    fe_sqr_i(t, rr);
    fe_sub(t, t, j);
    fe_sub(t, t, v);
    fe_sub(out.x, t, v);       // X3 = r^2 - J - 2V
    fe_sub(t, v, out.x);
    fe_mul_i(t, rr, t);
    fe_mul_i(j, a.y, j);
    fe_add(j, j, j);
    fe_sub(out.y, t, j);       // Y3 = r(V - X3) - 2 Y1 J
    fe_add(t, a.z, h);
    fe_sqr_i(t, t);
    fe_sub(t, t, z1z1);
    fe_sub(out.z, t, hh);      // Z3 = (Z1 + H)^2 - Z1Z1 - HH
    out.inf = 0;
    need_dbl = false;
    // Rare cases (a = inf, a = +-b) behind a wave-uniform test, so the common
    // path carries no per-lane selects.
    if (__any(a.inf | degen)) {
        if (a.inf) jac_from_aff(out, b);
        need_dbl = degen & same;
        if (degen) {
            if (same) {
                out = a;
            } else {
                jac_set_inf(out);
            }
        }
    }
}

// The same with b loaded from device memory.
UFS_DEV void jac_madd_mem_body(jac& out, const jac& a, const aff* bp, u32 flags, bool& need_dbl) {
    aff b;
    load_aff_g(b, bp);
    jac_madd_reg_body(out, a, b, flags, need_dbl);
}

// r = a + b (general Jacobian), exception-complete.
UFS_DEV void jac_add(jac& r, const jac& a, const jac& b) {
    if (a.inf) { r = b; return; }
    if (b.inf) { r = a; return; }
    fe z1z1, z2z2, u1, u2, s1, s2, h, i, j, rr, v, t;
    fe_sqr(z1z1, a.z);
    fe_sqr(z2z2, b.z);
    fe_mul(u1, a.x, z2z2);
    fe_mul(u2, b.x, z1z1);
    fe_mul(s1, b.z, z2z2);
    fe_mul(s1, a.y, s1);
    fe_mul(s2, a.z, z1z1);
    fe_mul(s2, b.y, s2);
    fe_sub(h, u2, u1);
    fe_sub(rr, s2, s1);
    if (fe_is_zero(h)) {
        if (fe_is_zero(rr)) {
            jac_dbl(r, a);
        } else {
            jac_set_inf(r);
        }
        return;
    }
    fe_add(i, h, h);
    fe_sqr(i, i);              // I = (2H)^2
    fe_mul(j, h, i);           // J = H*I
    fe_add(rr, rr, rr);        // r = 2(S2 - S1)
    fe_mul(v, u1, i);          // V = U1*I
    jac out;
    fe_sqr(t, rr);
    fe_sub(t, t, j);
    fe_sub(t, t, v);
    fe_sub(out.x, t, v);       // X3 = r^2 - J - 2V
    fe_sub(t, v, out.x);
    fe_mul(t, rr, t);
    fe_mul(s1, s1, j);
    fe_add(s1, s1, s1);
    fe_sub(out.y, t, s1);      // Y3 = r(V - X3) - 2 S1 J
    fe_add(t, a.z, b.z);
    fe_sqr(t, t);
    fe_sub(t, t, z1z1);
    fe_sub(t, t, z2z2);
    fe_mul(out.z, t, h);       // Z3 = ((Z1+Z2)^2 - Z1Z1 - Z2Z2) H
    out.inf = 0;
    r = out;
}

// Affine coordinates of a finite Jacobian point (one field inversion).
UFS_DEV void jac_to_aff(aff& r, const jac& a) {
    fe zi, zi2, zi3;
    fe_inv(zi, a.z);
    fe_sqr(zi2, zi);
    fe_mul(zi3, zi2, zi);
    fe_mul(r.x, a.x, zi2);
    fe_mul(r.y, a.y, zi3);
    fe_normalize(r.x);
    fe_normalize(r.y);
}

UFS_DEV void set_generator(aff& g) {
    const fe GX = {{UFS_GX_LIMBS}};
    const fe GY = {{UFS_GY_LIMBS}};
    g.x = GX;
    g.y = GY;
}

// Decompress x (raw limbs) to a point with the requested y parity.
// Returns false if x >= p or x^3 + 7 is not a square.
UFS_DEV bool lift_x(aff& r, const u32 xraw[8], u32 want_odd) {
    const bool lt = u256_lt_p(xraw);
#pragma unroll
    for (int i = 0; i < 8; ++i) r.x.v[i] = xraw[i];
    fe y2, t;
    fe_sqr(t, r.x);
    fe_mul(t, t, r.x);
    fe seven;
    fe_set_u32(seven, 7);
    fe_add(y2, t, seven);
    const bool ok = fe_sqrt(r.y, y2);
    fe_normalize(r.y);
    if ((r.y.v[0] & 1u) != want_odd) fe_neg(r.y, r.y);
    return lt & ok;
}

}  // namespace ufs
