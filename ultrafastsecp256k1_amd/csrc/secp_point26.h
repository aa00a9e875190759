// Jacobian doubling and mixed addition on 10 x 26-bit limbs, for the
// accumulators of the scalar-multiplication loops (secp_ecmult.h).  Same
// formulas and the same exception handling as the 8 x 32 versions in
// secp_point.h (dbl-2009-l, madd-2007-bl, a = 0, so they also run on the
// isomorphic curves y^2 = x^3 + 7 Z^6 of the point tables); the addend is
// read in the 8 x 32 storage form and converted.
//
// Lazy reduction: the constants K of fe26_sub<K> and the two weak
// normalisations (D in the doubling; h and r in the addition) come from
// worst-case limb-bound propagation through both formulas, iterated to a fixed
// point (X, Y, Z of magnitude at most 6.1, 10.1, 5.1 on entry and exit, every
// multiplication input at most 16).  Every accumulator value the loops feed
// in is within those bounds: affine entries (exact 26-bit limbs, y negated as
// 2p - y), multiplication outputs, and the formulas' own outputs.
#pragma once
#include "secp_field26.h"
#include "secp_point.h"

namespace ufs {

struct jac26 { fe26 x, y, z; u32 inf; };

UFS_DEV void fe26_set_u32(fe26& r, u32 x) {
    r.v[0] = x;
#pragma unroll
    for (int i = 1; i < 10; ++i) r.v[i] = 0;
}

UFS_DEV void jac26_set_inf(jac26& r) {
    fe26_set_u32(r.x, 1);
    fe26_set_u32(r.y, 1);
    fe26_set_u32(r.z, 0);
    r.inf = 1;
}

UFS_DEV void jac26_from_aff(jac26& r, const aff& a) {
    fe26_from_fe(r.x, a.x);
    fe26_from_fe(r.y, a.y);
    fe26_set_u32(r.z, 1);
    r.inf = 0;
}

UFS_DEV void jac26_to_jac(jac& r, const jac26& a) {
    fe26_to_fe(r.x, a.x);
    fe26_to_fe(r.y, a.y);
    fe26_to_fe(r.z, a.z);
    r.inf = a.inf;
}

// r = 2a (dbl-2009-l: 2M + 5S).  2a = inf iff a = inf.
UFS_DEV void jac26_dbl(jac26& r, const jac26& a) {
    fe26 A, B, C, D, E, F, t;
    fe26_sqr(A, a.x);          // A = X^2
    fe26_sqr(B, a.y);          // B = Y^2
    fe26_sqr(C, B);            // C = B^2
    fe26_add(t, a.x, B);
    fe26_sqr(t, t);
    fe26_sub<2>(t, t, A);
    fe26_sub<2>(t, t, C);
    fe26_add(t, t, t);
    fe26_norm(D, t);           // D = 2((X+B)^2 - A - C)
    fe26_mul3(E, A);           // E = 3A
    fe26_sqr(F, E);            // F = E^2
    fe26 z3;
    fe26_mul(z3, a.y, a.z);
    fe26_add(r.z, z3, z3);     // Z3 = 2YZ
    fe26_add(t, D, D);
    fe26_sub<3>(r.x, F, t);    // X3 = F - 2D
    fe26_sub<5>(t, D, r.x);
    fe26_mul(t, E, t);
    fe26_shl<3>(C, C);
    fe26_sub<9>(r.y, t, C);    // Y3 = E(D - X3) - 8C
    r.inf = a.inf;
}

// out = a + b for b affine (8 x 32, never infinity), flags kAddNeg / kAddBeta
// as in secp_point.h.  a == b (finite) leaves out = a and sets `need_dbl`
// (the caller doubles); a == -b gives infinity; a = inf gives b.
UFS_DEV void jac26_madd(jac26& out, const jac26& a, const aff& b8, u32 flags, bool& need_dbl) {
    fe26 bx, by;
    fe26_from_fe(bx, b8.x);
    fe26_from_fe(by, b8.y);
    if (flags & kAddBeta) {
        const fe BETA = {{UFS_BETA_LIMBS}};
        fe26 beta;
        fe26_from_fe(beta, BETA);
        fe26_mul(bx, bx, beta);
    }
    {
        fe26 ny;
        fe26_neg<2>(ny, by);
        const bool ng = (flags & kAddNeg) != 0;
#pragma unroll
        for (int l = 0; l < 10; ++l) by.v[l] = ng ? ny.v[l] : by.v[l];
    }
    fe26 z1z1, u2, s2, h, hh, i4, j, rr, v, t;
    fe26_sqr(z1z1, a.z);
    fe26_mul(u2, bx, z1z1);
    fe26_mul(s2, a.z, z1z1);
    fe26_mul(s2, by, s2);
    fe26_sub<7>(t, u2, a.x);
    fe26_norm(h, t);           // H = U2 - X1
    fe26_sub<11>(t, s2, a.y);
    fe26_add(t, t, t);
    fe26_norm(rr, t);          // r = 2(S2 - Y1)
    fe26_sqr(hh, h);           // HH = H^2
    fe26_shl<2>(i4, hh);       // I = 4HH
    fe26_mul(j, h, i4);        // J = H*I
    fe26_mul(v, a.x, i4);      // V = X1*I
    fe26_sqr(t, rr);
    fe26_sub<2>(t, t, j);
    fe26 v2;
    fe26_add(v2, v, v);
    fe26_sub<3>(out.x, t, v2); // X3 = r^2 - J - 2V
    fe26_sub<7>(t, v, out.x);
    fe26_mul(t, rr, t);
    fe26_mul(j, a.y, j);
    fe26_add(j, j, j);
    fe26_sub<3>(out.y, t, j);  // Y3 = r(V - X3) - 2 Y1 J
    fe26_add(t, a.z, h);
    fe26_sqr(t, t);
    fe26_sub<2>(t, t, z1z1);
    fe26_sub<2>(out.z, t, hh); // Z3 = (Z1 + H)^2 - Z1Z1 - HH
    out.inf = 0;
    need_dbl = false;
    // Rare cases (a = inf, H = 0) behind a wave-uniform test: H and r are
    // normalised, so a zero H has limb 2 equal to 0 or 2^26 - 1.
    if (__any(a.inf | fe26_normed_maybe_zero(h))) {
        const bool degen = fe26_normed_is_zero(h) & !a.inf;
        const bool same = fe26_normed_is_zero(rr);
        if (a.inf) {
            out.x = bx;
            out.y = by;
            fe26_set_u32(out.z, 1);
            out.inf = 0;
        }
        need_dbl = degen & same;
        if (degen) {
            if (same) {
                out = a;
            } else {
                jac26_set_inf(out);
            }
        }
    }
}

UFS_DEV void jac26_madd_mem(jac26& out, const jac26& a, const aff* bp, u32 flags, bool& need_dbl) {
    aff b;
    load_aff_g(b, bp);
    jac26_madd(out, a, b, flags, need_dbl);
}

UFS_DEV void jac26_from_jac(jac26& r, const jac& a) {
    fe26_from_fe(r.x, a.x);
    fe26_from_fe(r.y, a.y);
    fe26_from_fe(r.z, a.z);
    r.inf = a.inf;
}

// out = a + b for two Jacobian points (add-2007-bl: 11M + 5S, with
// Z3 = 2 Z1 Z2 H).  Every case is handled: a = inf gives b, b = inf gives a,
// a == b doubles (jac26_dbl), a == -b gives infinity.  `out` may alias an
// operand.  Sub constants from the same limb-bound fixed point as the
// doubling and the mixed addition, with both operands anywhere in it.
UFS_DEV void jac26_add(jac26& out, const jac26& a, const jac26& b) {
    fe26 z1z1, z2z2, u1, u2, s1, s2, h, rr, i4, j, v, t, t2;
    fe26_sqr(z1z1, a.z);
    fe26_sqr(z2z2, b.z);
    fe26_mul(u1, a.x, z2z2);
    fe26_mul(u2, b.x, z1z1);
    fe26_mul(t, b.z, z2z2);
    fe26_mul(s1, a.y, t);
    fe26_mul(t, a.z, z1z1);
    fe26_mul(s2, b.y, t);
    fe26_sub<2>(t, u2, u1);
    fe26_norm(h, t);           // H = U2 - U1
    fe26_sub<2>(t, s2, s1);
    fe26_add(t, t, t);
    fe26_norm(rr, t);          // r = 2(S2 - S1)
    fe26_sqr(t, h);
    fe26_shl<2>(i4, t);        // I = (2H)^2
    fe26_mul(j, h, i4);        // J = H I
    fe26_mul(v, u1, i4);       // V = U1 I
    jac26 r;
    fe26_sqr(t, rr);
    fe26_sub<2>(t, t, j);
    fe26_add(t2, v, v);
    fe26_sub<3>(r.x, t, t2);   // X3 = r^2 - J - 2V
    fe26_sub<7>(t, v, r.x);
    fe26_mul(t, rr, t);
    fe26_mul(j, s1, j);
    fe26_add(j, j, j);
    fe26_sub<3>(r.y, t, j);    // Y3 = r (V - X3) - 2 S1 J
    fe26_mul(t, a.z, b.z);
    fe26_mul(t, t, h);
    fe26_add(r.z, t, t);       // Z3 = 2 Z1 Z2 H
    r.inf = 0;
    // Rare and special cases behind a wave-uniform test (H and r are
    // normalised: a zero H has limb 2 equal to 0 or 2^26 - 1).
    if (__any(a.inf | b.inf | fe26_normed_maybe_zero(h))) {
        const bool degen = fe26_normed_is_zero(h) & !a.inf & !b.inf;
        const bool same = fe26_normed_is_zero(rr);
        if (__any(degen & same)) {
            jac26 d;
            jac26_dbl(d, a);
            if (degen & same) r = d;
        }
        if (degen & !same) jac26_set_inf(r);
        if (b.inf) r = a;
        if (a.inf) r = b;
    }
    out = r;
}

// r = lane `src`'s a on every lane (src wave-uniform: v_readlane_b32, no LDS
// round trip).
UFS_DEV void fe26_shfl(fe26& r, const fe26& a, int src) {
#pragma unroll
    for (int i = 0; i < 10; ++i) r.v[i] = (u32)__builtin_amdgcn_readlane((int)a.v[i], src);
}

UFS_DEV void fe26_select(fe26& r, bool c, const fe26& a, const fe26& b) {
#pragma unroll
    for (int i = 0; i < 10; ++i) r.v[i] = c ? a.v[i] : b.v[i];
}

// a = 2a with the products of the doubling spread over lanes 0..2 of the
// wave, for a single accumulator whose time is the latency of its dependent
// instructions (every lane of the wave calls it with the same a and gets the
// same result).  The operations and values are those of jac26_dbl:
//   step 1 (one multiply, per lane):  X^2 = A (lane 0), Y^2 = B (1), Y Z (2)
//   step 2 (one square, per lane):    (3A)^2 = F (0), B^2 = C (1), (X+B)^2 (2)
//   step 3 (one multiply):            E (D - X3) with D, X3 from steps 1-2
// so three multiplication latencies instead of seven, plus lane shuffles.
UFS_DEV void jac26_dbl_lanes(jac26& a) {
    const int lane = (int)(threadIdx.x & 63u);
    fe26 p, q, r1, r2, B, C, T, A3, D, E, F, X3, Y3, Z3, t;
    fe26_select(p, lane == 0, a.x, a.y);
    fe26_select(q, lane == 0, a.x, a.y);
    fe26_select(q, lane == 2, a.z, q);
    fe26_mul(r1, p, q);                 // A | B | YZ
    fe26_shfl(B, r1, 1);
    fe26_mul3(E, r1);                   // 3A on lane 0
    fe26_add(t, a.x, B);
    fe26_select(p, lane == 0, E, r1);
    fe26_select(p, lane == 2, t, p);
    fe26_sqr(r2, p);                    // F | C | (X+B)^2
    fe26_shfl(C, r2, 1);
    fe26_shfl(T, r2, 2);
    fe26_shfl(A3, r1, 0);               // A
    fe26_shfl(F, r2, 0);
    fe26_sub<2>(t, T, A3);
    fe26_sub<2>(t, t, C);
    fe26_add(t, t, t);
    fe26_norm(D, t);                    // D = 2((X+B)^2 - A - C)
    fe26_mul3(E, A3);                   // E = 3A
    fe26_add(t, D, D);
    fe26_sub<3>(X3, F, t);              // X3 = F - 2D
    fe26_sub<5>(t, D, X3);
    fe26_mul(t, E, t);
    fe26_shl<3>(C, C);
    fe26_sub<9>(Y3, t, C);              // Y3 = E(D - X3) - 8C
    fe26_shfl(Z3, r1, 2);
    fe26_add(Z3, Z3, Z3);               // Z3 = 2YZ
    a.x = X3;
    a.y = Y3;
    a.z = Z3;
}

// out = a + b, for operands held identically by every lane of the wave (the
// MSM final's Horner additions), with the sixteen products of jac26_add
// spread over lanes 0..3 in five dependent levels:
//   1: Z1^2 (lane 0), Z2^2 (1), Z1 Z2 (2)
//   2: U1 = X1 Z2^2 (0), U2 = X2 Z1^2 (1), Z2^3 (2), Z1^3 (3)
//   3: S1 = Y1 Z2^3 (0), S2 = Y2 Z1^3 (1), H^2 (2), Z1 Z2 H (3)
//   4: J = H I (0), V = U1 I (1), r^2 (2)          (I = 4 H^2)
//   5: r (V - X3) (0), S1 J (1)
// The operations, subtraction constants and special cases are those of
// jac26_add; every lane returns the same point.
UFS_DEV void jac26_add_lanes(jac26& out, const jac26& a, const jac26& b) {
    const int lane = (int)(threadIdx.x & 63u);
    fe26 p, q, r1, r2, r3, r4, r5, z1z1, z2z2, u1, u2, s1, s2, h, rr, i4, j, v, t, t2, z;
    // level 1
    fe26_select(p, lane == 1, b.z, a.z);
    fe26_select(q, lane == 0, a.z, b.z);
    fe26_mul(r1, p, q);                 // Z1^2 | Z2^2 | Z1 Z2
    fe26_shfl(z1z1, r1, 0);
    fe26_shfl(z2z2, r1, 1);
    // level 2
    fe26_select(p, lane == 0, a.x, b.x);
    fe26_select(p, lane == 2, b.z, p);
    fe26_select(p, lane == 3, a.z, p);
    fe26_select(q, (lane == 0) | (lane == 2), z2z2, z1z1);
    fe26_mul(r2, p, q);                 // U1 | U2 | Z2^3 | Z1^3
    fe26_shfl(u1, r2, 0);
    fe26_shfl(u2, r2, 1);
    fe26_sub<2>(t, u2, u1);
    fe26_norm(h, t);                    // H = U2 - U1
    // level 3
    fe26_shfl(t, r2, 2);                // Z2^3
    fe26_shfl(t2, r2, 3);               // Z1^3
    fe26_select(q, lane == 0, t, t2);
    fe26_select(q, lane == 2, h, q);
    fe26_shfl(z, r1, 2);                // Z1 Z2
    fe26_select(q, lane == 3, z, q);
    fe26_select(p, lane == 0, a.y, b.y);
    fe26_select(p, (lane == 2) | (lane == 3), h, p);
    fe26_mul(r3, p, q);                 // S1 | S2 | H^2 | Z1 Z2 H
    fe26_shfl(s1, r3, 0);
    fe26_shfl(s2, r3, 1);
    fe26_sub<2>(t, s2, s1);
    fe26_add(t, t, t);
    fe26_norm(rr, t);                   // r = 2(S2 - S1)
    fe26_shfl(t, r3, 2);
    fe26_shl<2>(i4, t);                 // I = (2H)^2
    // level 4
    fe26_select(p, lane == 1, u1, h);
    fe26_select(p, lane == 2, rr, p);
    fe26_select(q, lane == 2, rr, i4);
    fe26_mul(r4, p, q);                 // J | V | r^2
    fe26_shfl(j, r4, 0);
    fe26_shfl(v, r4, 1);
    fe26_shfl(t, r4, 2);
    jac26 r;
    fe26_sub<2>(t, t, j);
    fe26_add(t2, v, v);
    fe26_sub<3>(r.x, t, t2);            // X3 = r^2 - J - 2V
    // level 5
    fe26_sub<7>(t, v, r.x);
    fe26_select(p, lane == 0, rr, s1);
    fe26_select(q, lane == 0, t, j);
    fe26_mul(r5, p, q);                 // r (V - X3) | S1 J
    fe26_shfl(t, r5, 0);
    fe26_shfl(j, r5, 1);
    fe26_add(j, j, j);
    fe26_sub<3>(r.y, t, j);             // Y3 = r (V - X3) - 2 S1 J
    fe26_shfl(t, r3, 3);
    fe26_add(r.z, t, t);                // Z3 = 2 Z1 Z2 H
    r.inf = 0;
    // the special cases of jac26_add (uniform: every lane holds the same)
    if (a.inf | b.inf | fe26_normed_maybe_zero(h)) {
        const bool degen = fe26_normed_is_zero(h) & !a.inf & !b.inf;
        const bool same = fe26_normed_is_zero(rr);
        if (degen & same) jac26_dbl(r, a);
        if (degen & !same) jac26_set_inf(r);
        if (b.inf) r = a;
        if (a.inf) r = b;
    }
    out = r;
}

}  // namespace ufs
