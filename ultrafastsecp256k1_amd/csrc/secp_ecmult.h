// Double-scalar multiplication a*G + b*P for batch verification, one lane per
// signature, written for wave64 SIMT execution:
//
//  * b*P: GLV split b = k1 + lambda*k2 (|k1|,|k2| < 2^128) and a regular
//    signed-digit recoding with odd digits only, so every lane adds at the same
//    doubling positions (no wNAF divergence).  The odd multiples
//    {1,3,...,15}*P are built once per lane and rescaled to a common Z, i.e.
//    they are affine on an isomorphic curve y^2 = x^3 + 7 Z^6; the accumulator
//    runs on that curve with mixed additions and the result's Z is multiplied
//    by the common Z at the end.  phi(P) = (beta*x, y) reuses the same table.
//  * a*G: fixed-base comb over a context-resident table of odd multiples of
//    2^(w*i)*G, read from HBM (L2/MALL resident), no doublings.
//
// Reference counterpart: shamir_double_mul_glv (src/cuda/include/
// secp256k1.cuh:3575-3801), which GLV-splits both scalars and builds a joint
// 16-entry table with a field inversion per lane; the CPU oracle's
// Point::dual_scalar_mul_gen_point (src/cpu/src/point.cpp:4079-4234) uses fixed
// G tables like here.  The verdict depends only on the exact group element, so
// any exact algorithm yields identical results.
#pragma once
#include "secp_point.h"

namespace ufs {

constexpr int kPW = 4;                 // window of the variable-base part
constexpr int kPEntries = 1 << (kPW - 1);
constexpr int kPWindows = 128 / kPW;   // GLV halves are < 2^128

// ---------------------------------------------------------------------------
// GLV decomposition: k = k1 + lambda*k2 (mod n), returned as magnitudes with
// sign flags.  Follows the published libsecp256k1 split constants
// (reference: secp256k1.cuh:845-1064 glv_decompose).
// ---------------------------------------------------------------------------
UFS_DEV void glv_split(sc& k1, sc& k2, bool& neg1, bool& neg2, const sc& k) {
    const u32 G1[8] = {UFS_G1_LIMBS};
    const u32 G2[8] = {UFS_G2_LIMBS};
    const sc MB1 = {{UFS_MB1_LIMBS}};
    const sc MB2 = {{UFS_MB2_LIMBS}};
    const sc ML = {{UFS_MLAMBDA_LIMBS}};
    u32 t[16];
    sc c1, c2;
    mul_512(t, k.v, G1);
    {
        u32 c = t[11] >> 31;   // round(k*g1 / 2^384)
        c1.v[0] = addc(t[12], 0, c, c);
        c1.v[1] = addc(t[13], 0, c, c);
        c1.v[2] = addc(t[14], 0, c, c);
        c1.v[3] = addc(t[15], 0, c, c);
        c1.v[4] = c1.v[5] = c1.v[6] = c1.v[7] = 0;
    }
    mul_512(t, k.v, G2);
    {
        u32 c = t[11] >> 31;
        c2.v[0] = addc(t[12], 0, c, c);
        c2.v[1] = addc(t[13], 0, c, c);
        c2.v[2] = addc(t[14], 0, c, c);
        c2.v[3] = addc(t[15], 0, c, c);
        c2.v[4] = c2.v[5] = c2.v[6] = c2.v[7] = 0;
    }
    sc_mul(c1, c1, MB1);
    sc_mul(c2, c2, MB2);
    sc_add(k2, c1, c2);
    sc tmp;
    sc_mul(tmp, k2, ML);
    sc_add(k1, tmp, k);
    neg1 = sc_is_high(k1);
    neg2 = sc_is_high(k2);
    sc n1, n2;
    sc_neg(n1, k1);
    sc_neg(n2, k2);
#pragma unroll
    for (int i = 0; i < 8; ++i) {
        k1.v[i] = neg1 ? n1.v[i] : k1.v[i];
        k2.v[i] = neg2 ? n2.v[i] : k2.v[i];
    }
}

// k (< 2^128) -> odd k' = k | 1 (k' = k + 1 when k is even); returns whether
// a correction (subtract one base point) is needed.
UFS_DEV bool make_odd(sc& k) {
    const bool even = (k.v[0] & 1u) == 0;
    k.v[0] |= 1u;
    return even;
}

// Bit field [pos, pos + w) of a 256-bit value (w <= 16, pos may be dynamic).
UFS_DEV u32 bits_at(const u32 k[8], int pos, int w) {
    const int wi = pos >> 5, sh = pos & 31;
    u32 lo = k[0], hi = k[1];
#pragma unroll
    for (int i = 1; i < 8; ++i) {
        lo = (wi == i) ? k[i] : lo;
        hi = (wi + 1 == i) ? k[i] : hi;
    }
    hi = (wi >= 7) ? 0u : hi;
    lo = (wi >= 8) ? 0u : lo;
    const u64 v = ((u64)hi << 32) | lo;
    return (u32)(v >> sh) & ((1u << w) - 1u);
}

// Regular odd-digit recoding of an odd k: digit i (window width w) is
// d_i = (win_i | 1) - (bit w(i+1) of k ? 0 : 2^w); returns table index
// (|d|-1)/2 and the sign.
UFS_DEV void odd_digit(const u32 k[8], int i, int w, u32& idx, bool& neg) {
    const u32 win = bits_at(k, w * i, w) | 1u;
    const u32 nb = bits_at(k, w * (i + 1), 1);
    // d = win - 2^w (nb = 0, negative) or win (nb = 1, positive)
    const u32 mag = nb ? win : ((1u << w) - win);
    idx = (mag - 1u) >> 1;
    neg = nb == 0;
}

// ---------------------------------------------------------------------------
// Variable-base table: odd multiples {1,3,...,15}*P with a common Z.
// ---------------------------------------------------------------------------
struct ptable {
    fe x[kPEntries];
    fe y[kPEntries];
    fe zs;   // true Z of the accumulator = acc.z * zs
};

UFS_DEV void build_ptable(ptable& T, const aff& P) {
    jac Pj, D;
    jac_from_aff(Pj, P);
    jac_dbl(D, Pj);
    // Isomorphic curve scaled by D.z: D becomes affine (D.x, D.y).
    fe z2, z3;
    fe_sqr(z2, D.z);
    fe_mul(z3, z2, D.z);
    fe cx, cy, cz;
    fe_mul(cx, P.x, z2);
    fe_mul(cy, P.y, z3);
    fe_set_u32(cz, 1);
    T.x[0] = cx;
    T.y[0] = cy;
    fe zr[kPEntries];
#pragma unroll
    for (int j = 1; j < kPEntries; ++j) {
        // (cx,cy,cz) += (D.x, D.y) -- madd-2007-bl; no exceptional case can
        // arise: (2j-1)P = +-2P is impossible for a point of order n.
        fe z1z1, u2, s2, h, hh, i4, jj, rr, v, t;
        fe_sqr(z1z1, cz);
        fe_mul(u2, D.x, z1z1);
        fe_mul(s2, cz, z1z1);
        fe_mul(s2, D.y, s2);
        fe_sub(h, u2, cx);
        fe_sub(rr, s2, cy);
        fe_sqr(hh, h);
        fe_add(i4, hh, hh);
        fe_add(i4, i4, i4);
        fe_mul(jj, h, i4);
        fe_add(rr, rr, rr);
        fe_mul(v, cx, i4);
        fe nx, ny;
        fe_sqr(t, rr);
        fe_sub(t, t, jj);
        fe_sub(t, t, v);
        fe_sub(nx, t, v);
        fe_sub(t, v, nx);
        fe_mul(t, rr, t);
        fe_mul(jj, cy, jj);
        fe_add(jj, jj, jj);
        fe_sub(ny, t, jj);
        fe_add(zr[j], h, h);       // Z_j = Z_{j-1} * 2H
        fe_mul(cz, cz, zr[j]);
        cx = nx;
        cy = ny;
        T.x[j] = cx;
        T.y[j] = cy;
    }
    // Rescale entries 0..6 to the common Z = cz.
    fe s = zr[kPEntries - 1];
#pragma unroll
    for (int j = kPEntries - 2; j >= 0; --j) {
        fe s2, s3;
        fe_sqr(s2, s);
        fe_mul(s3, s2, s);
        fe_mul(T.x[j], T.x[j], s2);
        fe_mul(T.y[j], T.y[j], s3);
        if (j > 0) fe_mul(s, s, zr[j]);
    }
    fe_mul(T.zs, cz, D.z);
}

UFS_DEV void ptable_select(aff& r, const ptable& T, u32 idx) {
    r.x = T.x[0];
    r.y = T.y[0];
#pragma unroll
    for (int j = 1; j < kPEntries; ++j) {
        const bool c = idx == (u32)j;
#pragma unroll
        for (int l = 0; l < 8; ++l) {
            r.x.v[l] = c ? T.x[j].v[l] : r.x.v[l];
            r.y.v[l] = c ? T.y[j].v[l] : r.y.v[l];
        }
    }
}

UFS_DEV void fe_cond_neg(fe& y, bool c) {
    fe n;
    fe_neg(n, y);
#pragma unroll
    for (int l = 0; l < 8; ++l) y.v[l] = c ? n.v[l] : y.v[l];
}

// acc (on the isomorphic curve) = k1*P' + k2*phi(P') where P' = +-P by sign.
// Returns the true Jacobian point (Z multiplied by T.zs).
UFS_DEV void ecmult_var(jac& out, const ptable& T, sc k1, bool neg1, sc k2, bool neg2) {
    const fe BETA = {{UFS_BETA_LIMBS}};
    const bool fix1 = make_odd(k1);
    const bool fix2 = make_odd(k2);
    jac acc;
    aff e;
    // top digits are 1 for both streams (k < 2^128)
    e.x = T.x[0];
    e.y = T.y[0];
    fe_cond_neg(e.y, neg1);
    jac_from_aff(acc, e);
    fe_mul(e.x, T.x[0], BETA);
    e.y = T.y[0];
    fe_cond_neg(e.y, neg2);
    jac_madd(acc, acc, e);
    for (int i = kPWindows - 1; i >= 0; --i) {
        for (int d = 0; d < kPW; ++d) jac_dbl(acc, acc);
        u32 idx;
        bool ng;
        odd_digit(k1.v, i, kPW, idx, ng);
        ptable_select(e, T, idx);
        fe_cond_neg(e.y, ng != neg1);
        jac_madd(acc, acc, e);
        odd_digit(k2.v, i, kPW, idx, ng);
        ptable_select(e, T, idx);
        fe_mul(e.x, e.x, BETA);
        fe_cond_neg(e.y, ng != neg2);
        jac_madd(acc, acc, e);
    }
    // corrections for even scalars: subtract one P' (resp. phi(P'))
    {
        jac t;
        e.x = T.x[0];
        e.y = T.y[0];
        fe_cond_neg(e.y, !neg1);
        jac_madd(t, acc, e);
        jac_select(acc, fix1, t);
        fe_mul(e.x, T.x[0], BETA);
        e.y = T.y[0];
        fe_cond_neg(e.y, !neg2);
        jac_madd(t, acc, e);
        jac_select(acc, fix2, t);
    }
    out = acc;
    fe_mul(out.z, acc.z, T.zs);
}

// ---------------------------------------------------------------------------
// Fixed-base comb for G.  Table layout (affine, 64 B per entry):
//   entry(i, j) = (2j+1) * 2^(w*i) * G,  i < nwin, j < 2^(w-1)
//   entry(nwin, 0) = 2^(w*nwin) * G      (top term of the odd recoding)
// ---------------------------------------------------------------------------
struct gcomb {
    const aff* tab;
    int w;
    int nwin;
};

UFS_DEV void load_aff(aff& r, const aff* p) {
    const uint4* q = reinterpret_cast<const uint4*>(p);
    uint4 a = q[0], b = q[1], c = q[2], d = q[3];
    r.x.v[0] = a.x; r.x.v[1] = a.y; r.x.v[2] = a.z; r.x.v[3] = a.w;
    r.x.v[4] = b.x; r.x.v[5] = b.y; r.x.v[6] = b.z; r.x.v[7] = b.w;
    r.y.v[0] = c.x; r.y.v[1] = c.y; r.y.v[2] = c.z; r.y.v[3] = c.w;
    r.y.v[4] = d.x; r.y.v[5] = d.y; r.y.v[6] = d.z; r.y.v[7] = d.w;
}

// r = a*G for any a in [0, n).
UFS_DEV void ecmult_gen(jac& r, const gcomb& G, const sc& a) {
    const u32 N[8] = {UFS_N_LIMBS};
    // odd representative: a odd -> a, else n - a (odd), negating G.
    sc k;
    const bool even = (a.v[0] & 1u) == 0;
    {
        u32 bw = 0, t[8];
#pragma unroll
        for (int i = 0; i < 8; ++i) t[i] = subc(N[i], a.v[i], bw, bw);
#pragma unroll
        for (int i = 0; i < 8; ++i) k.v[i] = even ? t[i] : a.v[i];
    }
    const int half = 1 << (G.w - 1);
    aff e;
    load_aff(e, G.tab + (size_t)G.nwin * half);
    fe_cond_neg(e.y, even);
    jac acc;
    jac_from_aff(acc, e);
    u32 idx;
    bool ng;
    odd_digit(k.v, 0, G.w, idx, ng);
    aff nxt;
    load_aff(nxt, G.tab + idx);
    bool nxt_neg = ng;
    for (int i = 0; i < G.nwin; ++i) {
        e = nxt;
        const bool eneg = nxt_neg != even;
        if (i + 1 < G.nwin) {
            odd_digit(k.v, i + 1, G.w, idx, ng);
            load_aff(nxt, G.tab + (size_t)(i + 1) * half + idx);
            nxt_neg = ng;
        }
        fe_cond_neg(e.y, eneg);
        jac_madd(acc, acc, e);
    }
    r = acc;
}

}  // namespace ufs
