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
#include "secp_point26.h"

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
// Variable-base table: odd multiples {1,3,...,15}*P with a common Z, kept in
// a lane-private slot of device memory rather than in VGPRs (eight affine
// points are 128 VGPRs, which alone would cap the kernels at one wave per
// SIMD).  Each slot is 768 bytes, contiguous per lane, so one entry read is
// four 16-byte loads of a single 64-byte line segment:
//   slot[0..7]  (2j+1)*P' on the isomorphic curve (affine, 64 B each)
//   slot[8]     x = zs, the Z scale of the isomorphic curve
//   slot[9..11] Z ratios used while the table is built
// The entries for a window are loaded before its doublings, so the load
// latency hides behind four doublings.
// ---------------------------------------------------------------------------
constexpr int kSlotEntries = 13;   // 832 bytes per lane (entry 12: BIP-340 pair stash)

UFS_DEV void store_fe(fe* p, const fe& a) {
    uint4* q = reinterpret_cast<uint4*>(p);
    q[0] = make_uint4(a.v[0], a.v[1], a.v[2], a.v[3]);
    q[1] = make_uint4(a.v[4], a.v[5], a.v[6], a.v[7]);
}

UFS_DEV void load_fe(fe& a, const fe* p) {
    const uint4* q = reinterpret_cast<const uint4*>(p);
    const uint4 x = q[0], y = q[1];
    a.v[0] = x.x; a.v[1] = x.y; a.v[2] = x.z; a.v[3] = x.w;
    a.v[4] = y.x; a.v[5] = y.y; a.v[6] = y.z; a.v[7] = y.w;
}

UFS_DEV void store_aff(aff* p, const aff& a) {
    store_fe(&p->x, a.x);
    store_fe(&p->y, a.y);
}

UFS_DEV void load_aff(aff& r, const aff* p) {
    load_fe(r.x, &p->x);
    load_fe(r.y, &p->y);
}

// Z ratio zr[j] (j = 1..7) lives in the slot's spare entries.
UFS_DEV fe* slot_zr(aff* slot, int j) {
    fe* z = reinterpret_cast<fe*>(slot + 9);
    return z + (j - 1);
}

// The odd multiples by co-Z arithmetic (Meloni's addition of two points
// sharing Z; the table scheme of Longa and Miri; the formulas as XYCZ-ADD in
// Goundar, Joye, Miyaji, Rivain and Venelli, "Scalar multiplication on
// Weierstrass elliptic curves from Co-Z arithmetic", 2011).  A doubling of
// the affine P gives Q = 2P and P itself at the same Z = 2y; each step then
// adds Q to the last odd multiple, and the sum and Q come out sharing the
// new Z = Z (x_last - x_Q), whose factor (x_last - x_Q) is the step's Z
// ratio.  4M + 2S per step, against 7M + 4S for a mixed addition with a Z
// update.  No step can meet Q = +-(2j-1)P: that needs (2j+1)P or (2j-3)P to
// be infinity, impossible for a point of prime order n.
UFS_DEV void build_ptable(aff* slot, const aff& P) {
    // from Z = 1: B = x^2, E = y^2, L = E^2, S = 2((x + E)^2 - B - L) = 4xy^2,
    // M = 3B; 2P = (M^2 - 2S, M(S - X(2P)) - 8L) and P = (S, 8L), Z = 2y
    fe b, e, l, t, qx, qy, cx, cy;
    fe_sqr(b, P.x);
    fe_sqr(e, P.y);
    fe_sqr(l, e);
    fe_add(t, P.x, e);
    fe_sqr(t, t);
    fe_sub(t, t, b);
    fe_sub(t, t, l);
    fe_add(cx, t, t);          // S
    fe m;
    fe_add(m, b, b);
    fe_add(m, m, b);           // M = 3x^2 (a = 0)
    fe_sqr(t, m);
    fe_sub(t, t, cx);
    fe_sub(qx, t, cx);         // X(2P) = M^2 - 2S
    fe_sub(t, cx, qx);
    fe_mul(t, m, t);
    fe_add(cy, l, l);
    fe_add(cy, cy, cy);
    fe_add(cy, cy, cy);        // 8L
    fe_sub(qy, t, cy);         // Y(2P) = M(S - X(2P)) - 8L
    store_aff(slot, aff{cx, cy});
    fe zlast;
    for (int j = 1; j < kPEntries; ++j) {
        // (cx, cy) + (qx, qy) with a common Z; (qx, qy) moves to the new Z
        fe dx, a, bq, bc, dy, ee;
        fe_sub(dx, cx, qx);        // x_last - x_Q: the Z ratio
        fe_sqr(a, dx);
        fe_mul(bq, qx, a);         // Q's new X
        fe_mul(bc, cx, a);
        fe_sub(dy, cy, qy);
        fe_sub(t, bc, bq);
        fe_mul(ee, qy, t);         // Q's new Y = y_Q dx^3
        fe_sqr(t, dy);
        fe_sub(t, t, bq);
        fe_sub(cx, t, bc);         // X = dy^2 - x_Q dx^2 - x_last dx^2
        fe_sub(t, bq, cx);
        fe_mul(t, dy, t);
        fe_sub(cy, t, ee);         // Y = dy (x_Q dx^2 - X) - y_Q dx^3
        qx = bq;
        qy = ee;
        store_aff(slot + j, aff{cx, cy});
        if (j < kPEntries - 1) store_fe(slot_zr(slot, j), dx);
        zlast = dx;
    }
    // Rescale entries 0..6 to the common Z of entry 7: entry j gets the
    // product of the ratios after it, s = zr[j+1] * ... * zr[7].
    fe s = zlast;
    for (int j = kPEntries - 2; j >= 0; --j) {
        fe s2, s3;
        aff en;
        load_aff(en, slot + j);
        fe_sqr(s2, s);
        fe_mul(s3, s2, s);
        fe_mul(en.x, en.x, s2);
        fe_mul(en.y, en.y, s3);
        store_aff(slot + j, en);
        if (j > 0) {
            fe zr;
            load_fe(zr, slot_zr(slot, j));
            fe_mul(s, s, zr);
        }
    }
    // the common Z: 2y times every ratio (s now holds zr[1] * ... * zr[7])
    fe zs, y2;
    fe_add(y2, P.y, P.y);
    fe_mul(zs, s, y2);
    store_fe(&slot[8].x, zs);
}

UFS_DEV void fe_cond_neg(fe& y, bool c) {
    fe n;
    fe_neg(n, y);
#pragma unroll
    for (int l = 0; l < 8; ++l) y.v[l] = c ? n.v[l] : y.v[l];
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

// Per-lane digit store in LDS for ecmult_dual: words [0, 4) hold the
// variable-base digits of k1 and words [4, 8) those of k2 (4 bits each: table
// index 0..7 and the sign, 8 digits per word), words [8, 16) the odd comb
// scalar.  Word q of lane t sits at q * kDigitLanes + t, so a wave's reads hit
// 64 consecutive words.  The scalars leave the registers before the step loop:
// recoded once with constant bit positions instead of one dynamic bit-field
// selection per window, with nothing of them to spill across the loop.
constexpr int kDigitLanes = 256;   // the largest block that calls ecmult_dual
constexpr int kDigitWords = 16;

UFS_DEV u32* lane_digit_store() {
    __shared__ u32 store[kDigitWords * kDigitLanes];
    return store + threadIdx.x;
}

// Bit field [pos, pos + w) of the 256-bit value in words base.. of the store.
UFS_DEV u32 lds_bits_at(const u32* st, int base, int pos, int w) {
    const int wi = pos >> 5, sh = pos & 31;
    const u32 lo = wi < 8 ? st[(base + wi) * kDigitLanes] : 0u;
    const u32 hi = wi + 1 < 8 ? st[(base + wi + 1) * kDigitLanes] : 0u;
    const u64 v = ((u64)hi << 32) | lo;
    return (u32)(v >> sh) & ((1u << w) - 1u);
}


// ---------------------------------------------------------------------------
// Double-scalar multiplication u*G + (k1 + lambda k2)*P' as one flat loop of
// steps, so a single inline copy of the doubling and of the mixed addition
// serves the whole computation (small code, no calls on the hot path, free
// register allocation).  Every lane runs the same step sequence:
//   steps 0..191   window i = 31..0 of the GLV halves: 4 doublings, then the
//                  digit of k1 (+-P') and of k2 (+-phi(P')) from the slot;
//   steps 192,193  corrections for even k1, k2 (subtract P' / phi(P'));
//   step 194       rescale from the isomorphic curve: z *= zs;
//   steps 195..    the G comb: top entry, then one entry per window.
// A mixed addition that meets a == b (possible only for adversarial inputs)
// marks the lane; the next step is then a doubling for the marked lanes
// before the schedule continues.
// ---------------------------------------------------------------------------
UFS_DEV void ecmult_dual(jac& out, const aff* slot, sc k1, bool neg1, sc k2, bool neg2,
                         const gcomb& G, const sc& u) {
    const u32 N[8] = {UFS_N_LIMBS};
#pragma unroll
    for (int i = 4; i < 8; ++i) {
        k1.v[i] = 0;
        k2.v[i] = 0;
    }
    const bool fix1 = make_odd(k1);
    const bool fix2 = make_odd(k2);
    const u32 f1 = neg1 ? kAddNeg : 0u, f2 = (neg2 ? kAddNeg : 0u) | kAddBeta;
    // odd representative of u: u odd -> u, else n - u (odd), negating G
    sc ku;
    const bool ueven = (u.v[0] & 1u) == 0;
    {
        u32 bw = 0, t[8];
#pragma unroll
        for (int i = 0; i < 8; ++i) t[i] = subc(N[i], u.v[i], bw, bw);
#pragma unroll
        for (int i = 0; i < 8; ++i) ku.v[i] = ueven ? t[i] : u.v[i];
    }
    const u32 fg = ueven ? kAddNeg : 0u;
    const int half = 1 << (G.w - 1);
    constexpr int kVarSteps = kPWindows * (kPW + 2);
    const int nsteps = kVarSteps + 3 + 1 + G.nwin;
    u32* const dig = lane_digit_store();
    {
        static_assert(kPW == 4 && kPWindows % 8 == 0, "4-bit digits, 8 per word");
#pragma unroll
        for (int q = 0; q < kPWindows / 8; ++q) {
            u32 w1 = 0, w2 = 0;
#pragma unroll
            for (int e = 0; e < 8; ++e) {
                u32 idx;
                bool ng;
                odd_digit(k1.v, 8 * q + e, kPW, idx, ng);
                w1 |= (idx | (ng ? 8u : 0u)) << (4 * e);
                odd_digit(k2.v, 8 * q + e, kPW, idx, ng);
                w2 |= (idx | (ng ? 8u : 0u)) << (4 * e);
            }
            dig[q * kDigitLanes] = w1;
            dig[(4 + q) * kDigitLanes] = w2;
        }
#pragma unroll
        for (int l = 0; l < 8; ++l) dig[(8 + l) * kDigitLanes] = ku.v[l];
    }

    // accumulator (10 x 26-bit limbs) starts at the top digits,
    // P' + phi(P') (both 1: k < 2^128): P' itself here, phi(P') as the first
    // scheduled addition below.
    jac26 acc;
    {
        aff e;
        load_aff(e, slot);
        fe_cond_neg(e.y, neg1);
        jac26_from_aff(acc, e);
    }
    bool pending = false;
    int step = -1;   // step -1 adds phi(P') for the top digit of k2
    while (step < nsteps) {
        const bool fixup = __any(pending);
        bool is_dbl = fixup;
        int r = 0;
        if (!fixup && step >= 0 && step < kVarSteps) {
            r = step % (kPW + 2);
            is_dbl = r < kPW;
        }
        if (is_dbl) {
            jac26 t;
            jac26_dbl(t, acc);
            if (fixup) {
                if (pending) acc = t;
                pending = false;
                continue;
            }
            acc = t;
        } else if (step == kVarSteps + 2) {
            fe zs;
            load_fe(zs, &slot[8].x);
            fe26 zs26;
            fe26_from_fe(zs26, zs);
            fe26_mul(acc.z, acc.z, zs26);
        } else {
            const aff* bp;
            u32 fl;
            bool keep = true;
            if (step < 0) {
                bp = slot;
                fl = f2;
            } else if (step < kVarSteps) {
                const int w = kPWindows - 1 - step / (kPW + 2);
                const bool second = r != kPW;
                const u32 nib = (dig[((second ? 4 : 0) + (w >> 3)) * kDigitLanes] >> (4 * (w & 7))) & 15u;
                const u32 f = second ? f2 : f1;
                fl = (nib & 8u) ? (f ^ kAddNeg) : f;
                bp = slot + (nib & 7u);
            } else if (step < kVarSteps + 2) {
                const bool first = step == kVarSteps;
                bp = slot;
                fl = (first ? f1 : f2) ^ kAddNeg;
                keep = first ? fix1 : fix2;
            } else if (step == kVarSteps + 3) {
                bp = G.tab + (size_t)G.nwin * half;
                fl = fg;
            } else {
                const int i = step - (kVarSteps + 4);
                // odd_digit on the stored comb scalar
                const u32 win = lds_bits_at(dig, 8, G.w * i, G.w) | 1u;
                const u32 nb = lds_bits_at(dig, 8, G.w * (i + 1), 1);
                const u32 mag = nb ? win : ((1u << G.w) - win);
                bp = G.tab + (size_t)i * half + ((mag - 1u) >> 1);
                fl = nb ? fg : (fg ^ kAddNeg);
            }
            jac26 t;
            bool nd;
            jac26_madd_mem(t, acc, bp, fl, nd);
            if (keep) {
                acc = t;
                pending = nd;
            }
        }
        ++step;
    }
    jac26_to_jac(out, acc);
}

// u*G from the fixed-base comb alone (any u < n, u = 0 gives infinity): the
// odd representative of u (n - u with the sign folded into the entries when u
// is even), the top entry, then one entry per window.  A mixed addition that
// meets a == b doubles instead.
UFS_DEV void ecmult_gen(jac& out, const gcomb& G, const sc& u) {
    const u32 N[8] = {UFS_N_LIMBS};
    sc ku;
    const bool ueven = (u.v[0] & 1u) == 0;
    {
        u32 bw = 0, t[8];
#pragma unroll
        for (int i = 0; i < 8; ++i) t[i] = subc(N[i], u.v[i], bw, bw);
#pragma unroll
        for (int i = 0; i < 8; ++i) ku.v[i] = ueven ? t[i] : u.v[i];
    }
    const u32 fg = ueven ? kAddNeg : 0u;
    const int half = 1 << (G.w - 1);
    jac26 acc;
    {
        aff e;
        load_aff(e, G.tab + (size_t)G.nwin * half);
        fe_cond_neg(e.y, ueven);
        jac26_from_aff(acc, e);
    }
    for (int i = 0; i < G.nwin; ++i) {
        u32 idx;
        bool ng;
        odd_digit(ku.v, i, G.w, idx, ng);
        jac26 t;
        bool nd;
        jac26_madd_mem(t, acc, G.tab + (size_t)i * half + idx, ng ? (fg ^ kAddNeg) : fg, nd);
        if (__any(nd)) {
            jac26 d;
            jac26_dbl(d, acc);
            if (nd) t = d;
        }
        acc = t;
    }
    jac26_to_jac(out, acc);
}

}  // namespace ufs
