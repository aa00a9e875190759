// Wave-cooperative 10 x 26-bit field arithmetic for single-chain kernels
// (the Horner combination at the end of the MSM, gfx950).
//
// One accumulator's chain of doublings runs on a single wave with every other
// wave of the GPU idle, so its time is the instruction count of one lane.
// Here a field element is spread over the lanes instead: the wave is cut into
// three groups of 21 lanes, lane t of a group holds limb t (t < 10) and lanes
// 10..20 hold zero.  A product a*b puts column t of the schoolbook product in
// lane t: a is shifted one lane per step (DPP wave_shr:1, zero fill) and
// multiplied by b's limb j, broadcast within the group (ds_bpermute), so the
// ten steps give every lane its column at once.  Carries move one lane per
// round, the bits at or above 2^260 fold back through 2^260 = R1 2^26 + R0
// (secp_field26.h) by two cross-lane reads, and the bits above 2^256 in limb 9
// fold through 2^256 = 2^32 + 977.  The three groups run three different
// products at once, which is the parallelism of the doubling formula.
//
// Additions and subtractions are one lane-wise instruction (K p in each lane's
// limb), so a doubling issues about a third of the instructions of the
// one-lane form.  Bounds match secp_field26.h: the product accepts inputs of
// magnitude up to 20 (column sums below 2^64) and returns limbs 0..8 below
// 2^26 + 2^10 and limb 9 at most 2^22; wnorm returns limbs 0..8 at most 2^26
// and limb 9 at most 2^22 + 33.  Those stay within the output bounds the
// lazy-reduction constants of the doubling were derived for
// (tests/test_field26_bounds.py: multiply outputs with every limb below
// 2^26 + 2^22 and limb 9 below 2^22 + 64).
//
// Every lane of the wave must execute these functions (cross-lane reads).
#pragma once
#include "secp_point26.h"

namespace ufs {

constexpr u32 kWideGroup = 21;

struct Wide {
    u32 lane;   // 0..63
    u32 g;      // group (0..2; lane 63 is a fourth, unused group)
    u32 t;      // position in the group: limb / column index
    u32 base;   // first lane of the group
    u32 p;      // limb t of p (0 for t >= 10)
};

UFS_DEV Wide wide_lane() {
    Wide W;
    W.lane = threadIdx.x & 63u;
    W.g = W.lane / kWideGroup;
    W.t = W.lane - W.g * kWideGroup;
    W.base = W.g * kWideGroup;
    W.p = W.t == 0 ? kP26[0] : (W.t == 1 ? kP26[1] : (W.t < 9 ? kM26 : (W.t == 9 ? kM22 : 0u)));
    return W;
}

// v of lane - 1 (0 into lane 0): the whole wave shifts right by one lane.
UFS_DEV u32 w_shr1(u32 v) { return (u32)__builtin_amdgcn_update_dpp(0, (int)v, 0x138, 0xF, 0xF, false); }
UFS_DEV u64 w_shr1(u64 v) { return ((u64)w_shr1((u32)(v >> 32)) << 32) | w_shr1((u32)v); }
// v of lane `src` (0..63)
UFS_DEV u32 w_at(u32 v, u32 src) { return (u32)__builtin_amdgcn_ds_bpermute((int)(src << 2), (int)v); }
UFS_DEV u64 w_at(u64 v, u32 src) { return ((u64)w_at((u32)(v >> 32), src) << 32) | w_at((u32)v, src); }

UFS_DEV u32 wadd(u32 a, u32 b) { return a + b; }
template <u32 K>
UFS_DEV u32 wsub(const Wide& W, u32 a, u32 b) { return a + (K * W.p - b); }

// a * b mod p (weakly reduced), every group with its own operands.
UFS_DEV u32 wmul(const Wide& W, u32 a, u32 b) {
    u32 bj[10];
#pragma unroll
    for (int j = 0; j < 10; ++j) bj[j] = w_at(b, W.base + (u32)j);
    // column t: sum_j a[t - j] b[j]; lanes 10..20 of a are zero, so shifting
    // by up to 9 never carries one group's limbs into the next group
    u64 acc = 0;
    u32 as = a;
#pragma unroll
    for (int j = 0; j < 10; ++j) {
        acc += (u64)as * bj[j];
        if (j < 9) as = w_shr1(as);
    }
    // two carry rounds: columns below 2^62 -> lanes 0..20 below 2^26 + 2^11
#pragma unroll
    for (int r = 0; r < 2; ++r) {
        u64 in = w_shr1(acc >> 26);
        if (W.t == 0) in = 0;
        acc = (acc & kM26) + in;
    }
    const u32 L = (u32)acc;
    // fold lanes 10..20: position s + 10 adds R0 to s and R1 to s + 1;
    // positions 19 and 20 land on 10 and 11 and fold once more
    const u32 l10 = w_at(L, W.lane + 10), l9 = w_at(L, W.lane + 9);
    const u32 L19 = w_at(L, W.base + 19), L20 = w_at(L, W.base + 20);
    const u64 V = (u64)kR1 * L19 + (u64)kR0 * L20;   // at position 10
    const u64 V11 = (u64)kR1 * L20;                  // at position 11
    u64 m = 0;
    if (W.t < 10) {
        m = (u64)L + (u64)kR0 * l10;
        if (W.t >= 1) m += (u64)kR1 * l9;
        if (W.t == 0) m += (u64)kR0 * V;
        if (W.t == 1) m += (u64)kR1 * V + (u64)kR0 * V11;
        if (W.t == 2) m += (u64)kR1 * V11;
    }
    // two carry rounds on lanes 0..9 (below 2^51): limb 9 keeps 22 bits and
    // its excess folds as x (2^32 + 977) into lanes 0 and 1
#pragma unroll
    for (int r = 0; r < 2; ++r) {
        const u32 sh = W.t == 9 ? 22u : 26u;
        const u64 hi = m >> sh;
        u64 in = w_shr1(hi);
        if (W.t == 0 || W.t >= 10) in = 0;
        const u64 h9 = w_at(hi, W.base + 9);
        u64 wrap = 0;
        if (W.t == 0) wrap = h9 * 977u;
        if (W.t == 1) wrap = h9 << 6;
        m = (W.t < 10 ? (m & (((u64)1 << sh) - 1)) : 0) + in + wrap;
    }
    return (u32)m;
}

// Weak normalisation (fe26_norm): input limbs below 2^31; limb 9's bits at or
// above 2^22 fold as x (2^32 + 977), then two carry rounds.
UFS_DEV u32 wnorm(const Wide& W, u32 a) {
    const u32 x = w_at(a >> 22, W.base + 9);
    u32 v = a;
    if (W.t == 9) v &= kM22;
    if (W.t == 0) v += x * 977u;
    if (W.t == 1) v += x << 6;
#pragma unroll
    for (int r = 0; r < 2; ++r) {
        u32 in = w_shr1(W.t == 9 ? 0u : v >> 26);
        if (W.t == 0 || W.t >= 10) in = 0;
        v = (W.t == 9 ? v : (v & kM26)) + in;
    }
    return v;
}

// 2(X, Y, Z) by dbl-2009-l, with the operations and subtraction constants
// of jac26_dbl_lanes; the three products of each of the first two levels
// run in groups 0, 1, 2.
UFS_DEV void wdbl(const Wide& W, u32& X, u32& Y, u32& Z) {
    const u32 a1 = W.g == 0 ? X : Y;
    const u32 b1 = W.g == 0 ? X : (W.g == 1 ? Y : Z);
    const u32 r1 = wmul(W, a1, b1);                          // A | B | YZ
    const u32 A = w_at(r1, W.t), B = w_at(r1, kWideGroup + W.t), YZ = w_at(r1, 2 * kWideGroup + W.t);
    const u32 E = (A << 1) + A;                              // 3A
    const u32 s2 = W.g == 0 ? E : (W.g == 1 ? B : X + B);
    const u32 r2 = wmul(W, s2, s2);                          // F | C | (X+B)^2
    const u32 F = w_at(r2, W.t), C = w_at(r2, kWideGroup + W.t), T = w_at(r2, 2 * kWideGroup + W.t);
    u32 t = wsub<2>(W, T, A);
    t = wsub<2>(W, t, C);
    const u32 D = wnorm(W, t + t);                           // D = 2((X+B)^2 - A - C)
    const u32 X3 = wsub<3>(W, F, D + D);                     // X3 = F - 2D
    const u32 ey = wmul(W, E, wsub<5>(W, D, X3));
    const u32 Y3 = wsub<9>(W, ey, C << 3);                   // Y3 = E(D - X3) - 8C
    X = X3;
    Y = Y3;
    Z = YZ + YZ;                                             // Z3 = 2YZ
}

// a (the same value in every lane) -> wide form.
UFS_DEV u32 wide_from_fe26(const Wide& W, const fe26& a) {
    u32 v = 0;
#pragma unroll
    for (int i = 0; i < 10; ++i) v = W.t == (u32)i ? a.v[i] : v;
    return v;
}

// wide form (group 0) -> the same fe26 in every lane.
UFS_DEV void fe26_from_wide(fe26& r, u32 v) {
#pragma unroll
    for (int i = 0; i < 10; ++i) r.v[i] = (u32)__builtin_amdgcn_readlane((int)v, i);
}

// a = 2^n a for an accumulator held identically by every lane of the wave
// (infinity stays infinity: Z3 = 2YZ = 0).
UFS_DEV void jac26_dbl_wide_n(jac26& a, u32 n) {
    const Wide W = wide_lane();
    u32 X = wide_from_fe26(W, a.x), Y = wide_from_fe26(W, a.y), Z = wide_from_fe26(W, a.z);
    for (u32 d = 0; d < n; ++d) wdbl(W, X, Y, Z);
    fe26_from_wide(a.x, X);
    fe26_from_wide(a.y, Y);
    fe26_from_wide(a.z, Z);
}

}  // namespace ufs
