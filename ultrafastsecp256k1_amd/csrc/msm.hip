// Pippenger MSM kernels and host engine.  See msm.hpp for the pipeline.
#include <algorithm>
#include <cstring>
#include <mutex>

#include <hipcub/hipcub.hpp>

#include "msm.hpp"
#include "secp_ecmult.h"
#include "secp_wide26.h"

namespace ufs {

constexpr int kMsmBlock = 256;
constexpr int kReduceBlock = 256;
constexpr uint32_t kSegChunkMax = 64;   // entries per lane in the first accumulation stage: 32 or 64
constexpr uint32_t kLevelChunk = 16;  // slots per lane in the levels after the joins
// a level writes two slots per chunk: chunks of four or more slots shrink
// every slot count above one chunk (with three, four slots stay four)
static_assert(kLevelChunk >= 4, "level chunks of at least four slots");

UFS_DEV void msm_load_scalar(u32 k[8], const u8* p) {
    const uint4* q = reinterpret_cast<const uint4*>(p);
    const uint4 a = q[0], b = q[1];
    const u32 w[8] = {a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w};
#pragma unroll
    for (int i = 0; i < 8; ++i) k[i] = __builtin_bswap32(w[7 - i]);
}

UFS_DEV u32 msm_bits(const u32 k[8], int pos, int w) {
    const int wi = pos >> 5, sh = pos & 31;
    u32 lo = 0, hi = 0;
#pragma unroll
    for (int i = 0; i < 8; ++i) {
        lo = (wi == i) ? k[i] : lo;
        hi = (wi + 1 == i) ? k[i] : hi;
    }
    const u64 v = ((u64)hi << 32) | lo;
    return (u32)(v >> sh) & ((1u << w) - 1u);
}

// Digit of window w given the running carry.  Windows below the top one
// are signed, |d| <= 2^(c-1), carrying into the next; the top window (W - 1,
// tb = 128 - c(W-1) bits of a GLV half) takes its bits plus the carry unsigned, a value in
// [0, 2^tb], so no extra carry window is needed (a carry window would put
// about half of all terms into a single bucket).
UFS_DEV int msm_digit(const u32 k[8], int w, int c, int W, u32& carry) {
    const u32 raw = msm_bits(k, w * c, c) + carry;
    if (w == W - 1) {
        carry = 0;
        return (int)raw;
    }
    const u32 half = 1u << (c - 1);
    if (raw > half) {
        carry = 1;
        return (int)raw - (1 << c);
    }
    carry = 0;
    return (int)raw;
}

// Scalars are split by GLV (k = k1 + lambda k2, |k1|, |k2| < 2^128): every
// term becomes two of 128 bits, on P and phi(P) = (beta x, y) = lambda P.
constexpr int kScalarBits = 128;

// Bases: one term per lane, a grid of at most kPointBlocksPerCu blocks per CU
// striding over the terms.
constexpr int kPointBlocksPerCu = 16;
// Scalars: one term per lane, likewise.
constexpr int kEntryBlocksPerCu = 16;

// A base point given as 64 bytes x || y (big-endian affine coordinates):
// valid when x, y < p and y^2 == x^3 + 7 (two squarings and a
// multiplication instead of a square root).
UFS_DEV bool load_affine64(aff& P, const u8* pp) {
    u32 x[8], y[8];
#pragma unroll
    for (int l = 0; l < 8; ++l) {
        const u8* q = pp + 28 - 4 * l;
        x[l] = ((u32)q[0] << 24) | ((u32)q[1] << 16) | ((u32)q[2] << 8) | (u32)q[3];
        q += 32;
        y[l] = ((u32)q[0] << 24) | ((u32)q[1] << 16) | ((u32)q[2] << 8) | (u32)q[3];
    }
    const bool lt = u256_lt_p(x) & u256_lt_p(y);
#pragma unroll
    for (int l = 0; l < 8; ++l) {
        P.x.v[l] = x[l];
        P.y.v[l] = y[l];
    }
    fe rhs, lhs, seven;
    fe_sqr(rhs, P.x);
    fe_mul(rhs, rhs, P.x);
    fe_set_u32(seven, 7);
    fe_add(rhs, rhs, seven);
    fe_sqr(lhs, P.y);
    return lt & fe_equal(lhs, rhs);
}

// Bases of the MSM: every point is parsed (strict SEC1 compressed, or
// AFFINE: 64-byte x || y checked on the curve) and stored as P at pts[2i] and
// phi(P) = (beta x, y) = lambda P at pts[2i + 1], the bases of the two GLV
// halves of term i.  The halves' signs are applied by their bucket entries
// (k_msm_entries), so this kernel does not read the scalars: it runs on the
// engine's second stream while the scalars are split and sorted.  A failing
// point reports 2i + 1 through atomicMin (scalars report 2i), so the smallest
// failing term decides the error.  Its entries stay in the buckets with the
// unchecked coordinates, which only costs arithmetic: the call then returns
// the error and discards the sums.
template <bool AFFINE>
__global__ void __launch_bounds__(kMsmBlock)
k_msm_points(const u8* __restrict__ points, u32 n, aff* __restrict__ pts, u32* __restrict__ err) {
    for (u32 i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        aff P;
        bool pok;
        if (AFFINE) {
            pok = load_affine64(P, points + 64ull * i);
        } else {
            const u8* pp = points + 33ull * i;
            const u32 prefix = pp[0];
            u32 x[8];
#pragma unroll
            for (int l = 0; l < 8; ++l) {
                const u8* q = pp + 1 + 28 - 4 * l;
                x[l] = ((u32)q[0] << 24) | ((u32)q[1] << 16) | ((u32)q[2] << 8) | (u32)q[3];
            }
            pok = lift_x(P, x, prefix & 1u);
            pok &= (prefix == 2u) | (prefix == 3u);
        }
        if (!pok) atomicMin(err, 2u * i + 1u);
        aff Q;
        const fe BETA = {{UFS_BETA_LIMBS}};
        fe_mul(Q.x, P.x, BETA);
        Q.y = P.y;
        pts[2 * i] = P;
        pts[2 * i + 1] = Q;
    }
}

// Entry: bucket id in the high word, term index | sign bit (31) | skip bit
// (30) in the low word, one 8-byte store.  Term indices stay below 2^29 (the
// enqueue bounds 2N * W), so the flag bits never meet them.
constexpr u32 kEntrySign = 0x80000000u;
constexpr u32 kEntrySkip = 0x40000000u;
constexpr u32 kEntryTerm = 0x3FFFFFFFu;

UFS_DEV u64 msm_entry(u32 bucket, u32 term, bool neg, bool skip) {
    return ((u64)bucket << 32) | (u64)(term | (neg ? kEntrySign : 0u) | (skip ? kEntrySkip : 0u));
}

// The W entries of half-term i (value v below 2^128, sign flip hneg): one
// per window, window-major, zero digits as skip entries.
UFS_DEV void msm_half_entries(const u32 v[8], bool hneg, u32 i, u32 n2, u64* __restrict__ entries, int c, int W,
                              u32* __restrict__ range_err) {
    const u32 B = 1u << (c - 1);
    u32 carry = 0;
    for (int w = 0; w < W; ++w) {
        const int d = msm_digit(v, w, c, W, carry);
        const u32 mag = (u32)(d < 0 ? -d : d);
        const u32 lim = w == W - 1 ? 2u * B : B;
        const bool out = mag > lim;
        if (out) atomicOr(range_err, 1u);
        const bool skip = d == 0 || out;
        const u32 b = skip ? (i & (B - 1u)) : mag - 1u;
        entries[(size_t)w * n2 + i] = msm_entry((u32)w * B + b, i, (d < 0) != hneg, skip);
    }
}

// Scalars -> unsorted bucket entries, one term per lane: the scalar is
// checked (< n, else 2i into err[0] through atomicMin and no entries but
// skips), split by GLV into k = +-k1 +- lambda k2 with halves below 2^128, and
// each half's signed c-bit digits become its entries: half-term 2i on P,
// 2i + 1 on phi(P), an entry's sign the digit's sign flipped by its half's.
//
// Entries are window-major (window w's entries at [w 2n, (w + 1) 2n)).  A
// zero digit becomes a skip entry in one of its own window's buckets (chosen
// by the half-term index, so skips spread evenly and never form long runs),
// which the accumulation passes over.
//
// Order: the sort keys on the low c bits of the bucket id only, which are
// distinct for distinct buckets of one window ((w & 1) * B + b below the top
// window; the top window's 2B ids are consecutive).  The sort is stable and
// the input is window-major, so every bucket's entries still end up
// contiguous -- buckets of different windows interleave, which nothing
// downstream depends on -- and c bits (two 8-bit passes at c = 16) replace the
// log2 of the bucket count (three passes).
//
// Bound: window w < W - 1 owns ids [wB, (w + 1)B), the top window [(W - 1)B,
// (W + 1)B).  The halves are cut to 128 bits (a scalar >= n would split into
// halves of any size) and are below 2^128 for every scalar below n, so every
// digit fits (checked on the CPU by replaying this arithmetic for any input);
// a digit that would not becomes a skip entry and raises err[1], which the
// final kernel reports as an internal error, so no later kernel indexes
// outside bsum.
__global__ void __launch_bounds__(kMsmBlock)
k_msm_entries(const u8* __restrict__ scalars32, u32 n, u64* __restrict__ entries, int c, int W,
              u32* __restrict__ err) {
    const u32 n2 = 2 * n;
    for (u32 t = blockIdx.x * blockDim.x + threadIdx.x; t < n; t += gridDim.x * blockDim.x) {
        u32 k[8];
        msm_load_scalar(k, scalars32 + 32ull * t);
        const bool sok = u256_lt_n(k);
        if (!sok) atomicMin(err, 2u * t);
        sc s, k1, k2;
        bool n1, n2h;
#pragma unroll
        for (int l = 0; l < 8; ++l) s.v[l] = k[l];
        glv_split(k1, k2, n1, n2h, s);
#pragma unroll
        for (int l = 4; l < 8; ++l) {
            k1.v[l] = 0;
            k2.v[l] = 0;
        }
        if (!sok) {
            sc_set_u32(k1, 0);
            sc_set_u32(k2, 0);
        }
        msm_half_entries(k1.v, n1, 2 * t, n2, entries, c, W, err + 1);
        msm_half_entries(k2.v, n2h, 2 * t + 1, n2, entries, c, W, err + 1);
    }
}

constexpr u32 kEmpty = 0xFFFFFFFFu;
// Level-0 tail slot of a chunk that lies inside one bucket's run (its head
// partial continues into the next chunk); read as an empty slot by k_level.
constexpr u32 kCont = 0xFFFFFFFEu;

// A bucket sum or partial as the accumulators hold it: 10 x 26-bit limbs, not
// normalised (x, y, z, then the infinity flag), 128 bytes.  Storing it takes
// eight 16-byte stores.  Normalising to 8 x 32 limbs at every bucket end
// would cost about as much as a quarter of a mixed addition, and in the
// first stage some lane of a wave ends a bucket at most steps, so the whole
// wave would pay it nearly every step.  Every reader adds these values with
// jac26_add, whose limb bounds allow any output of the doubling, the mixed
// and the general addition as either operand.
struct __align__(16) jac26s { u32 w[32]; };

UFS_DEV void store_j26(jac26s* p, const jac26& a) {
    uint4* q = reinterpret_cast<uint4*>(p->w);
    q[0] = make_uint4(a.x.v[0], a.x.v[1], a.x.v[2], a.x.v[3]);
    q[1] = make_uint4(a.x.v[4], a.x.v[5], a.x.v[6], a.x.v[7]);
    q[2] = make_uint4(a.x.v[8], a.x.v[9], a.y.v[0], a.y.v[1]);
    q[3] = make_uint4(a.y.v[2], a.y.v[3], a.y.v[4], a.y.v[5]);
    q[4] = make_uint4(a.y.v[6], a.y.v[7], a.y.v[8], a.y.v[9]);
    q[5] = make_uint4(a.z.v[0], a.z.v[1], a.z.v[2], a.z.v[3]);
    q[6] = make_uint4(a.z.v[4], a.z.v[5], a.z.v[6], a.z.v[7]);
    q[7] = make_uint4(a.z.v[8], a.z.v[9], a.inf, 0u);
}

UFS_DEV void load_j26(jac26& a, const jac26s* p) {
    const uint4* q = reinterpret_cast<const uint4*>(p->w);
    u32 w[32];
#pragma unroll
    for (int i = 0; i < 8; ++i) {
        const uint4 v = q[i];
        w[4 * i] = v.x;
        w[4 * i + 1] = v.y;
        w[4 * i + 2] = v.z;
        w[4 * i + 3] = v.w;
    }
#pragma unroll
    for (int i = 0; i < 10; ++i) {
        a.x.v[i] = w[i];
        a.y.v[i] = w[10 + i];
        a.z.v[i] = w[20 + i];
    }
    a.inf = w[30];
}

__global__ void __launch_bounds__(kMsmBlock) k_j26_fill_inf(jac26s* __restrict__ p, u32 n) {
    const u32 i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    jac26 v;
    jac26_set_inf(v);
    store_j26(p + i, v);
}

// Bucket accumulation, level 0: the bucket-sorted entries are cut into chunks
// of K; lane j sums chunk j run by run.  A bucket that starts and ends inside
// the chunk is written to bsum directly; the chunk's first bucket, if it began
// in an earlier chunk, becomes the chunk's head partial, and its last bucket,
// if it continues into the next chunk, its tail partial.  Partials are merged
// by the higher levels, so every lane does at most K additions whatever the
// bucket sizes (the top window's buckets can hold half of all terms).
__global__ void __launch_bounds__(kMsmBlock)
k_seg0(const aff* __restrict__ pts, const u64* __restrict__ entries, u32 E, u32 K, u32 max_chunks,
       jac26s* __restrict__ bsum, u32* __restrict__ okeys, jac26s* __restrict__ opart) {
    const u32 j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= max_chunks) return;
    const u32 beg = j * K;
    jac26 hv, tv;
    jac26_set_inf(hv);
    jac26_set_inf(tv);
    u32 hk = kEmpty, tk = kEmpty;
    if (beg < E) {
        const u32 end = min(beg + K, E);
        u64 ent = entries[beg];
        const bool before = beg > 0 && (u32)(entries[beg - 1] >> 32) == (u32)(ent >> 32);
        const bool after = end < E && (u32)(entries[end] >> 32) == (u32)(entries[end - 1] >> 32);
        u32 cur = (u32)(ent >> 32);
        bool first = true;
        // the running sum on 10 x 26-bit limbs, stored as it is when a bucket ends
        jac26 acc;
        jac26_set_inf(acc);
        aff nxt = pts[(u32)ent & kEntryTerm];
        for (u32 e = beg; e < end; ++e) {
            const u32 k = (u32)(ent >> 32);
            aff P = nxt;
            const bool neg = ((u32)ent & kEntrySign) != 0;
            const bool skip = ((u32)ent & kEntrySkip) != 0;
            if (e + 1 < end) {
                ent = entries[e + 1];
                nxt = pts[(u32)ent & kEntryTerm];
            }
            if (k != cur) {
                if (first && before) { hk = cur; hv = acc; } else { store_j26(bsum + cur, acc); }
                first = false;
                cur = k;
                jac26_set_inf(acc);
            }
            // inline mixed addition; a == P (possible for repeated points)
            // doubles instead; a skip entry (zero digit) leaves acc as it is
            jac26 t;
            bool nd;
            jac26_madd(t, acc, P, neg ? kAddNeg : 0u, nd);
            nd &= !skip;
            if (__any(nd)) {
                jac26 dd;
                jac26_dbl(dd, acc);
                if (nd) t = dd;
            }
            if (__any(skip)) {
                if (skip) t = acc;   // rare: only zero digits
            }
            acc = t;
        }
        if (first && before) {
            hk = cur;
            hv = acc;
            if (after) tk = kCont;   // the whole chunk is inside one run
        } else if (after) {
            tk = cur;
            tv = acc;
        } else {
            store_j26(bsum + cur, acc);
        }
    }
    okeys[2 * j] = hk;
    store_j26(opart + 2 * j, hv);
    okeys[2 * j + 1] = tk;
    store_j26(opart + 2 * j + 1, tv);
}

// Join of level-0 partials: a bucket that continues past its first chunk j
// left a tail partial in slot 2j + 1 and a head partial in each following
// chunk it reaches (slot 2(j + m)); the chunks strictly inside its run mark
// their tail slot kCont.  Lane j sums such a chain when it spans at most
// kJoinSpan chunks after j, writes the bucket and empties the chain's slots;
// longer chains (buckets holding a large share of the terms) stay for the
// levels.  Each head belongs to one chain, so no slot is written by two
// lanes.  The one slot two lanes touch is the tail of a chain's last chunk:
// the chain's lane reads it while that chunk's own lane may empty it (its own
// chain joined).  Both accesses are relaxed atomics, and either value reads as
// "not kCont", so the result does not depend on the order.
constexpr u32 kJoinSpan = 8;

__global__ void __launch_bounds__(kMsmBlock)
k_join(u32* __restrict__ keys, const jac26s* __restrict__ part, u32 chunks, jac26s* __restrict__ bsum) {
    const u32 j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= chunks) return;
    const u32 tk = __hip_atomic_load(&keys[2 * j + 1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (tk >= kCont) return;
    u32 m = 1;
    for (; m <= kJoinSpan && j + m < chunks; ++m) {
        if (keys[2 * (j + m)] != tk) return;   // not a chain of this bucket: leave it
        if (__hip_atomic_load(&keys[2 * (j + m) + 1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != kCont) break;
    }
    if (m > kJoinSpan || j + m >= chunks) return;
    jac26 acc, b;
    load_j26(acc, part + 2 * j + 1);
    for (u32 i = 1; i <= m; ++i) {
        load_j26(b, part + 2 * (j + i));
        jac26_add(acc, acc, b);
    }
    store_j26(bsum + tk, acc);
    __hip_atomic_store(&keys[2 * j + 1], kEmpty, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    for (u32 i = 1; i <= m; ++i) keys[2 * (j + i)] = kEmpty;
}

// Higher levels: the same chunked run reduction over the previous level's
// slots, read in place (no compaction; kCont and kEmpty are both empty).  Slot pairs (head, tail) of
// consecutive chunks keep bucket order; a slot is empty (kEmpty) when its
// chunk had no partial there.  Within the partials of one bucket two
// consecutive ones are at most one empty slot apart: a chunk that lies inside
// a bucket's run emits its head only, a chunk where the run starts emits a
// tail, one where it ends a head.  A chunk therefore finds the neighbour of
// its first and last partial at most two slots away, and a level costs one
// launch whatever is left; with nothing left every lane only reads its keys.
__global__ void __launch_bounds__(kMsmBlock)
k_level(const u32* __restrict__ ikeys, const jac26s* __restrict__ ipart, u32 n, u32 K, u32 max_chunks,
        jac26s* __restrict__ bsum, u32* __restrict__ okeys, jac26s* __restrict__ opart) {
    const u32 j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= max_chunks) return;
    const u32 beg = j * K;
    const u32 end = min(beg + K, n);
    u32 f = end, l = end;   // first and last non-empty slot of the chunk
    for (u32 e = beg; e < end; ++e)
        if (ikeys[e] < kCont) {
            if (f == end) f = e;
            l = e;
        }
    u32 hk = kEmpty, tk = kEmpty;
    jac26 hv, tv;
    jac26_set_inf(hv);
    jac26_set_inf(tv);
    if (f < end) {
        const u32 kf = ikeys[f], kl = ikeys[l];
        // the previous partial of kf's bucket, if any, is at f - 1 or f - 2
        bool before = false;
        if (f >= 1) {
            u32 q = f - 1;
            if (ikeys[q] >= kCont && q >= 1) --q;
            before = ikeys[q] == kf;
        }
        bool after = false;
        if (l + 1 < n) {
            u32 q = l + 1;
            if (ikeys[q] >= kCont && q + 1 < n) ++q;
            after = ikeys[q] == kl;
        }
        u32 cur = kf;
        bool first = true;
        jac26 acc;
        jac26_set_inf(acc);
        for (u32 e = f; e <= l; ++e) {
            const u32 k = ikeys[e];
            if (k >= kCont) continue;
            if (k != cur) {
                if (first && before) { hk = cur; hv = acc; } else { store_j26(bsum + cur, acc); }
                first = false;
                cur = k;
                jac26_set_inf(acc);
            }
            jac26 b;
            load_j26(b, ipart + e);
            jac26_add(acc, acc, b);
        }
        if (first && before) { hk = cur; hv = acc; }
        else if (after) { tk = cur; tv = acc; }
        else { store_j26(bsum + cur, acc); }
    }
    okeys[2 * j] = hk;
    store_j26(opart + 2 * j, hv);
    okeys[2 * j + 1] = tk;
    store_j26(opart + 2 * j + 1, tv);
}

// r = m * a for a small non-negative integer m (double-and-add).
UFS_DEV void jac26_mul_small(jac26& r, const jac26& a, u32 m) {
    jac26 acc;
    jac26_set_inf(acc);
    if (m != 0) {
        const int top = 31 - __builtin_clz(m);
        acc = a;
        for (int b = top - 1; b >= 0; --b) {
            jac26_dbl(acc, acc);
            if ((m >> b) & 1u) jac26_add(acc, acc, a);
        }
    }
    r = acc;
}

// Lane (w, s): P = sum_{j<L} (s*L + j + 1) * B[w][s*L + j] for the groups
// w0 .. w0 + Gn - 1; group gtop holds magnitudes B + 1 .. 2B of the top
// window.  The lanes are few (under one wave per SIMD) and each runs a chain
// of dependent point additions, so the time follows the instruction count:
// 10 x 26-bit limbs.  These waves raise their issue priority: when the high
// windows' sums run beside the low windows' bucket accumulation (a second
// stream), the short chains go first on a shared SIMD.
__global__ void __launch_bounds__(kMsmBlock)
k_msm_segments(const jac26s* __restrict__ bsum, jac* __restrict__ seg, u32 w0, u32 Gn, u32 gtop, u32 B, u32 L) {
    __builtin_amdgcn_s_setprio(3);
    const u32 T = B / L;
    const u32 id = blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= Gn * T) return;
    const u32 w = w0 + id / T, s = id % T;
    const u32 off = s * L + (w == gtop ? B : 0u);
    const jac26s* bb = bsum + (size_t)w * B + (size_t)s * L;
    jac26 R, S;
    jac26_set_inf(R);
    jac26_set_inf(S);
    for (int j = (int)L - 1; j >= 0; --j) {
        jac26 Bj;
        load_j26(Bj, bb + j);
        jac26_add(R, R, Bj);
        jac26_add(S, S, R);
    }
    jac26 M;
    jac26_mul_small(M, R, off);
    jac26_add(S, S, M);
    jac out;
    jac26_to_jac(out, S);
    seg[id] = out;
}

// Block tree sum of `len` points per group into out[g * out_len + blockIdx.x].
__global__ void __launch_bounds__(kReduceBlock)
k_msm_reduce(const jac* __restrict__ in, jac* __restrict__ out, u32 len, u32 out_len) {
    __builtin_amdgcn_s_setprio(3);
    __shared__ jac26 sh[kReduceBlock];
    const u32 g = blockIdx.y, q = blockIdx.x, t = threadIdx.x;
    const u32 idx = q * kReduceBlock + t;
    jac26 v;
    if (idx < len) {
        jac26_from_jac(v, in[(size_t)g * len + idx]);
    } else {
        jac26_set_inf(v);
    }
    sh[t] = v;
    __syncthreads();
    // tree depth from this block's live inputs: a block of 16 takes 4 levels
    const u32 live = min((u32)kReduceBlock, len - q * kReduceBlock);
    u32 top = 1;
    while (top < live) top <<= 1;
    for (u32 stride = top / 2; stride > 0; stride >>= 1) {
        if (t < stride) {
            jac26 a = sh[t];
            const jac26 b = sh[t + stride];
            jac26_add(a, a, b);
            sh[t] = a;
        }
        __syncthreads();
    }
    if (t == 0) {
        jac r;
        jac26_to_jac(r, sh[0]);
        out[(size_t)g * out_len + q] = r;
    }
}

// Horner combination of the window sums rows[0 .. nwin - 1] (row w weighs
// 2^(c w)); with `top`, row nwin is the top window's second bucket group and
// joins row nwin - 1.  The sum is then doubled `shift` times and `extra`, if
// given, is added.  With `part` set the kernel writes that Jacobian point
// (the high windows' share, computed beside the low windows' buckets) and
// nothing else; otherwise it checks the error words and writes the
// compressed result and the status.
__global__ void __launch_bounds__(64) k_msm_final(const jac* __restrict__ rows, u32 nwin, u32 top, u32 c, u32 shift,
                            const jac* __restrict__ extra, const u32* __restrict__ err,
                            u8* __restrict__ out33, int32_t* __restrict__ status, jac* __restrict__ part) {
    // one wave: every lane holds the same accumulator; the doublings spread
    // each field product over 21 lanes, three products at a time
    // (jac26_dbl_wide_n), the additions one product per lane
    // (jac26_add_lanes), lane 0 writes
    __builtin_amdgcn_s_setprio(3);
    if (blockIdx.x != 0) return;
    const bool writer = threadIdx.x == 0;
    if (!part) {
        const u32 e = err[0];
        if (e != 0xFFFFFFFFu || err[1] != 0u) {
            if (writer) {
                for (int i = 0; i < 33; ++i) out33[i] = 0;
                *status = e != 0xFFFFFFFFu ? ((e & 1u) ? 3 /* BAD_PUBKEY */ : 5 /* BAD_INPUT */)
                                           : 9 /* INTERNAL: a bucket id out of range */;
            }
            return;
        }
    }
    // The time is the dependent-instruction latency of the doublings: they
    // run on 10 x 26-bit limbs, whose additions and subtractions are
    // independent per limb instead of carry chains.
    jac26 a26, b26;
    jac26_from_jac(a26, rows[nwin - 1]);
    if (top) {
        jac26_from_jac(b26, rows[nwin]);
        jac26_add_lanes(a26, a26, b26);         // the top window's two bucket groups
    }
    for (int w = (int)nwin - 2; w >= 0; --w) {
        jac26_dbl_wide_n(a26, c);
        jac26_from_jac(b26, rows[w]);
        jac26_add_lanes(a26, a26, b26);
    }
    jac26_dbl_wide_n(a26, shift);
    if (extra) {
        jac26_from_jac(b26, extra[0]);
        jac26_add_lanes(a26, a26, b26);
    }
    jac acc;
    jac26_to_jac(acc, a26);
    if (!writer) return;
    if (part) {
        part[0] = acc;
        return;
    }
    if (acc.inf) {
        for (int i = 0; i < 33; ++i) out33[i] = 0;
        *status = 7;   // ERR_ARITH: result is the point at infinity
        return;
    }
    aff A;
    jac_to_aff(A, acc);
    out33[0] = 0x02 | (u8)(A.y.v[0] & 1u);
    u256_to_be(out33 + 1, A.x.v);
    *status = 0;
}

__global__ void k_u32_fill(u32* p, u32 v, size_t n) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// ---------------------------------------------------------------------------
// Launchers
// ---------------------------------------------------------------------------
static inline unsigned blocks_for(size_t n, unsigned b) { return (unsigned)((n + b - 1) / b); }

hipError_t launch_u32_fill(uint32_t* p, uint32_t v, size_t n, hipStream_t st) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_u32_fill, dim3(blocks_for(n, 256)), dim3(256), 0, st, p, v, n);
    return hipGetLastError();
}

hipError_t launch_msm_points(const uint8_t* points, bool affine, uint32_t n, void* pts, uint32_t* err,
                             unsigned cus, hipStream_t st) {
    if (n == 0) return hipSuccess;
    unsigned g = blocks_for(n, kMsmBlock);
    if (g > cus * kPointBlocksPerCu) g = cus * kPointBlocksPerCu;
    if (affine)
        hipLaunchKernelGGL(k_msm_points<true>, dim3(g), dim3(kMsmBlock), 0, st, points, n, (aff*)pts, err);
    else
        hipLaunchKernelGGL(k_msm_points<false>, dim3(g), dim3(kMsmBlock), 0, st, points, n, (aff*)pts, err);
    return hipGetLastError();
}

hipError_t launch_msm_segments(const void* bsum, void* seg, const MsmPlan& p, uint32_t w0, uint32_t gn,
                               hipStream_t st) {
    const uint32_t T = (uint32_t)(p.buckets / p.seg_len);
    const uint32_t lanes = gn * T;
    hipLaunchKernelGGL(k_msm_segments, dim3(blocks_for(lanes, kMsmBlock)), dim3(kMsmBlock), 0, st,
                       (const jac26s*)bsum, (jac*)seg, w0, gn, (uint32_t)p.groups - 1u, (uint32_t)p.buckets,
                       (uint32_t)p.seg_len);
    return hipGetLastError();
}

hipError_t launch_msm_reduce(const void* in, void* out, uint32_t groups, uint32_t len, uint32_t out_len,
                             hipStream_t st) {
    hipLaunchKernelGGL(k_msm_reduce, dim3(out_len, groups), dim3(kReduceBlock), 0, st, (const jac*)in, (jac*)out,
                       len, out_len);
    return hipGetLastError();
}

hipError_t launch_msm_final(const void* rows, uint32_t nwin, bool top, uint32_t c, uint32_t shift, const void* extra,
                            const uint32_t* err, uint8_t* out33, int32_t* status, void* part, hipStream_t st) {
    hipLaunchKernelGGL(k_msm_final, dim3(1), dim3(64), 0, st, (const jac*)rows, nwin, top ? 1u : 0u, c, shift,
                       (const jac*)extra, err, out33, status, (jac*)part);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Host engine
// ---------------------------------------------------------------------------
int msm_default_window(size_t n) {
    int lg = 0;
    while (((size_t)1 << (lg + 1)) <= n) ++lg;
    int c = lg - 4;
    if (c < 4) c = 4;
    if (c > 16) c = 16;
    return c;
}

MsmPlan msm_plan(size_t n, int c, const Tuning* tun) {
    if (c <= 0 && tun) {
        int lg = 0;
        while (((size_t)1 << (lg + 1)) <= n) ++lg;
        if (lg < 40) c = tun->msm_window[lg];
    }
    if (c <= 0) c = msm_default_window(n);
    MsmPlan p;
    p.c = c;
    p.windows = (kScalarBits + c - 1) / c;
    p.groups = p.windows + 1;
    p.buckets = 1 << (c - 1);
    // Segments of 8 buckets, shorter (down to 2) while that leaves fewer than
    // 16384 lanes for the window sums: their chains of point additions are
    // latency-bound, so small bucket counts want more, shorter chains
    // (measured 0.75 -> 0.68 ms at 2^10 terms, 0.78 -> 0.72 at 2^12, 0.88 ->
    // 0.83 at 2^14 and 2^16; 2^18 and up keep 8).
    p.seg_len = p.buckets < 8 ? p.buckets : 8;
    while (p.seg_len > 2 && (size_t)p.groups * p.buckets / p.seg_len < 16384) p.seg_len /= 2;
    return p;
}

// Entries per lane of k_seg0 for E entries on a GPU of `cus` compute units.
// The stage runs two waves per SIMD (512 lanes per CU).  Chunks of 64 halve
// the partials to join but pay only while the grid fills the GPU twice over
// (2^20 terms at c = 16 measured 3.26 -> 3.15 ms; at 2^18 and 2^16 the wider
// chunks were slower).  Narrower chunks than 32 for small E (8 or 16, so
// that one round of resident lanes is filled) were faster at 2^10 and 2^14
// terms but slower at 2^12 and 2^16 (1.05 and 1.46 against 0.78 and 0.88
// ms): more partials for the join and the levels.
static uint32_t msm_seg_chunk(size_t E, size_t cus) {
    return E >= (size_t)kSegChunkMax * 2 * cus * 512 ? kSegChunkMax : kSegChunkMax / 2;
}

// ---------------------------------------------------------------------------
// MSM without allocation: the buffers of one MSM of at most
// kNoAllocMaxTerms terms are carved from `scratch`, device memory that every
// GPU the library runs on holds from module load to process end.  In host
// code `scratch` names a host-side stand-in, never a device address: the
// device address comes from hipGetSymbolAddress, once per device at context
// creation (msm_noalloc_setup), and only that address is used.
// ---------------------------------------------------------------------------
constexpr size_t kNoAllocScratchBytes = size_t(32) << 20;
constexpr size_t kNoAllocMaxTerms = size_t(1) << 14;
static __device__ __align__(256) uint8_t scratch[kNoAllocScratchBytes];

namespace {
constexpr int kMaxScratchDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_dev[kMaxScratchDevices] = {};   // device address of `scratch`, per device
}  // namespace

hipError_t msm_noalloc_setup(int device) {
    if (device < 0 || device >= kMaxScratchDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (g_scratch_dev[device]) return hipSuccess;
    hipError_t e = hipSetDevice(device);
    if (e != hipSuccess) return e;
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    if (!p) return hipErrorInvalidSymbol;
    g_scratch_dev[device] = p;
    return hipSuccess;
}

static void* msm_noalloc_scratch(int device) {
    if (device < 0 || device >= kMaxScratchDevices) return nullptr;
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    return g_scratch_dev[device];
}

GpuError MsmEngine::init() {
    hipError_t e = hipSetDevice(be_->device());
    if (e == hipSuccess && !done_) e = hipEventCreateWithFlags(&done_, hipEventDisableTiming);
    if (e != hipSuccess) return be_->fail(GpuError::Backend, "MSM completion event", e);
    if ((e = msm_noalloc_setup(be_->device())) != hipSuccess)
        return be_->fail(GpuError::Device, "MSM scratch symbol address", e);
    if ((e = ensure_aux()) != hipSuccess) return be_->fail(GpuError::Backend, "MSM second stream", e);
    return GpuError::Ok;
}

// The second stream (default priority) and its events: the bases'
// decompression and the split form's high windows run on it.
hipError_t MsmEngine::ensure_aux() {
    hipError_t e = hipSuccess;
    if (!aux_) e = hipStreamCreateWithFlags(&aux_, hipStreamNonBlocking);
    if (e == hipSuccess && !fork_) e = hipEventCreateWithFlags(&fork_, hipEventDisableTiming);
    if (e == hipSuccess && !join_) e = hipEventCreateWithFlags(&join_, hipEventDisableTiming);
    if (e == hipSuccess && !ready_) e = hipEventCreateWithFlags(&ready_, hipEventDisableTiming);
    return e;
}

GpuError MsmEngine::msm_device_noalloc(const uint8_t* d_scalars32, const uint8_t* d_points33, size_t n,
                                       uint8_t* d_out33, int32_t* d_status, int c, hipStream_t st) {
    if (!be_->is_ready()) return be_->fail(GpuError::Device, "context not initialised");
    if (n > kNoAllocMaxTerms) return be_->fail(GpuError::BadInput, "no-allocation MSM: more than 16384 terms");
    if (c != 0 && (c < 2 || c > 20)) return be_->fail(GpuError::BadInput, "MSM window width out of range");
    uint8_t* base = (uint8_t*)msm_noalloc_scratch(be_->device());
    if (!base || !done_) return be_->fail(GpuError::Device, "no-allocation MSM scratch not set up on this device");
    st = st ? st : be_->stream();
    const MsmPlan p = msm_plan(n, c, &be_->tuning());
    MsmLayout L;
    GpuError r = msm_layout(n, p, st, L);
    if (r != GpuError::Ok) return r;
    // wider first-stage chunks need fewer partial slots: take the narrowest
    // that fits the scratch
    for (uint32_t K = L.seg_chunk * 2; L.total > kNoAllocScratchBytes && K <= kSegChunkMax; K *= 2)
        if ((r = msm_layout(n, p, st, L, K)) != GpuError::Ok) return r;
    if (L.total > kNoAllocScratchBytes)
        return be_->fail(GpuError::BadInput,
                         "no-allocation MSM: this term count and window width need more than the 32 MiB scratch "
                         "(up to 16384 terms with windows of at most 13 bits fit)");
    void* ptr[MsmLayout::kParts];
    for (int i = 0; i < MsmLayout::kParts; ++i) ptr[i] = base + L.off[i];
    hipError_t e = order(st);
    if (e != hipSuccess) return be_->fail(GpuError::Backend, "MSM stream order", e);
    r = run_msm(L, p, ptr, d_scalars32, d_points33, false, n, d_out33, d_status, false, st);
    e = finish(st);
    if (r == GpuError::Ok && e != hipSuccess) return be_->fail(GpuError::Backend, "MSM completion event", e);
    if (r == GpuError::Ok) be_->clear_error();
    return r;
}

hipError_t MsmEngine::order(hipStream_t st) {
    if (!done_) {
        hipError_t e = hipEventCreateWithFlags(&done_, hipEventDisableTiming);
        if (e != hipSuccess) return e;
    }
    if (pending_ && last_stream_ != st) return hipStreamWaitEvent(st, done_, 0);
    return hipSuccess;
}

hipError_t MsmEngine::grow(DeviceBuffer& b, size_t bytes) {
    if (bytes > b.cap && pending_) {
        hipError_t e = hipEventSynchronize(done_);
        if (e != hipSuccess) return e;
        pending_ = false;
    }
    return b.ensure(bytes);
}

hipError_t MsmEngine::finish(hipStream_t st) {
    last_stream_ = st;
    pending_ = true;
    return hipEventRecord(done_, st);
}

MsmEngine::~MsmEngine() {
    if (done_) {
        (void)hipEventSynchronize(done_);
        (void)hipEventDestroy(done_);
    }
    if (aux_) {
        (void)hipStreamSynchronize(aux_);
        (void)hipStreamDestroy(aux_);
    }
    if (fork_) (void)hipEventDestroy(fork_);
    if (join_) (void)hipEventDestroy(join_);
    if (ready_) (void)hipEventDestroy(ready_);
    DeviceBuffer* all[] = {&in_s_, &in_p_, &pts_, &entries_,
                           &bsum_, &seg_, &red_a_, &red_b_, &err_, &out_, &slot_keys_[0],
                           &slot_keys_[1], &slot_part_[0], &slot_part_[1],
                           &rlc_in_, &rlc_terms_, &rlc_part_, &sorted_, &sort_tmp_};
    for (DeviceBuffer* b : all) b->release();
}

GpuError MsmEngine::enqueue(const uint8_t* d_scalars32, const uint8_t* d_points, bool affine, size_t n,
                            uint8_t* d_out33, int32_t* d_status, int c, hipStream_t st) {
    if (c != 0 && (c < 2 || c > 20)) return be_->fail(GpuError::BadInput, "MSM window width out of range");
    hipError_t eo = order(st);
    if (eo != hipSuccess) return be_->fail(GpuError::Backend, "MSM stream order", eo);
    const GpuError r = enqueue_work(d_scalars32, d_points, affine, n, d_out33, d_status, c, st);
    // whatever was enqueued, later calls on other streams wait for it
    eo = finish(st);
    if (r == GpuError::Ok && eo != hipSuccess) return be_->fail(GpuError::Backend, "MSM completion event", eo);
    return r;
}

// Byte layout of one MSM's device buffers for n terms at plan p; the
// hipcub temporary-storage sizes are queried for this n (host only, no
// allocation).  Offsets are 256-byte aligned from a common base.
GpuError MsmEngine::msm_layout(size_t n, const MsmPlan& p, hipStream_t st, MsmLayout& L, uint32_t seg_chunk) {
    L = MsmLayout{};
    L.max_entries = 2 * n * (size_t)p.windows;   // two GLV terms per term
    if (L.max_entries > 0x7FFFFFFFu)
        return be_->fail(GpuError::BadInput, "MSM: too many (term, window) pairs for this window width");
    const size_t cus = (size_t)(be_->compute_units() > 0 ? be_->compute_units() : 256);
    L.seg_chunk = seg_chunk ? seg_chunk : msm_seg_chunk(L.max_entries, cus);
    // every bucket phase covers at most max_entries with this chunk width
    L.chunks0 = (L.max_entries + L.seg_chunk - 1) / L.seg_chunk;
    L.slots0 = 2 * L.chunks0;
    // the levels write two slots per chunk of kLevelChunk: the second slot
    // buffer holds level 1's output, and every later level's fits the buffer
    // it alternates into
    const size_t S1 = 2 * ((L.slots0 + kLevelChunk - 1) / kLevelChunk);
    const size_t nbk = (size_t)p.groups * p.buckets;
    const size_t nseg = (size_t)p.groups * (p.buckets / p.seg_len);
    const size_t jsz = sizeof(jac), bsz = sizeof(jac26s), S = L.slots0;
    // hipcub temporary storage of the bucket sort
    size_t sort_bytes = 0;
    hipError_t e = hipcub::DeviceRadixSort::SortKeys(nullptr, sort_bytes, (const uint64_t*)nullptr,
                                                     (uint64_t*)nullptr, (int)L.max_entries, 32, 32 + p.c, st);
    if (e != hipSuccess) return be_->fail(GpuError::Launch, "MSM temporary-storage size query", e);
    L.sort_bytes = sort_bytes;
    const size_t sizes[MsmLayout::kParts] = {
        2 * n * sizeof(aff), L.max_entries * 8, L.max_entries * 8, nbk * bsz, nseg * jsz,
        nseg * jsz, nseg * jsz, 64 + jsz, S * 4, S1 * 4, S * bsz, S1 * bsz, L.sort_bytes + 16};
    size_t off = 0;
    for (int i = 0; i < MsmLayout::kParts; ++i) {
        L.size[i] = sizes[i];
        L.off[i] = off;
        off += (sizes[i] + 255) & ~(size_t)255;
    }
    L.total = off;
    return GpuError::Ok;
}

GpuError MsmEngine::enqueue_work(const uint8_t* d_scalars32, const uint8_t* d_points, bool affine, size_t n,
                                 uint8_t* d_out33, int32_t* d_status, int c, hipStream_t st) {
    const MsmPlan p = msm_plan(n, c, &be_->tuning());
    MsmLayout L;
    GpuError r = msm_layout(n, p, st, L);
    if (r != GpuError::Ok) return r;
    // the engine's own grow-only buffers, one per part of the layout
    DeviceBuffer* bufs[MsmLayout::kParts] = {&pts_, &entries_, &sorted_, &bsum_, &seg_, &red_a_, &red_b_,
                                              &err_, &slot_keys_[0], &slot_keys_[1], &slot_part_[0], &slot_part_[1],
                                              &sort_tmp_};
    void* ptr[MsmLayout::kParts];
    for (int i = 0; i < MsmLayout::kParts; ++i) {
        hipError_t e = grow(*bufs[i], L.size[i]);
        if (e != hipSuccess) return be_->fail(GpuError::Memory, "hipMalloc(MSM buffers)", e);
        ptr[i] = bufs[i]->ptr;
    }
    // from 2^msm_split_from terms on (tuning file), the high windows' sums and
    // Horner share run on a second stream beside the low windows' buckets
    const int from = be_->tuning().msm_split_from;
    const bool split = from < 40 && n >= ((size_t)1 << from) && p.windows >= 2 && ensure_aux() == hipSuccess;
    return run_msm(L, p, ptr, d_scalars32, d_points, affine, n, d_out33, d_status, split, st);
}

namespace {
// the parts of MsmLayout, in order
enum MsmPart { kPts, kEntries, kSorted, kBsum, kSeg, kRedA, kRedB, kErr, kKeys0, kKeys1, kPart0, kPart1,
               kSortTmp };
// the high windows' share sits after the two error words in the kErr part
constexpr size_t kPartOffset = 64;
}  // namespace

// Buckets of windows [wa, wb): their entries are the window-major range
// [wa 2n, wb 2n) of the unsorted list; they are sorted into the same range,
// accumulated by chunks (level 0), joins and levels into bsum.  The slot and
// sort buffers are reused by the next call on the same stream.  With
// `points` set, st waits for that event (the bases) between the sort and the
// accumulation.
GpuError MsmEngine::bucket_phase(const MsmLayout& L, const MsmPlan& p, void* const* ptr, size_t n, int wa, int wb,
                                 hipEvent_t points, hipStream_t st) {
    const size_t n2 = 2 * n;
    const size_t off = (size_t)wa * n2, E = (size_t)(wb - wa) * n2;
    const uint32_t K = L.seg_chunk;
    const size_t chunks0 = (E + K - 1) / K, slots0 = 2 * chunks0;
    hipError_t e;
    const uint64_t* entries = (const uint64_t*)ptr[kEntries] + off;
    uint64_t* sorted = (uint64_t*)ptr[kSorted] + off;
    jac26s* bsum = (jac26s*)ptr[kBsum];
    uint32_t* slot_keys[2] = {(uint32_t*)ptr[kKeys0], (uint32_t*)ptr[kKeys1]};
    jac26s* slot_part[2] = {(jac26s*)ptr[kPart0], (jac26s*)ptr[kPart1]};
    // stable radix sort on the c low bits of the bucket id
    {
        size_t sort_bytes = L.sort_bytes;
        e = hipcub::DeviceRadixSort::SortKeys(ptr[kSortTmp], sort_bytes, entries, sorted, (int)E, 32, 32 + p.c, st);
        if (e != hipSuccess) return be_->fail(GpuError::Launch, "MSM bucket sort", e);
    }
    if (points && (e = hipStreamWaitEvent(st, points, 0)) != hipSuccess)
        return be_->fail(GpuError::Backend, "MSM wait for the bases", e);
    hipLaunchKernelGGL(k_seg0, dim3(blocks_for(chunks0, kMsmBlock)), dim3(kMsmBlock), 0, st, (const aff*)ptr[kPts],
                       (const uint64_t*)sorted, (uint32_t)E, K, (uint32_t)chunks0, bsum, slot_keys[0],
                       slot_part[0]);
    if ((e = hipGetLastError()) != hipSuccess) return be_->fail(GpuError::Launch, "MSM bucket level 0", e);
    // buckets spanning a few chunks (all of them for uniform scalars) are
    // completed by joining their chunk partials
    hipLaunchKernelGGL(k_join, dim3(blocks_for(chunks0, kMsmBlock)), dim3(kMsmBlock), 0, st, slot_keys[0],
                       (const jac26s*)slot_part[0], (uint32_t)chunks0, bsum);
    if ((e = hipGetLastError()) != hipSuccess) return be_->fail(GpuError::Launch, "MSM partial join", e);
    // levels over what is left, read in place (one launch each; with every
    // bucket complete a level only reads its keys), until a single chunk
    // holds every remaining slot
    size_t slots = slots0;
    int lv = 0;
    for (;;) {
        const size_t chunks = (slots + kLevelChunk - 1) / kLevelChunk;
        hipLaunchKernelGGL(k_level, dim3(blocks_for(chunks, kMsmBlock)), dim3(kMsmBlock), 0, st,
                           (const uint32_t*)slot_keys[lv], (const jac26s*)slot_part[lv], (uint32_t)slots,
                           kLevelChunk, (uint32_t)chunks, bsum, slot_keys[lv ^ 1], slot_part[lv ^ 1]);
        if ((e = hipGetLastError()) != hipSuccess) return be_->fail(GpuError::Launch, "MSM bucket level", e);
        lv ^= 1;
        if (chunks == 1) break;
        slots = 2 * chunks;
    }
    return GpuError::Ok;
}

// Window sums of windows [wa, wb): segments, then block trees until one point
// per bucket group remains.  Groups wa .. wb - 1 (and W, the top window's
// second group, when wb = W) use rows wa.. of the segment and tree buffers,
// so the two halves of the split form never share a row.
GpuError MsmEngine::window_phase(const MsmPlan& p, void* const* ptr, int wa, int wb, hipStream_t st,
                                 const void** rows) {
    const uint32_t gn = (uint32_t)(wb - wa) + (wb == p.windows ? 1u : 0u);
    const size_t T = (size_t)(p.buckets / p.seg_len);
    jac* seg = (jac*)ptr[kSeg] + (size_t)wa * T;
    hipError_t e = launch_msm_segments(ptr[kBsum], seg, p, (uint32_t)wa, gn, st);
    if (e != hipSuccess) return be_->fail(GpuError::Launch, "MSM window sums", e);
    uint32_t len = (uint32_t)T;
    const void* cur = seg;
    void* bufs[2] = {(jac*)ptr[kRedA] + (size_t)wa * T, (jac*)ptr[kRedB] + (size_t)wa * T};
    int which = 0;
    while (len > 1) {
        const uint32_t out_len = (len + kReduceBlock - 1) / kReduceBlock;
        e = launch_msm_reduce(cur, bufs[which], gn, len, out_len, st);
        if (e != hipSuccess) return be_->fail(GpuError::Launch, "MSM reduce", e);
        cur = bufs[which];
        which ^= 1;
        len = out_len;
    }
    *rows = cur;
    return GpuError::Ok;
}

// The kernel sequence of one MSM over the buffers ptr[] of layout L.  Every
// buffer a kernel reads is written earlier in the same sequence (the error
// words and bucket sums by fill kernels), so no buffer needs to start zeroed.
//
// Split form: the window sums and the final Horner chain are latency-bound
// (a few waves, chains of dependent point operations) and would leave the
// GPU idle.  The high windows [W/2, W) therefore go first; their sums and
// their Horner share, doubled c W/2 times, run on the second stream while the
// low windows' bucket phase fills the GPU on st; the last kernel adds that
// share to the low windows' Horner sum.  The two halves touch disjoint bsum
// ranges and disjoint rows of the segment and tree buffers.
GpuError MsmEngine::run_msm(const MsmLayout& L, const MsmPlan& p, void* const* ptr, const uint8_t* d_scalars32,
                            const uint8_t* d_points, bool affine, size_t n, uint8_t* d_out33, int32_t* d_status,
                            bool split, hipStream_t st) {
    const size_t nbk = (size_t)p.groups * p.buckets;
    hipError_t e;
    const unsigned cus = (unsigned)(be_->compute_units() > 0 ? be_->compute_units() : 256);
    uint32_t* err = (uint32_t*)ptr[kErr];
    void* part = (uint8_t*)ptr[kErr] + kPartOffset;
    jac26s* bsum = (jac26s*)ptr[kBsum];

    // err[0]: first failing term (2i: scalar, 2i + 1: point); err[1]: bucket range flag
    if ((e = launch_u32_fill(err, 0xFFFFFFFFu, 1, st)) != hipSuccess ||
        (e = launch_u32_fill(err + 1, 0u, 1, st)) != hipSuccess)
        return be_->fail(GpuError::Launch, "MSM error words", e);
    // The bases are decompressed on the second stream (integer VALU work)
    // while st splits the scalars into bucket entries and sorts them (memory
    // and latency work); st waits for the bases just before the first
    // accumulation.  Without the second stream the two run in turn on st.
    hipStream_t ps = st;
    if (aux_ && ready_ && hipEventRecord(fork_, st) == hipSuccess && hipStreamWaitEvent(aux_, fork_, 0) == hipSuccess)
        ps = aux_;
    if ((e = launch_msm_points(d_points, affine, (uint32_t)n, ptr[kPts], err, cus, ps)) == hipSuccess && ps != st)
        e = hipEventRecord(ready_, ps);
    if (e != hipSuccess) {
        if (ps != st) (void)hipStreamSynchronize(ps);
        return be_->fail(GpuError::Launch, "MSM bases", e);
    }
    const hipEvent_t points = ps != st ? ready_ : nullptr;
    // every return from here leaves st waiting for the bases (after the
    // bucket phase's own wait this adds nothing), so no later work on st
    // or on another stream ordered after st's completion event runs beside
    // the decompression
    struct PointsJoin {
        hipStream_t st;
        hipEvent_t ev;
        ~PointsJoin() {
            // if the stream cannot wait, the host does
            if (ev && hipStreamWaitEvent(st, ev, 0) != hipSuccess) (void)hipEventSynchronize(ev);
        }
    } join_points{st, points};
    // unsorted window-major entries of every window (zero digits are skip entries)
    {
        unsigned g = blocks_for(n, kMsmBlock);
        if (g > cus * kEntryBlocksPerCu) g = cus * kEntryBlocksPerCu;
        hipLaunchKernelGGL(k_msm_entries, dim3(g), dim3(kMsmBlock), 0, st, d_scalars32, (uint32_t)n,
                           (uint64_t*)ptr[kEntries], p.c, p.windows, err);
        if ((e = hipGetLastError()) != hipSuccess) return be_->fail(GpuError::Launch, "MSM entries", e);
    }
    // empty buckets stay at infinity
    hipLaunchKernelGGL(k_j26_fill_inf, dim3(blocks_for(nbk, kMsmBlock)), dim3(kMsmBlock), 0, st, bsum,
                       (uint32_t)nbk);
    if ((e = hipGetLastError()) != hipSuccess) return be_->fail(GpuError::Launch, "MSM bucket fill", e);
    const void* rows = nullptr;
    GpuError r;
    if (!split) {
        if ((r = bucket_phase(L, p, ptr, n, 0, p.windows, points, st)) != GpuError::Ok) return r;
        if ((r = window_phase(p, ptr, 0, p.windows, st, &rows)) != GpuError::Ok) return r;
        e = launch_msm_final(rows, (uint32_t)p.windows, true, (uint32_t)p.c, 0, nullptr, err, d_out33, d_status,
                             nullptr, st);
        if (e != hipSuccess) return be_->fail(GpuError::Launch, "MSM final", e);
        return GpuError::Ok;
    }
    const int ws = p.windows / 2;
    if ((r = bucket_phase(L, p, ptr, n, ws, p.windows, points, st)) != GpuError::Ok) return r;
    if ((e = hipEventRecord(fork_, st)) != hipSuccess || (e = hipStreamWaitEvent(aux_, fork_, 0)) != hipSuccess)
        return be_->fail(GpuError::Backend, "MSM fork", e);
    // the high windows' share, on the second stream
    const void* hrows = nullptr;
    GpuError rh = window_phase(p, ptr, ws, p.windows, aux_, &hrows);
    if (rh == GpuError::Ok) {
        e = launch_msm_final(hrows, (uint32_t)(p.windows - ws), true, (uint32_t)p.c, (uint32_t)(p.c * ws), nullptr,
                             err, nullptr, nullptr, part, aux_);
        if (e != hipSuccess) rh = be_->fail(GpuError::Launch, "MSM high windows' share", e);
    }
    if ((e = hipEventRecord(join_, aux_)) != hipSuccess) return be_->fail(GpuError::Backend, "MSM join", e);
    // the low windows on st; st then waits for the share (also when a launch
    // failed, so that nothing later on st runs beside the second stream)
    r = rh;
    if (r == GpuError::Ok) r = bucket_phase(L, p, ptr, n, 0, ws, nullptr, st);
    if (r == GpuError::Ok) r = window_phase(p, ptr, 0, ws, st, &rows);
    if ((e = hipStreamWaitEvent(st, join_, 0)) != hipSuccess) return be_->fail(GpuError::Backend, "MSM join", e);
    if (r != GpuError::Ok) return r;
    e = launch_msm_final(rows, (uint32_t)ws, false, (uint32_t)p.c, 0, part, err, d_out33, d_status, nullptr, st);
    if (e != hipSuccess) return be_->fail(GpuError::Launch, "MSM final", e);
    return GpuError::Ok;
}

GpuError MsmEngine::msm_device(const uint8_t* d_scalars32, const uint8_t* d_points, bool affine, size_t n,
                               uint8_t* d_out33, int32_t* d_status, int c, hipStream_t st) {
    if (!be_->is_ready()) return be_->fail(GpuError::Device, "context not initialised");
    GpuError r = enqueue(d_scalars32, d_points, affine, n, d_out33, d_status, c, st ? st : be_->stream());
    if (r == GpuError::Ok) be_->clear_error();
    return r;
}

GpuError MsmEngine::msm_host(const uint8_t* scalars32, const uint8_t* points33, size_t n, uint8_t* out33, int c) {
    if (!be_->is_ready()) return be_->fail(GpuError::Device, "context not initialised");
    hipStream_t st = be_->stream();
    hipError_t e;
    if ((e = order(st)) != hipSuccess) return be_->fail(GpuError::Backend, "MSM stream order", e);
    if ((e = grow(in_s_, n * 32)) != hipSuccess || (e = grow(in_p_, n * 33)) != hipSuccess ||
        (e = grow(out_, 64)) != hipSuccess)
        return be_->fail(GpuError::Memory, "hipMalloc(MSM inputs)", e);
    e = hipMemcpyAsync(in_s_.ptr, scalars32, n * 32, hipMemcpyHostToDevice, st);
    if (e == hipSuccess) e = hipMemcpyAsync(in_p_.ptr, points33, n * 33, hipMemcpyHostToDevice, st);
    if (e != hipSuccess) return be_->fail(GpuError::Memory, "copy MSM inputs", e);
    uint8_t* d_out = (uint8_t*)out_.ptr;
    int32_t* d_status = (int32_t*)(d_out + 48);
    GpuError r = enqueue((const uint8_t*)in_s_.ptr, (const uint8_t*)in_p_.ptr, false, n, d_out, d_status, c, st);
    if (r != GpuError::Ok) return r;
    uint8_t host[64];
    e = hipMemcpyAsync(host, d_out, 64, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess) return be_->fail(GpuError::Backend, "MSM synchronize", e);
    int32_t status;
    std::memcpy(&status, host + 48, 4);
    if (status != 0) {
        const char* m = status == 7   ? "MSM result is infinity"
                        : status == 3 ? "invalid point in array"
                        : status == 9 ? "MSM bucket id out of range (internal)"
                                      : "scalar >= n";
        return be_->fail((GpuError)status, m);
    }
    std::memcpy(out33, host, 33);
    be_->clear_error();
    return GpuError::Ok;
}

}  // namespace ufs
