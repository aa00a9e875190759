// Batch ECDSA public-key recovery, one lane per item: Q = u1*G + u2*R with
// u1 = -z/r, u2 = s/r and R lifted from r (plus n when bit 1 of the recovery
// id is set) with the y parity of bit 0.  Semantics of the reference CPU
// ufsecp_ecdsa_recover (src/cpu/src/impl/ufsecp_ecdsa.cpp:710-739 and
// recovery.cpp:158-235): r, s strict in [1, n) (high s allowed), recovery id
// 0..3, r < p - n required for ids 2 and 3, R must lift, Q finite; a failed
// item writes 33 zero bytes and a 0 flag.
//
// The double multiplication is the verify kernel's (GLV split of u2, the
// lane-private table of R in a slot of device memory, the resident G comb
// for u1), so the kernel runs persistent over the verify grid's slots.
#include "kernels.h"
#include "secp_ecmult.h"

#ifndef UFS_VERIFY_MIN_BLOCKS
#define UFS_VERIFY_MIN_BLOCKS 3
#endif

namespace ufs {

__global__ void __launch_bounds__(kVerifyBlock, UFS_VERIFY_MIN_BLOCKS)
k_ecrecover(const u8* __restrict__ msgs, const u8* __restrict__ sigs, const int* __restrict__ recids, u32 n,
            u8* __restrict__ out33, u8* __restrict__ valid, void* slots, gcomb G) {
    const u32 gtid = blockIdx.x * blockDim.x + threadIdx.x;
    aff* slot = reinterpret_cast<aff*>(slots) + (size_t)gtid * kSlotEntries;
    const u32 PMN[8] = {UFS_PMN_LIMBS};
    const u32 NL[8] = {UFS_N_LIMBS};
    for (u32 i = gtid; i < n; i += gridDim.x * blockDim.x) {
        u32 z[8], r[8], s[8];
        u256_from_be(z, msgs + 32ull * i);
        u256_from_be(r, sigs + 64ull * i);
        u256_from_be(s, sigs + 64ull * i + 32);
        const int recid = recids[i];
        bool ok = !u256_is_zero(r) && u256_lt_n(r) && !u256_is_zero(s) && u256_lt_n(s) && recid >= 0 && recid <= 3;
        u32 rx[8];
        const bool hi = (recid & 2) != 0;
        {
            u32 c = 0;
#pragma unroll
            for (int k = 0; k < 8; ++k) rx[k] = hi ? addc(r[k], NL[k], c, c) : r[k];
        }
        if (hi) ok &= u256_lt(r, PMN);
        aff R;
        ok &= lift_x(R, rx, (u32)recid & 1u);
        sc rs, ss, zs;
#pragma unroll
        for (int k = 0; k < 8; ++k) { rs.v[k] = r[k]; ss.v[k] = s[k]; zs.v[k] = z[k]; }
        if (!ok) {
            // keep the lane on the common path with harmless values
            set_generator(R);
            sc_set_u32(rs, 1);
            sc_set_u32(ss, 1);
        }
        sc_reduce_once(zs);
        sc ri, nz, u1, u2;
        sc_inv(ri, rs);
        sc_neg(nz, zs);
        sc_mul(u1, nz, ri);
        sc_mul(u2, ss, ri);
        sc k1, k2;
        bool n1, n2;
        glv_split(k1, k2, n1, n2, u2);
        build_ptable(slot, R);
        jac Q;
        ecmult_dual(Q, slot, k1, n1, k2, n2, G, u1);
        const bool v = ok && !Q.inf;
        u32 qx[8];
        u32 odd = 0;
        if (v) {
            aff q;
            jac_to_aff(q, Q);
#pragma unroll
            for (int k = 0; k < 8; ++k) qx[k] = q.x.v[k];
            odd = q.y.v[0] & 1u;
        } else {
#pragma unroll
            for (int k = 0; k < 8; ++k) qx[k] = 0;
        }
        u8* o = out33 + 33ull * i;
        o[0] = v ? (u8)(2u | odd) : 0;
        u256_to_be(o + 1, qx);
        valid[i] = v ? 1 : 0;
    }
}

hipError_t launch_ecrecover(const uint8_t* msgs, const uint8_t* sigs, const int* recids, uint32_t n, uint8_t* out33,
                            uint8_t* valid, void* slots, const GcombDesc& g, unsigned grid, hipStream_t st) {
    if (n == 0 || grid == 0) return hipSuccess;
    gcomb G{reinterpret_cast<const aff*>(g.table), g.w, g.nwin};
    hipLaunchKernelGGL(k_ecrecover, dim3(grid), dim3(kVerifyBlock), 0, st, msgs, sigs, recids, n, out33, valid, slots,
                       G);
    return hipGetLastError();
}

}  // namespace ufs
