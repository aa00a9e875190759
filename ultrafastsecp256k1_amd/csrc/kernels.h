// Host-visible declarations of the HIP kernel launchers.  Included by the
// kernel translation units and by the host backend; contains no device code.
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

namespace ufs {

constexpr int kVerifyBlock = 256;

// Fixed-base comb table for G resident in device memory.
struct GcombDesc {
    const void* table;   // (nwin * 2^(w-1) + 1) affine points, 64 B each
    int w;
    int nwin;
};

inline int gcomb_nwin(int w) { return (256 + w - 1) / w; }
inline size_t gcomb_entries(int w) { return (size_t)gcomb_nwin(w) * ((size_t)1 << (w - 1)) + 1; }

// Self-test operation codes for launch_selftest (inputs are 16-word records,
// outputs 32-word records; see k_selftest).
enum SelftestOp : int {
    kOpFeMul = 0,
    kOpFeSqr = 1,
    kOpFeAdd = 2,
    kOpFeSub = 3,
    kOpFeInv = 4,
    kOpFeSqrt = 5,
    kOpFeMulChain = 6,
    kOpFeSqrChain = 7,
    kOpFeInvFermat = 8,
    kOpScMul = 10,
    kOpScInv = 11,
    kOpGlv = 12,
    kOpScInvFermat = 13,
    kOpGenMul = 20,
    kOpVarMul = 21,
    // 10 x 26-bit limbs: inputs are 10 limbs (any values the op allows),
    // outputs the 8 x 32 result of fe26_to_fe, or the 10 limbs of fe26_norm.
    kOpFe26Mul = 30,
    kOpFe26Sqr = 31,
    kOpFe26Norm = 32,
    kOpFe26FromTo = 33,
    // Point formulas on 10 x 26-bit limbs.  Inputs: affine A (a = x || y) and
    // B (b = x || y).  Item i uses A in Jacobian form with Z = i + 3, A at
    // infinity when i % 11 == 10, and the addend flags i & 3 (kAddNeg,
    // kAddBeta).  Outputs: affine x || y of A + B' (or 2A), [16] infinity,
    // [17] the mixed addition's need-doubling flag.
    kOpJac26Madd = 34,
    kOpJac26Dbl = 35,
    kOpJac26Add = 36,
    // The wave-cooperative doubling (secp_wide26.h): item i is A in Jacobian
    // form with Z = i + 3 (infinity when i % 11 == 10), doubled 1 + i % 7
    // times by one whole wave; output affine x || y, [16] infinity.
    kOpJac26DblWide = 37,
    // The lane-spread general addition (jac26_add_lanes): as kOpJac26Add,
    // one whole wave per item.
    kOpJac26AddLanes = 38,
};

// Signature kinds of a verify batch.
//   kKindEcdsa        compact big-endian r||s, r,s in [1,n), low-S required
//   kKindSchnorr      BIP-340
//   kKindEcdsaOpaque  libbitcoin opaque form: r, s as 256-bit little-endian
//                     integers, strict [1,n), high-S accepted
constexpr uint32_t kKindEcdsa = 0;
constexpr uint32_t kKindSchnorr = 1;
constexpr uint32_t kKindEcdsaOpaque = 2;

// One verify batch of a multi-batch launch (device pointers).  With
// stride == 0 the three inputs are packed columns (msgs 32 B, pubkeys 33 B for
// ECDSA or 32 B for BIP-340, sigs 64 B per item); a non-zero stride (opaque
// ECDSA only) is the byte distance between consecutive items of all three,
// as in interleaved 32 | 33 | 64-byte rows.
struct BatchDesc {
    uint32_t kind;
    uint32_t count;
    const uint8_t* msgs;
    const uint8_t* pubs;
    const uint8_t* sigs;
    uint8_t* out;
    uint32_t stride;
    uint32_t reserved;
};

// 64-item tile of one batch.
struct TileDesc {
    uint32_t batch;
    uint32_t first;
};

constexpr uint32_t kTileItems = 64;

// Bytes of lane-private table space the verify kernel needs per grid lane
// (thirteen 64-byte entries: the table, its Z scale and ratios, a stash).
constexpr size_t kSlotBytes = 832;

// Arguments of the persistent verify kernel.  Single mode (tiles == nullptr):
// the tiles of `single`, single_tiles of them, taken through `counter` when it
// is set (zero at launch), else by a fixed stride.  List mode: tiles[0 .. *ntiles_dev)
// over descs, distributed through `counter`, which must read zero at launch
// (ntiles_dev and counter are device words so a captured graph stays valid
// when the tile list changes between replays).
struct VerifyLaunch {
    BatchDesc single;
    const BatchDesc* descs;
    const TileDesc* tiles;
    const uint32_t* ntiles_dev;
    uint32_t* counter;
    void* slots;                // kSlotBytes per lane of the grid
    uint32_t single_tiles;
    uint32_t collect;           // 1: valid -> zero the caller's byte, else keep
    // Variant 3 only: the serving loop's key cache (two-way, 64-byte entries;
    // mask = entries - 1).  Each item's key is looked up there; a miss is
    // decompressed in the kernel and inserted.  key_hits counts the items
    // found (zero at launch).
    void* key_cache;
    uint32_t key_cache_mask;
    uint32_t* key_hits;
};

// Kernel variant: 0 = ECDSA only, 1 = BIP-340 only, 2 = kind read per batch
// (list mode), 3 = as 2 with the keys through the key cache (list mode),
// 4 = single batches of the opaque ECDSA kind (the lane's items paired as
// in 0; same grid as 0).
hipError_t launch_verify(int kind, const VerifyLaunch& a, const GcombDesc& g, const uint32_t* mid, unsigned grid,
                         hipStream_t st);
int verify_blocks_per_cu(int kind);
hipError_t launch_u32_fill(uint32_t* p, uint32_t v, size_t n, hipStream_t st);

// Public-data key operations (kernels_pubdata.hip): one verdict byte per item.
hipError_t launch_pubkey_validate(const uint8_t* pk33, uint32_t n, uint8_t* res, unsigned grid, hipStream_t st);
hipError_t launch_xonly_validate(const uint8_t* x32, uint32_t n, uint8_t* res, unsigned grid, hipStream_t st);
hipError_t launch_commitment_verify(const uint8_t* ix, const uint8_t* tw, const uint8_t* tx, const uint8_t* par,
                                    uint32_t n, uint8_t* res, const GcombDesc& g, unsigned grid, hipStream_t st);
// ECDSA public-key recovery (kernels_recover.hip): 33-byte compressed key
// (zeros on failure) and a flag per item; `slots` as for launch_verify, for
// at least grid * kVerifyBlock lanes.
hipError_t launch_ecrecover(const uint8_t* msgs, const uint8_t* sigs, const int* recids, uint32_t n, uint8_t* out33,
                            uint8_t* valid, void* slots, const GcombDesc& g, unsigned grid, hipStream_t st);
// BIP-340 batch check by a random linear combination (kernels_rlc.hip):
// writes the 2n + 1 MSM terms (scalars 32 B big-endian, points 33 B
// compressed) for n signatures and the weight seed; *err |= 1 when some s is
// 0 or >= n.  `partial` holds rlc_partial_words(n) words.
struct RlcSeed {
    uint32_t w[8];   // big-endian words of the 32-byte seed
};
hipError_t launch_rlc_terms(const uint8_t* msgs, const uint8_t* pubx, const uint8_t* sigs, uint32_t n,
                            const RlcSeed& seed, const uint32_t* mid, uint8_t* scalars, uint8_t* points,
                            uint32_t* partial, uint32_t* err, hipStream_t st);
uint32_t rlc_partial_words(uint32_t n);
hipError_t launch_build_gcomb(void* table, int w, int nwin, hipStream_t st);
hipError_t launch_bip340_midstate(uint32_t* out, hipStream_t st);
hipError_t launch_selftest(int op, const uint32_t* a, const uint32_t* b, uint32_t* out, uint32_t count,
                           const GcombDesc& g, void* slots, hipStream_t st);

}  // namespace ufs
