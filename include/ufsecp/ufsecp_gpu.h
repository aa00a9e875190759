/* ufsecp GPU batch C ABI -- MI355X (gfx950) implementation.
 *
 * Drop-in for the batch-verify / MSM part of the reference's
 * include/ufsecp/ufsecp_gpu.h.  Every declaration below keeps the reference's
 * name, argument order, strides, ownership and error behaviour; the
 * reference location is cited on each entry point.  Conventions shared by all
 * batch calls (reference src/cpu/src/ufsecp_gpu_impl.cpp:76-140, 317-558):
 *
 *   - ctx == NULL                      -> UFSECP_ERR_NULL_ARG
 *   - count == 0                       -> UFSECP_OK, nothing written
 *   - count > 2^26                     -> UFSECP_ERR_BAD_INPUT
 *   - result buffers are zeroed before dispatch and again on any non-OK return
 *     (the in-place *_collect calls excepted)
 *   - a malformed item is a per-item 0 verdict, never an error code
 *   - host buffers are owned by the caller; device memory is owned by ctx
 *   - one ctx per thread, or external locking
 *   - public data only: no secret key ever reaches the device.
 *
 * Backend id 4 ("ROCm") names this implementation; ids 1-3 (CUDA, OpenCL,
 * Metal) are the reference's other backends and are not compiled here. */
#ifndef UFSECP_GPU_H
#define UFSECP_GPU_H

#include <stddef.h>
#include <stdint.h>

#include "ufsecp_error.h"

#ifdef __cplusplus
#  define UFSECP_NOEXCEPT noexcept
extern "C" {
#else
#  define UFSECP_NOEXCEPT
#endif

#define UFSECP_GPU_BACKEND_NONE   0
#define UFSECP_GPU_BACKEND_CUDA   1
#define UFSECP_GPU_BACKEND_OPENCL 2
#define UFSECP_GPU_BACKEND_METAL  3
#define UFSECP_GPU_BACKEND_ROCM   4

typedef struct ufsecp_gpu_ctx ufsecp_gpu_ctx;

/* Device description (reference ufsecp_gpu.h:121-129, same layout). */
typedef struct {
    char     name[128];
    uint64_t global_mem_bytes;
    uint32_t compute_units;
    uint32_t max_clock_mhz;
    uint32_t max_threads_per_block;
    uint32_t backend_id;
    uint32_t device_index;
} ufsecp_gpu_device_info_t;

/* ---- discovery and lifecycle (reference ufsecp_gpu.h:109-167) ----------- */

/* Number of compiled-in backends; fills backend_ids[0..max_ids) if non-NULL.
 * Replaces ufsecp_gpu_backend_count (ufsecp_gpu.h:109). */
UFSECP_API uint32_t ufsecp_gpu_backend_count(uint32_t* backend_ids, uint32_t max_ids) UFSECP_NOEXCEPT;

/* "ROCm" for 4, "CUDA"/"OpenCL"/"Metal" for 1-3, "none" otherwise
 * (ufsecp_gpu.h:112). */
UFSECP_API const char* ufsecp_gpu_backend_name(uint32_t backend_id);

/* 1 iff the backend is compiled in and a device exists (ufsecp_gpu.h:115). */
UFSECP_API int ufsecp_gpu_is_available(uint32_t backend_id) UFSECP_NOEXCEPT;

/* Device count for a backend, 0 if unavailable (ufsecp_gpu.h:118). */
UFSECP_API uint32_t ufsecp_gpu_device_count(uint32_t backend_id) UFSECP_NOEXCEPT;

/* Fill *info_out for (backend, device) (ufsecp_gpu.h:132). */
UFSECP_API ufsecp_error_t ufsecp_gpu_device_info(uint32_t backend_id, uint32_t device_index,
                                                 ufsecp_gpu_device_info_t* info_out);

/* Create a context on one device; builds the device-resident G table
 * (ufsecp_gpu.h:146). */
UFSECP_API ufsecp_error_t ufsecp_gpu_ctx_create(ufsecp_gpu_ctx** ctx_out, uint32_t backend_id,
                                                uint32_t device_index);

/* Release a context and all its device memory; NULL is a no-op (ufsecp_gpu.h:152). */
UFSECP_API void ufsecp_gpu_ctx_destroy(ufsecp_gpu_ctx* ctx);

/* 1 while ctx accepts batch work, 0 for NULL (ufsecp_gpu.h:158). */
UFSECP_API int ufsecp_gpu_is_ready(const ufsecp_gpu_ctx* ctx) UFSECP_NOEXCEPT;

/* Last error code of ctx (ufsecp_gpu.h:161). */
UFSECP_API ufsecp_error_t ufsecp_gpu_last_error(const ufsecp_gpu_ctx* ctx);

/* Last error message, borrowed until the next mutating call on ctx
 * (ufsecp_gpu.h:167). */
UFSECP_API const char* ufsecp_gpu_last_error_msg(const ufsecp_gpu_ctx* ctx);

/* Static description of an error code (ufsecp_gpu.h:607). */
UFSECP_API const char* ufsecp_gpu_error_str(ufsecp_error_t err);

/* ---- batch verification ------------------------------------------------- */

/* ECDSA: msg_hashes32 N*32, pubkeys33 N*33 (SEC1 compressed), sigs64 N*64
 * compact big-endian r||s; out_results N bytes, 1 valid / 0 invalid.
 * Verdicts equal the reference CPU ufsecp_ecdsa_verify: strict r,s in [1,n),
 * low-S required, strict pubkey.  Replaces ufsecp_gpu_ecdsa_verify_batch
 * (ufsecp_gpu.h:200-206). */
UFSECP_API ufsecp_error_t ufsecp_gpu_ecdsa_verify_batch(ufsecp_gpu_ctx* ctx, const uint8_t* msg_hashes32,
                                                        const uint8_t* pubkeys33, const uint8_t* sigs64,
                                                        size_t count, uint8_t* out_results);

/* BIP-340: msg_hashes32 N*32, pubkeys_x32 N*32, sigs64 N*64 (r||s);
 * out_results N bytes.  Verdicts equal the reference CPU
 * ufsecp_schnorr_verify (r < p, 0 < s < n, x-only key < p).  Replaces
 * ufsecp_gpu_schnorr_verify_batch (ufsecp_gpu.h:254-260). */
UFSECP_API ufsecp_error_t ufsecp_gpu_schnorr_verify_batch(ufsecp_gpu_ctx* ctx, const uint8_t* msg_hashes32,
                                                          const uint8_t* pubkeys_x32, const uint8_t* sigs64,
                                                          size_t count, uint8_t* out_results);

/* libbitcoin opaque-signature ECDSA rows: item i is the `stride`-byte row at
 * rows + i*stride holding 32-byte digest | 33-byte compressed pubkey |
 * 64-byte opaque signature (r, s as 256-bit little-endian integers) | tail.
 * stride < 129 or count > 2^26 -> BAD_INPUT.  Verdicts equal the reference
 * CPU ufsecp_ecdsa_verify_opaque_rows (impl/ufsecp_ecdsa.cpp:412-452): r, s
 * strict in [1,n), high-S accepted, strict pubkey.  Replaces
 * ufsecp_gpu_ecdsa_verify_opaque_rows (ufsecp_gpu.h:228-233). */
UFSECP_API ufsecp_error_t ufsecp_gpu_ecdsa_verify_opaque_rows(ufsecp_gpu_ctx* ctx, const uint8_t* rows,
                                                              size_t stride, size_t count, uint8_t* out_results);

/* Alias of the above.  Replaces ufsecp_gpu_ecdsa_verify_lbtc_rows
 * (ufsecp_gpu.h:236-241). */
UFSECP_API ufsecp_error_t ufsecp_gpu_ecdsa_verify_lbtc_rows(ufsecp_gpu_ctx* ctx, const uint8_t* rows, size_t stride,
                                                            size_t count, uint8_t* out_results);

/* The same opaque ECDSA verification over three packed columns: digests32
 * N*32, pubkeys33 N*33, sigs64 N*64 opaque.  Verdicts equal the reference CPU
 * ufsecp_ecdsa_verify_opaque_batch (impl/ufsecp_ecdsa.cpp:368-410).  Replaces
 * ufsecp_gpu_ecdsa_verify_lbtc_columns (ufsecp_gpu.h:278-284). */
UFSECP_API ufsecp_error_t ufsecp_gpu_ecdsa_verify_lbtc_columns(ufsecp_gpu_ctx* ctx, const uint8_t* digests32,
                                                               const uint8_t* pubkeys33, const uint8_t* sigs64,
                                                               size_t count, uint8_t* out_results);

/* BIP-340 over columns digests32 | xonly32 | sigs64: the verdicts of
 * ufsecp_gpu_schnorr_verify_batch (strict s in [1,n), r < p).  Replaces
 * ufsecp_gpu_schnorr_verify_lbtc_columns (ufsecp_gpu.h:299-305). */
UFSECP_API ufsecp_error_t ufsecp_gpu_schnorr_verify_lbtc_columns(ufsecp_gpu_ctx* ctx, const uint8_t* digests32,
                                                                 const uint8_t* xonly32, const uint8_t* sigs64,
                                                                 size_t count, uint8_t* out_results);

/* In-place verdict collapse: key_buffer bytes are caller-seeded non-zero
 * markers; a VALID item zeroes its byte, an invalid one leaves it.  Not
 * pre-cleared.  Replaces ufsecp_gpu_ecdsa_verify_collect (ufsecp_gpu.h:326-332). */
UFSECP_API ufsecp_error_t ufsecp_gpu_ecdsa_verify_collect(ufsecp_gpu_ctx* ctx, const uint8_t* msg_hashes32,
                                                          const uint8_t* pubkeys33, const uint8_t* sigs64,
                                                          size_t count, uint8_t* key_buffer);

/* Schnorr sibling of the above.  Replaces ufsecp_gpu_schnorr_verify_collect
 * (ufsecp_gpu.h:345-351). */
UFSECP_API ufsecp_error_t ufsecp_gpu_schnorr_verify_collect(ufsecp_gpu_ctx* ctx, const uint8_t* msg_hashes32,
                                                            const uint8_t* pubkeys_x32, const uint8_t* sigs64,
                                                            size_t count, uint8_t* key_buffer);

/* ---- public-data key operations (libbitcoin surface) -------------------- */

/* 1 per key that is a valid compressed point (prefix 02/03, x < p, on the
 * curve), as the reference CPU ufsecp_pubkey_parse.  Replaces
 * ufsecp_gpu_pubkey_validate (ufsecp_gpu.h:477-481). */
UFSECP_API ufsecp_error_t ufsecp_gpu_pubkey_validate(ufsecp_gpu_ctx* ctx, const uint8_t* pubkeys33, size_t n,
                                                     uint8_t* results);

/* 1 per x-only key with x < p and x^3 + 7 a square (lift_x succeeds).
 * Replaces ufsecp_gpu_xonly_validate (ufsecp_gpu.h:419-423). */
UFSECP_API ufsecp_error_t ufsecp_gpu_xonly_validate(ufsecp_gpu_ctx* ctx, const uint8_t* keys32, size_t n,
                                                    uint8_t* results);

/* BIP-341 tweak-add check per item: Q = lift_x(internal_x, even y) + t*G;
 * 1 iff t < n, Q is finite, x(Q) == tweaked_x and y(Q) is odd exactly when
 * parity[i] != 0 (the reference CPU two-term ufsecp_multi_scalar_mul check,
 * compat/libbitcoin_bridge/src/ufsecp_libbitcoin.cpp:1138-1149).  Replaces
 * ufsecp_gpu_commitment_verify (ufsecp_gpu.h:439-446). */
UFSECP_API ufsecp_error_t ufsecp_gpu_commitment_verify(ufsecp_gpu_ctx* ctx, const uint8_t* internal_x32,
                                                       const uint8_t* tweak32, const uint8_t* tweaked_x32,
                                                       const uint8_t* parity, size_t n, uint8_t* results);

/* ---- public-key recovery ------------------------------------------------ */

/* Per item, the compressed public key recovered from (msg_hash, compact
 * big-endian r||s, recovery id 0..3); a failed item gets 33 zero bytes and
 * out_valid[i] = 0.  Any recovery id outside 0..3 -> BAD_INPUT for the whole
 * call.  Per-item results equal the reference CPU ufsecp_ecdsa_recover
 * (impl/ufsecp_ecdsa.cpp:710-739): r, s strict in [1,n), ids 2/3 need
 * r < p - n.  Replaces ufsecp_gpu_ecrecover_batch (ufsecp_gpu.h:596-603). */
UFSECP_API ufsecp_error_t ufsecp_gpu_ecrecover_batch(ufsecp_gpu_ctx* ctx, const uint8_t* msg_hashes32,
                                                     const uint8_t* sigs64, const int* recids, size_t count,
                                                     uint8_t* out_pubkeys33, uint8_t* out_valid);

/* ---- multi-scalar multiplication ---------------------------------------- */

/* out_result33 = compressed sum(scalars[i] * points[i]); scalars32 N*32
 * big-endian, points33 N*33 compressed.  Scalars >= n -> BAD_INPUT, invalid
 * point -> BAD_PUBKEY (the reference CPU ufsecp_multi_scalar_mul semantics,
 * impl/ufsecp_taproot.cpp:957-991); result at infinity -> ERR_ARITH with the
 * output zeroed.  Replaces ufsecp_gpu_msm (ufsecp_gpu.h:397-402). */
UFSECP_API ufsecp_error_t ufsecp_gpu_msm(ufsecp_gpu_ctx* ctx, const uint8_t* scalars32, const uint8_t* points33,
                                         size_t n, uint8_t* out_result33);

/* ---- operations this backend does not implement ------------------------- *
 *
 * The rest of the reference header, off the batch-verify / MSM path: secret-
 * bearing operations (ECDH, BIP-324 AEAD, BIP-352 scanning, generator
 * multiplication, which the reference runs on its constant-time kernel),
 * byte hashing and the ZK / FROST verifiers.  Each is exported with the
 * reference's signature so that callers written against the reference link
 * and bind unchanged.  Each runs the reference's argument checks in the
 * reference's order (src/cpu/src/ufsecp_gpu_impl.cpp:299-313, 495-538,
 * 600-1127), so count == 0 is still OK with nothing written and a NULL buffer
 * still NULL_ARG; a call that passes them clears its outputs and returns
 * UFSECP_ERR_GPU_UNSUPPORTED -- the reference's code for an operation a
 * backend lacks (src/gpu/include/gpu_backend.hpp:139-231). */

#define UFSECP_ECDSA_SNARK_WITNESS_BYTES 760
#define UFSECP_SCHNORR_SNARK_WITNESS_BYTES 472
#define UFSECP_BIP352_SCAN_PLAN_BYTES 264

/* Replaces ufsecp_gpu_generator_mul_batch (ufsecp_gpu.h:182-186). */
UFSECP_API ufsecp_error_t ufsecp_gpu_generator_mul_batch(ufsecp_gpu_ctx* ctx, const uint8_t* scalars32,
                                                         size_t count, uint8_t* out_pubkeys33);

/* Replaces ufsecp_gpu_ecdh_batch (ufsecp_gpu.h:364-369). */
UFSECP_API ufsecp_error_t ufsecp_gpu_ecdh_batch(ufsecp_gpu_ctx* ctx, const uint8_t* privkeys32,
                                                const uint8_t* peer_pubkeys33, size_t count,
                                                uint8_t* out_secrets32);

/* Replaces ufsecp_gpu_hash160_pubkey_batch (ufsecp_gpu.h:380-384). */
UFSECP_API ufsecp_error_t ufsecp_gpu_hash160_pubkey_batch(ufsecp_gpu_ctx* ctx, const uint8_t* pubkeys33,
                                                          size_t count, uint8_t* out_hash160);

/* Replaces ufsecp_gpu_tagged_hash (ufsecp_gpu.h:460-466). */
UFSECP_API ufsecp_error_t ufsecp_gpu_tagged_hash(ufsecp_gpu_ctx* ctx, const uint8_t* tag_hash32,
                                                 const uint8_t* msgs, size_t msg_len, size_t n, uint8_t* out32);

/* Replaces ufsecp_gpu_tagged_hash_var (ufsecp_gpu.h:496-503). */
UFSECP_API ufsecp_error_t ufsecp_gpu_tagged_hash_var(ufsecp_gpu_ctx* ctx, const uint8_t* tag_hash32,
                                                     const uint8_t* msgs, const uint32_t* msg_lens, size_t stride,
                                                     size_t n, uint8_t* out32);

/* Replaces ufsecp_gpu_hash256 (ufsecp_gpu.h:516-521). */
UFSECP_API ufsecp_error_t ufsecp_gpu_hash256(ufsecp_gpu_ctx* ctx, const uint8_t* inputs, size_t input_len,
                                             size_t n, uint8_t* out32);

/* Replaces ufsecp_gpu_hash256_var (ufsecp_gpu.h:538-544). */
UFSECP_API ufsecp_error_t ufsecp_gpu_hash256_var(ufsecp_gpu_ctx* ctx, const uint8_t* inputs,
                                                 const uint32_t* input_lens, size_t stride, size_t n,
                                                 uint8_t* out32);

/* Replaces ufsecp_gpu_frost_verify_partial_batch (ufsecp_gpu.h:569-580). */
UFSECP_API ufsecp_error_t ufsecp_gpu_frost_verify_partial_batch(
    ufsecp_gpu_ctx* ctx, const uint8_t* z_i32, const uint8_t* D_i33, const uint8_t* E_i33, const uint8_t* Y_i33,
    const uint8_t* rho_i32, const uint8_t* lambda_ie32, const uint8_t* negate_R, const uint8_t* negate_key,
    size_t count, uint8_t* out_results);

/* Replaces ufsecp_gpu_zk_knowledge_verify_batch (ufsecp_gpu.h:629-635). */
UFSECP_API ufsecp_error_t ufsecp_gpu_zk_knowledge_verify_batch(ufsecp_gpu_ctx* ctx, const uint8_t* proofs64,
                                                               const uint8_t* pubkeys65, const uint8_t* messages32,
                                                               size_t count, uint8_t* out_results);

/* Replaces ufsecp_gpu_zk_dleq_verify_batch (ufsecp_gpu.h:655-663). */
UFSECP_API ufsecp_error_t ufsecp_gpu_zk_dleq_verify_batch(ufsecp_gpu_ctx* ctx, const uint8_t* proofs64,
                                                          const uint8_t* G_pts65, const uint8_t* H_pts65,
                                                          const uint8_t* P_pts65, const uint8_t* Q_pts65,
                                                          size_t count, uint8_t* out_results);

/* Replaces ufsecp_gpu_bulletproof_verify_batch (ufsecp_gpu.h:681-687). */
UFSECP_API ufsecp_error_t ufsecp_gpu_bulletproof_verify_batch(ufsecp_gpu_ctx* ctx, const uint8_t* proofs324,
                                                              const uint8_t* commitments65,
                                                              const uint8_t* H_generator65, size_t count,
                                                              uint8_t* out_results);

/* Replaces ufsecp_gpu_bip324_aead_encrypt_batch (ufsecp_gpu.h:710-718). */
UFSECP_API ufsecp_error_t ufsecp_gpu_bip324_aead_encrypt_batch(ufsecp_gpu_ctx* ctx, const uint8_t* keys32,
                                                               const uint8_t* nonces12, const uint8_t* plaintexts,
                                                               const uint32_t* sizes, uint32_t max_payload,
                                                               size_t count, uint8_t* wire_out);

/* Replaces ufsecp_gpu_bip324_aead_decrypt_batch (ufsecp_gpu.h:736-745). */
UFSECP_API ufsecp_error_t ufsecp_gpu_bip324_aead_decrypt_batch(ufsecp_gpu_ctx* ctx, const uint8_t* keys32,
                                                               const uint8_t* nonces12, const uint8_t* wire_in,
                                                               const uint32_t* sizes, uint32_t max_payload,
                                                               size_t count, uint8_t* plaintext_out,
                                                               uint8_t* out_valid);

/* Replaces ufsecp_gpu_zk_ecdsa_snark_witness_batch (ufsecp_gpu.h:767-773). */
UFSECP_API ufsecp_error_t ufsecp_gpu_zk_ecdsa_snark_witness_batch(ufsecp_gpu_ctx* ctx, const uint8_t* msg_hashes32,
                                                                  const uint8_t* pubkeys33, const uint8_t* sigs64,
                                                                  size_t count, uint8_t* out_witnesses);

/* Replaces ufsecp_gpu_zk_schnorr_snark_witness_batch (ufsecp_gpu.h:794-800). */
UFSECP_API ufsecp_error_t ufsecp_gpu_zk_schnorr_snark_witness_batch(ufsecp_gpu_ctx* ctx, const uint8_t* msgs32,
                                                                    const uint8_t* pubkeys_x32, const uint8_t* sigs64,
                                                                    size_t count, uint8_t* out_witnesses);

/* CPU-side plan for ufsecp_gpu_bip352_scan_batch: NULL_ARG on NULL buffers,
 * otherwise the plan is zeroed and UFSECP_ERR_GPU_UNSUPPORTED returned (no
 * secret key is read).  Replaces ufsecp_bip352_prepare_scan_plan
 * (ufsecp_gpu.h:825-827). */
UFSECP_API ufsecp_error_t ufsecp_bip352_prepare_scan_plan(const uint8_t scan_privkey32[32],
                                                          uint8_t plan264_out[264]);

/* Replaces ufsecp_gpu_bip352_scan_batch (ufsecp_gpu.h:861-867). */
UFSECP_API ufsecp_error_t ufsecp_gpu_bip352_scan_batch(ufsecp_gpu_ctx* ctx, const uint8_t scan_privkey32[32],
                                                       const uint8_t spend_pubkey33[33],
                                                       const uint8_t* tweak_pubkeys33, size_t n_tweaks,
                                                       uint64_t* prefix64_out);

#ifdef __cplusplus
}
#endif

#endif /* UFSECP_GPU_H */
