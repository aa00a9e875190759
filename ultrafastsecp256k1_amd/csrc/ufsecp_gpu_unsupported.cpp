// The reference header's entry points that are off the batch-verify / MSM
// path (secret-bearing operations, byte hashing, ZK / FROST verifiers).  Each
// keeps the reference's signature and runs its argument checks in the order of
// src/cpu/src/ufsecp_gpu_impl.cpp (cited per function), so callers see the
// same codes for empty, oversized, NULL and malformed calls; a call that
// passes them clears its outputs and returns UFSECP_ERR_GPU_UNSUPPORTED, the
// code the reference returns for an operation a backend does not implement
// (src/gpu/include/gpu_backend.hpp:139-231).  No input byte, secret or
// public, is read beyond those checks and nothing reaches the device.
#include <cstddef>
#include <cstdint>
#include <cstring>
#include <limits>

#include "../../include/ufsecp/ufsecp_gpu.h"

ufsecp_error_t ufs_abi_unsupported(ufsecp_gpu_ctx* ctx, const char* op);

namespace {

constexpr size_t kMaxGpuBatchN = size_t{1} << 26;

bool clear_out(void* out, size_t count, size_t stride) {
    if (!out || count == 0 || stride == 0) return true;
    if (count > std::numeric_limits<size_t>::max() / stride) return false;
    std::memset(out, 0, count * stride);
    return true;
}

bool prefixes_ok(const uint8_t* keys, size_t count, size_t stride, uint8_t a, uint8_t b) {
    for (size_t i = 0; i < count; ++i) {
        const uint8_t p = keys[i * stride];
        if (p != a && p != b) return false;
    }
    return true;
}

bool compressed_ok(const uint8_t* keys33, size_t count) { return prefixes_ok(keys33, count, 33, 2, 3); }
bool uncompressed_ok(const uint8_t* keys65, size_t count) { return prefixes_ok(keys65, count, 65, 4, 4); }

}  // namespace

extern "C" {

// ufsecp_gpu_impl.cpp:299-313
ufsecp_error_t ufsecp_gpu_generator_mul_batch(ufsecp_gpu_ctx* ctx, const uint8_t* scalars32, size_t count,
                                              uint8_t* out_pubkeys33) {
    if (!ctx) return UFSECP_ERR_NULL_ARG;
    if (count == 0) return UFSECP_OK;
    if (count > kMaxGpuBatchN) return UFSECP_ERR_BAD_INPUT;
    if (!clear_out(out_pubkeys33, count, 33)) return UFSECP_ERR_BAD_INPUT;
    if (!scalars32 || !out_pubkeys33) return UFSECP_ERR_NULL_ARG;
    return ufs_abi_unsupported(ctx, "generator_mul_batch");
}

// ufsecp_gpu_impl.cpp:495-518
ufsecp_error_t ufsecp_gpu_ecdh_batch(ufsecp_gpu_ctx* ctx, const uint8_t* privkeys32, const uint8_t* peer_pubkeys33,
                                     size_t count, uint8_t* out_secrets32) {
    if (!ctx) return UFSECP_ERR_NULL_ARG;
    if (count == 0) return UFSECP_OK;
    if (count > kMaxGpuBatchN) return UFSECP_ERR_BAD_INPUT;
    if (!clear_out(out_secrets32, count, 32)) return UFSECP_ERR_BAD_INPUT;
    if (!privkeys32 || !peer_pubkeys33 || !out_secrets32) return UFSECP_ERR_NULL_ARG;
    if (!compressed_ok(peer_pubkeys33, count)) return UFSECP_ERR_BAD_PUBKEY;
    return ufs_abi_unsupported(ctx, "ecdh_batch");
}

// ufsecp_gpu_impl.cpp:520-538
ufsecp_error_t ufsecp_gpu_hash160_pubkey_batch(ufsecp_gpu_ctx* ctx, const uint8_t* pubkeys33, size_t count,
                                               uint8_t* out_hash160) {
    if (!ctx) return UFSECP_ERR_NULL_ARG;
    if (count == 0) return UFSECP_OK;
    if (count > kMaxGpuBatchN) return UFSECP_ERR_BAD_INPUT;
    if (!clear_out(out_hash160, count, 20)) return UFSECP_ERR_BAD_INPUT;
    if (!pubkeys33 || !out_hash160) return UFSECP_ERR_NULL_ARG;
    if (!compressed_ok(pubkeys33, count)) return UFSECP_ERR_BAD_PUBKEY;
    return ufs_abi_unsupported(ctx, "hash160_pubkey_batch");
}

// ufsecp_gpu_impl.cpp:600-618
ufsecp_error_t ufsecp_gpu_tagged_hash(ufsecp_gpu_ctx* ctx, const uint8_t* tag_hash32, const uint8_t* msgs,
                                      size_t msg_len, size_t n, uint8_t* out32) {
    if (!ctx) return UFSECP_ERR_NULL_ARG;
    if (n == 0) return UFSECP_OK;
    if (n > kMaxGpuBatchN) return UFSECP_ERR_BAD_INPUT;
    if (msg_len == 0 || msg_len > 256) return UFSECP_ERR_BAD_INPUT;
    if (!clear_out(out32, n, 32)) return UFSECP_ERR_BAD_INPUT;
    if (!tag_hash32 || !msgs || !out32) return UFSECP_ERR_NULL_ARG;
    return ufs_abi_unsupported(ctx, "tagged_hash");
}

// ufsecp_gpu_impl.cpp:639-659
ufsecp_error_t ufsecp_gpu_tagged_hash_var(ufsecp_gpu_ctx* ctx, const uint8_t* tag_hash32, const uint8_t* msgs,
                                          const uint32_t* msg_lens, size_t stride, size_t n, uint8_t* out32) {
    if (!ctx) return UFSECP_ERR_NULL_ARG;
    if (n == 0) return UFSECP_OK;
    if (n > kMaxGpuBatchN) return UFSECP_ERR_BAD_INPUT;
    if (stride == 0 || stride > 256) return UFSECP_ERR_BAD_INPUT;
    if (!clear_out(out32, n, 32)) return UFSECP_ERR_BAD_INPUT;
    if (!tag_hash32 || !msgs || !msg_lens || !out32) return UFSECP_ERR_NULL_ARG;
    return ufs_abi_unsupported(ctx, "tagged_hash_var");
}

// ufsecp_gpu_impl.cpp:661-679
ufsecp_error_t ufsecp_gpu_hash256(ufsecp_gpu_ctx* ctx, const uint8_t* inputs, size_t input_len, size_t n,
                                  uint8_t* out32) {
    if (!ctx) return UFSECP_ERR_NULL_ARG;
    if (n == 0) return UFSECP_OK;
    if (n > kMaxGpuBatchN) return UFSECP_ERR_BAD_INPUT;
    if (input_len == 0 || input_len > 320) return UFSECP_ERR_BAD_INPUT;
    if (!clear_out(out32, n, 32)) return UFSECP_ERR_BAD_INPUT;
    if (!inputs || !out32) return UFSECP_ERR_NULL_ARG;
    return ufs_abi_unsupported(ctx, "hash256");
}

// ufsecp_gpu_impl.cpp:681-719 (row stride cap 4 MiB; every row length in [1, stride])
ufsecp_error_t ufsecp_gpu_hash256_var(ufsecp_gpu_ctx* ctx, const uint8_t* inputs, const uint32_t* input_lens,
                                      size_t stride, size_t n, uint8_t* out32) {
    if (!ctx) return UFSECP_ERR_NULL_ARG;
    if (n == 0) return UFSECP_OK;
    if (n > kMaxGpuBatchN) return UFSECP_ERR_BAD_INPUT;
    if (stride == 0 || stride > (size_t{4} << 20)) return UFSECP_ERR_BAD_INPUT;
    if (!clear_out(out32, n, 32)) return UFSECP_ERR_BAD_INPUT;
    if (!inputs || !input_lens || !out32) return UFSECP_ERR_NULL_ARG;
    for (size_t i = 0; i < n; ++i)
        if (input_lens[i] == 0 || input_lens[i] > stride) return UFSECP_ERR_BAD_INPUT;
    return ufs_abi_unsupported(ctx, "hash256_var");
}

// ufsecp_gpu_impl.cpp:721-755
ufsecp_error_t ufsecp_gpu_frost_verify_partial_batch(ufsecp_gpu_ctx* ctx, const uint8_t* z_i32,
                                                     const uint8_t* D_i33, const uint8_t* E_i33,
                                                     const uint8_t* Y_i33, const uint8_t* rho_i32,
                                                     const uint8_t* lambda_ie32, const uint8_t* negate_R,
                                                     const uint8_t* negate_key, size_t count, uint8_t* out_results) {
    if (!ctx) return UFSECP_ERR_NULL_ARG;
    if (count == 0) return UFSECP_OK;
    if (count > kMaxGpuBatchN) return UFSECP_ERR_BAD_INPUT;
    if (!clear_out(out_results, count, 1)) return UFSECP_ERR_BAD_INPUT;
    if (!z_i32 || !D_i33 || !E_i33 || !Y_i33 || !rho_i32 || !lambda_ie32 || !negate_R || !negate_key ||
        !out_results)
        return UFSECP_ERR_NULL_ARG;
    if (!compressed_ok(D_i33, count) || !compressed_ok(E_i33, count) || !compressed_ok(Y_i33, count))
        return UFSECP_ERR_BAD_PUBKEY;
    return ufs_abi_unsupported(ctx, "frost_verify_partial_batch");
}

// ufsecp_gpu_impl.cpp:792-815
ufsecp_error_t ufsecp_gpu_zk_knowledge_verify_batch(ufsecp_gpu_ctx* ctx, const uint8_t* proofs64,
                                                    const uint8_t* pubkeys65, const uint8_t* messages32,
                                                    size_t count, uint8_t* out_results) {
    if (!ctx) return UFSECP_ERR_NULL_ARG;
    if (count == 0) return UFSECP_OK;
    if (count > kMaxGpuBatchN) return UFSECP_ERR_BAD_INPUT;
    if (!clear_out(out_results, count, 1)) return UFSECP_ERR_BAD_INPUT;
    if (!proofs64 || !pubkeys65 || !messages32 || !out_results) return UFSECP_ERR_NULL_ARG;
    if (!uncompressed_ok(pubkeys65, count)) return UFSECP_ERR_BAD_PUBKEY;
    return ufs_abi_unsupported(ctx, "zk_knowledge_verify_batch");
}

// ufsecp_gpu_impl.cpp:817-845
ufsecp_error_t ufsecp_gpu_zk_dleq_verify_batch(ufsecp_gpu_ctx* ctx, const uint8_t* proofs64,
                                               const uint8_t* G_pts65, const uint8_t* H_pts65,
                                               const uint8_t* P_pts65, const uint8_t* Q_pts65, size_t count,
                                               uint8_t* out_results) {
    if (!ctx) return UFSECP_ERR_NULL_ARG;
    if (count == 0) return UFSECP_OK;
    if (count > kMaxGpuBatchN) return UFSECP_ERR_BAD_INPUT;
    if (!clear_out(out_results, count, 1)) return UFSECP_ERR_BAD_INPUT;
    if (!proofs64 || !G_pts65 || !H_pts65 || !P_pts65 || !Q_pts65 || !out_results) return UFSECP_ERR_NULL_ARG;
    if (!uncompressed_ok(G_pts65, count) || !uncompressed_ok(H_pts65, count) || !uncompressed_ok(P_pts65, count) ||
        !uncompressed_ok(Q_pts65, count))
        return UFSECP_ERR_BAD_PUBKEY;
    return ufs_abi_unsupported(ctx, "zk_dleq_verify_batch");
}

// ufsecp_gpu_impl.cpp:847-872 (A, S, T1, T2 of each 324-byte proof carry prefix 04)
ufsecp_error_t ufsecp_gpu_bulletproof_verify_batch(ufsecp_gpu_ctx* ctx, const uint8_t* proofs324,
                                                   const uint8_t* commitments65, const uint8_t* H_generator65,
                                                   size_t count, uint8_t* out_results) {
    if (!ctx) return UFSECP_ERR_NULL_ARG;
    if (count == 0) return UFSECP_OK;
    if (count > kMaxGpuBatchN) return UFSECP_ERR_BAD_INPUT;
    if (!clear_out(out_results, count, 1)) return UFSECP_ERR_BAD_INPUT;
    if (!proofs324 || !commitments65 || !H_generator65 || !out_results) return UFSECP_ERR_NULL_ARG;
    for (size_t i = 0; i < count; ++i) {
        const uint8_t* pr = proofs324 + i * 324;
        if (pr[0] != 4 || pr[65] != 4 || pr[130] != 4 || pr[195] != 4) return UFSECP_ERR_BAD_PUBKEY;
    }
    if (!uncompressed_ok(commitments65, count) || !uncompressed_ok(H_generator65, 1)) return UFSECP_ERR_BAD_PUBKEY;
    return ufs_abi_unsupported(ctx, "bulletproof_verify_batch");
}

// ufsecp_gpu_impl.cpp:878-907 (wire stride max_payload + 19)
ufsecp_error_t ufsecp_gpu_bip324_aead_encrypt_batch(ufsecp_gpu_ctx* ctx, const uint8_t* keys32,
                                                    const uint8_t* nonces12, const uint8_t* plaintexts,
                                                    const uint32_t* sizes, uint32_t max_payload, size_t count,
                                                    uint8_t* wire_out) {
    if (!ctx) return UFSECP_ERR_NULL_ARG;
    if (count == 0) return UFSECP_OK;
    if (count > kMaxGpuBatchN) return UFSECP_ERR_BAD_INPUT;
    const size_t wire_stride = (size_t)max_payload + 19u;
    if (!clear_out(wire_out, count, wire_stride)) return UFSECP_ERR_BAD_INPUT;
    if (!keys32 || !nonces12 || !plaintexts || !sizes || !wire_out) return UFSECP_ERR_NULL_ARG;
    for (size_t i = 0; i < count; ++i)
        if (sizes[i] > max_payload) return UFSECP_ERR_BAD_INPUT;
    return ufs_abi_unsupported(ctx, "bip324_aead_encrypt_batch");
}

// ufsecp_gpu_impl.cpp:909-943
ufsecp_error_t ufsecp_gpu_bip324_aead_decrypt_batch(ufsecp_gpu_ctx* ctx, const uint8_t* keys32,
                                                    const uint8_t* nonces12, const uint8_t* wire_in,
                                                    const uint32_t* sizes, uint32_t max_payload, size_t count,
                                                    uint8_t* plaintext_out, uint8_t* out_valid) {
    if (!ctx) return UFSECP_ERR_NULL_ARG;
    if (count == 0) return UFSECP_OK;
    if (count > kMaxGpuBatchN) return UFSECP_ERR_BAD_INPUT;
    if (!clear_out(plaintext_out, count, (size_t)max_payload)) return UFSECP_ERR_BAD_INPUT;
    if (!clear_out(out_valid, count, 1)) return UFSECP_ERR_BAD_INPUT;
    if (!keys32 || !nonces12 || !wire_in || !sizes || !plaintext_out || !out_valid) return UFSECP_ERR_NULL_ARG;
    for (size_t i = 0; i < count; ++i)
        if (sizes[i] > max_payload) return UFSECP_ERR_BAD_INPUT;
    return ufs_abi_unsupported(ctx, "bip324_aead_decrypt_batch");
}

// ufsecp_gpu_impl.cpp:949-975 (count checked before ctx, as there)
ufsecp_error_t ufsecp_gpu_zk_ecdsa_snark_witness_batch(ufsecp_gpu_ctx* ctx, const uint8_t* msg_hashes32,
                                                       const uint8_t* pubkeys33, const uint8_t* sigs64, size_t count,
                                                       uint8_t* out_witnesses) {
    if (count == 0) return UFSECP_OK;
    if (count > kMaxGpuBatchN) return UFSECP_ERR_BAD_INPUT;
    if (!clear_out(out_witnesses, count, UFSECP_ECDSA_SNARK_WITNESS_BYTES)) return UFSECP_ERR_BAD_INPUT;
    if (!ctx || !msg_hashes32 || !pubkeys33 || !sigs64 || !out_witnesses) return UFSECP_ERR_NULL_ARG;
    if (!compressed_ok(pubkeys33, count)) return UFSECP_ERR_BAD_PUBKEY;
    return ufs_abi_unsupported(ctx, "zk_ecdsa_snark_witness_batch");
}

// ufsecp_gpu_impl.cpp:978-1010 (an all-zero s is BAD_SIG before dispatch)
ufsecp_error_t ufsecp_gpu_zk_schnorr_snark_witness_batch(ufsecp_gpu_ctx* ctx, const uint8_t* msgs32,
                                                         const uint8_t* pubkeys_x32, const uint8_t* sigs64,
                                                         size_t count, uint8_t* out_witnesses) {
    if (count == 0) return UFSECP_OK;
    if (count > kMaxGpuBatchN) return UFSECP_ERR_BAD_INPUT;
    if (!clear_out(out_witnesses, count, UFSECP_SCHNORR_SNARK_WITNESS_BYTES)) return UFSECP_ERR_BAD_INPUT;
    if (!ctx || !msgs32 || !pubkeys_x32 || !sigs64 || !out_witnesses) return UFSECP_ERR_NULL_ARG;
    for (size_t i = 0; i < count; ++i) {
        const uint8_t* s = sigs64 + i * 64 + 32;
        uint8_t any = 0;
        for (int b = 0; b < 32; ++b) any |= s[b];
        if (!any) return UFSECP_ERR_BAD_SIG;
    }
    return ufs_abi_unsupported(ctx, "zk_schnorr_snark_witness_batch");
}

// ufsecp_gpu_impl.cpp:1015-1020: NULL checks and a cleared plan; the plan is
// for ufsecp_gpu_bip352_scan_batch, which this backend does not run, so the
// secret scan key is not read.
ufsecp_error_t ufsecp_bip352_prepare_scan_plan(const uint8_t scan_privkey32[32], uint8_t plan264_out[264]) {
    if (!scan_privkey32 || !plan264_out) return UFSECP_ERR_NULL_ARG;
    std::memset(plan264_out, 0, UFSECP_BIP352_SCAN_PLAN_BYTES);
    return UFSECP_ERR_GPU_UNSUPPORTED;
}

// ufsecp_gpu_impl.cpp:1093-1127.  The reference's strict scan-key parse
// (BAD_KEY for 0 or >= n) reads the secret key; it is not repeated here, the
// call ends UNSUPPORTED after the public-input checks instead.
ufsecp_error_t ufsecp_gpu_bip352_scan_batch(ufsecp_gpu_ctx* ctx, const uint8_t scan_privkey32[32],
                                            const uint8_t spend_pubkey33[33], const uint8_t* tweak_pubkeys33,
                                            size_t n_tweaks, uint64_t* prefix64_out) {
    if (!ctx || !scan_privkey32 || !spend_pubkey33 || !tweak_pubkeys33 || !prefix64_out)
        return UFSECP_ERR_NULL_ARG;
    if (n_tweaks > kMaxGpuBatchN) return UFSECP_ERR_BAD_INPUT;
    if (!clear_out(prefix64_out, n_tweaks, sizeof(uint64_t))) return UFSECP_ERR_BAD_INPUT;
    if (spend_pubkey33[0] != 2 && spend_pubkey33[0] != 3) return UFSECP_ERR_BAD_INPUT;
    if (!compressed_ok(tweak_pubkeys33, n_tweaks)) return UFSECP_ERR_BAD_INPUT;
    return ufs_abi_unsupported(ctx, "bip352_scan_batch");
}

}  // extern "C"
