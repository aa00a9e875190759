// extern "C" ufsecp GPU ABI over RocmBackend.
//
// Argument checking, output clearing and exception mapping follow the
// reference C ABI shim src/cpu/src/ufsecp_gpu_impl.cpp:76-140 (helpers),
// 182-338 (lifecycle + ECDSA), 378-399 (Schnorr), 453-493 (collect),
// 541-558 (MSM), 1136-1158 (error strings).
#include <cstddef>
#include <cstring>
#include <limits>
#include <memory>
#include <new>
#include <string>

#include "../../include/ufsecp/ufsecp_gpu.h"
#include "../../include/ufsecp/ufsecp_gpu_ext.h"
#include "backend.hpp"
#include "msm.hpp"
#include "server.hpp"

using ufs::GpuError;
using ufs::RocmBackend;

namespace {

constexpr size_t kMaxGpuBatchN = size_t{1} << 26;

ufsecp_error_t to_abi(GpuError e) { return static_cast<ufsecp_error_t>(static_cast<int>(e)); }

bool clear_out(void* out, size_t count, size_t stride) {
    if (!out || count == 0 || stride == 0) return true;
    if (count > std::numeric_limits<size_t>::max() / stride) return false;
    std::memset(out, 0, count * stride);
    return true;
}

ufsecp_error_t to_abi_clear(GpuError e, void* out, size_t count, size_t stride) {
    const ufsecp_error_t r = to_abi(e);
    if (r != UFSECP_OK) (void)clear_out(out, count, stride);
    return r;
}

}  // namespace

struct ufsecp_gpu_ctx {
    std::unique_ptr<RocmBackend> backend;
    std::unique_ptr<ufs::MsmEngine> msm;
    uint32_t backend_id;
    uint32_t device_index;
};

struct ufsecp_gpu_server {
    std::unique_ptr<ufs::VerifyServer> impl;
};

#define UFSECP_GPU_CATCH                                           \
    catch (const std::bad_alloc&) { return UFSECP_ERR_GPU_MEMORY; } \
    catch (...) { return UFSECP_ERR_INTERNAL; }

// Records "<op>: not supported by the ROCm backend" as ctx's last error and
// returns UFSECP_ERR_GPU_UNSUPPORTED (used by ufsecp_gpu_unsupported.cpp).
ufsecp_error_t ufs_abi_unsupported(ufsecp_gpu_ctx* ctx, const char* op) {
    std::string m = op;
    m += ": not supported by the ROCm backend";
    ctx->backend->fail(GpuError::Unsupported, m.c_str());
    return UFSECP_ERR_GPU_UNSUPPORTED;
}

extern "C" {

uint32_t ufsecp_gpu_backend_count(uint32_t* backend_ids, uint32_t max_ids) noexcept {
    if (backend_ids && max_ids > 0) backend_ids[0] = UFSECP_GPU_BACKEND_ROCM;
    return 1;
}

const char* ufsecp_gpu_backend_name(uint32_t backend_id) {
    switch (backend_id) {
    case UFSECP_GPU_BACKEND_CUDA: return "CUDA";
    case UFSECP_GPU_BACKEND_OPENCL: return "OpenCL";
    case UFSECP_GPU_BACKEND_METAL: return "Metal";
    case UFSECP_GPU_BACKEND_ROCM: return "ROCm";
    default: return "none";
    }
}

int ufsecp_gpu_is_available(uint32_t backend_id) noexcept {
    if (backend_id != UFSECP_GPU_BACKEND_ROCM) return 0;
    return RocmBackend::device_count() > 0 ? 1 : 0;
}

uint32_t ufsecp_gpu_device_count(uint32_t backend_id) noexcept {
    if (backend_id != UFSECP_GPU_BACKEND_ROCM) return 0;
    const int n = RocmBackend::device_count();
    return n > 0 ? (uint32_t)n : 0u;
}

ufsecp_error_t ufsecp_gpu_device_info(uint32_t backend_id, uint32_t device_index,
                                      ufsecp_gpu_device_info_t* info_out) {
    if (!info_out) return UFSECP_ERR_NULL_ARG;
    if (backend_id != UFSECP_GPU_BACKEND_ROCM) return UFSECP_ERR_GPU_UNAVAILABLE;
    try {
        ufs::DeviceInfoData d;
        const GpuError e = RocmBackend::device_info(device_index, d);
        if (e != GpuError::Ok) return to_abi(e);
        std::memcpy(info_out->name, d.name, sizeof(info_out->name));
        info_out->global_mem_bytes = d.global_mem_bytes;
        info_out->compute_units = d.compute_units;
        info_out->max_clock_mhz = d.max_clock_mhz;
        info_out->max_threads_per_block = d.max_threads_per_block;
        info_out->backend_id = backend_id;
        info_out->device_index = device_index;
        return UFSECP_OK;
    }
    UFSECP_GPU_CATCH
}

ufsecp_error_t ufsecp_gpu_ctx_create(ufsecp_gpu_ctx** ctx_out, uint32_t backend_id, uint32_t device_index) {
    if (!ctx_out) return UFSECP_ERR_NULL_ARG;
    *ctx_out = nullptr;
    try {
        if (backend_id != UFSECP_GPU_BACKEND_ROCM) return UFSECP_ERR_GPU_UNAVAILABLE;
        auto backend = std::make_unique<RocmBackend>();
        const GpuError e = backend->init(device_index);
        if (e != GpuError::Ok) return to_abi(e);
        auto* ctx = new (std::nothrow) ufsecp_gpu_ctx;
        if (!ctx) return UFSECP_ERR_INTERNAL;
        auto msm = std::make_unique<ufs::MsmEngine>(backend.get());
        if (msm->init() != GpuError::Ok) {
            const ufsecp_error_t r = to_abi(backend->last_error());
            delete ctx;
            return r;
        }
        ctx->msm = std::move(msm);
        ctx->backend = std::move(backend);
        ctx->backend_id = backend_id;
        ctx->device_index = device_index;
        *ctx_out = ctx;
        return UFSECP_OK;
    }
    UFSECP_GPU_CATCH
}

void ufsecp_gpu_ctx_destroy(ufsecp_gpu_ctx* ctx) {
    if (!ctx) return;
    ctx->msm.reset();
    ctx->backend.reset();
    delete ctx;
}

int ufsecp_gpu_is_ready(const ufsecp_gpu_ctx* ctx) noexcept {
    if (!ctx || !ctx->backend) return 0;
    return ctx->backend->is_ready() ? 1 : 0;
}

ufsecp_error_t ufsecp_gpu_last_error(const ufsecp_gpu_ctx* ctx) {
    if (!ctx) return UFSECP_ERR_NULL_ARG;
    return to_abi(ctx->backend->last_error());
}

const char* ufsecp_gpu_last_error_msg(const ufsecp_gpu_ctx* ctx) {
    if (!ctx) return "NULL GPU context";
    return ctx->backend->last_error_msg();
}

const char* ufsecp_gpu_error_str(ufsecp_error_t err) {
    switch (err) {
    case UFSECP_OK: return "OK";
    case UFSECP_ERR_NULL_ARG: return "NULL argument";
    case UFSECP_ERR_BAD_KEY: return "invalid private key";
    case UFSECP_ERR_BAD_PUBKEY: return "invalid public key";
    case UFSECP_ERR_BAD_SIG: return "invalid signature";
    case UFSECP_ERR_BAD_INPUT: return "malformed input";
    case UFSECP_ERR_VERIFY_FAIL: return "verification failed";
    case UFSECP_ERR_ARITH: return "arithmetic error";
    case UFSECP_ERR_SELFTEST: return "self-test failed";
    case UFSECP_ERR_INTERNAL: return "internal error";
    case UFSECP_ERR_BUF_TOO_SMALL: return "buffer too small";
    case UFSECP_ERR_GPU_UNAVAILABLE: return "GPU backend unavailable";
    case UFSECP_ERR_GPU_DEVICE: return "GPU device error";
    case UFSECP_ERR_GPU_LAUNCH: return "GPU kernel launch failed";
    case UFSECP_ERR_GPU_MEMORY: return "GPU memory error";
    case UFSECP_ERR_GPU_UNSUPPORTED: return "operation not supported on this GPU backend";
    case UFSECP_ERR_GPU_BACKEND: return "GPU backend driver error";
    case UFSECP_ERR_GPU_QUEUE: return "GPU command queue error";
    default: return "unknown error";
    }
}

ufsecp_error_t ufsecp_gpu_ecdsa_verify_batch(ufsecp_gpu_ctx* ctx, const uint8_t* msg_hashes32,
                                             const uint8_t* pubkeys33, const uint8_t* sigs64, size_t count,
                                             uint8_t* out_results) {
    if (!ctx) return UFSECP_ERR_NULL_ARG;
    if (count == 0) return UFSECP_OK;
    if (count > kMaxGpuBatchN) return UFSECP_ERR_BAD_INPUT;
    if (!clear_out(out_results, count, 1)) return UFSECP_ERR_BAD_INPUT;
    if (!msg_hashes32 || !pubkeys33 || !sigs64 || !out_results) return UFSECP_ERR_NULL_ARG;
    try {
        return to_abi_clear(
            ctx->backend->ecdsa_verify_batch(msg_hashes32, pubkeys33, sigs64, count, out_results, false),
            out_results, count, 1);
    }
    UFSECP_GPU_CATCH
}

ufsecp_error_t ufsecp_gpu_schnorr_verify_batch(ufsecp_gpu_ctx* ctx, const uint8_t* msg_hashes32,
                                               const uint8_t* pubkeys_x32, const uint8_t* sigs64, size_t count,
                                               uint8_t* out_results) {
    if (!ctx) return UFSECP_ERR_NULL_ARG;
    if (count == 0) return UFSECP_OK;
    if (count > kMaxGpuBatchN) return UFSECP_ERR_BAD_INPUT;
    if (!clear_out(out_results, count, 1)) return UFSECP_ERR_BAD_INPUT;
    if (!msg_hashes32 || !pubkeys_x32 || !sigs64 || !out_results) return UFSECP_ERR_NULL_ARG;
    try {
        return to_abi_clear(
            ctx->backend->schnorr_verify_batch(msg_hashes32, pubkeys_x32, sigs64, count, out_results, false),
            out_results, count, 1);
    }
    UFSECP_GPU_CATCH
}

// Argument checks in the reference's order (ufsecp_gpu_impl.cpp:340-359):
// NULL ctx, count 0, count cap and stride, NULL buffers, then the cleared
// output.
ufsecp_error_t ufsecp_gpu_ecdsa_verify_opaque_rows(ufsecp_gpu_ctx* ctx, const uint8_t* rows, size_t stride,
                                                   size_t count, uint8_t* out_results) {
    if (!ctx) return UFSECP_ERR_NULL_ARG;
    if (count == 0) return UFSECP_OK;
    if (count > kMaxGpuBatchN || stride < 129u) return UFSECP_ERR_BAD_INPUT;
    if (!rows || !out_results) return UFSECP_ERR_NULL_ARG;
    if (!clear_out(out_results, count, 1)) return UFSECP_ERR_BAD_INPUT;
    try {
        return to_abi_clear(ctx->backend->ecdsa_verify_opaque_rows(rows, stride, count, out_results), out_results,
                            count, 1);
    }
    UFSECP_GPU_CATCH
}

ufsecp_error_t ufsecp_gpu_ecdsa_verify_lbtc_rows(ufsecp_gpu_ctx* ctx, const uint8_t* rows, size_t stride,
                                                 size_t count, uint8_t* out_results) {
    if (!ctx) return UFSECP_ERR_NULL_ARG;
    if (count == 0) return UFSECP_OK;
    if (count > kMaxGpuBatchN || stride < 129u) return UFSECP_ERR_BAD_INPUT;
    if (!rows || !out_results) return UFSECP_ERR_NULL_ARG;
    return ufsecp_gpu_ecdsa_verify_opaque_rows(ctx, rows, stride, count, out_results);
}

ufsecp_error_t ufsecp_gpu_ecdsa_verify_lbtc_columns(ufsecp_gpu_ctx* ctx, const uint8_t* digests32,
                                                    const uint8_t* pubkeys33, const uint8_t* sigs64, size_t count,
                                                    uint8_t* out_results) {
    if (!ctx) return UFSECP_ERR_NULL_ARG;
    if (count == 0) return UFSECP_OK;
    if (count > kMaxGpuBatchN) return UFSECP_ERR_BAD_INPUT;
    if (!clear_out(out_results, count, 1)) return UFSECP_ERR_BAD_INPUT;
    if (!digests32 || !pubkeys33 || !sigs64 || !out_results) return UFSECP_ERR_NULL_ARG;
    try {
        return to_abi_clear(ctx->backend->ecdsa_verify_opaque_columns(digests32, pubkeys33, sigs64, count, out_results),
                            out_results, count, 1);
    }
    UFSECP_GPU_CATCH
}

ufsecp_error_t ufsecp_gpu_schnorr_verify_lbtc_columns(ufsecp_gpu_ctx* ctx, const uint8_t* digests32,
                                                      const uint8_t* xonly32, const uint8_t* sigs64, size_t count,
                                                      uint8_t* out_results) {
    if (!ctx) return UFSECP_ERR_NULL_ARG;
    if (count == 0) return UFSECP_OK;
    if (count > kMaxGpuBatchN) return UFSECP_ERR_BAD_INPUT;
    if (!clear_out(out_results, count, 1)) return UFSECP_ERR_BAD_INPUT;
    if (!digests32 || !xonly32 || !sigs64 || !out_results) return UFSECP_ERR_NULL_ARG;
    try {
        return to_abi_clear(ctx->backend->schnorr_verify_batch(digests32, xonly32, sigs64, count, out_results, false),
                            out_results, count, 1);
    }
    UFSECP_GPU_CATCH
}

// Public-data key operations; argument checks as ufsecp_gpu_impl.cpp:560-640.
ufsecp_error_t ufsecp_gpu_pubkey_validate(ufsecp_gpu_ctx* ctx, const uint8_t* pubkeys33, size_t n,
                                          uint8_t* results) {
    if (!ctx) return UFSECP_ERR_NULL_ARG;
    if (n == 0) return UFSECP_OK;
    if (n > kMaxGpuBatchN) return UFSECP_ERR_BAD_INPUT;
    if (!clear_out(results, n, 1)) return UFSECP_ERR_BAD_INPUT;
    if (!pubkeys33 || !results) return UFSECP_ERR_NULL_ARG;
    try {
        return to_abi_clear(ctx->backend->pubkey_validate(pubkeys33, n, results), results, n, 1);
    }
    UFSECP_GPU_CATCH
}

ufsecp_error_t ufsecp_gpu_xonly_validate(ufsecp_gpu_ctx* ctx, const uint8_t* keys32, size_t n, uint8_t* results) {
    if (!ctx) return UFSECP_ERR_NULL_ARG;
    if (n == 0) return UFSECP_OK;
    if (n > kMaxGpuBatchN) return UFSECP_ERR_BAD_INPUT;
    if (!clear_out(results, n, 1)) return UFSECP_ERR_BAD_INPUT;
    if (!keys32 || !results) return UFSECP_ERR_NULL_ARG;
    try {
        return to_abi_clear(ctx->backend->xonly_validate(keys32, n, results), results, n, 1);
    }
    UFSECP_GPU_CATCH
}

ufsecp_error_t ufsecp_gpu_commitment_verify(ufsecp_gpu_ctx* ctx, const uint8_t* internal_x32, const uint8_t* tweak32,
                                            const uint8_t* tweaked_x32, const uint8_t* parity, size_t n,
                                            uint8_t* results) {
    if (!ctx) return UFSECP_ERR_NULL_ARG;
    if (n == 0) return UFSECP_OK;
    if (n > kMaxGpuBatchN) return UFSECP_ERR_BAD_INPUT;
    if (!clear_out(results, n, 1)) return UFSECP_ERR_BAD_INPUT;
    if (!internal_x32 || !tweak32 || !tweaked_x32 || !parity || !results) return UFSECP_ERR_NULL_ARG;
    try {
        return to_abi_clear(ctx->backend->commitment_verify(internal_x32, tweak32, tweaked_x32, parity, n, results),
                            results, n, 1);
    }
    UFSECP_GPU_CATCH
}

// Argument checks as ufsecp_gpu_impl.cpp:756-786 (recovery ids checked for
// the whole batch before dispatch; outputs cleared on any non-OK).
ufsecp_error_t ufsecp_gpu_ecrecover_batch(ufsecp_gpu_ctx* ctx, const uint8_t* msg_hashes32, const uint8_t* sigs64,
                                          const int* recids, size_t count, uint8_t* out_pubkeys33,
                                          uint8_t* out_valid) {
    if (!ctx) return UFSECP_ERR_NULL_ARG;
    if (count == 0) return UFSECP_OK;
    if (count > kMaxGpuBatchN) return UFSECP_ERR_BAD_INPUT;
    if (!clear_out(out_pubkeys33, count, 33)) return UFSECP_ERR_BAD_INPUT;
    if (!clear_out(out_valid, count, 1)) return UFSECP_ERR_BAD_INPUT;
    if (!msg_hashes32 || !sigs64 || !recids || !out_pubkeys33 || !out_valid) return UFSECP_ERR_NULL_ARG;
    for (size_t i = 0; i < count; ++i)
        if (recids[i] < 0 || recids[i] > 3) return UFSECP_ERR_BAD_INPUT;
    try {
        const ufsecp_error_t r =
            to_abi(ctx->backend->ecrecover_batch(msg_hashes32, sigs64, recids, count, out_pubkeys33, out_valid));
        if (r != UFSECP_OK) {
            (void)clear_out(out_pubkeys33, count, 33);
            (void)clear_out(out_valid, count, 1);
        }
        return r;
    }
    UFSECP_GPU_CATCH
}

ufsecp_error_t ufsecp_gpu_ecdsa_verify_collect(ufsecp_gpu_ctx* ctx, const uint8_t* msg_hashes32,
                                               const uint8_t* pubkeys33, const uint8_t* sigs64, size_t count,
                                               uint8_t* key_buffer) {
    if (!ctx) return UFSECP_ERR_NULL_ARG;
    if (count == 0) return UFSECP_OK;
    if (!msg_hashes32 || !pubkeys33 || !sigs64 || !key_buffer) return UFSECP_ERR_NULL_ARG;
    if (count > kMaxGpuBatchN) return UFSECP_ERR_BAD_INPUT;
    try {
        return to_abi(ctx->backend->ecdsa_verify_batch(msg_hashes32, pubkeys33, sigs64, count, key_buffer, true));
    }
    UFSECP_GPU_CATCH
}

ufsecp_error_t ufsecp_gpu_schnorr_verify_collect(ufsecp_gpu_ctx* ctx, const uint8_t* msg_hashes32,
                                                 const uint8_t* pubkeys_x32, const uint8_t* sigs64, size_t count,
                                                 uint8_t* key_buffer) {
    if (!ctx) return UFSECP_ERR_NULL_ARG;
    if (count == 0) return UFSECP_OK;
    if (!msg_hashes32 || !pubkeys_x32 || !sigs64 || !key_buffer) return UFSECP_ERR_NULL_ARG;
    if (count > kMaxGpuBatchN) return UFSECP_ERR_BAD_INPUT;
    try {
        return to_abi(
            ctx->backend->schnorr_verify_batch(msg_hashes32, pubkeys_x32, sigs64, count, key_buffer, true));
    }
    UFSECP_GPU_CATCH
}

ufsecp_error_t ufsecp_gpu_msm(ufsecp_gpu_ctx* ctx, const uint8_t* scalars32, const uint8_t* points33, size_t n,
                              uint8_t* out_result33) {
    if (!ctx) return UFSECP_ERR_NULL_ARG;
    if (n == 0) return UFSECP_OK;
    if (n > kMaxGpuBatchN) return UFSECP_ERR_BAD_INPUT;
    if (!clear_out(out_result33, 1, 33)) return UFSECP_ERR_BAD_INPUT;
    if (!scalars32 || !points33 || !out_result33) return UFSECP_ERR_NULL_ARG;
    try {
        return to_abi_clear(ctx->msm->msm_host(scalars32, points33, n, out_result33, 0), out_result33, 1, 33);
    }
    UFSECP_GPU_CATCH
}

// ---- extensions ------------------------------------------------------------

ufsecp_error_t ufsecp_gpu_ext_ecdsa_verify_device(ufsecp_gpu_ctx* ctx, const uint8_t* d_msg_hashes32,
                                                  const uint8_t* d_pubkeys33, const uint8_t* d_sigs64, size_t count,
                                                  uint8_t* d_out_results, void* stream) {
    if (!ctx) return UFSECP_ERR_NULL_ARG;
    if (count == 0) return UFSECP_OK;
    if (count > kMaxGpuBatchN) return UFSECP_ERR_BAD_INPUT;
    if (!d_msg_hashes32 || !d_pubkeys33 || !d_sigs64 || !d_out_results) return UFSECP_ERR_NULL_ARG;
    try {
        return to_abi(ctx->backend->ecdsa_verify_device(d_msg_hashes32, d_pubkeys33, d_sigs64, count, d_out_results,
                                                        (hipStream_t)stream));
    }
    UFSECP_GPU_CATCH
}

ufsecp_error_t ufsecp_gpu_ext_schnorr_verify_device(ufsecp_gpu_ctx* ctx, const uint8_t* d_msg_hashes32,
                                                    const uint8_t* d_pubkeys_x32, const uint8_t* d_sigs64,
                                                    size_t count, uint8_t* d_out_results, void* stream) {
    if (!ctx) return UFSECP_ERR_NULL_ARG;
    if (count == 0) return UFSECP_OK;
    if (count > kMaxGpuBatchN) return UFSECP_ERR_BAD_INPUT;
    if (!d_msg_hashes32 || !d_pubkeys_x32 || !d_sigs64 || !d_out_results) return UFSECP_ERR_NULL_ARG;
    try {
        return to_abi(ctx->backend->schnorr_verify_device(d_msg_hashes32, d_pubkeys_x32, d_sigs64, count,
                                                          d_out_results, (hipStream_t)stream));
    }
    UFSECP_GPU_CATCH
}

static_assert(sizeof(ufsecp_gpu_verify_batch_t) == sizeof(ufs::BatchDesc), "batch descriptor layout");
static_assert(offsetof(ufsecp_gpu_verify_batch_t, out_results) == offsetof(ufs::BatchDesc, out),
              "batch descriptor layout");
static_assert(offsetof(ufsecp_gpu_verify_batch_t, row_stride) == offsetof(ufs::BatchDesc, stride),
              "batch descriptor layout");
static_assert(UFSECP_GPU_KIND_ECDSA_OPAQUE == ufs::kKindEcdsaOpaque, "batch kinds");

static ufsecp_error_t check_batches(const ufsecp_gpu_verify_batch_t* b, size_t nb) {
    if (nb > 0 && !b) return UFSECP_ERR_NULL_ARG;
    size_t total = 0;
    for (size_t i = 0; i < nb; ++i) {
        if (b[i].kind > UFSECP_GPU_KIND_ECDSA_OPAQUE || b[i].reserved != 0) return UFSECP_ERR_BAD_INPUT;
        if (b[i].row_stride != 0 && (b[i].kind != UFSECP_GPU_KIND_ECDSA_OPAQUE || b[i].row_stride < 129))
            return UFSECP_ERR_BAD_INPUT;
        if (b[i].count == 0) continue;
        if (!b[i].msg_hashes32 || !b[i].pubkeys || !b[i].sigs64 || !b[i].out_results) return UFSECP_ERR_NULL_ARG;
        total += b[i].count;
        if (total > kMaxGpuBatchN) return UFSECP_ERR_BAD_INPUT;
    }
    return UFSECP_OK;
}

ufsecp_error_t ufsecp_gpu_ext_verify_multi_device(ufsecp_gpu_ctx* ctx, const ufsecp_gpu_verify_batch_t* batches,
                                                  size_t nbatches, int use_graph, void* stream) {
    if (!ctx) return UFSECP_ERR_NULL_ARG;
    const ufsecp_error_t c = check_batches(batches, nbatches);
    if (c != UFSECP_OK) return c;
    try {
        return to_abi(ctx->backend->verify_multi_device(reinterpret_cast<const ufs::BatchDesc*>(batches), nbatches,
                                                        use_graph != 0, (hipStream_t)stream));
    }
    UFSECP_GPU_CATCH
}

ufsecp_error_t ufsecp_gpu_ext_verify_each_device(ufsecp_gpu_ctx* ctx, const ufsecp_gpu_verify_batch_t* batches,
                                                 size_t nbatches, void* stream) {
    if (!ctx) return UFSECP_ERR_NULL_ARG;
    const ufsecp_error_t c = check_batches(batches, nbatches);
    if (c != UFSECP_OK) return c;
    try {
        return to_abi(ctx->backend->verify_each_device(reinterpret_cast<const ufs::BatchDesc*>(batches), nbatches,
                                                       (hipStream_t)stream));
    }
    UFSECP_GPU_CATCH
}

ufsecp_error_t ufsecp_gpu_ext_msm(ufsecp_gpu_ctx* ctx, const uint8_t* scalars32, const uint8_t* points33, size_t n,
                                  uint8_t* out_result33, int window_bits) {
    if (!ctx) return UFSECP_ERR_NULL_ARG;
    if (n == 0) return UFSECP_OK;
    if (n > kMaxGpuBatchN || window_bits < 0 || window_bits == 1 || window_bits > 20) return UFSECP_ERR_BAD_INPUT;
    if (!clear_out(out_result33, 1, 33)) return UFSECP_ERR_BAD_INPUT;
    if (!scalars32 || !points33 || !out_result33) return UFSECP_ERR_NULL_ARG;
    try {
        return to_abi_clear(ctx->msm->msm_host(scalars32, points33, n, out_result33, window_bits), out_result33, 1,
                            33);
    }
    UFSECP_GPU_CATCH
}

ufsecp_error_t ufsecp_gpu_ext_msm_device(ufsecp_gpu_ctx* ctx, const uint8_t* d_scalars32, const uint8_t* d_points33,
                                         size_t n, uint8_t* d_out33, int32_t* d_status, int window_bits,
                                         void* stream) {
    if (!ctx) return UFSECP_ERR_NULL_ARG;
    if (n == 0) return UFSECP_OK;
    if (n > kMaxGpuBatchN || window_bits < 0 || window_bits == 1 || window_bits > 20) return UFSECP_ERR_BAD_INPUT;
    if (!d_scalars32 || !d_points33 || !d_out33 || !d_status) return UFSECP_ERR_NULL_ARG;
    try {
        return to_abi(ctx->msm->msm_device(d_scalars32, d_points33, false, n, d_out33, d_status, window_bits,
                                           (hipStream_t)stream));
    }
    UFSECP_GPU_CATCH
}

ufsecp_error_t ufsecp_gpu_ext_msm_device_noalloc(ufsecp_gpu_ctx* ctx, const uint8_t* d_scalars32,
                                                 const uint8_t* d_points33, size_t n, uint8_t* d_out33,
                                                 int32_t* d_status, int window_bits, void* stream) {
    if (!ctx) return UFSECP_ERR_NULL_ARG;
    if (n == 0) return UFSECP_OK;
    if (window_bits < 0 || window_bits == 1 || window_bits > 20) return UFSECP_ERR_BAD_INPUT;
    if (!d_scalars32 || !d_points33 || !d_out33 || !d_status) return UFSECP_ERR_NULL_ARG;
    try {
        return to_abi(ctx->msm->msm_device_noalloc(d_scalars32, d_points33, n, d_out33, d_status, window_bits,
                                                   (hipStream_t)stream));
    }
    UFSECP_GPU_CATCH
}

uint64_t ufsecp_gpu_ext_buffer_alloc_count(void) { return ufs::buffer_alloc_count(); }

int64_t ufsecp_gpu_ext_buffer_live_count(void) { return ufs::buffer_live_count(); }

ufsecp_error_t ufsecp_gpu_ext_msm_affine_device(ufsecp_gpu_ctx* ctx, const uint8_t* d_scalars32,
                                                const uint8_t* d_points64, size_t n, uint8_t* d_out33,
                                                int32_t* d_status, int window_bits, void* stream) {
    if (!ctx) return UFSECP_ERR_NULL_ARG;
    if (n == 0) return UFSECP_OK;
    if (n > kMaxGpuBatchN || window_bits < 0 || window_bits == 1 || window_bits > 20) return UFSECP_ERR_BAD_INPUT;
    if (!d_scalars32 || !d_points64 || !d_out33 || !d_status) return UFSECP_ERR_NULL_ARG;
    try {
        return to_abi(ctx->msm->msm_device(d_scalars32, d_points64, true, n, d_out33, d_status, window_bits,
                                           (hipStream_t)stream));
    }
    UFSECP_GPU_CATCH
}

ufsecp_error_t ufsecp_gpu_ext_schnorr_batch_check(ufsecp_gpu_ctx* ctx, const uint8_t* msg_hashes32,
                                                 const uint8_t* pubkeys_x32, const uint8_t* sigs64, size_t count,
                                                 int* all_valid) {
    if (!ctx) return UFSECP_ERR_NULL_ARG;
    if (!all_valid) return UFSECP_ERR_NULL_ARG;
    *all_valid = 0;
    if (count == 0) return UFSECP_OK;
    if (count > kMaxGpuBatchN) return UFSECP_ERR_BAD_INPUT;
    if (!msg_hashes32 || !pubkeys_x32 || !sigs64) return UFSECP_ERR_NULL_ARG;
    try {
        return to_abi(ctx->msm->schnorr_batch_check(msg_hashes32, pubkeys_x32, sigs64, count, all_valid));
    }
    UFSECP_GPU_CATCH
}

ufsecp_error_t ufsecp_gpu_ext_alloc(ufsecp_gpu_ctx* ctx, size_t bytes, void** d_ptr_out) {
    if (!ctx || !d_ptr_out) return UFSECP_ERR_NULL_ARG;
    *d_ptr_out = nullptr;
    if (bytes == 0) return UFSECP_ERR_BAD_INPUT;
    if (ufs::dev_malloc(d_ptr_out, bytes) != hipSuccess) {
        *d_ptr_out = nullptr;
        return UFSECP_ERR_GPU_MEMORY;
    }
    return UFSECP_OK;
}

ufsecp_error_t ufsecp_gpu_ext_free(ufsecp_gpu_ctx* ctx, void* d_ptr) {
    if (!ctx) return UFSECP_ERR_NULL_ARG;
    if (!d_ptr) return UFSECP_OK;
    (void)hipStreamSynchronize(ctx->backend->stream());
    return ufs::dev_free(d_ptr) == hipSuccess ? UFSECP_OK : UFSECP_ERR_GPU_MEMORY;
}

ufsecp_error_t ufsecp_gpu_ext_copy_to_device(ufsecp_gpu_ctx* ctx, void* d_dst, const void* h_src, size_t bytes) {
    if (!ctx || (bytes && (!d_dst || !h_src))) return UFSECP_ERR_NULL_ARG;
    hipStream_t st = ctx->backend->stream();
    if (hipMemcpyAsync(d_dst, h_src, bytes, hipMemcpyHostToDevice, st) != hipSuccess ||
        hipStreamSynchronize(st) != hipSuccess)
        return UFSECP_ERR_GPU_MEMORY;
    return UFSECP_OK;
}

ufsecp_error_t ufsecp_gpu_ext_copy_to_host(ufsecp_gpu_ctx* ctx, void* h_dst, const void* d_src, size_t bytes) {
    if (!ctx || (bytes && (!d_src || !h_dst))) return UFSECP_ERR_NULL_ARG;
    hipStream_t st = ctx->backend->stream();
    if (hipMemcpyAsync(h_dst, d_src, bytes, hipMemcpyDeviceToHost, st) != hipSuccess ||
        hipStreamSynchronize(st) != hipSuccess)
        return UFSECP_ERR_GPU_MEMORY;
    return UFSECP_OK;
}

ufsecp_error_t ufsecp_gpu_ext_event_create(ufsecp_gpu_ctx* ctx, void** event_out) {
    if (!ctx || !event_out) return UFSECP_ERR_NULL_ARG;
    hipEvent_t ev = nullptr;
    if (hipEventCreate(&ev) != hipSuccess) return UFSECP_ERR_GPU_BACKEND;
    *event_out = (void*)ev;
    return UFSECP_OK;
}

ufsecp_error_t ufsecp_gpu_ext_event_record(ufsecp_gpu_ctx* ctx, void* event, void* stream) {
    if (!ctx || !event) return UFSECP_ERR_NULL_ARG;
    hipStream_t st = stream ? (hipStream_t)stream : ctx->backend->stream();
    return hipEventRecord((hipEvent_t)event, st) == hipSuccess ? UFSECP_OK : UFSECP_ERR_GPU_BACKEND;
}

ufsecp_error_t ufsecp_gpu_ext_event_elapsed(void* start, void* stop, float* ms_out) {
    if (!start || !stop || !ms_out) return UFSECP_ERR_NULL_ARG;
    if (hipEventSynchronize((hipEvent_t)stop) != hipSuccess) return UFSECP_ERR_GPU_BACKEND;
    return hipEventElapsedTime(ms_out, (hipEvent_t)start, (hipEvent_t)stop) == hipSuccess ? UFSECP_OK
                                                                                          : UFSECP_ERR_GPU_BACKEND;
}

void ufsecp_gpu_ext_event_destroy(void* event) {
    if (event) (void)hipEventDestroy((hipEvent_t)event);
}

void* ufsecp_gpu_ext_stream(ufsecp_gpu_ctx* ctx) {
    if (!ctx) return nullptr;
    return (void*)ctx->backend->stream();
}

uint32_t ufsecp_gpu_ext_verify_lanes(ufsecp_gpu_ctx* ctx, uint32_t kind) {
    if (!ctx || kind > 2) return 0;
    const int k = kind == 0 ? 0 : kind == 1 ? 1 : 2;
    return ctx->backend->verify_grid(k) * (uint32_t)ufs::kVerifyBlock;
}

ufsecp_error_t ufsecp_gpu_ext_stream_create(ufsecp_gpu_ctx* ctx, void** stream_out) {
    if (!ctx || !stream_out) return UFSECP_ERR_NULL_ARG;
    *stream_out = nullptr;
    if (hipSetDevice(ctx->backend->device()) != hipSuccess) return UFSECP_ERR_GPU_DEVICE;
    hipStream_t st = nullptr;
    if (hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess) return UFSECP_ERR_GPU_QUEUE;
    *stream_out = (void*)st;
    return UFSECP_OK;
}

ufsecp_error_t ufsecp_gpu_ext_stream_synchronize(void* stream) {
    if (!stream) return UFSECP_ERR_NULL_ARG;
    return hipStreamSynchronize((hipStream_t)stream) == hipSuccess ? UFSECP_OK : UFSECP_ERR_GPU_BACKEND;
}

void ufsecp_gpu_ext_stream_destroy(void* stream) {
    if (!stream) return;
    (void)hipStreamSynchronize((hipStream_t)stream);
    (void)hipStreamDestroy((hipStream_t)stream);
}

ufsecp_error_t ufsecp_gpu_ext_synchronize(ufsecp_gpu_ctx* ctx) {
    if (!ctx) return UFSECP_ERR_NULL_ARG;
    try {
        return to_abi(ctx->backend->synchronize());
    }
    UFSECP_GPU_CATCH
}

ufsecp_error_t ufsecp_gpu_ext_selftest(ufsecp_gpu_ctx* ctx, int op, const uint32_t* a, const uint32_t* b,
                                       uint32_t* out, size_t count) {
    if (!ctx || !a || !b || !out) return UFSECP_ERR_NULL_ARG;
    if (count == 0) return UFSECP_OK;
    if (count > (size_t{1} << 20)) return UFSECP_ERR_BAD_INPUT;
    try {
        return to_abi(ctx->backend->selftest(op, a, b, out, count));
    }
    UFSECP_GPU_CATCH
}

ufsecp_error_t ufsecp_gpu_ext_host_alloc(ufsecp_gpu_ctx* ctx, size_t bytes, void** h_ptr_out) {
    if (!ctx || !h_ptr_out) return UFSECP_ERR_NULL_ARG;
    *h_ptr_out = nullptr;
    if (bytes == 0) return UFSECP_ERR_BAD_INPUT;
    if (ufs::pinned_malloc(h_ptr_out, bytes, hipHostMallocMapped | hipHostMallocPortable) != hipSuccess) {
        *h_ptr_out = nullptr;
        return UFSECP_ERR_GPU_MEMORY;
    }
    // kernels read this memory through the host address itself
    void* dptr = nullptr;
    if (hipHostGetDevicePointer(&dptr, *h_ptr_out, 0) != hipSuccess || dptr != *h_ptr_out) {
        (void)ufs::pinned_free(*h_ptr_out);
        *h_ptr_out = nullptr;
        return UFSECP_ERR_GPU_UNSUPPORTED;
    }
    return UFSECP_OK;
}

ufsecp_error_t ufsecp_gpu_ext_host_free(ufsecp_gpu_ctx* ctx, void* h_ptr) {
    if (!ctx) return UFSECP_ERR_NULL_ARG;
    if (!h_ptr) return UFSECP_OK;
    return ufs::pinned_free(h_ptr) == hipSuccess ? UFSECP_OK : UFSECP_ERR_GPU_MEMORY;
}

ufsecp_error_t ufsecp_gpu_ext_server_create(ufsecp_gpu_ctx* ctx, const ufsecp_gpu_server_config_t* cfg,
                                            ufsecp_gpu_server** server_out) {
    if (!ctx || !server_out) return UFSECP_ERR_NULL_ARG;
    *server_out = nullptr;
    try {
        if (!ctx->backend->is_ready()) return UFSECP_ERR_GPU_DEVICE;
        ufs::ServerConfig c;
        c.groups = (unsigned)ctx->backend->tuning().server_groups;
        if (cfg) {
            if (cfg->hbm_budget_bytes) c.hbm_budget = (size_t)cfg->hbm_budget_bytes;
            if (cfg->inflight_groups) c.groups = cfg->inflight_groups;
            c.cache_log2 = cfg->key_cache_log2;
            if (cfg->max_group_items) c.max_group_items = cfg->max_group_items;
            if (c.groups > 4 || c.cache_log2 > 26 || c.max_group_items > (1u << 24)) return UFSECP_ERR_BAD_INPUT;
        }
        auto s = std::make_unique<ufsecp_gpu_server>();
        s->impl = std::make_unique<ufs::VerifyServer>(ctx->backend.get());
        const GpuError e = s->impl->start(c);
        if (e != GpuError::Ok) return to_abi(e);
        *server_out = s.release();
        return UFSECP_OK;
    }
    UFSECP_GPU_CATCH
}

void ufsecp_gpu_ext_server_destroy(ufsecp_gpu_server* server) { delete server; }

ufsecp_error_t ufsecp_gpu_ext_server_submit(ufsecp_gpu_server* server, uint32_t kind, const uint8_t* msg_hashes32,
                                            const uint8_t* pubkeys, const uint8_t* sigs64, size_t count,
                                            uint8_t* out_results, uint64_t* request_id) {
    if (!server || !request_id) return UFSECP_ERR_NULL_ARG;
    *request_id = 0;
    if (kind > 1 || count == 0) return UFSECP_ERR_BAD_INPUT;
    if (!msg_hashes32 || !pubkeys || !sigs64 || !out_results) return UFSECP_ERR_NULL_ARG;
    try {
        return to_abi(
            server->impl->submit(kind, msg_hashes32, pubkeys, sigs64, count, out_results, false, request_id));
    }
    UFSECP_GPU_CATCH
}

ufsecp_error_t ufsecp_gpu_ext_server_submit_mapped(ufsecp_gpu_server* server, uint32_t kind,
                                                   const uint8_t* msg_hashes32, const uint8_t* pubkeys,
                                                   const uint8_t* sigs64, size_t count, uint8_t* out_results,
                                                   uint64_t* request_id) {
    if (!server || !request_id) return UFSECP_ERR_NULL_ARG;
    *request_id = 0;
    if (kind > 1 || count == 0) return UFSECP_ERR_BAD_INPUT;
    if (!msg_hashes32 || !pubkeys || !sigs64 || !out_results) return UFSECP_ERR_NULL_ARG;
    try {
        return to_abi(
            server->impl->submit(kind, msg_hashes32, pubkeys, sigs64, count, out_results, true, request_id));
    }
    UFSECP_GPU_CATCH
}

int ufsecp_gpu_ext_server_poll(ufsecp_gpu_server* server, uint64_t request_id) {
    if (!server) return -1;
    return server->impl->poll(request_id);
}

ufsecp_error_t ufsecp_gpu_ext_server_wait(ufsecp_gpu_server* server, uint64_t request_id) {
    if (!server) return UFSECP_ERR_NULL_ARG;
    try {
        return to_abi(server->impl->wait(request_id));
    }
    UFSECP_GPU_CATCH
}

ufsecp_error_t ufsecp_gpu_ext_server_drain(ufsecp_gpu_server* server) {
    if (!server) return UFSECP_ERR_NULL_ARG;
    try {
        return to_abi(server->impl->drain());
    }
    UFSECP_GPU_CATCH
}

ufsecp_error_t ufsecp_gpu_ext_server_stats(ufsecp_gpu_server* server, ufsecp_gpu_server_stats_t* out, int reset) {
    if (!server || !out) return UFSECP_ERR_NULL_ARG;
    ufs::ServerStats s;
    server->impl->stats(s, reset != 0);
    out->requests_completed = s.requests_completed;
    out->items_completed = s.items_completed;
    out->groups_launched = s.groups_launched;
    out->key_lookups = s.key_lookups;
    out->key_cache_hits = s.key_cache_hits;
    out->latency_p50_us = s.latency_p50_us;
    out->latency_p99_us = s.latency_p99_us;
    out->latency_mean_us = s.latency_mean_us;
    out->latency_max_us = s.latency_max_us;
    out->busy_seconds = s.busy_seconds;
    return UFSECP_OK;
}

}  // extern "C"
