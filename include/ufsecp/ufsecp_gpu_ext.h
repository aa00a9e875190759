/* MI355X extensions to the ufsecp GPU C ABI.
 *
 * None of these entry points has a counterpart in the reference
 * (include/ufsecp/ufsecp_gpu.h); they add what its synchronous,
 * one-batch-per-call surface cannot express:
 *   - device-resident variants that take device pointers and a HIP stream and
 *     return without synchronising (inputs already in HBM, caller-owned
 *     ordering), used by graph capture and by callers that keep data on the
 *     device;
 *   - arithmetic self-test hooks used by the parity tests.
 * The existing ufsecp_gpu_* entry points are unchanged by this header. */
#ifndef UFSECP_GPU_EXT_H
#define UFSECP_GPU_EXT_H

#include "ufsecp_gpu.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Device-pointer ECDSA verify: same strides and verdict bytes as
 * ufsecp_gpu_ecdsa_verify_batch, enqueued on `stream` (a hipStream_t, NULL =
 * the context's own stream).  Does not synchronise. */
UFSECP_API ufsecp_error_t ufsecp_gpu_ext_ecdsa_verify_device(ufsecp_gpu_ctx* ctx, const uint8_t* d_msg_hashes32,
                                                             const uint8_t* d_pubkeys33, const uint8_t* d_sigs64,
                                                             size_t count, uint8_t* d_out_results, void* stream);

/* Device-pointer BIP-340 verify, as above. */
UFSECP_API ufsecp_error_t ufsecp_gpu_ext_schnorr_verify_device(ufsecp_gpu_ctx* ctx, const uint8_t* d_msg_hashes32,
                                                               const uint8_t* d_pubkeys_x32, const uint8_t* d_sigs64,
                                                               size_t count, uint8_t* d_out_results, void* stream);

/* The context's HIP stream (hipStream_t) for callers that want to order their
 * own work after the library's. */
UFSECP_API void* ufsecp_gpu_ext_stream(ufsecp_gpu_ctx* ctx);

/* Synchronise the context's stream. */
UFSECP_API ufsecp_error_t ufsecp_gpu_ext_synchronize(ufsecp_gpu_ctx* ctx);

/* Arithmetic self-test: runs device operation `op` (field/scalar/point
 * primitives, codes in the library's test suite) on `count` records of 16
 * 32-bit words from a and b, writing 32-word records to out.  Host buffers. */
UFSECP_API ufsecp_error_t ufsecp_gpu_ext_selftest(ufsecp_gpu_ctx* ctx, int op, const uint32_t* a, const uint32_t* b,
                                                  uint32_t* out, size_t count);

#ifdef __cplusplus
}
#endif

#endif /* UFSECP_GPU_EXT_H */
