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

/* One batch of a multi-batch verification (all pointers are device
 * addresses).  Kinds:
 *   UFSECP_GPU_KIND_ECDSA         as ufsecp_gpu_ecdsa_verify_batch
 *   UFSECP_GPU_KIND_SCHNORR       as ufsecp_gpu_schnorr_verify_batch
 *   UFSECP_GPU_KIND_ECDSA_OPAQUE  as ufsecp_gpu_ecdsa_verify_lbtc_columns
 *                                 (opaque little-endian r, s; high-S accepted)
 * With row_stride 0 the inputs are packed columns: msgs 32 B, pubkeys 33 B
 * (ECDSA kinds) or 32 B (BIP-340), sigs 64 B, out 1 B per item.  A non-zero
 * row_stride (>= 129, opaque kind only) is the byte distance between items of
 * all three inputs, as for interleaved rows (msg_hashes32 = rows,
 * pubkeys = rows + 32, sigs64 = rows + 65).  `reserved` must be zero. */
#define UFSECP_GPU_KIND_ECDSA        0u
#define UFSECP_GPU_KIND_SCHNORR      1u
#define UFSECP_GPU_KIND_ECDSA_OPAQUE 2u
typedef struct {
    uint32_t kind;
    uint32_t count;
    const uint8_t* msg_hashes32;
    const uint8_t* pubkeys;
    const uint8_t* sigs64;
    uint8_t* out_results;
    uint32_t row_stride;
    uint32_t reserved;
} ufsecp_gpu_verify_batch_t;

/* Verify `nbatches` batches of uneven sizes and mixed kinds in ONE persistent
 * launch: every wave pulls 64-item tiles from a device work queue.  With
 * use_graph != 0 the launch is replayed from a captured hipGraph (captured on
 * first use, re-captured only when internal tables are reallocated).
 * Asynchronous on `stream` (NULL = context stream).  Verdicts are identical to
 * the per-batch calls. */
UFSECP_API ufsecp_error_t ufsecp_gpu_ext_verify_multi_device(ufsecp_gpu_ctx* ctx,
                                                             const ufsecp_gpu_verify_batch_t* batches,
                                                             size_t nbatches, int use_graph, void* stream);

/* The plain form of the above: one kernel launch per batch (for comparison). */
UFSECP_API ufsecp_error_t ufsecp_gpu_ext_verify_each_device(ufsecp_gpu_ctx* ctx,
                                                            const ufsecp_gpu_verify_batch_t* batches,
                                                            size_t nbatches, void* stream);

/* ufsecp_gpu_msm with an explicit Pippenger window width c (2..20; 0 = the
 * size-based default).  Same inputs, outputs and error codes. */
UFSECP_API ufsecp_error_t ufsecp_gpu_ext_msm(ufsecp_gpu_ctx* ctx, const uint8_t* scalars32, const uint8_t* points33,
                                             size_t n, uint8_t* out_result33, int window_bits);

/* Device-pointer MSM, asynchronous on `stream`: writes the 33-byte result to
 * d_out33 and the ABI status code (0, BAD_INPUT, BAD_PUBKEY or ERR_ARITH; or
 * INTERNAL if a device-side bucket bound check ever fired) to *d_status, both
 * device addresses.  Calls on different streams are ordered one after
 * another (they share the context's MSM scratch). */
UFSECP_API ufsecp_error_t ufsecp_gpu_ext_msm_device(ufsecp_gpu_ctx* ctx, const uint8_t* d_scalars32,
                                                    const uint8_t* d_points33, size_t n, uint8_t* d_out33,
                                                    int32_t* d_status, int window_bits, void* stream);

/* ufsecp_gpu_ext_msm_device that allocates nothing: no hipMalloc, hipFree or
 * allocating library routine on any call.  No counterpart in the reference;
 * it is for callers that cannot allocate (code under HIP graph capture,
 * tight serving loops, harnesses that forbid allocation).  The MSM's
 * buffers come from a static 32 MiB device array that every GPU the library
 * runs on holds until the process ends; n <= 16384 terms, and the window
 * (window_bits, 0 = tuned) must fit that array: up to 13 bits at 16384
 * terms.  Larger requests -> BAD_INPUT with nothing enqueued.  Same inputs,
 * outputs and status codes as ufsecp_gpu_ext_msm_device.  Calls through one
 * context are ordered; calls on one GPU from different contexts share the
 * array and race if they overlap. */
UFSECP_API ufsecp_error_t ufsecp_gpu_ext_msm_device_noalloc(ufsecp_gpu_ctx* ctx, const uint8_t* d_scalars32,
                                                            const uint8_t* d_points33, size_t n, uint8_t* d_out33,
                                                            int32_t* d_status, int window_bits, void* stream);

/* Device and pinned allocations the library has made since it was loaded:
 * its own buffers and those of ufsecp_gpu_ext_alloc / _host_alloc.  A test
 * hook for allocation-free entry points and for failure injection: with
 * UFSECP_GPU_FAIL_ALLOC=K in the environment (read at each allocation), the
 * allocation that brings this count to K fails as if the device were out of
 * memory, and the call that made it returns UFSECP_ERR_GPU_MEMORY. */
UFSECP_API uint64_t ufsecp_gpu_ext_buffer_alloc_count(void);

/* Device and pinned allocations of the library that are still live (made and
 * not yet released): a leak check for tests of the error paths. */
UFSECP_API int64_t ufsecp_gpu_ext_buffer_live_count(void);

/* ufsecp_gpu_ext_msm_device over bases already decompressed: d_points64 holds
 * n * 64 bytes, x || y big-endian affine coordinates.  A point with x or y >=
 * p or off the curve is BAD_PUBKEY, like an invalid compressed point. */
UFSECP_API ufsecp_error_t ufsecp_gpu_ext_msm_affine_device(ufsecp_gpu_ctx* ctx, const uint8_t* d_scalars32,
                                                           const uint8_t* d_points64, size_t n, uint8_t* d_out33,
                                                           int32_t* d_status, int window_bits, void* stream);

/* Device memory and timing helpers for callers without their own HIP runtime
 * handle (tests, benchmarks, language bindings).  Copies are synchronous with
 * respect to the context stream. */
UFSECP_API ufsecp_error_t ufsecp_gpu_ext_alloc(ufsecp_gpu_ctx* ctx, size_t bytes, void** d_ptr_out);
UFSECP_API ufsecp_error_t ufsecp_gpu_ext_free(ufsecp_gpu_ctx* ctx, void* d_ptr);
UFSECP_API ufsecp_error_t ufsecp_gpu_ext_copy_to_device(ufsecp_gpu_ctx* ctx, void* d_dst, const void* h_src,
                                                        size_t bytes);
UFSECP_API ufsecp_error_t ufsecp_gpu_ext_copy_to_host(ufsecp_gpu_ctx* ctx, void* h_dst, const void* d_src,
                                                      size_t bytes);
UFSECP_API ufsecp_error_t ufsecp_gpu_ext_event_create(ufsecp_gpu_ctx* ctx, void** event_out);
UFSECP_API ufsecp_error_t ufsecp_gpu_ext_event_record(ufsecp_gpu_ctx* ctx, void* event, void* stream);
/* Milliseconds between two recorded events (waits for `stop`). */
UFSECP_API ufsecp_error_t ufsecp_gpu_ext_event_elapsed(void* start, void* stop, float* ms_out);
UFSECP_API void ufsecp_gpu_ext_event_destroy(void* event);

/* The context's HIP stream (hipStream_t) for callers that want to order their
 * own work after the library's. */
UFSECP_API void* ufsecp_gpu_ext_stream(ufsecp_gpu_ctx* ctx);

/* Synchronise the context's stream. */
UFSECP_API ufsecp_error_t ufsecp_gpu_ext_synchronize(ufsecp_gpu_ctx* ctx);

/* Lanes of the resident verification grid for a single batch of `kind` (0
 * ECDSA, 1 BIP-340, 2 opaque ECDSA): lane l of the grid verifies items l,
 * l + lanes, l + 2 lanes, ...; 0 for a NULL ctx or an unknown kind. */
UFSECP_API uint32_t ufsecp_gpu_ext_verify_lanes(ufsecp_gpu_ctx* ctx, uint32_t kind);

/* A new non-blocking HIP stream at the default priority on ctx's device, for
 * callers that issue device-pointer calls on streams of their own;
 * stream_synchronize waits for it, stream_destroy waits and releases it. */
UFSECP_API ufsecp_error_t ufsecp_gpu_ext_stream_create(ufsecp_gpu_ctx* ctx, void** stream_out);
UFSECP_API ufsecp_error_t ufsecp_gpu_ext_stream_synchronize(void* stream);
UFSECP_API void ufsecp_gpu_ext_stream_destroy(void* stream);

/* Arithmetic self-test: runs device operation `op` (field/scalar/point
 * primitives, codes in the library's test suite) on `count` records of 16
 * 32-bit words from a and b, writing 32-word records to out.  Host buffers. */
UFSECP_API ufsecp_error_t ufsecp_gpu_ext_selftest(ufsecp_gpu_ctx* ctx, int op, const uint32_t* a, const uint32_t* b,
                                                  uint32_t* out, size_t count);

/* Pinned (page-locked) host memory, mapped into the device's address space at
 * the same address: kernels can read it in place (see
 * ufsecp_gpu_ext_server_submit_mapped).  Not a device allocation. */
UFSECP_API ufsecp_error_t ufsecp_gpu_ext_host_alloc(ufsecp_gpu_ctx* ctx, size_t bytes, void** h_ptr_out);
UFSECP_API ufsecp_error_t ufsecp_gpu_ext_host_free(ufsecp_gpu_ctx* ctx, void* h_ptr);

/* ---------------------------------------------------------------------------
 * Serving loop: continuous batching of verify requests.
 *
 * Requests of either kind and any size (1 .. max_group_items) are submitted
 * from any thread and complete asynchronously.  Each request is admitted when
 * its inputs fit in the server's device memory budget (FIFO; waiting requests
 * are admitted as earlier ones complete), and every admitted request joins the
 * next launch group: one persistent launch over all of them, with public keys
 * resolved through a device-resident decompression cache.  Up to
 * `inflight_groups` groups are on the device at once.  Verdicts (1 byte per
 * item, as ufsecp_gpu_*_verify_batch) are in out_results once poll/wait
 * report completion; the input and output buffers must stay valid until then.
 * The context must outlive the server.  Public data only.
 * ------------------------------------------------------------------------- */
typedef struct ufsecp_gpu_server ufsecp_gpu_server;

typedef struct {
    uint64_t hbm_budget_bytes;   /* device bytes for in-flight inputs (0 = 256 MiB) */
    uint32_t inflight_groups;    /* launch groups on the device at once, 1..4 (0 = 2) */
    uint32_t key_cache_log2;     /* decompression cache of 2^k 64-byte entries; 0 = none */
    uint32_t max_group_items;    /* items coalesced into one launch (0 = 2^21) */
    uint32_t reserved;
} ufsecp_gpu_server_config_t;

typedef struct {
    uint64_t requests_completed;
    uint64_t items_completed;
    uint64_t groups_launched;
    uint64_t key_lookups;
    uint64_t key_cache_hits;
    double latency_p50_us;       /* submit -> verdicts in the caller's buffer */
    double latency_p99_us;
    double latency_mean_us;
    double latency_max_us;
    double busy_seconds;         /* wall time with at least one group in flight */
} ufsecp_gpu_server_stats_t;

UFSECP_API ufsecp_error_t ufsecp_gpu_ext_server_create(ufsecp_gpu_ctx* ctx, const ufsecp_gpu_server_config_t* cfg,
                                                       ufsecp_gpu_server** server_out);
/* Stops the server; call after ufsecp_gpu_ext_server_drain. */
UFSECP_API void ufsecp_gpu_ext_server_destroy(ufsecp_gpu_server* server);
/* kind: UFSECP_GPU_KIND_ECDSA (pubkeys N x 33) or UFSECP_GPU_KIND_SCHNORR
 * (x-only pubkeys N x 32); msgs N x 32, sigs N x 64, out_results N bytes. */
UFSECP_API ufsecp_error_t ufsecp_gpu_ext_server_submit(ufsecp_gpu_server* server, uint32_t kind,
                                                       const uint8_t* msg_hashes32, const uint8_t* pubkeys,
                                                       const uint8_t* sigs64, size_t count, uint8_t* out_results,
                                                       uint64_t* request_id);
/* As ufsecp_gpu_ext_server_submit, for inputs the GPU can read in place:
 * buffers from ufsecp_gpu_ext_host_alloc or device memory (msgs and sigs, and
 * x-only pubkeys, 16-byte aligned).  No staging copy is made. */
UFSECP_API ufsecp_error_t ufsecp_gpu_ext_server_submit_mapped(ufsecp_gpu_server* server, uint32_t kind,
                                                              const uint8_t* msg_hashes32, const uint8_t* pubkeys,
                                                              const uint8_t* sigs64, size_t count,
                                                              uint8_t* out_results, uint64_t* request_id);
/* 1: done (the id is then forgotten), 0: pending, -1: unknown id. */
UFSECP_API int ufsecp_gpu_ext_server_poll(ufsecp_gpu_server* server, uint64_t request_id);
UFSECP_API ufsecp_error_t ufsecp_gpu_ext_server_wait(ufsecp_gpu_server* server, uint64_t request_id);
/* Waits for every submitted request and forgets all completed ids. */
UFSECP_API ufsecp_error_t ufsecp_gpu_ext_server_drain(ufsecp_gpu_server* server);
/* Counters and latency percentiles since creation or the last reset. */
UFSECP_API ufsecp_error_t ufsecp_gpu_ext_server_stats(ufsecp_gpu_server* server, ufsecp_gpu_server_stats_t* out,
                                                      int reset);

/* BIP-340 batch verification by a random linear combination (the
 * reference CPU engine's schnorr_batch_verify, src/cpu/src/batch_verify.cpp:
 * 324-424, on the device): one MSM of 2N + 1 terms with fresh random weights
 * per call.  *all_valid = 1 when every signature would verify under
 * ufsecp_gpu_schnorr_verify_batch, 0 otherwise (a wrong 1 has probability
 * about 2^-128).  No per-item verdicts: on 0, run the per-item call to find
 * the invalid ones.  count == 0 -> OK with *all_valid = 0. */
UFSECP_API ufsecp_error_t ufsecp_gpu_ext_schnorr_batch_check(ufsecp_gpu_ctx* ctx, const uint8_t* msg_hashes32,
                                                            const uint8_t* pubkeys_x32, const uint8_t* sigs64,
                                                            size_t count, int* all_valid);

/* ---- reference CPU engine hook ------------------------------------------ */

/* The reference engine's GPU column-verify hook
 * (src/cpu/include/secp256k1/batch_verify.hpp:149-168), served by this
 * library: kind 0 = opaque ECDSA columns (keys = pubkeys33, the verdicts of
 * ufsecp_gpu_ecdsa_verify_lbtc_columns), kind 1 = BIP-340 columns (keys =
 * xonly32).  Returns 1 (handled, all valid), 0 (handled, some invalid) or -1
 * (declined: no device, an operational error, count > 2^26).  One device
 * context per process, created on first use; calls are serialised. */
UFSECP_API int ufsecp_gpu_ext_engine_columns_hook(int kind, const uint8_t* digests32, const uint8_t* keys,
                                                  const uint8_t* sigs64, size_t count,
                                                  uint8_t* out_results) UFSECP_NOEXCEPT;
/* Installs the hook above into the reference engine when its installer
 * (secp256k1::install_gpu_columns_verify_hook) is visible in the process.
 * Runs by itself when the library loads (UFSECP_GPU_ENGINE_HOOK=0 disables
 * that); returns 1 if installed, 0 if no engine was found. */
UFSECP_API int ufsecp_gpu_ext_engine_hook_install(void);
typedef struct {
    uint64_t calls;
    uint64_t handled;
    uint64_t declined;
    uint64_t items;    /* items verified on the device */
} ufsecp_gpu_engine_hook_stats_t;
UFSECP_API void ufsecp_gpu_ext_engine_hook_stats(ufsecp_gpu_engine_hook_stats_t* out);

#ifdef __cplusplus
}
#endif

#endif /* UFSECP_GPU_EXT_H */
