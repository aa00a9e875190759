// RocmBackend: host orchestration of the gfx950 batch-verify / MSM kernels.
//
// Mirrors the role of the reference's GpuBackend / CudaBackend
// (src/gpu/include/gpu_backend.hpp:79-607, src/gpu/src/gpu_backend_cuda.cu:
// 833-2622) for the verify and MSM operations, redesigned for MI355X:
//   - one HIP stream per context, every copy and launch asynchronous on it;
//   - grow-only device buffers and pinned host staging kept for the context's
//     lifetime (no per-call hipMalloc/hipFree);
//   - the fixed-base G comb table and the BIP-340 challenge midstate are built
//     on the device once at init and stay resident in HBM.
#pragma once
#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <string>

#include "kernels.h"
#include "tuning.hpp"

namespace ufs {

// Error codes: numerically identical to the C ABI (ufsecp_error.h).
enum class GpuError : int {
    Ok = 0,
    NullArg = 1,
    BadPubkey = 3,
    BadInput = 5,
    Arith = 7,
    Internal = 9,
    Unavailable = 100,
    Device = 101,
    Launch = 102,
    Memory = 103,
    Unsupported = 104,
    Backend = 105,
    Queue = 106,
};

// Every device and pinned host allocation of the library goes through
// dev_malloc / pinned_malloc.  buffer_alloc_count() is the number made since
// the library was loaded.  Failure injection for tests of the error paths:
// with UFSECP_GPU_FAIL_ALLOC=K in the environment (read at each allocation),
// the allocation that would bring the count to K returns hipErrorOutOfMemory
// without calling the runtime (it still counts).
uint64_t buffer_alloc_count();
// Allocations made through dev_malloc / pinned_malloc and not yet released
// through dev_free / pinned_free: a leak check for the error paths.
int64_t buffer_live_count();
hipError_t dev_malloc_raw(void** p, size_t bytes);
hipError_t pinned_malloc_raw(void** p, size_t bytes, unsigned flags);
hipError_t dev_free(void* p);      // NULL is a no-op
hipError_t pinned_free(void* p);   // NULL is a no-op
template <class T>
hipError_t dev_malloc(T** p, size_t bytes) {
    return dev_malloc_raw(reinterpret_cast<void**>(p), bytes);
}
template <class T>
hipError_t pinned_malloc(T** p, size_t bytes, unsigned flags) {
    return pinned_malloc_raw(reinterpret_cast<void**>(p), bytes, flags);
}

struct DeviceBuffer {
    void* ptr = nullptr;
    size_t cap = 0;
    hipError_t ensure(size_t bytes);
    void release();
};

struct PinnedBuffer {
    void* ptr = nullptr;
    size_t cap = 0;
    hipError_t ensure(size_t bytes);
    void release();
};

struct DeviceInfoData {
    char name[128];
    uint64_t global_mem_bytes;
    uint32_t compute_units;
    uint32_t max_clock_mhz;
    uint32_t max_threads_per_block;
};

class RocmBackend {
public:
    RocmBackend() = default;
    ~RocmBackend();
    RocmBackend(const RocmBackend&) = delete;
    RocmBackend& operator=(const RocmBackend&) = delete;

    static int device_count();
    static GpuError device_info(uint32_t device_index, DeviceInfoData& out);

    GpuError init(uint32_t device_index);
    bool is_ready() const { return ready_; }
    GpuError last_error() const { return last_err_; }
    const char* last_error_msg() const { return last_msg_.c_str(); }
    hipStream_t stream() const { return stream_; }

    // Host-buffer batch operations (synchronous, like the reference).
    GpuError ecdsa_verify_batch(const uint8_t* msgs, const uint8_t* pubs33, const uint8_t* sigs, size_t count,
                                uint8_t* out, bool collect);
    GpuError schnorr_verify_batch(const uint8_t* msgs, const uint8_t* pubx, const uint8_t* sigs, size_t count,
                                  uint8_t* out, bool collect);
    // libbitcoin opaque-signature ECDSA (kKindEcdsaOpaque): interleaved rows of
    // 32-byte digest | 33-byte pubkey | 64-byte opaque sig, `stride` >= 129
    // bytes apart, or the same three fields as packed columns.
    GpuError ecdsa_verify_opaque_rows(const uint8_t* rows, size_t stride, size_t count, uint8_t* out);
    GpuError ecdsa_verify_opaque_columns(const uint8_t* msgs, const uint8_t* pubs33, const uint8_t* sigs,
                                         size_t count, uint8_t* out);
    // Public-data key operations of the libbitcoin surface (host buffers,
    // synchronous): compressed-key and x-only-key validation, BIP-341
    // tweak-add commitment check.
    GpuError pubkey_validate(const uint8_t* pk33, size_t n, uint8_t* out);
    GpuError xonly_validate(const uint8_t* x32, size_t n, uint8_t* out);
    GpuError commitment_verify(const uint8_t* internal_x32, const uint8_t* tweak32, const uint8_t* tweaked_x32,
                               const uint8_t* parity, size_t n, uint8_t* out);
    // ECDSA public-key recovery (host buffers, synchronous); recids must be 0..3.
    GpuError ecrecover_batch(const uint8_t* msgs, const uint8_t* sigs, const int* recids, size_t n,
                             uint8_t* out33, uint8_t* valid);

    // Device-pointer variants (asynchronous on `st`, or the context stream).
    GpuError ecdsa_verify_device(const uint8_t* d_msgs, const uint8_t* d_pubs33, const uint8_t* d_sigs,
                                 size_t count, uint8_t* d_out, hipStream_t st);
    GpuError schnorr_verify_device(const uint8_t* d_msgs, const uint8_t* d_pubx, const uint8_t* d_sigs,
                                   size_t count, uint8_t* d_out, hipStream_t st);

    // Many uneven batches in one persistent launch (optionally replayed from a
    // captured hipGraph); `each` issues one plain launch per batch instead.
    GpuError verify_multi_device(const BatchDesc* batches, size_t nbatches, bool use_graph, hipStream_t st);
    GpuError verify_each_device(const BatchDesc* batches, size_t nbatches, hipStream_t st);

    GpuError selftest(int op, const uint32_t* a, const uint32_t* b, uint32_t* out, size_t count);
    GpuError synchronize();

    const GcombDesc& gcomb() const { return gdesc_; }
    const uint32_t* bip340_midstate() const { return d_mid_; }
    int device() const { return device_; }
    const Tuning& tuning() const { return tuning_; }
    int compute_units() const { return cus_; }
    unsigned verify_grid(int kind) const { return verify_grid_[kind]; }

    GpuError fail(GpuError e, const char* what, hipError_t he = hipSuccess);
    void clear_error() {
        last_err_ = GpuError::Ok;
        last_msg_ = "OK";
    }

private:
    GpuError verify_host(uint32_t kind, const uint8_t* msgs, const uint8_t* pubs, const uint8_t* sigs,
                         size_t count, uint8_t* out, bool collect);
    // Copies `nin` host inputs to device buffers, runs `launch` on the context
    // stream and copies the n verdict bytes back (the public-data ops).
    template <class Launch>
    GpuError keyop_host(const char* what, size_t n, uint8_t* out, const uint8_t* const* in, const size_t* bytes,
                        int nin, Launch&& launch);
    // Host-to-device upload of a batch pipelined with its verification: the
    // batch is cut into pieces of r, 2r, 4r, ... items (r = lanes of the
    // resident grid, so every launch but the last runs whole rounds), each
    // piece is uploaded on the copy stream and verified on the context stream
    // as soon as its inputs are in place.  Region k is item-strided
    // (`item_bytes[k]`; the batch's last item has `last_bytes[k]`); `whole`
    // describes the full batch in device memory.
    hipError_t upload_and_verify(const uint8_t* const* src, void* const* dst, const size_t* item_bytes,
                                 const size_t* last_bytes, int nregions, const BatchDesc& whole, size_t ms,
                                 size_t ps, size_t ss, bool collect);
    // One batch through the persistent kernel (single mode).
    hipError_t launch_single(const BatchDesc& d, bool collect, hipStream_t st);
    // Verify launches share the lane slots and the work counter: a launch on
    // a stream other than the previous one waits for the previous launch.
    hipError_t verify_begin(hipStream_t st);
    hipError_t verify_end(hipStream_t st);

    int device_ = -1;
    bool ready_ = false;
    hipStream_t stream_ = nullptr;
    hipStream_t copy_stream_ = nullptr;        // uploads of pipelined host calls
    static constexpr int kMaxPieces = 12;
    hipEvent_t piece_ev_[kMaxPieces] = {};
    void* d_gcomb_ = nullptr;
    uint32_t* d_mid_ = nullptr;
    GcombDesc gdesc_{nullptr, 0, 0};
    DeviceBuffer d_in_[3];
    DeviceBuffer d_out_;
    // multi-batch state: descriptor/tile tables, work counter, captured graph
    DeviceBuffer d_descs_, d_tiles_, d_words_;
    PinnedBuffer h_tables_;
    hipGraph_t graph_ = nullptr;
    hipGraphExec_t graph_exec_ = nullptr;
    hipStream_t graph_stream_ = nullptr;
    const void* graph_descs_ = nullptr;
    const void* graph_tiles_ = nullptr;
    Tuning tuning_;
    unsigned verify_grid_[4] = {0, 0, 0, 0};   // blocks: ECDSA, BIP-340, mixed, mixed with key cache
    DeviceBuffer d_slots_;                   // kSlotBytes per lane of the largest grid
    hipEvent_t verify_done_ = nullptr;
    hipStream_t verify_stream_ = nullptr;
    int cus_ = 0;
    GpuError last_err_ = GpuError::Ok;
    std::string last_msg_ = "OK";
};

}  // namespace ufs
