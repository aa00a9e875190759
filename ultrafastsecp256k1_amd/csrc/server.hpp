// Continuous-batching verification server on one GPU.
//
// Requests (one kind, any size) are submitted from any thread and complete
// asynchronously.  The server
//   * admits a request when its bytes fit in a fixed device input ring (sized
//     by an HBM budget given at creation) and the matching pinned host rings;
//     the submitting thread copies the inputs into the pinned input ring, so a
//     launch group's inputs are one contiguous range (one or two DMA copies
//     per group, not per request);
//   * coalesces every admitted request into the next launch group as soon as
//     one of its group slots (stream + lane slots + tables) is free: with two
//     or more slots the next group is queued while the current one runs, and
//     its waves fill compute units as the current group's waves retire;
//   * per group: runs the persistent verify kernel over all requests, which
//     looks each public key up in the device-resident decompression cache
//     (hits skip the square root) and decompresses and inserts the misses;
//   * writes verdicts into a pinned ring the kernel stores to directly, and on
//     completion copies them to each request's output buffer;
//   * records per-request latency (submit -> verdicts in the caller's buffer).
// No allocation happens per request: every buffer is created with the server.
//
// The reference has no serving loop; its batch entry points are synchronous
// per call (src/cpu/src/ufsecp_gpu_impl.cpp:317-558).  Verdicts are the same
// as those of the per-batch entry points.
#pragma once
#include <hip/hip_runtime.h>

#include <atomic>
#include <condition_variable>
#include <cstdint>
#include <deque>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <vector>

#include "backend.hpp"

namespace ufs {

struct ServerConfig {
    size_t hbm_budget = 256ull << 20;   // device bytes for in-flight inputs
    unsigned groups = 2;                // launch groups in flight (1..4)
    unsigned cache_log2 = 21;           // decompression cache entries (0 = none)
    uint32_t max_group_items = 1u << 21;
};

struct ServerStats {
    uint64_t requests_completed = 0;
    uint64_t items_completed = 0;
    uint64_t groups_launched = 0;
    uint64_t key_lookups = 0;
    uint64_t key_cache_hits = 0;
    double latency_p50_us = 0, latency_p99_us = 0, latency_mean_us = 0, latency_max_us = 0;
    double busy_seconds = 0;   // wall time with at least one group in flight
};

// FIFO byte ring: allocations in submission order, released in order.
struct FifoRing {
    size_t cap = 0, head = 0, tail = 0, used = 0;
    // returns the offset, or SIZE_MAX if `len` does not fit now
    size_t alloc(size_t len, size_t& charged);
    void release(size_t charged);
};

class VerifyServer {
public:
    explicit VerifyServer(RocmBackend* be) : be_(be) {}
    ~VerifyServer();
    GpuError start(const ServerConfig& cfg);
    // mapped: inputs are device-accessible (pinned host memory from the
    // library's allocator, or device memory) and are read in place.
    GpuError submit(uint32_t kind, const uint8_t* msgs, const uint8_t* pubs, const uint8_t* sigs, size_t count,
                    uint8_t* out, bool mapped, uint64_t* id);
    int poll(uint64_t id);   // 1 done (forgotten), 0 pending, -1 unknown
    GpuError wait(uint64_t id);
    GpuError drain();
    void stats(ServerStats& out, bool reset);
    GpuError error() const { return (GpuError)err_.load(); }

private:
    struct Request {
        uint64_t id;
        uint32_t kind, count;
        const uint8_t *msgs, *pubs, *sigs;
        uint8_t* out;
        size_t in_off, in_len, in_charge;    // input rings (pinned + device mirror)
        size_t out_off, out_charge;          // pinned verdict ring
        double t_submit;
        bool mapped = false;                  // inputs read in place, no input ring
        bool staged = false, done = false;
        bool released = false, observed = false;   // freed when both
    };
    struct Group {
        hipStream_t stream = nullptr;
        hipEvent_t done = nullptr;
        void* d_slots = nullptr;             // lane slots of the verify grid
        uint32_t* d_words = nullptr;         // [16] tile counter, [32] key-cache hits
        uint8_t* d_tables = nullptr;         // descriptors | tiles | tile count
        uint8_t* h_tables = nullptr;         // pinned staging of the same
        uint32_t* h_key_hits = nullptr;      // pinned copy-back of the key-cache hits
        std::vector<Request*> reqs;
        uint64_t items = 0;
        bool busy = false;
    };

    void scheduler_loop();
    void completion_loop();
    bool try_admit_locked(Request* r);       // ring space + staging copy
    void stage_inputs(Request* r);
    GpuError launch_group(Group& g, std::vector<Request*>& batch);
    void finish_group(Group& g);
    void set_error(GpuError e);
    static double now_us();

    RocmBackend* be_;
    ServerConfig cfg_;
    int device_ = 0;
    int verify_variant_ = 2;            // 3 with the key cache, 2 without
    unsigned verify_grid_ = 0;
    size_t max_batches_ = 0, max_tiles_ = 0, table_bytes_ = 0;
    // rings
    uint8_t* h_in_ = nullptr;     // pinned input ring
    uint8_t* d_in_ = nullptr;     // device mirror of the input ring
    uint8_t* h_out_ = nullptr;    // pinned verdict ring (written by the kernel)
    uint8_t* d_out_view_ = nullptr;   // device address of h_out_
    FifoRing in_ring_, out_ring_;
    void* d_cache_ = nullptr;
    uint32_t cache_mask_ = 0;
    std::vector<Group> groups_;

    std::mutex mu_;
    std::condition_variable cv_sched_, cv_done_, cv_wait_;
    std::deque<Request*> pending_;    // waiting for ring space (FIFO)
    std::deque<Request*> to_launch_;  // admitted, not yet launched (ring order)
    std::deque<Request*> admitted_;   // admitted, not yet released (ring order)
    uint64_t outstanding_ = 0;        // submitted, not completed
    std::unordered_map<uint64_t, Request*> reqs_;
    std::deque<Group*> inflight_;
    uint64_t next_id_ = 1;
    bool stop_ = false;
    std::atomic<int> err_{0};
    std::thread sched_thr_, done_thr_;

    // statistics (guarded by mu_)
    std::vector<float> lat_us_;
    uint64_t st_requests_ = 0, st_items_ = 0, st_groups_ = 0, st_lookups_ = 0, st_hits_ = 0;
    double busy_since_ = -1, busy_us_ = 0;
};

}  // namespace ufs
