// RocmBackend implementation (host side).  See backend.hpp.
#include "backend.hpp"

#include <atomic>
#include <mutex>
#include <vector>

#include <cstdio>
#include <cstdlib>
#include <cstring>

namespace ufs {

static std::atomic<uint64_t> g_buffer_allocs{0};
static std::atomic<int64_t> g_buffer_live{0};

uint64_t buffer_alloc_count() { return g_buffer_allocs.load(std::memory_order_relaxed); }
int64_t buffer_live_count() { return g_buffer_live.load(std::memory_order_relaxed); }

static bool injected_failure(uint64_t ordinal) {
    const char* s = std::getenv("UFSECP_GPU_FAIL_ALLOC");
    if (!s || !*s) return false;
    char* end = nullptr;
    const unsigned long long k = std::strtoull(s, &end, 10);
    return end && *end == 0 && k == ordinal;
}

hipError_t dev_malloc_raw(void** p, size_t bytes) {
    *p = nullptr;
    if (injected_failure(g_buffer_allocs.fetch_add(1, std::memory_order_relaxed) + 1)) return hipErrorOutOfMemory;
    const hipError_t e = hipMalloc(p, bytes);
    if (e != hipSuccess) *p = nullptr;
    else g_buffer_live.fetch_add(1, std::memory_order_relaxed);
    return e;
}

hipError_t pinned_malloc_raw(void** p, size_t bytes, unsigned flags) {
    *p = nullptr;
    if (injected_failure(g_buffer_allocs.fetch_add(1, std::memory_order_relaxed) + 1)) return hipErrorOutOfMemory;
    const hipError_t e = hipHostMalloc(p, bytes, flags);
    if (e != hipSuccess) *p = nullptr;
    else g_buffer_live.fetch_add(1, std::memory_order_relaxed);
    return e;
}

hipError_t dev_free(void* p) {
    if (!p) return hipSuccess;
    const hipError_t e = hipFree(p);
    if (e == hipSuccess) g_buffer_live.fetch_sub(1, std::memory_order_relaxed);
    return e;
}

hipError_t pinned_free(void* p) {
    if (!p) return hipSuccess;
    const hipError_t e = hipHostFree(p);
    if (e == hipSuccess) g_buffer_live.fetch_sub(1, std::memory_order_relaxed);
    return e;
}

hipError_t DeviceBuffer::ensure(size_t bytes) {
    if (bytes <= cap) return hipSuccess;
    release();
    size_t want = 1;
    while (want < bytes) want <<= 1;   // grow-only, power of two
    hipError_t e = dev_malloc(&ptr, want);
    if (e != hipSuccess) {
        ptr = nullptr;
        cap = 0;
        return e;
    }
    cap = want;
    return hipSuccess;
}

void DeviceBuffer::release() {
    (void)dev_free(ptr);
    ptr = nullptr;
    cap = 0;
}

hipError_t PinnedBuffer::ensure(size_t bytes) {
    if (bytes <= cap) return hipSuccess;
    release();
    size_t want = 1;
    while (want < bytes) want <<= 1;
    hipError_t e = pinned_malloc(&ptr, want, hipHostMallocDefault);
    if (e != hipSuccess) {
        ptr = nullptr;
        cap = 0;
        return e;
    }
    cap = want;
    return hipSuccess;
}

void PinnedBuffer::release() {
    (void)pinned_free(ptr);
    ptr = nullptr;
    cap = 0;
}

int RocmBackend::device_count() {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

GpuError RocmBackend::device_info(uint32_t device_index, DeviceInfoData& out) {
    const int n = device_count();
    if (n <= 0) return GpuError::Unavailable;
    if ((int)device_index >= n) return GpuError::Device;
    hipDeviceProp_t p;
    if (hipGetDeviceProperties(&p, (int)device_index) != hipSuccess) return GpuError::Device;
    std::memset(&out, 0, sizeof(out));
    std::snprintf(out.name, sizeof(out.name), "%s (%s)", p.name, p.gcnArchName);
    out.global_mem_bytes = p.totalGlobalMem;
    out.compute_units = (uint32_t)p.multiProcessorCount;
    out.max_clock_mhz = (uint32_t)(p.clockRate / 1000);
    out.max_threads_per_block = (uint32_t)p.maxThreadsPerBlock;
    return GpuError::Ok;
}

GpuError RocmBackend::fail(GpuError e, const char* what, hipError_t he) {
    last_err_ = e;
    last_msg_ = what;
    if (he != hipSuccess) {
        last_msg_ += ": ";
        last_msg_ += hipGetErrorString(he);
    }
    return e;
}

// The fixed-base comb for G is the same for every context on a device (436 MB
// at width 20): one table per (device, width), built by the first context
// that needs it, shared read-only, freed with the last context using it.
namespace {
struct SharedComb {
    int device;
    int w;
    void* ptr;
    int refs;
};
std::mutex g_comb_mu;
std::vector<SharedComb> g_combs;
}  // namespace

// The caller's device is current; the build finishes before the table is
// handed to any context.
// *alloc_failed tells an allocation failure from a failed build.
static hipError_t acquire_comb(int device, int w, hipStream_t st, void** out, bool* alloc_failed) {
    std::lock_guard<std::mutex> lk(g_comb_mu);
    for (SharedComb& c : g_combs)
        if (c.device == device && c.w == w) {
            ++c.refs;
            *out = c.ptr;
            return hipSuccess;
        }
    void* p = nullptr;
    hipError_t e = dev_malloc(&p, gcomb_entries(w) * 64);
    *alloc_failed = e != hipSuccess;
    if (e != hipSuccess) return e;
    e = launch_build_gcomb(p, w, gcomb_nwin(w), st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess) {
        (void)dev_free(p);
        return e;
    }
    g_combs.push_back(SharedComb{device, w, p, 1});
    *out = p;
    return hipSuccess;
}

static void release_comb(int device, void* ptr) {
    if (!ptr) return;
    std::lock_guard<std::mutex> lk(g_comb_mu);
    for (size_t i = 0; i < g_combs.size(); ++i)
        if (g_combs[i].device == device && g_combs[i].ptr == ptr) {
            if (--g_combs[i].refs == 0) {
                (void)dev_free(ptr);
                g_combs.erase(g_combs.begin() + (std::ptrdiff_t)i);
            }
            return;
        }
}

RocmBackend::~RocmBackend() {
    if (device_ >= 0) (void)hipSetDevice(device_);
    if (stream_) (void)hipStreamSynchronize(stream_);
    if (graph_exec_) (void)hipGraphExecDestroy(graph_exec_);
    if (graph_) (void)hipGraphDestroy(graph_);
    if (graph_stream_) (void)hipStreamDestroy(graph_stream_);
    d_descs_.release();
    d_tiles_.release();
    d_words_.release();
    d_slots_.release();
    if (verify_done_) (void)hipEventDestroy(verify_done_);
    if (copy_stream_) (void)hipStreamSynchronize(copy_stream_);
    for (hipEvent_t& ev : piece_ev_)
        if (ev) (void)hipEventDestroy(ev);
    if (copy_stream_) (void)hipStreamDestroy(copy_stream_);
    h_tables_.release();
    for (auto& b : d_in_) b.release();
    d_out_.release();
    release_comb(device_, d_gcomb_);
    (void)dev_free(d_mid_);
    if (stream_) (void)hipStreamDestroy(stream_);
}

static int env_int(const char* name, int dflt, int lo, int hi) {
    const char* s = std::getenv(name);
    if (!s || !*s) return dflt;
    int v = std::atoi(s);
    if (v < lo || v > hi) return dflt;
    return v;
}

GpuError RocmBackend::init(uint32_t device_index) {
    const int n = device_count();
    if (n <= 0) return fail(GpuError::Unavailable, "no HIP device");
    if ((int)device_index >= n) return fail(GpuError::Device, "device index out of range");
    hipError_t e = hipSetDevice((int)device_index);
    if (e != hipSuccess) return fail(GpuError::Device, "hipSetDevice", e);
    device_ = (int)device_index;
    e = hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking);
    if (e != hipSuccess) return fail(GpuError::Queue, "hipStreamCreate", e);

    // Fixed-base comb table for G: width from the tuning file, overridden by
    // UFSECP_GCOMB_W (4..20); shared by the contexts on this device.
    tuning_ = load_tuning();
    const int w = env_int("UFSECP_GCOMB_W", tuning_.gcomb_w, 4, 20);
    bool alloc_failed = false;
    e = acquire_comb(device_, w, stream_, &d_gcomb_, &alloc_failed);
    if (e != hipSuccess)
        return fail(alloc_failed ? GpuError::Memory : GpuError::Launch,
                    alloc_failed ? "hipMalloc(G table)" : "build G table", e);
    gdesc_ = GcombDesc{d_gcomb_, w, gcomb_nwin(w)};

    e = dev_malloc(&d_mid_, 8 * sizeof(uint32_t));
    if (e != hipSuccess) return fail(GpuError::Memory, "hipMalloc(midstate)", e);
    e = launch_bip340_midstate(d_mid_, stream_);
    if (e != hipSuccess) return fail(GpuError::Launch, "BIP-340 midstate", e);

    // Persistent verify grids (one resident wave set) and their lane slots.
    hipDeviceProp_t p;
    if ((e = hipGetDeviceProperties(&p, device_)) != hipSuccess) return fail(GpuError::Device, "properties", e);
    cus_ = p.multiProcessorCount;
    unsigned max_grid = 0;
    for (int k = 0; k < 4; ++k) {
        int bpc = verify_blocks_per_cu(k);
        if (tuning_.verify_blocks_per_cu > 0 && tuning_.verify_blocks_per_cu < bpc) bpc = tuning_.verify_blocks_per_cu;
        verify_grid_[k] = (unsigned)(cus_ * bpc);
        if (verify_grid_[k] > max_grid) max_grid = verify_grid_[k];
    }
    if ((e = d_slots_.ensure((size_t)max_grid * kVerifyBlock * kSlotBytes)) != hipSuccess)
        return fail(GpuError::Memory, "hipMalloc(lane slots)", e);
    if ((e = d_words_.ensure(64 * sizeof(uint32_t))) != hipSuccess)
        return fail(GpuError::Memory, "hipMalloc(work counter)", e);
    if ((e = hipEventCreateWithFlags(&verify_done_, hipEventDisableTiming)) != hipSuccess)
        return fail(GpuError::Backend, "hipEventCreate", e);
    if ((e = hipStreamCreateWithFlags(&copy_stream_, hipStreamNonBlocking)) != hipSuccess)
        return fail(GpuError::Queue, "hipStreamCreate(copy)", e);
    for (hipEvent_t& ev : piece_ev_)
        if ((e = hipEventCreateWithFlags(&ev, hipEventDisableTiming)) != hipSuccess)
            return fail(GpuError::Backend, "hipEventCreate", e);
    e = hipStreamSynchronize(stream_);
    if (e != hipSuccess) return fail(GpuError::Backend, "init synchronize", e);
    ready_ = true;
    clear_error();
    return GpuError::Ok;
}

GpuError RocmBackend::synchronize() {
    hipError_t e = hipStreamSynchronize(stream_);
    if (e != hipSuccess) return fail(GpuError::Backend, "hipStreamSynchronize", e);
    return GpuError::Ok;
}

hipError_t RocmBackend::verify_begin(hipStream_t st) {
    if (verify_stream_ && verify_stream_ != st) return hipStreamWaitEvent(st, verify_done_, 0);
    return hipSuccess;
}

hipError_t RocmBackend::verify_end(hipStream_t st) {
    verify_stream_ = st;
    return hipEventRecord(verify_done_, st);
}

hipError_t RocmBackend::launch_single(const BatchDesc& d, bool collect, hipStream_t st) {
    if (d.count == 0) return hipSuccess;
    const uint32_t tiles = (d.count + kTileItems - 1) / kTileItems;
    // kernel variant: ECDSA, BIP-340 or opaque ECDSA (its grid is ECDSA's)
    const int kind = d.kind == kKindEcdsa ? 0 : d.kind == kKindSchnorr ? 1 : 4;
    const unsigned waves_per_block = kVerifyBlock / 64;
    unsigned grid = (tiles + waves_per_block - 1) / waves_per_block;
    const unsigned gmax = verify_grid_[kind == 4 ? 0 : kind];
    if (grid > gmax) grid = gmax;
    VerifyLaunch a{};
    a.single = d;
    a.slots = d_slots_.ptr;
    a.single_tiles = tiles;
    a.collect = collect ? 1u : 0u;
    // more tiles than waves: the waves take tiles from a work counter, zeroed
    // on st before the launch (the verify launches of a context are ordered,
    // so they share the word); otherwise one tile per wave by position
    const bool dyn = tiles > (size_t)grid * waves_per_block;
    if (dyn) a.counter = (uint32_t*)d_words_.ptr + 48;
    hipError_t e = verify_begin(st);
    if (e == hipSuccess && dyn) e = launch_u32_fill(a.counter, 0u, 1, st);
    if (e == hipSuccess) e = launch_verify(kind, a, gdesc_, d_mid_, grid, st);
    if (e == hipSuccess) e = verify_end(st);
    return e;
}

hipError_t RocmBackend::upload_and_verify(const uint8_t* const* src, void* const* dst, const size_t* item_bytes,
                                          const size_t* last_bytes, int nregions, const BatchDesc& whole, size_t ms,
                                          size_t ps, size_t ss, bool collect) {
    const size_t count = whole.count;
    const int kv = whole.kind == kKindSchnorr ? 1 : 0;
    const size_t round = (size_t)verify_grid_[kv] * kVerifyBlock;
    size_t a = 0, len = round;
    hipError_t e = hipSuccess;
    for (int p = 0; a < count && e == hipSuccess; ++p) {
        // a piece shorter than 2 rounds before the end takes the rest
        size_t b = a + len;
        if (p == kMaxPieces - 1 || b + round > count) b = count;
        for (int k = 0; k < nregions && e == hipSuccess; ++k) {
            const size_t bytes = (b - a) * item_bytes[k] - (b == count ? item_bytes[k] - last_bytes[k] : 0);
            e = hipMemcpyAsync((uint8_t*)dst[k] + a * item_bytes[k], src[k] + a * item_bytes[k], bytes,
                               hipMemcpyHostToDevice, copy_stream_);
        }
        if (e == hipSuccess) e = hipEventRecord(piece_ev_[p], copy_stream_);
        if (e == hipSuccess) e = hipStreamWaitEvent(stream_, piece_ev_[p], 0);
        if (e != hipSuccess) break;
        BatchDesc d = whole;
        d.count = (uint32_t)(b - a);
        d.msgs += a * ms;
        d.pubs += a * ps;
        d.sigs += a * ss;
        d.out += a;
        e = launch_single(d, collect, stream_);
        a = b;
        len *= 2;
    }
    return e;
}

GpuError RocmBackend::verify_host(uint32_t kind, const uint8_t* msgs, const uint8_t* pubs, const uint8_t* sigs,
                                  size_t count, uint8_t* out, bool collect) {
    if (!ready_) return fail(GpuError::Device, "context not initialised");
    hipError_t e = hipSetDevice(device_);
    if (e != hipSuccess) return fail(GpuError::Device, "hipSetDevice", e);
    const size_t pub_stride = kind == kKindSchnorr ? 32 : 33;
    const size_t sz_m = count * 32, sz_p = count * pub_stride, sz_s = count * 64;
    if ((e = d_in_[0].ensure(sz_m)) != hipSuccess || (e = d_in_[1].ensure(sz_p)) != hipSuccess ||
        (e = d_in_[2].ensure(sz_s)) != hipSuccess || (e = d_out_.ensure(count)) != hipSuccess)
        return fail(GpuError::Memory, "hipMalloc(batch buffers)", e);
    if (collect && (e = hipMemcpyAsync(d_out_.ptr, out, count, hipMemcpyHostToDevice, stream_)) != hipSuccess)
        return fail(GpuError::Memory, "copy inputs to device", e);
    const uint8_t* dm = (const uint8_t*)d_in_[0].ptr;
    const uint8_t* dp = (const uint8_t*)d_in_[1].ptr;
    const uint8_t* ds = (const uint8_t*)d_in_[2].ptr;
    uint8_t* dout = (uint8_t*)d_out_.ptr;
    const uint8_t* src[3] = {msgs, pubs, sigs};
    void* dst[3] = {d_in_[0].ptr, d_in_[1].ptr, d_in_[2].ptr};
    const size_t item[3] = {32, pub_stride, 64};
    e = upload_and_verify(src, dst, item, item, 3, BatchDesc{kind, (uint32_t)count, dm, dp, ds, dout, 0, 0}, 32,
                          pub_stride, 64, collect);
    if (e != hipSuccess) return fail(GpuError::Launch, "verify upload / launch", e);
    e = hipMemcpyAsync(out, dout, count, hipMemcpyDeviceToHost, stream_);
    if (e != hipSuccess) return fail(GpuError::Memory, "copy verdicts to host", e);
    e = hipStreamSynchronize(stream_);
    if (e != hipSuccess) return fail(GpuError::Backend, "verify synchronize", e);
    clear_error();
    return GpuError::Ok;
}

GpuError RocmBackend::ecdsa_verify_batch(const uint8_t* msgs, const uint8_t* pubs33, const uint8_t* sigs,
                                         size_t count, uint8_t* out, bool collect) {
    return verify_host(kKindEcdsa, msgs, pubs33, sigs, count, out, collect);
}

GpuError RocmBackend::schnorr_verify_batch(const uint8_t* msgs, const uint8_t* pubx, const uint8_t* sigs,
                                           size_t count, uint8_t* out, bool collect) {
    return verify_host(kKindSchnorr, msgs, pubx, sigs, count, out, collect);
}

GpuError RocmBackend::ecdsa_verify_opaque_columns(const uint8_t* msgs, const uint8_t* pubs33, const uint8_t* sigs,
                                                  size_t count, uint8_t* out) {
    return verify_host(kKindEcdsaOpaque, msgs, pubs33, sigs, count, out, false);
}

// Row layout of the reference (gpu_backend_cuda.cu:1030-1082): rows are read
// at `stride`, only the first 129 bytes of each row are used.  Rows up to
// twice that long travel as one contiguous copy and are read in place at
// their stride; longer rows are gathered by a 2-D copy into 129-byte rows, so
// the transfer never carries much more than the used bytes.  Only
// (count - 1) * stride + 129 bytes of the caller's buffer are read.
GpuError RocmBackend::ecdsa_verify_opaque_rows(const uint8_t* rows, size_t stride, size_t count, uint8_t* out) {
    constexpr size_t kRow = 32 + 33 + 64;
    if (!ready_) return fail(GpuError::Device, "context not initialised");
    if (count == 0) {
        clear_error();
        return GpuError::Ok;
    }
    if (!rows || !out) return fail(GpuError::NullArg, "NULL buffer");
    if (stride < kRow) return fail(GpuError::BadInput, "opaque row stride < 129");
    if (count > SIZE_MAX / stride) return fail(GpuError::BadInput, "row byte size overflow");
    if (count > UINT32_MAX) return fail(GpuError::BadInput, "batch count too large");
    hipError_t e = hipSetDevice(device_);
    if (e != hipSuccess) return fail(GpuError::Device, "hipSetDevice", e);
    const bool dense = stride <= 2 * kRow;
    const size_t pitch = dense ? stride : kRow;
    const size_t bytes = (count - 1) * pitch + kRow;
    if ((e = d_in_[0].ensure(bytes)) != hipSuccess || (e = d_out_.ensure(count)) != hipSuccess)
        return fail(GpuError::Memory, "hipMalloc(row buffers)", e);
    uint8_t* d = (uint8_t*)d_in_[0].ptr;
    uint8_t* dout = (uint8_t*)d_out_.ptr;
    const BatchDesc whole{kKindEcdsaOpaque, (uint32_t)count, d, d + 32, d + 65, dout, (uint32_t)pitch, 0};
    if (dense) {
        const uint8_t* src[1] = {rows};
        void* dst[1] = {d};
        const size_t item[1] = {stride}, last[1] = {kRow};
        e = upload_and_verify(src, dst, item, last, 1, whole, stride, stride, stride, false);
        if (e != hipSuccess) return fail(GpuError::Launch, "opaque-row upload / launch", e);
    } else {
        e = hipMemcpy2DAsync(d, kRow, rows, stride, kRow, count, hipMemcpyHostToDevice, stream_);
        if (e != hipSuccess) return fail(GpuError::Memory, "copy rows to device", e);
        e = launch_single(whole, false, stream_);
        if (e != hipSuccess) return fail(GpuError::Launch, "opaque-row verify launch", e);
    }
    e = hipMemcpyAsync(out, dout, count, hipMemcpyDeviceToHost, stream_);
    if (e != hipSuccess) return fail(GpuError::Memory, "copy verdicts to host", e);
    e = hipStreamSynchronize(stream_);
    if (e != hipSuccess) return fail(GpuError::Backend, "verify synchronize", e);
    clear_error();
    return GpuError::Ok;
}

GpuError RocmBackend::ecdsa_verify_device(const uint8_t* d_msgs, const uint8_t* d_pubs33, const uint8_t* d_sigs,
                                          size_t count, uint8_t* d_out, hipStream_t st) {
    if (!ready_) return fail(GpuError::Device, "context not initialised");
    const BatchDesc d{kKindEcdsa, (uint32_t)count, d_msgs, d_pubs33, d_sigs, d_out, 0, 0};
    hipError_t e = launch_single(d, false, st ? st : stream_);
    if (e != hipSuccess) return fail(GpuError::Launch, "ECDSA verify launch", e);
    clear_error();
    return GpuError::Ok;
}

GpuError RocmBackend::schnorr_verify_device(const uint8_t* d_msgs, const uint8_t* d_pubx, const uint8_t* d_sigs,
                                            size_t count, uint8_t* d_out, hipStream_t st) {
    if (!ready_) return fail(GpuError::Device, "context not initialised");
    const BatchDesc d{kKindSchnorr, (uint32_t)count, d_msgs, d_pubx, d_sigs, d_out, 0, 0};
    hipError_t e = launch_single(d, false, st ? st : stream_);
    if (e != hipSuccess) return fail(GpuError::Launch, "Schnorr verify launch", e);
    clear_error();
    return GpuError::Ok;
}

template <class Launch>
GpuError RocmBackend::keyop_host(const char* what, size_t n, uint8_t* out, const uint8_t* const* in,
                                 const size_t* bytes, int nin, Launch&& launch) {
    if (!ready_) return fail(GpuError::Device, "context not initialised");
    if (n == 0) {
        clear_error();
        return GpuError::Ok;
    }
    if (n > UINT32_MAX) return fail(GpuError::BadInput, "batch count too large");
    hipError_t e = hipSetDevice(device_);
    if (e != hipSuccess) return fail(GpuError::Device, "hipSetDevice", e);
    const uint8_t* d[3] = {nullptr, nullptr, nullptr};
    for (int k = 0; k < nin; ++k) {
        if ((e = d_in_[k].ensure(bytes[k])) != hipSuccess) return fail(GpuError::Memory, "hipMalloc(inputs)", e);
        if ((e = hipMemcpyAsync(d_in_[k].ptr, in[k], bytes[k], hipMemcpyHostToDevice, stream_)) != hipSuccess)
            return fail(GpuError::Memory, "copy inputs to device", e);
        d[k] = (const uint8_t*)d_in_[k].ptr;
    }
    if ((e = d_out_.ensure(n)) != hipSuccess) return fail(GpuError::Memory, "hipMalloc(results)", e);
    uint8_t* dout = (uint8_t*)d_out_.ptr;
    unsigned grid = (unsigned)((n + 255) / 256);
    const unsigned cap = (unsigned)cus_ * 8u;
    if (grid > cap) grid = cap;
    if ((e = launch(d, (uint32_t)n, dout, grid)) != hipSuccess) return fail(GpuError::Launch, what, e);
    if ((e = hipMemcpyAsync(out, dout, n, hipMemcpyDeviceToHost, stream_)) != hipSuccess)
        return fail(GpuError::Memory, "copy results to host", e);
    if ((e = hipStreamSynchronize(stream_)) != hipSuccess) return fail(GpuError::Backend, "synchronize", e);
    clear_error();
    return GpuError::Ok;
}

GpuError RocmBackend::pubkey_validate(const uint8_t* pk33, size_t n, uint8_t* out) {
    const uint8_t* in[1] = {pk33};
    const size_t bytes[1] = {n * 33};
    return keyop_host("pubkey validate launch", n, out, in, bytes, 1,
                      [&](const uint8_t* const* d, uint32_t cnt, uint8_t* res, unsigned grid) {
                          return launch_pubkey_validate(d[0], cnt, res, grid, stream_);
                      });
}

GpuError RocmBackend::xonly_validate(const uint8_t* x32, size_t n, uint8_t* out) {
    const uint8_t* in[1] = {x32};
    const size_t bytes[1] = {n * 32};
    return keyop_host("x-only validate launch", n, out, in, bytes, 1,
                      [&](const uint8_t* const* d, uint32_t cnt, uint8_t* res, unsigned grid) {
                          return launch_xonly_validate(d[0], cnt, res, grid, stream_);
                      });
}

// internal_x and tweak in two buffers; tweaked_x and the parity bytes share
// the third (parity after the n x-coordinates).
GpuError RocmBackend::commitment_verify(const uint8_t* internal_x32, const uint8_t* tweak32,
                                        const uint8_t* tweaked_x32, const uint8_t* parity, size_t n, uint8_t* out) {
    if (!ready_) return fail(GpuError::Device, "context not initialised");
    hipError_t e = hipSetDevice(device_);
    if (e != hipSuccess) return fail(GpuError::Device, "hipSetDevice", e);
    if ((e = d_in_[2].ensure(n * 33)) != hipSuccess) return fail(GpuError::Memory, "hipMalloc(inputs)", e);
    if (n && (e = hipMemcpyAsync((uint8_t*)d_in_[2].ptr + n * 32, parity, n, hipMemcpyHostToDevice, stream_)) !=
                 hipSuccess)
        return fail(GpuError::Memory, "copy inputs to device", e);
    const uint8_t* in[3] = {internal_x32, tweak32, tweaked_x32};
    const size_t bytes[3] = {n * 32, n * 32, n * 32};
    return keyop_host("commitment verify launch", n, out, in, bytes, 3,
                      [&](const uint8_t* const* d, uint32_t cnt, uint8_t* res, unsigned grid) {
                          return launch_commitment_verify(d[0], d[1], d[2], d[2] + (size_t)cnt * 32, cnt, res,
                                                          gdesc_, grid, stream_);
                      });
}

// Inputs in d_in_[0..2] (digests, signatures, recovery ids), keys and flags
// in d_out_ (33 n bytes of keys, then n flags).  The kernel takes the lane
// slots, so it is ordered after any verify launch like another verify launch.
GpuError RocmBackend::ecrecover_batch(const uint8_t* msgs, const uint8_t* sigs, const int* recids, size_t n,
                                      uint8_t* out33, uint8_t* valid) {
    if (!ready_) return fail(GpuError::Device, "context not initialised");
    if (n == 0) {
        clear_error();
        return GpuError::Ok;
    }
    if (n > UINT32_MAX) return fail(GpuError::BadInput, "batch count too large");
    hipError_t e = hipSetDevice(device_);
    if (e != hipSuccess) return fail(GpuError::Device, "hipSetDevice", e);
    if ((e = d_in_[0].ensure(n * 32)) != hipSuccess || (e = d_in_[1].ensure(n * 64)) != hipSuccess ||
        (e = d_in_[2].ensure(n * sizeof(int))) != hipSuccess || (e = d_out_.ensure(n * 34)) != hipSuccess)
        return fail(GpuError::Memory, "hipMalloc(recover buffers)", e);
    e = hipMemcpyAsync(d_in_[0].ptr, msgs, n * 32, hipMemcpyHostToDevice, stream_);
    if (e == hipSuccess) e = hipMemcpyAsync(d_in_[1].ptr, sigs, n * 64, hipMemcpyHostToDevice, stream_);
    if (e == hipSuccess) e = hipMemcpyAsync(d_in_[2].ptr, recids, n * sizeof(int), hipMemcpyHostToDevice, stream_);
    if (e != hipSuccess) return fail(GpuError::Memory, "copy inputs to device", e);
    uint8_t* dk = (uint8_t*)d_out_.ptr;
    uint8_t* dv = dk + n * 33;
    unsigned grid = (unsigned)((n + kVerifyBlock - 1) / kVerifyBlock);
    if (grid > verify_grid_[0]) grid = verify_grid_[0];
    e = verify_begin(stream_);
    if (e == hipSuccess)
        e = launch_ecrecover((const uint8_t*)d_in_[0].ptr, (const uint8_t*)d_in_[1].ptr, (const int*)d_in_[2].ptr,
                             (uint32_t)n, dk, dv, d_slots_.ptr, gdesc_, grid, stream_);
    if (e == hipSuccess) e = verify_end(stream_);
    if (e != hipSuccess) return fail(GpuError::Launch, "recover launch", e);
    e = hipMemcpyAsync(out33, dk, n * 33, hipMemcpyDeviceToHost, stream_);
    if (e == hipSuccess) e = hipMemcpyAsync(valid, dv, n, hipMemcpyDeviceToHost, stream_);
    if (e != hipSuccess) return fail(GpuError::Memory, "copy results to host", e);
    if ((e = hipStreamSynchronize(stream_)) != hipSuccess) return fail(GpuError::Backend, "synchronize", e);
    clear_error();
    return GpuError::Ok;
}

// Why a caller's batch descriptor is unusable, or nullptr.
static const char* batch_desc_error(const BatchDesc& d) {
    if (d.kind > kKindEcdsaOpaque) return "batch kind must be 0 (ECDSA), 1 (BIP-340) or 2 (opaque ECDSA)";
    if (d.stride != 0 && (d.kind != kKindEcdsaOpaque || d.stride < 129))
        return "row stride is for opaque ECDSA batches only and must be >= 129";
    return nullptr;
}

GpuError RocmBackend::verify_each_device(const BatchDesc* batches, size_t nbatches, hipStream_t st) {
    if (!ready_) return fail(GpuError::Device, "context not initialised");
    st = st ? st : stream_;
    for (size_t b = 0; b < nbatches; ++b) {
        const BatchDesc& d = batches[b];
        if (const char* why = batch_desc_error(d)) return fail(GpuError::BadInput, why);
        hipError_t e = launch_single(d, false, st);
        if (e != hipSuccess) return fail(GpuError::Launch, "per-batch verify launch", e);
    }
    clear_error();
    return GpuError::Ok;
}

GpuError RocmBackend::verify_multi_device(const BatchDesc* batches, size_t nbatches, bool use_graph,
                                          hipStream_t st) {
    if (!ready_) return fail(GpuError::Device, "context not initialised");
    st = st ? st : stream_;
    // tile list: 64-item tiles per batch, in batch order
    size_t ntiles = 0;
    for (size_t b = 0; b < nbatches; ++b) {
        if (const char* why = batch_desc_error(batches[b])) return fail(GpuError::BadInput, why);
        ntiles += (batches[b].count + kTileItems - 1) / kTileItems;
    }
    if (ntiles == 0) {
        clear_error();
        return GpuError::Ok;
    }
    if (ntiles > 0xFFFFFFFFull) return fail(GpuError::BadInput, "too many tiles");
    const size_t desc_bytes = nbatches * sizeof(BatchDesc), tile_bytes = ntiles * sizeof(TileDesc);
    hipError_t e;
    if ((e = h_tables_.ensure(desc_bytes + tile_bytes + 64)) != hipSuccess ||
        (e = d_descs_.ensure(desc_bytes)) != hipSuccess || (e = d_tiles_.ensure(tile_bytes)) != hipSuccess)
        return fail(GpuError::Memory, "multi-batch tables", e);
    // The descriptor, tile and word tables (device and pinned staging) may
    // still feed the previous multi-batch launch, on this stream or another:
    // order this stream after the previous verify launch, then wait on the
    // host for both, before either table is rewritten.
    if ((e = verify_begin(st)) != hipSuccess) return fail(GpuError::Backend, "verify stream order", e);
    if ((e = hipStreamSynchronize(st)) != hipSuccess) return fail(GpuError::Backend, "staging wait", e);
    uint8_t* h = (uint8_t*)h_tables_.ptr;
    std::memcpy(h, batches, desc_bytes);
    TileDesc* tl = (TileDesc*)(h + desc_bytes);
    size_t t = 0;
    for (size_t b = 0; b < nbatches; ++b)
        for (uint32_t f = 0; f < batches[b].count; f += kTileItems) tl[t++] = TileDesc{(uint32_t)b, f};
    uint32_t* words = (uint32_t*)(h + desc_bytes + tile_bytes);
    words[0] = (uint32_t)ntiles;
    uint32_t* d_words = (uint32_t*)d_words_.ptr;   // [0] = ntiles, [16] = work counter
    if ((e = hipMemcpyAsync(d_descs_.ptr, h, desc_bytes, hipMemcpyHostToDevice, st)) != hipSuccess ||
        (e = hipMemcpyAsync(d_tiles_.ptr, tl, tile_bytes, hipMemcpyHostToDevice, st)) != hipSuccess ||
        (e = hipMemcpyAsync(d_words, words, 4, hipMemcpyHostToDevice, st)) != hipSuccess)
        return fail(GpuError::Memory, "copy multi-batch tables", e);
    const BatchDesc* dd = (const BatchDesc*)d_descs_.ptr;
    const TileDesc* dt = (const TileDesc*)d_tiles_.ptr;
    VerifyLaunch a{};
    a.descs = dd;
    a.tiles = dt;
    a.ntiles_dev = d_words;
    a.counter = d_words + 16;
    a.slots = d_slots_.ptr;
    if (!use_graph) {
        if ((e = launch_u32_fill(d_words + 16, 0u, 1, st)) != hipSuccess ||
            (e = launch_verify(2, a, gdesc_, d_mid_, verify_grid_[2], st)) != hipSuccess)
            return fail(GpuError::Launch, "multi-batch verify launch", e);
        if ((e = verify_end(st)) != hipSuccess) return fail(GpuError::Backend, "verify event", e);
        clear_error();
        return GpuError::Ok;
    }
    // Graph: [counter reset (fill kernel), persistent verify].  Captured once and
    // re-captured only when the table buffers move.
    if (!graph_exec_ || graph_descs_ != dd || graph_tiles_ != dt) {
        if (graph_exec_) (void)hipGraphExecDestroy(graph_exec_);
        if (graph_) (void)hipGraphDestroy(graph_);
        graph_exec_ = nullptr;
        graph_ = nullptr;
        if (!graph_stream_ && (e = hipStreamCreateWithFlags(&graph_stream_, hipStreamNonBlocking)) != hipSuccess)
            return fail(GpuError::Queue, "capture stream", e);
        if ((e = hipStreamBeginCapture(graph_stream_, hipStreamCaptureModeThreadLocal)) != hipSuccess)
            return fail(GpuError::Backend, "hipStreamBeginCapture", e);
        hipError_t e1 = launch_u32_fill(d_words + 16, 0u, 1, graph_stream_);
        hipError_t e2 = launch_verify(2, a, gdesc_, d_mid_, verify_grid_[2], graph_stream_);
        e = hipStreamEndCapture(graph_stream_, &graph_);
        if (e1 != hipSuccess || e2 != hipSuccess || e != hipSuccess)
            return fail(GpuError::Backend, "graph capture", e != hipSuccess ? e : (e1 != hipSuccess ? e1 : e2));
        if ((e = hipGraphInstantiate(&graph_exec_, graph_, nullptr, nullptr, 0)) != hipSuccess)
            return fail(GpuError::Backend, "hipGraphInstantiate", e);
        graph_descs_ = dd;
        graph_tiles_ = dt;
    }
    if ((e = hipGraphLaunch(graph_exec_, st)) != hipSuccess) return fail(GpuError::Launch, "hipGraphLaunch", e);
    if ((e = verify_end(st)) != hipSuccess) return fail(GpuError::Backend, "verify event", e);
    clear_error();
    return GpuError::Ok;
}

GpuError RocmBackend::selftest(int op, const uint32_t* a, const uint32_t* b, uint32_t* out, size_t count) {
    if (!ready_) return fail(GpuError::Device, "context not initialised");
    DeviceBuffer da, db, dout, dslot;
    hipError_t e;
    if ((e = da.ensure(count * 64)) != hipSuccess || (e = db.ensure(count * 64)) != hipSuccess ||
        (e = dout.ensure(count * 128)) != hipSuccess || (e = dslot.ensure(count * kSlotBytes)) != hipSuccess) {
        da.release(); db.release(); dout.release(); dslot.release();
        return fail(GpuError::Memory, "selftest buffers", e);
    }
    e = hipMemcpyAsync(da.ptr, a, count * 64, hipMemcpyHostToDevice, stream_);
    if (e == hipSuccess) e = hipMemcpyAsync(db.ptr, b, count * 64, hipMemcpyHostToDevice, stream_);
    if (e == hipSuccess)
        e = launch_selftest(op, (const uint32_t*)da.ptr, (const uint32_t*)db.ptr, (uint32_t*)dout.ptr,
                            (uint32_t)count, gdesc_, dslot.ptr, stream_);
    if (e == hipSuccess) e = hipMemcpyAsync(out, dout.ptr, count * 128, hipMemcpyDeviceToHost, stream_);
    if (e == hipSuccess) e = hipStreamSynchronize(stream_);
    da.release(); db.release(); dout.release(); dslot.release();
    if (e != hipSuccess) return fail(GpuError::Backend, "selftest", e);
    clear_error();
    return GpuError::Ok;
}

}  // namespace ufs
