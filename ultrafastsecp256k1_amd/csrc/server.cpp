// VerifyServer implementation (host side).  See server.hpp.
#include "server.hpp"

#include <algorithm>
#include <chrono>
#include <cstring>
#include <unordered_set>

namespace ufs {

namespace {
constexpr size_t kAlign = 256;
constexpr size_t kNone = SIZE_MAX;
size_t align_up(size_t x, size_t a) { return (x + a - 1) / a * a; }

// Input block of a request: msgs | pubs | sigs, each 16-byte aligned.
struct InLayout {
    size_t msgs, pubs, sigs, len;
};
InLayout in_layout(uint32_t kind, size_t n) {
    InLayout l;
    l.msgs = 0;
    l.pubs = align_up(32 * n, 16);
    l.sigs = align_up(l.pubs + (kind == 0 ? 33 : 32) * n, 16);
    l.len = align_up(l.sigs + 64 * n, kAlign);
    return l;
}
}  // namespace

size_t FifoRing::alloc(size_t len, size_t& charged) {
    if (len > cap) return kNone;
    if (used == 0) head = tail = 0;
    else if (head == tail) return kNone;   // full
    size_t off;
    if (head >= tail) {
        if (len <= cap - head) {
            off = head;
            charged = len;
            head += len;
        } else if (len <= tail) {   // wrap, leaving the end unused
            off = 0;
            charged = (cap - head) + len;
            head = len;
        } else {
            return kNone;
        }
    } else {
        if (len > tail - head) return kNone;
        off = head;
        charged = len;
        head += len;
    }
    if (head == cap) head = 0;
    used += charged;
    return off;
}

void FifoRing::release(size_t charged) {
    used -= charged;
    tail = (tail + charged) % cap;
}

double VerifyServer::now_us() {
    return std::chrono::duration<double, std::micro>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

void VerifyServer::set_error(GpuError e) {
    int z = 0;
    err_.compare_exchange_strong(z, (int)e);
    cv_sched_.notify_all();
    cv_done_.notify_all();
    cv_wait_.notify_all();
}

GpuError VerifyServer::start(const ServerConfig& cfg) {
    cfg_ = cfg;
    if (cfg_.groups < 1) cfg_.groups = 1;
    if (cfg_.groups > 4) cfg_.groups = 4;
    if (cfg_.hbm_budget < (16ull << 20)) cfg_.hbm_budget = 16ull << 20;
    if (cfg_.max_group_items < 64) cfg_.max_group_items = 64;
    if (cfg_.cache_log2 > 26) cfg_.cache_log2 = 26;
    device_ = be_->device();
    hipError_t e = hipSetDevice(device_);
    if (e != hipSuccess) return GpuError::Device;
    verify_variant_ = cfg_.cache_log2 ? 3 : 2;
    verify_grid_ = be_->verify_grid(verify_variant_);

    // Ring sizes: the HBM budget holds the inputs (<= 129 B per item plus
    // alignment); the verdict ring one byte per item.
    const size_t in_cap = cfg_.hbm_budget / kAlign * kAlign;
    const size_t items_cap = in_cap / 129;
    const size_t out_cap = align_up(items_cap + (1 << 20), kAlign);
    in_ring_.cap = in_cap;
    out_ring_.cap = out_cap;
    if ((e = pinned_malloc(&h_in_, in_cap, hipHostMallocDefault)) != hipSuccess) return GpuError::Memory;
    if ((e = pinned_malloc(&h_out_, out_cap, hipHostMallocMapped)) != hipSuccess) return GpuError::Memory;
    if ((e = hipHostGetDevicePointer((void**)&d_out_view_, h_out_, 0)) != hipSuccess) return GpuError::Memory;
    if ((e = dev_malloc(&d_in_, in_cap)) != hipSuccess) return GpuError::Memory;
    if (cfg_.cache_log2) {
        const size_t entries = (size_t)1 << cfg_.cache_log2;
        if ((e = dev_malloc(&d_cache_, entries * 64)) != hipSuccess) return GpuError::Memory;
        if ((e = hipMemset(d_cache_, 0, entries * 64)) != hipSuccess) return GpuError::Memory;
        cache_mask_ = (uint32_t)(entries - 1);
    }

    max_batches_ = 65536;
    max_tiles_ = cfg_.max_group_items / kTileItems + max_batches_;
    table_bytes_ = align_up(max_batches_ * sizeof(BatchDesc), kAlign) + align_up(max_tiles_ * sizeof(TileDesc), kAlign) +
                   align_up(max_batches_ * sizeof(void*), kAlign) + kAlign;
    groups_.resize(cfg_.groups);
    for (Group& g : groups_) {
        if ((e = hipStreamCreateWithFlags(&g.stream, hipStreamNonBlocking)) != hipSuccess) return GpuError::Queue;
        if ((e = hipEventCreateWithFlags(&g.done, hipEventDisableTiming)) != hipSuccess) return GpuError::Backend;
        if ((e = dev_malloc(&g.d_slots, (size_t)verify_grid_ * kVerifyBlock * kSlotBytes)) != hipSuccess)
            return GpuError::Memory;
        if ((e = dev_malloc(&g.d_words, 64 * sizeof(uint32_t))) != hipSuccess) return GpuError::Memory;
        if ((e = dev_malloc(&g.d_tables, table_bytes_)) != hipSuccess) return GpuError::Memory;
        if ((e = pinned_malloc(&g.h_tables, table_bytes_, hipHostMallocDefault)) != hipSuccess)
            return GpuError::Memory;
        if ((e = pinned_malloc(&g.h_key_hits, 64, hipHostMallocDefault)) != hipSuccess)
            return GpuError::Memory;
    }
    sched_thr_ = std::thread([this] { scheduler_loop(); });
    done_thr_ = std::thread([this] { completion_loop(); });
    return GpuError::Ok;
}

VerifyServer::~VerifyServer() {
    {
        std::lock_guard<std::mutex> lk(mu_);
        stop_ = true;
    }
    cv_sched_.notify_all();
    cv_done_.notify_all();
    if (sched_thr_.joinable()) sched_thr_.join();
    if (done_thr_.joinable()) done_thr_.join();
    (void)hipSetDevice(device_);
    for (Group& g : groups_) {
        if (g.stream) (void)hipStreamSynchronize(g.stream);
        if (g.done) (void)hipEventDestroy(g.done);
        if (g.stream) (void)hipStreamDestroy(g.stream);
        (void)dev_free(g.d_slots);
        (void)dev_free(g.d_words);
        (void)dev_free(g.d_tables);
        (void)pinned_free(g.h_tables);
        (void)pinned_free(g.h_key_hits);
    }
    (void)dev_free(d_cache_);
    (void)dev_free(d_in_);
    (void)pinned_free(h_out_);
    (void)pinned_free(h_in_);
    std::unordered_set<Request*> all;
    for (auto& kv : reqs_) all.insert(kv.second);
    for (Request* r : admitted_) all.insert(r);
    for (Request* r : pending_) all.insert(r);
    for (Request* r : all) delete r;
}

bool VerifyServer::try_admit_locked(Request* r) {
    const size_t in_len = r->mapped ? 0 : in_layout(r->kind, r->count).len;
    const size_t out_len = align_up(r->count, kAlign);
    size_t c_in = 0, c_out = 0;
    FifoRing in_try = in_ring_, out_try = out_ring_;
    size_t o_in = 0;
    if (!r->mapped) {
        o_in = in_try.alloc(in_len, c_in);
        if (o_in == kNone) return false;
    }
    const size_t o_out = out_try.alloc(out_len, c_out);
    if (o_out == kNone) return false;
    in_ring_ = in_try;
    out_ring_ = out_try;
    r->in_off = o_in;
    r->in_len = in_len;
    r->in_charge = c_in;
    r->out_off = o_out;
    r->out_charge = c_out;
    if (r->mapped) r->staged = true;
    admitted_.push_back(r);
    to_launch_.push_back(r);
    return true;
}

void VerifyServer::stage_inputs(Request* r) {
    if (r->mapped) return;
    const InLayout l = in_layout(r->kind, r->count);
    uint8_t* base = h_in_ + r->in_off;
    std::memcpy(base + l.msgs, r->msgs, 32ull * r->count);
    std::memcpy(base + l.pubs, r->pubs, (r->kind == 0 ? 33ull : 32ull) * r->count);
    std::memcpy(base + l.sigs, r->sigs, 64ull * r->count);
}

GpuError VerifyServer::submit(uint32_t kind, const uint8_t* msgs, const uint8_t* pubs, const uint8_t* sigs,
                              size_t count, uint8_t* out, bool mapped, uint64_t* id) {
    if (err_.load()) return error();
    if (kind > 1 || count == 0 || count > cfg_.max_group_items) return GpuError::BadInput;
    if (in_layout(kind, count).len > in_ring_.cap) return GpuError::BadInput;
    // in-place reads use 16-byte loads of the 32-byte fields
    if (mapped && ((((uintptr_t)msgs | (uintptr_t)sigs) & 15) || (kind == 1 && ((uintptr_t)pubs & 15))))
        return GpuError::BadInput;
    Request* r = new Request();
    r->mapped = mapped;
    r->kind = kind;
    r->count = (uint32_t)count;
    r->msgs = msgs;
    r->pubs = pubs;
    r->sigs = sigs;
    r->out = out;
    r->t_submit = now_us();
    bool admitted;
    {
        std::lock_guard<std::mutex> lk(mu_);
        r->id = next_id_++;
        reqs_[r->id] = r;
        ++outstanding_;
        admitted = pending_.empty() && try_admit_locked(r);
        if (!admitted) pending_.push_back(r);
    }
    *id = r->id;
    if (admitted && !mapped) {
        stage_inputs(r);
        {
            std::lock_guard<std::mutex> lk(mu_);
            r->staged = true;
        }
    }
    if (admitted) cv_sched_.notify_one();
    return GpuError::Ok;
}

GpuError VerifyServer::launch_group(Group& g, std::vector<Request*>& batch) {
    hipStream_t st = g.stream;
    // inputs: the staged input blocks of a group are contiguous in ring order
    // except for at most one wrap; mapped requests are read in place
    hipError_t e = hipSuccess;
    bool have_run = false;
    size_t run_start = 0, run_end = 0;
    for (Request* r : batch) {
        if (r->mapped) continue;
        if (have_run && r->in_off != run_end) {
            e = hipMemcpyAsync(d_in_ + run_start, h_in_ + run_start, run_end - run_start, hipMemcpyHostToDevice, st);
            if (e != hipSuccess) return GpuError::Memory;
            have_run = false;
        }
        if (!have_run) {
            run_start = r->in_off;
            have_run = true;
        }
        run_end = r->in_off + r->in_len;
    }
    if (have_run) {
        e = hipMemcpyAsync(d_in_ + run_start, h_in_ + run_start, run_end - run_start, hipMemcpyHostToDevice, st);
        if (e != hipSuccess) return GpuError::Memory;
    }

    // tables: descriptors | tiles | tile count
    const size_t nb = batch.size();
    const size_t off_tiles = align_up(nb * sizeof(BatchDesc), kAlign);
    BatchDesc* hd = reinterpret_cast<BatchDesc*>(g.h_tables);
    TileDesc* ht = reinterpret_cast<TileDesc*>(g.h_tables + off_tiles);
    size_t nt = 0;
    for (size_t b = 0; b < nb; ++b) {
        const Request* r = batch[b];
        if (r->mapped) {
            hd[b] = BatchDesc{r->kind, r->count, r->msgs, r->pubs, r->sigs, d_out_view_ + r->out_off};
        } else {
            const InLayout l = in_layout(r->kind, r->count);
            const uint8_t* base = d_in_ + r->in_off;
            hd[b] = BatchDesc{r->kind, r->count, base + l.msgs, base + l.pubs, base + l.sigs,
                              d_out_view_ + r->out_off};
        }
        for (uint32_t f = 0; f < r->count; f += kTileItems) ht[nt++] = TileDesc{(uint32_t)b, f};
    }
    const size_t off_cnt = off_tiles + align_up(nt * sizeof(TileDesc), kAlign);
    *reinterpret_cast<uint32_t*>(g.h_tables + off_cnt) = (uint32_t)nt;
    const size_t tbytes = off_cnt + sizeof(uint32_t);
    if ((e = hipMemcpyAsync(g.d_tables, g.h_tables, tbytes, hipMemcpyHostToDevice, st)) != hipSuccess)
        return GpuError::Memory;
    const BatchDesc* dd = reinterpret_cast<const BatchDesc*>(g.d_tables);
    const TileDesc* dt = reinterpret_cast<const TileDesc*>(g.d_tables + off_tiles);
    const uint32_t* dn = reinterpret_cast<const uint32_t*>(g.d_tables + off_cnt);

    // counters: [16] work counter, [32] key-cache hits
    if ((e = launch_u32_fill(g.d_words, 0u, 48, st)) != hipSuccess) return GpuError::Launch;
    VerifyLaunch a{};
    a.descs = dd;
    a.tiles = dt;
    a.ntiles_dev = dn;
    a.counter = g.d_words + 16;
    a.slots = g.d_slots;
    // Keys are resolved inside the verify kernel, through the cache when
    // there is one (variant 3), else each decompressed there (variant 2).
    // A separate lookup pass with dense decompression of the misses was
    // slower at every measured hit rate, up to 96%: a 64-item wave with any
    // miss waits for its square root either way, and two more launches per
    // group cost more than the square roots skipped.
    if (d_cache_) {
        a.key_cache = d_cache_;
        a.key_cache_mask = cache_mask_;
        a.key_hits = g.d_words + 32;
    }
    if ((e = launch_verify(verify_variant_, a, be_->gcomb(), be_->bip340_midstate(), verify_grid_, st)) !=
        hipSuccess)
        return GpuError::Launch;
    if ((e = hipMemcpyAsync(g.h_key_hits, g.d_words + 32, sizeof(uint32_t), hipMemcpyDeviceToHost, st)) !=
        hipSuccess)
        return GpuError::Memory;
    if ((e = hipEventRecord(g.done, st)) != hipSuccess) return GpuError::Backend;
    return GpuError::Ok;
}

void VerifyServer::scheduler_loop() {
    (void)hipSetDevice(device_);
    std::unique_lock<std::mutex> lk(mu_);
    std::vector<Request*> batch;
    for (;;) {
        Group* g = nullptr;
        for (Group& c : groups_)
            if (!c.busy) {
                g = &c;
                break;
            }
        const bool ready = !to_launch_.empty() && to_launch_.front()->staged;
        if (stop_ || err_.load()) break;
        if (!g || !ready) {
            cv_sched_.wait(lk);
            continue;
        }
        batch.clear();
        uint64_t items = 0;
        while (!to_launch_.empty() && to_launch_.front()->staged && batch.size() < max_batches_) {
            Request* r = to_launch_.front();
            if (!batch.empty() && items + r->count > cfg_.max_group_items) break;
            batch.push_back(r);
            items += r->count;
            to_launch_.pop_front();
        }
        g->busy = true;
        g->reqs = batch;
        g->items = items;
        lk.unlock();
        const GpuError ge = launch_group(*g, batch);
        lk.lock();
        if (ge != GpuError::Ok) {
            set_error(ge);
            break;
        }
        if (inflight_.empty()) busy_since_ = now_us();
        inflight_.push_back(g);
        ++st_groups_;
        cv_done_.notify_all();
    }
}

void VerifyServer::finish_group(Group& g) {
    for (Request* r : g.reqs) std::memcpy(r->out, h_out_ + r->out_off, r->count);
    const double t = now_us();
    std::vector<Request*> to_stage;
    {
        std::lock_guard<std::mutex> lk(mu_);
        for (Request* r : g.reqs) {
            r->done = true;
            lat_us_.push_back((float)(t - r->t_submit));
            ++st_requests_;
            st_items_ += r->count;
            --outstanding_;
        }
        st_lookups_ += g.items;
        st_hits_ += d_cache_ ? std::min<uint64_t>(g.items, *g.h_key_hits) : 0;
        g.reqs.clear();
        g.busy = false;
        inflight_.erase(std::find(inflight_.begin(), inflight_.end(), &g));
        if (inflight_.empty() && busy_since_ >= 0) {
            busy_us_ += t - busy_since_;
            busy_since_ = -1;
        }
        // release ring space in admission order
        while (!admitted_.empty() && admitted_.front()->done) {
            Request* r = admitted_.front();
            admitted_.pop_front();
            in_ring_.release(r->in_charge);
            out_ring_.release(r->out_charge);
            r->released = true;
            if (r->observed) delete r;
        }
        // admit waiting requests into the freed space
        while (!pending_.empty() && try_admit_locked(pending_.front())) {
            to_stage.push_back(pending_.front());
            pending_.pop_front();
        }
    }
    for (Request* r : to_stage) stage_inputs(r);
    {
        std::lock_guard<std::mutex> lk(mu_);
        for (Request* r : to_stage) r->staged = true;
    }
    cv_sched_.notify_all();
    cv_wait_.notify_all();
}

void VerifyServer::completion_loop() {
    (void)hipSetDevice(device_);
    std::vector<Group*> snap;
    for (;;) {
        {
            std::unique_lock<std::mutex> lk(mu_);
            cv_done_.wait(lk, [&] { return stop_ || !inflight_.empty() || err_.load(); });
            if (inflight_.empty() && (stop_ || err_.load())) break;
            snap.assign(inflight_.begin(), inflight_.end());
        }
        bool any = false;
        for (Group* g : snap) {
            const hipError_t q = hipEventQuery(g->done);
            if (q == hipSuccess) {
                finish_group(*g);
                any = true;
            } else if (q != hipErrorNotReady) {
                set_error(GpuError::Launch);
                std::lock_guard<std::mutex> lk(mu_);
                inflight_.clear();
                break;
            }
        }
        if (!any) std::this_thread::sleep_for(std::chrono::microseconds(20));
    }
    cv_wait_.notify_all();
}

int VerifyServer::poll(uint64_t id) {
    std::lock_guard<std::mutex> lk(mu_);
    auto it = reqs_.find(id);
    if (it == reqs_.end()) return -1;
    Request* r = it->second;
    if (!r->done) return 0;
    reqs_.erase(it);
    r->observed = true;
    if (r->released) delete r;
    return 1;
}

GpuError VerifyServer::wait(uint64_t id) {
    std::unique_lock<std::mutex> lk(mu_);
    auto it = reqs_.find(id);
    if (it == reqs_.end()) return GpuError::BadInput;
    Request* r = it->second;
    cv_wait_.wait(lk, [&] { return r->done || err_.load(); });
    if (!r->done) return error();
    reqs_.erase(id);
    r->observed = true;
    if (r->released) delete r;
    return GpuError::Ok;
}

GpuError VerifyServer::drain() {
    std::unique_lock<std::mutex> lk(mu_);
    cv_wait_.wait(lk, [&] { return outstanding_ == 0 || err_.load(); });
    if (outstanding_ != 0) return error();
    for (auto it = reqs_.begin(); it != reqs_.end();) {
        Request* r = it->second;
        it = reqs_.erase(it);
        r->observed = true;
        if (r->released) delete r;
    }
    return GpuError::Ok;
}

void VerifyServer::stats(ServerStats& out, bool reset) {
    std::lock_guard<std::mutex> lk(mu_);
    out = ServerStats{};
    out.requests_completed = st_requests_;
    out.items_completed = st_items_;
    out.groups_launched = st_groups_;
    out.key_lookups = st_lookups_;
    out.key_cache_hits = st_hits_;
    double busy = busy_us_;
    if (busy_since_ >= 0) busy += now_us() - busy_since_;
    out.busy_seconds = busy * 1e-6;
    if (!lat_us_.empty()) {
        std::vector<float> v = lat_us_;
        const size_t n = v.size();
        auto pct = [&](double q) {
            size_t k = (size_t)(q * (double)(n - 1) + 0.5);
            std::nth_element(v.begin(), v.begin() + k, v.end());
            return (double)v[k];
        };
        out.latency_p50_us = pct(0.50);
        out.latency_p99_us = pct(0.99);
        double s = 0, mx = 0;
        for (float x : lat_us_) {
            s += x;
            mx = std::max(mx, (double)x);
        }
        out.latency_mean_us = s / (double)n;
        out.latency_max_us = mx;
    }
    if (reset) {
        lat_us_.clear();
        st_requests_ = st_items_ = st_groups_ = st_lookups_ = st_hits_ = 0;
        busy_us_ = 0;
        if (busy_since_ >= 0) busy_since_ = now_us();
    }
}

}  // namespace ufs
