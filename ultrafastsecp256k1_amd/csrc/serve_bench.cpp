// Serving benchmark: Poisson arrivals of verify requests against the
// continuous-batching server, compared with a plain server that makes one
// synchronous batch call (one launch) per request.
//
// Workload: request sizes log-uniform in [1, max_size] (floor(exp(u ln(max+1)))),
// kind ECDSA with probability 0.7 else BIP-340; items are taken in order,
// cyclically, from two pre-generated serving workloads in which 20% (by default) of the
// keys come from a seeded pool of 2^16 keys (so the device decompression cache
// has something to keep) and 1% of the items are corrupted.  For each server
// the sustained peak is measured first (a saturating backlog); then arrivals at
// 50%, 80% and 95% of that peak run for a fixed time and the latency of every
// request (arrival -> verdicts in the caller's buffer) is recorded.  With
// --load-rate R the batched loads are fractions of R sigs/s instead of the
// measured peak (the same offered load for every server configuration).
//
//   serve_bench [--seconds S] [--loads 0.5,0.8,0.95] [--ecdsa-items N]
//               [--schnorr-items N] [--max-size N] [--cache-log2 K]
//               [--groups G] [--budget-mb M] [--no-plain] [--peak-only]
//               [--load-rate R]
// Prints one JSON object.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "ufsecp/ufsecp_gpu.h"
#include "ufsecp/ufsecp_gpu_ext.h"

extern "C" int ufsecp_wl_ecdsa_pool(uint64_t seed, uint64_t first, uint64_t n, uint64_t pool_seed, uint64_t pool_size,
                                    uint32_t pool_permille, uint8_t* msgs, uint8_t* pubs, uint8_t* sigs,
                                    uint8_t* corrupted, int threads);
extern "C" int ufsecp_wl_schnorr_pool(uint64_t seed, uint64_t first, uint64_t n, uint64_t pool_seed,
                                      uint64_t pool_size, uint32_t pool_permille, uint8_t* msgs, uint8_t* pubs,
                                      uint8_t* sigs, uint8_t* corrupted, int threads);

namespace {

using clk = std::chrono::steady_clock;
double now_s() { return std::chrono::duration<double>(clk::now().time_since_epoch()).count(); }

struct SplitMix {
    uint64_t s;
    explicit SplitMix(uint64_t seed) : s(seed) {}
    uint64_t next() {
        s += 0x9E3779B97F4A7C15ULL;
        uint64_t z = s;
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
        return z ^ (z >> 31);
    }
    double uniform() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }
};

struct Pool {
    uint32_t kind = 0;
    size_t n = 0;
    uint8_t *msgs = nullptr, *pubs = nullptr, *sigs = nullptr, *corr = nullptr;
    size_t pub_stride() const { return kind == 0 ? 33 : 32; }
};

struct Req {
    uint32_t kind, size;
    size_t first;     // first item in the kind's pool
    size_t out_off;   // offset in the run's verdict buffer
    double t_arrive, t_done;
};

struct Args {
    double seconds = 3.0;
    std::vector<double> loads = {0.5, 0.8, 0.95};
    size_t ecdsa_items = 4u << 20, schnorr_items = 2u << 20;
    uint32_t max_size = 8192;
    uint32_t cache_log2 = 20;
    uint32_t pool_permille = 200;   // share of items whose key comes from the 2^16-key pool
    uint32_t groups = 0;
    uint64_t budget_mb = 512;
    bool plain = true;
    bool peak_only = false;
    double load_rate = 0;   // > 0: batched loads are fractions of this rate
};

[[noreturn]] void die(const char* what, int rc) {
    std::fprintf(stderr, "serve_bench: %s failed (%d)\n", what, rc);
    std::exit(1);
}

// Request stream: sizes log-uniform in [1, max], 70% ECDSA, consecutive
// items of the kind's pool (wrapping to the start when a request would cross
// the end).
struct Plan {
    SplitMix rng;
    uint32_t max_size;
    size_t cursor[2] = {0, 0};
    Plan(uint64_t seed, uint32_t m) : rng(seed), max_size(m) {}
    Req next(const Pool* pools) {
        Req r{};
        uint32_t size = (uint32_t)std::floor(std::exp(rng.uniform() * std::log((double)max_size + 1.0)));
        r.size = std::min(std::max(size, 1u), max_size);
        r.kind = (rng.next() % 1000) < 700 ? 0u : 1u;
        const Pool& p = pools[r.kind];
        if (cursor[r.kind] + r.size > p.n) cursor[r.kind] = 0;
        r.first = cursor[r.kind];
        cursor[r.kind] += r.size;
        return r;
    }
};

double mean_size(uint32_t m) { return ((double)m + 1.0 - 1.0) / std::log((double)m + 1.0); }

struct RunResult {
    double offered = 0, achieved = 0, p50 = 0, p99 = 0, mean = 0, max = 0, seconds = 0, late_max = 0;
    uint64_t requests = 0, items = 0, mismatches = 0, groups = 0, lookups = 0, hits = 0;
};

void latency_stats(std::vector<Req>& reqs, RunResult& r) {
    std::vector<double> lat;
    lat.reserve(reqs.size());
    for (const Req& q : reqs) lat.push_back((q.t_done - q.t_arrive) * 1e3);
    if (lat.empty()) return;
    std::sort(lat.begin(), lat.end());
    auto pct = [&](double f) { return lat[(size_t)std::min<double>(lat.size() - 1, std::floor(f * (lat.size() - 1) + 0.5))]; };
    r.p50 = pct(0.50);
    r.p99 = pct(0.99);
    r.max = lat.back();
    double s = 0;
    for (double x : lat) s += x;
    r.mean = s / (double)lat.size();
}

uint64_t check_verdicts(const std::vector<Req>& reqs, const uint8_t* out, const Pool* pools) {
    uint64_t bad = 0;
    for (const Req& q : reqs) {
        const Pool& p = pools[q.kind];
        for (uint32_t i = 0; i < q.size; ++i)
            bad += out[q.out_off + i] != (uint8_t)(1 - p.corr[q.first + i]);
    }
    return bad;
}

// Continuous-batching server.  rate = requests/s (0 = saturating backlog of
// `backlog_items` items kept in flight).
RunResult run_batched(ufsecp_gpu_server* srv, const Pool* pools, Plan& plan, double rate, double seconds,
                      std::vector<uint8_t>& outbuf) {
    RunResult res;
    std::vector<Req> reqs;
    std::vector<uint64_t> ids;
    SplitMix arr(0xA7A7 + (uint64_t)(rate * 1000));
    ufsecp_gpu_server_stats_t st;
    ufsecp_gpu_ext_server_stats(srv, &st, 1);
    // plan the run
    const double t0 = now_s() + 0.002;
    double t = t0;
    size_t out_off = 0;
    const size_t cap_items = outbuf.size();
    const size_t backlog_items = 6u << 20;
    while (true) {
        Req q = plan.next(pools);
        if (rate > 0) {
            t += -std::log(1.0 - arr.uniform()) / rate;
            if (t - t0 > seconds) break;
        } else if (out_off + q.size > std::min(cap_items, backlog_items * 8)) {
            break;
        }
        if (out_off + q.size > cap_items) break;
        q.t_arrive = rate > 0 ? t : 0.0;
        q.out_off = out_off;
        out_off += q.size;
        reqs.push_back(q);
    }
    ids.resize(reqs.size());
    // submit at the arrival times (saturating: keep backlog_items in flight)
    size_t inflight_items = 0, oldest = 0;
    const double start = now_s();
    for (size_t k = 0; k < reqs.size(); ++k) {
        Req& q = reqs[k];
        if (rate > 0) {
            for (;;) {
                const double d = q.t_arrive - now_s();
                if (d <= 0) break;
                if (d > 300e-6) std::this_thread::sleep_for(std::chrono::duration<double>(d - 200e-6));
            }
        } else {
            while (inflight_items + q.size > backlog_items && oldest < k) {
                if (ufsecp_gpu_ext_server_wait(srv, ids[oldest]) != UFSECP_OK) die("server_wait", -1);
                reqs[oldest].t_done = now_s();
                inflight_items -= reqs[oldest].size;
                ++oldest;
            }
            q.t_arrive = now_s();
        }
        const double ts = now_s();
        res.late_max = std::max(res.late_max, ts - q.t_arrive);
        const Pool& p = pools[q.kind];
        const int rc = ufsecp_gpu_ext_server_submit_mapped(srv, q.kind, p.msgs + 32 * q.first,
                                                           p.pubs + p.pub_stride() * q.first, p.sigs + 64 * q.first,
                                                           q.size, outbuf.data() + q.out_off, &ids[k]);
        if (rc != UFSECP_OK) die("server_submit_mapped", rc);
        inflight_items += q.size;
    }
    // completion times: poll in submission order
    for (size_t k = oldest; k < reqs.size(); ++k) {
        for (;;) {
            const int r = ufsecp_gpu_ext_server_poll(srv, ids[k]);
            if (r == 1) break;
            if (r < 0) die("server_poll", r);
            std::this_thread::sleep_for(std::chrono::microseconds(20));
        }
        reqs[k].t_done = now_s();
    }
    const double end = now_s();
    // completion is observed in submission order above; the server's own
    // per-request times (submit -> done) are reported beside
    ufsecp_gpu_ext_server_stats(srv, &st, 0);
    res.seconds = end - start;
    res.requests = reqs.size();
    for (const Req& q : reqs) res.items += q.size;
    res.achieved = (double)res.items / res.seconds;
    res.offered = rate > 0 ? rate * mean_size(plan.max_size) : 0;
    res.groups = st.groups_launched;
    res.lookups = st.key_lookups;
    res.hits = st.key_cache_hits;
    // latency percentiles from the server (arrival == submit when on time)
    res.p50 = st.latency_p50_us / 1e3;
    res.p99 = st.latency_p99_us / 1e3;
    res.mean = st.latency_mean_us / 1e3;
    res.max = st.latency_max_us / 1e3;
    res.mismatches = check_verdicts(reqs, outbuf.data(), pools);
    return res;
}

// Plain server: one worker verifies requests FIFO with one synchronous batch
// call (one launch) each.
RunResult run_plain(ufsecp_gpu_ctx* ctx, const Pool* pools, Plan& plan, double rate, double seconds,
                    std::vector<uint8_t>& outbuf, size_t backlog_requests) {
    RunResult res;
    std::vector<Req> reqs;
    SplitMix arr(0xB7B7 + (uint64_t)(rate * 1000));
    const double t0 = now_s() + 0.002;
    double t = t0;
    size_t out_off = 0;
    while (true) {
        Req q = plan.next(pools);
        if (rate > 0) {
            t += -std::log(1.0 - arr.uniform()) / rate;
            if (t - t0 > seconds) break;
        } else if (reqs.size() >= backlog_requests) {
            break;
        }
        if (out_off + q.size > outbuf.size()) break;
        q.t_arrive = rate > 0 ? t : t0;
        q.out_off = out_off;
        out_off += q.size;
        reqs.push_back(q);
    }
    std::mutex mu;
    std::condition_variable cv;
    std::deque<size_t> queue;
    bool done_arrivals = false;
    std::atomic<int> err{0};
    std::thread worker([&] {
        for (;;) {
            size_t k;
            {
                std::unique_lock<std::mutex> lk(mu);
                cv.wait(lk, [&] { return !queue.empty() || done_arrivals; });
                if (queue.empty()) return;
                k = queue.front();
                queue.pop_front();
            }
            Req& q = reqs[k];
            const Pool& p = pools[q.kind];
            const int rc = q.kind == 0
                ? ufsecp_gpu_ecdsa_verify_batch(ctx, p.msgs + 32 * q.first, p.pubs + 33 * q.first,
                                                p.sigs + 64 * q.first, q.size, outbuf.data() + q.out_off)
                : ufsecp_gpu_schnorr_verify_batch(ctx, p.msgs + 32 * q.first, p.pubs + 32 * q.first,
                                                  p.sigs + 64 * q.first, q.size, outbuf.data() + q.out_off);
            if (rc != UFSECP_OK) err = rc;
            q.t_done = now_s();
        }
    });
    const double start = now_s();
    for (size_t k = 0; k < reqs.size(); ++k) {
        if (rate > 0) {
            for (;;) {
                const double d = reqs[k].t_arrive - now_s();
                if (d <= 0) break;
                if (d > 300e-6) std::this_thread::sleep_for(std::chrono::duration<double>(d - 200e-6));
            }
        } else {
            reqs[k].t_arrive = start;
        }
        {
            std::lock_guard<std::mutex> lk(mu);
            queue.push_back(k);
        }
        cv.notify_one();
    }
    {
        std::lock_guard<std::mutex> lk(mu);
        done_arrivals = true;
    }
    cv.notify_one();
    worker.join();
    if (err) die("plain verify", err);
    const double end = now_s();
    res.seconds = end - start;
    res.requests = reqs.size();
    for (const Req& q : reqs) res.items += q.size;
    res.achieved = (double)res.items / res.seconds;
    res.offered = rate > 0 ? rate * mean_size(plan.max_size) : 0;
    latency_stats(reqs, res);
    res.mismatches = check_verdicts(reqs, outbuf.data(), pools);
    return res;
}

void print_run(const char* name, const RunResult& r, double frac, bool last) {
    std::printf("{\"load\": %s, \"frac\": %.2f, \"offered_sigs_per_s\": %.0f, \"achieved_sigs_per_s\": %.0f, "
                "\"requests\": %llu, \"items\": %llu, \"seconds\": %.3f, \"p50_ms\": %.3f, \"p99_ms\": %.3f, "
                "\"mean_ms\": %.3f, \"max_ms\": %.3f, \"groups\": %llu, \"key_lookups\": %llu, "
                "\"key_cache_hits\": %llu, \"max_submit_lateness_ms\": %.3f, \"mismatches\": %llu}%s",
                name, frac, r.offered, r.achieved, (unsigned long long)r.requests, (unsigned long long)r.items,
                r.seconds, r.p50, r.p99, r.mean, r.max, (unsigned long long)r.groups, (unsigned long long)r.lookups,
                (unsigned long long)r.hits, r.late_max * 1e3, (unsigned long long)r.mismatches, last ? "" : ", ");
}

}  // namespace

int main(int argc, char** argv) {
    Args a;
    for (int i = 1; i < argc; ++i) {
        std::string s = argv[i];
        auto val = [&]() -> const char* {
            if (i + 1 >= argc) die("missing value", 2);
            return argv[++i];
        };
        if (s == "--seconds") a.seconds = std::atof(val());
        else if (s == "--ecdsa-items") a.ecdsa_items = std::strtoull(val(), nullptr, 10);
        else if (s == "--schnorr-items") a.schnorr_items = std::strtoull(val(), nullptr, 10);
        else if (s == "--max-size") a.max_size = (uint32_t)std::atoi(val());
        else if (s == "--cache-log2") a.cache_log2 = (uint32_t)std::atoi(val());
        else if (s == "--pool-permille") a.pool_permille = (uint32_t)std::atoi(val());
        else if (s == "--groups") a.groups = (uint32_t)std::atoi(val());
        else if (s == "--budget-mb") a.budget_mb = std::strtoull(val(), nullptr, 10);
        else if (s == "--no-plain") a.plain = false;
        else if (s == "--load-rate") a.load_rate = std::atof(val());
        else if (s == "--peak-only") {
            a.peak_only = true;
            a.plain = false;
            a.loads.clear();
        }
        else if (s == "--loads") {
            a.loads.clear();
            std::string v = val();
            size_t p = 0;
            while (p < v.size()) {
                size_t q = v.find(',', p);
                if (q == std::string::npos) q = v.size();
                a.loads.push_back(std::atof(v.substr(p, q - p).c_str()));
                p = q + 1;
            }
        } else {
            std::fprintf(stderr, "unknown argument %s\n", s.c_str());
            return 2;
        }
    }
    ufsecp_gpu_ctx* ctx = nullptr;
    int rc = ufsecp_gpu_ctx_create(&ctx, UFSECP_GPU_BACKEND_ROCM, 0);
    if (rc != UFSECP_OK) die("ctx_create", rc);

    // item pools in pinned, device-mapped host memory
    Pool pools[2];
    const size_t counts[2] = {a.ecdsa_items, a.schnorr_items};
    const double tg = now_s();
    for (uint32_t k = 0; k < 2; ++k) {
        Pool& p = pools[k];
        p.kind = k;
        p.n = counts[k];
        void* mem = nullptr;
        const size_t bytes = p.n * (32 + p.pub_stride() + 64 + 1) + 64;
        if ((rc = ufsecp_gpu_ext_host_alloc(ctx, bytes, &mem)) != UFSECP_OK) die("host_alloc", rc);
        uint8_t* b = (uint8_t*)mem;
        p.msgs = b;
        p.sigs = b + 32 * p.n;
        p.pubs = p.sigs + 64 * p.n;
        p.corr = p.pubs + p.pub_stride() * p.n;
        rc = k == 0 ? ufsecp_wl_ecdsa_pool(0x5E4A, 0, p.n, 0x9001, 1u << 16, a.pool_permille, p.msgs, p.pubs, p.sigs,
                                           p.corr, 0)
                    : ufsecp_wl_schnorr_pool(0x5E4B, 0, p.n, 0x9001, 1u << 16, a.pool_permille, p.msgs, p.pubs, p.sigs,
                                             p.corr,
                                             0);
        if (rc) die("workload", rc);
    }
    const double gen_s = now_s() - tg;
    std::fprintf(stderr, "[serve] generated %zu ECDSA + %zu BIP-340 items in %.1f s\n", pools[0].n, pools[1].n,
                 gen_s);

    ufsecp_gpu_server_config_t cfg{};
    cfg.hbm_budget_bytes = a.budget_mb << 20;
    cfg.inflight_groups = a.groups;   // 0: the tuning file's value
    cfg.key_cache_log2 = a.cache_log2;
    ufsecp_gpu_server* srv = nullptr;
    if ((rc = ufsecp_gpu_ext_server_create(ctx, &cfg, &srv)) != UFSECP_OK) die("server_create", rc);

    const double msize = mean_size(a.max_size);
    std::vector<uint8_t> outbuf((size_t)1 << 28);
    Plan plan(0x5EB7, a.max_size);
    // warm-up (fills the key cache), then the saturated peak
    (void)run_batched(srv, pools, plan, 0, 0, outbuf);
    RunResult peak = run_batched(srv, pools, plan, 0, 0, outbuf);
    std::fprintf(stderr, "[serve] batched peak %.2f M sigs/s (%llu requests)\n", peak.achieved / 1e6,
                 (unsigned long long)peak.requests);
    std::vector<RunResult> runs;
    const double load_base = a.load_rate > 0 ? a.load_rate : peak.achieved;
    for (double f : a.loads) {
        RunResult r = run_batched(srv, pools, plan, f * load_base / msize, a.seconds, outbuf);
        std::fprintf(stderr, "[serve] batched %.0f%%: %.2f M sigs/s p50 %.2f ms p99 %.2f ms (%llu req, %llu bad)\n",
                     f * 100, r.achieved / 1e6, r.p50, r.p99, (unsigned long long)r.requests,
                     (unsigned long long)r.mismatches);
        runs.push_back(r);
    }
    ufsecp_gpu_ext_server_drain(srv);
    ufsecp_gpu_ext_server_destroy(srv);

    RunResult ppeak;
    std::vector<RunResult> pruns;
    if (a.plain) {
        Plan pplan(0x5EB8, a.max_size);
        (void)run_plain(ctx, pools, pplan, 0, 0, outbuf, 50);
        ppeak = run_plain(ctx, pools, pplan, 0, 0, outbuf, 600);
        std::fprintf(stderr, "[serve] plain peak %.3f M sigs/s\n", ppeak.achieved / 1e6);
        for (double f : a.loads) {
            RunResult r = run_plain(ctx, pools, pplan, f * ppeak.achieved / msize, a.seconds, outbuf, 0);
            std::fprintf(stderr, "[serve] plain %.0f%%: %.3f M sigs/s p50 %.2f ms p99 %.2f ms\n", f * 100,
                         r.achieved / 1e6, r.p50, r.p99);
            pruns.push_back(r);
        }
    }

    uint64_t bad = peak.mismatches + ppeak.mismatches;
    for (auto& r : runs) bad += r.mismatches;
    for (auto& r : pruns) bad += r.mismatches;
    std::printf("{\"workload\": \"Poisson arrivals, sizes log-uniform [1,%u] (mean %.0f), 70/30 ECDSA/BIP-340, "
                "%.1f%% of keys from a 2^16 pool, 1%% corrupted\", \"mean_request_size\": %.1f, "
                "\"pool_items\": [%zu, %zu], \"generate_seconds\": %.1f, "
                "\"config\": {\"hbm_budget_mb\": %llu, \"inflight_groups\": %u, \"key_cache_log2\": %u}, ",
                a.max_size, msize, a.pool_permille / 10.0, msize, pools[0].n, pools[1].n, gen_s,
                (unsigned long long)a.budget_mb, a.groups, a.cache_log2);
    std::printf("\"batched\": {\"peak_sigs_per_s\": %.0f, \"peak_run\": ", peak.achieved);
    print_run("\"peak\"", peak, 1.0, true);
    std::printf(", \"loads\": [");
    for (size_t i = 0; i < runs.size(); ++i) print_run("\"poisson\"", runs[i], a.loads[i], i + 1 == runs.size());
    std::printf("]}");
    if (a.plain) {
        std::printf(", \"plain\": {\"peak_sigs_per_s\": %.0f, \"loads\": [", ppeak.achieved);
        for (size_t i = 0; i < pruns.size(); ++i)
            print_run("\"poisson\"", pruns[i], a.loads[i], i + 1 == pruns.size());
        std::printf("]}");
    }
    std::printf(", \"parity\": %s, \"mismatches\": %llu}\n", bad == 0 ? "true" : "false", (unsigned long long)bad);
    for (uint32_t k = 0; k < 2; ++k) ufsecp_gpu_ext_host_free(ctx, pools[k].msgs);
    ufsecp_gpu_ctx_destroy(ctx);
    return bad == 0 ? 0 : 1;
}
