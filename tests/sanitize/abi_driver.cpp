// Host-side exercise of the ufsecp GPU C ABI under AddressSanitizer and
// UndefinedBehaviorSanitizer.
//
// Built with -fsanitize=address,undefined and linked against
// libufsecp_gpu_san.so, a copy of the library whose host code carries the
// same instrumentation (device code is unchanged).  Every buffer handed to the
// library is a heap block of exactly the size the ABI documents, so a host
// read or write past an input or output is reported by ASan; any undefined
// behaviour on the host aborts (-fno-sanitize-recover).
//
//   abi_driver --host     argument checks and the no-device paths: NULL
//                         arguments, counts of 0 and above 2^26, backend and
//                         device ids out of range, every unsupported entry
//                         point, the engine hook.  Runs without a GPU.
//   abi_driver --device   the above, then real work on device 0: batch verify
//                         (all forms), key operations, recovery, MSM (host,
//                         device, affine, no-allocation), batch check,
//                         multi-batch (direct and graph), streams and events,
//                         the serving loop from four submitting threads, and
//                         allocation failures injected through
//                         UFSECP_GPU_FAIL_ALLOC at every allocation of context
//                         creation, a verify call, an MSM and server creation;
//                         at the end no allocation of the library is live.
//
// Prints "abi driver: <checks> checks, <failed> failed" and exits 1 if any
// check failed.  Inputs come from libufsecp_workload (seeded synthetic
// signatures whose expected verdict is 1 unless the item was corrupted).
#include <atomic>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <string>
#include <thread>
#include <vector>

#include "ufsecp/ufsecp_gpu.h"
#include "ufsecp/ufsecp_gpu_ext.h"

extern "C" {
int ufsecp_wl_ecdsa(uint64_t seed, uint64_t first, uint64_t n, uint8_t* msgs, uint8_t* pubs, uint8_t* sigs,
                    uint8_t* corrupted, int threads);
int ufsecp_wl_schnorr(uint64_t seed, uint64_t first, uint64_t n, uint8_t* msgs, uint8_t* pubs, uint8_t* sigs,
                      uint8_t* corrupted, int threads);
int ufsecp_wl_msm(uint64_t seed, uint64_t first, uint64_t n, uint8_t* scalars, uint8_t* points, uint8_t* points64,
                  uint8_t* ks, int threads);
}

namespace {

std::atomic<int> g_checks{0}, g_failed{0};

void check(bool ok, const char* what, int line) {
    ++g_checks;
    if (!ok) {
        ++g_failed;
        std::fprintf(stderr, "FAIL line %d: %s\n", line, what);
    }
}
#define CHECK(cond) check((cond), #cond, __LINE__)

// A heap block of exactly n bytes (at least one, so that a zero-length
// buffer is still a valid, non-NULL pointer whose every byte is out of range).
struct Buf {
    std::unique_ptr<uint8_t[]> p;
    size_t n;
    explicit Buf(size_t bytes, uint8_t fill = 0) : p(new uint8_t[bytes ? bytes : 1]), n(bytes) {
        std::memset(p.get(), fill, bytes ? bytes : 1);
    }
    uint8_t* get() const { return p.get(); }
    bool all(uint8_t v) const {
        for (size_t i = 0; i < n; ++i)
            if (p[i] != v) return false;
        return true;
    }
};

struct Items {
    size_t n;
    Buf msgs, pubs, sigs, corrupted;
    Items(bool ecdsa, size_t count, uint64_t first)
        : n(count), msgs(count * 32), pubs(count * (ecdsa ? 33 : 32)), sigs(count * 64), corrupted(count) {
        const int rc = ecdsa ? ufsecp_wl_ecdsa(0x5A17, first, count, msgs.get(), pubs.get(), sigs.get(),
                                               corrupted.get(), 4)
                             : ufsecp_wl_schnorr(0x5A18, first, count, msgs.get(), pubs.get(), sigs.get(),
                                                 corrupted.get(), 4);
        CHECK(rc == 0);
    }
    bool verdicts_ok(const uint8_t* out) const {
        for (size_t i = 0; i < n; ++i)
            if (out[i] != (corrupted.p[i] ? 0 : 1)) return false;
        return true;
    }
};

constexpr size_t kTooMany = (size_t(1) << 26) + 1;

// ---- no device needed ------------------------------------------------------

void host_contract() {
    uint32_t ids[8] = {};
    const uint32_t nb = ufsecp_gpu_backend_count(ids, 8);
    CHECK(nb == 1 && ids[0] == UFSECP_GPU_BACKEND_ROCM);
    CHECK(ufsecp_gpu_backend_count(nullptr, 0) == nb);
    for (uint32_t b = 0; b < 8; ++b) CHECK(ufsecp_gpu_backend_name(b) != nullptr);
    CHECK(ufsecp_gpu_backend_name(0xFFFFFFFFu) != nullptr);
    CHECK(ufsecp_gpu_is_available(UFSECP_GPU_BACKEND_CUDA) == 0);
    CHECK(ufsecp_gpu_device_count(UFSECP_GPU_BACKEND_METAL) == 0);
    ufsecp_gpu_device_info_t info;
    CHECK(ufsecp_gpu_device_info(UFSECP_GPU_BACKEND_ROCM, 0, nullptr) == UFSECP_ERR_NULL_ARG);
    CHECK(ufsecp_gpu_device_info(UFSECP_GPU_BACKEND_ROCM, 4096, &info) != UFSECP_OK);
    CHECK(ufsecp_gpu_device_info(UFSECP_GPU_BACKEND_OPENCL, 0, &info) != UFSECP_OK);
    for (int e = -3; e < 120; ++e) CHECK(ufsecp_gpu_error_str((ufsecp_error_t)e) != nullptr);

    CHECK(ufsecp_gpu_ctx_create(nullptr, UFSECP_GPU_BACKEND_ROCM, 0) == UFSECP_ERR_NULL_ARG);
    ufsecp_gpu_ctx* c = reinterpret_cast<ufsecp_gpu_ctx*>(0x1);
    CHECK(ufsecp_gpu_ctx_create(&c, UFSECP_GPU_BACKEND_CUDA, 0) != UFSECP_OK && c == nullptr);
    c = reinterpret_cast<ufsecp_gpu_ctx*>(0x1);
    CHECK(ufsecp_gpu_ctx_create(&c, UFSECP_GPU_BACKEND_ROCM, 4096) != UFSECP_OK && c == nullptr);
    ufsecp_gpu_ctx_destroy(nullptr);
    CHECK(ufsecp_gpu_is_ready(nullptr) == 0);
    CHECK(ufsecp_gpu_last_error(nullptr) == UFSECP_ERR_NULL_ARG);
    CHECK(ufsecp_gpu_last_error_msg(nullptr) != nullptr);

    // every verify / key / MSM entry point with a NULL context
    Buf in(4096, 2), out(4096, 0xEE);
    int recid = 0;
    CHECK(ufsecp_gpu_ecdsa_verify_batch(nullptr, in.get(), in.get(), in.get(), 1, out.get()) == UFSECP_ERR_NULL_ARG);
    CHECK(ufsecp_gpu_schnorr_verify_batch(nullptr, in.get(), in.get(), in.get(), 1, out.get()) == UFSECP_ERR_NULL_ARG);
    CHECK(ufsecp_gpu_ecdsa_verify_opaque_rows(nullptr, in.get(), 129, 1, out.get()) == UFSECP_ERR_NULL_ARG);
    CHECK(ufsecp_gpu_ecdsa_verify_lbtc_rows(nullptr, in.get(), 129, 1, out.get()) == UFSECP_ERR_NULL_ARG);
    CHECK(ufsecp_gpu_ecdsa_verify_lbtc_columns(nullptr, in.get(), in.get(), in.get(), 1, out.get()) ==
          UFSECP_ERR_NULL_ARG);
    CHECK(ufsecp_gpu_schnorr_verify_lbtc_columns(nullptr, in.get(), in.get(), in.get(), 1, out.get()) ==
          UFSECP_ERR_NULL_ARG);
    CHECK(ufsecp_gpu_ecdsa_verify_collect(nullptr, in.get(), in.get(), in.get(), 1, out.get()) == UFSECP_ERR_NULL_ARG);
    CHECK(ufsecp_gpu_schnorr_verify_collect(nullptr, in.get(), in.get(), in.get(), 1, out.get()) ==
          UFSECP_ERR_NULL_ARG);
    CHECK(ufsecp_gpu_pubkey_validate(nullptr, in.get(), 1, out.get()) == UFSECP_ERR_NULL_ARG);
    CHECK(ufsecp_gpu_xonly_validate(nullptr, in.get(), 1, out.get()) == UFSECP_ERR_NULL_ARG);
    CHECK(ufsecp_gpu_commitment_verify(nullptr, in.get(), in.get(), in.get(), in.get(), 1, out.get()) ==
          UFSECP_ERR_NULL_ARG);
    CHECK(ufsecp_gpu_ecrecover_batch(nullptr, in.get(), in.get(), &recid, 1, out.get(), out.get()) ==
          UFSECP_ERR_NULL_ARG);
    CHECK(ufsecp_gpu_msm(nullptr, in.get(), in.get(), 1, out.get()) == UFSECP_ERR_NULL_ARG);
    CHECK(ufsecp_gpu_ext_msm(nullptr, in.get(), in.get(), 1, out.get(), 8) == UFSECP_ERR_NULL_ARG);
    int all_valid = 7;
    CHECK(ufsecp_gpu_ext_schnorr_batch_check(nullptr, in.get(), in.get(), in.get(), 1, &all_valid) ==
          UFSECP_ERR_NULL_ARG);
    CHECK(ufsecp_gpu_ext_verify_multi_device(nullptr, nullptr, 0, 0, nullptr) == UFSECP_ERR_NULL_ARG);
    CHECK(ufsecp_gpu_ext_verify_lanes(nullptr, 0) == 0);
    void* p = reinterpret_cast<void*>(0x1);
    CHECK(ufsecp_gpu_ext_alloc(nullptr, 64, &p) == UFSECP_ERR_NULL_ARG);
    CHECK(ufsecp_gpu_ext_host_alloc(nullptr, 64, &p) == UFSECP_ERR_NULL_ARG);
    CHECK(ufsecp_gpu_ext_stream_create(nullptr, &p) == UFSECP_ERR_NULL_ARG);
    CHECK(ufsecp_gpu_ext_event_create(nullptr, &p) == UFSECP_ERR_NULL_ARG);
    float ms = 0;
    CHECK(ufsecp_gpu_ext_event_elapsed(nullptr, nullptr, &ms) != UFSECP_OK);
    ufsecp_gpu_ext_event_destroy(nullptr);
    ufsecp_gpu_ext_stream_destroy(nullptr);
    ufsecp_gpu_server* s = reinterpret_cast<ufsecp_gpu_server*>(0x1);
    CHECK(ufsecp_gpu_ext_server_create(nullptr, nullptr, &s) == UFSECP_ERR_NULL_ARG);
    uint64_t rid = 0;
    CHECK(ufsecp_gpu_ext_server_submit(nullptr, 0, in.get(), in.get(), in.get(), 1, out.get(), &rid) ==
          UFSECP_ERR_NULL_ARG);
    CHECK(ufsecp_gpu_ext_server_poll(nullptr, 1) == -1);
    CHECK(ufsecp_gpu_ext_server_drain(nullptr) == UFSECP_ERR_NULL_ARG);
    ufsecp_gpu_server_stats_t st;
    CHECK(ufsecp_gpu_ext_server_stats(nullptr, &st, 0) == UFSECP_ERR_NULL_ARG);
    ufsecp_gpu_ext_server_destroy(nullptr);

    // the unsupported entry points: NULL context first, then the plan helper
    CHECK(ufsecp_gpu_generator_mul_batch(nullptr, in.get(), 1, out.get()) == UFSECP_ERR_NULL_ARG);
    CHECK(ufsecp_gpu_ecdh_batch(nullptr, in.get(), in.get(), 1, out.get()) == UFSECP_ERR_NULL_ARG);
    CHECK(ufsecp_gpu_hash160_pubkey_batch(nullptr, in.get(), 1, out.get()) == UFSECP_ERR_NULL_ARG);
    CHECK(ufsecp_gpu_tagged_hash(nullptr, in.get(), in.get(), 32, 1, out.get()) == UFSECP_ERR_NULL_ARG);
    CHECK(ufsecp_gpu_hash256(nullptr, in.get(), 64, 1, out.get()) == UFSECP_ERR_NULL_ARG);
    Buf plan(UFSECP_BIP352_SCAN_PLAN_BYTES, 0xEE), key(32, 1);
    CHECK(ufsecp_bip352_prepare_scan_plan(nullptr, plan.get()) == UFSECP_ERR_NULL_ARG);
    CHECK(ufsecp_bip352_prepare_scan_plan(key.get(), nullptr) == UFSECP_ERR_NULL_ARG);
    CHECK(ufsecp_bip352_prepare_scan_plan(key.get(), plan.get()) == UFSECP_ERR_GPU_UNSUPPORTED && plan.all(0));

    // the engine hook declines NULL buffers and unknown kinds before it looks
    // for a device, and counts each call
    ufsecp_gpu_engine_hook_stats_t h0, h1;
    ufsecp_gpu_ext_engine_hook_stats(&h0);
    Buf hout(1, 0xEE);
    CHECK(ufsecp_gpu_ext_engine_columns_hook(0, nullptr, in.get(), in.get(), 1, hout.get()) == -1);
    CHECK(ufsecp_gpu_ext_engine_columns_hook(5, in.get(), in.get(), in.get(), 1, hout.get()) == -1);
    ufsecp_gpu_ext_engine_hook_stats(&h1);
    CHECK(h1.calls == h0.calls + 2 && h1.declined == h0.declined + 2 && hout.all(0xEE));
}

// ---- with a device -----------------------------------------------------------

void unsupported_on_live_context(ufsecp_gpu_ctx* ctx) {
    Buf in(4096, 2);
    struct Op {
        const char* name;
        size_t width;
    };
    const size_t count = 3;
    auto run = [&](const char* name, size_t width, auto&& call) {
        Buf out(count * width, 0xEE);    // exactly the documented output size
        const ufsecp_error_t rc = call(out.get());
        check(rc == UFSECP_ERR_GPU_UNSUPPORTED && out.all(0), name, __LINE__);
    };
    run("generator_mul", 33, [&](uint8_t* o) { return ufsecp_gpu_generator_mul_batch(ctx, in.get(), count, o); });
    run("ecdh", 32, [&](uint8_t* o) { return ufsecp_gpu_ecdh_batch(ctx, in.get(), in.get(), count, o); });
    run("hash160", 20, [&](uint8_t* o) { return ufsecp_gpu_hash160_pubkey_batch(ctx, in.get(), count, o); });
    run("tagged_hash", 32, [&](uint8_t* o) { return ufsecp_gpu_tagged_hash(ctx, in.get(), in.get(), 32, count, o); });
    run("hash256", 32, [&](uint8_t* o) { return ufsecp_gpu_hash256(ctx, in.get(), 64, count, o); });
    run("frost", 1, [&](uint8_t* o) {
        return ufsecp_gpu_frost_verify_partial_batch(ctx, in.get(), in.get(), in.get(), in.get(), in.get(), in.get(),
                                                     in.get(), in.get(), count, o);
    });
    // 65-byte points must carry the uncompressed prefix 04 to pass the checks;
    // inputs of exactly count items
    Buf u65(count * 65, 4), p64(count * 64, 1), p324(count * 324, 4), m32(count * 32, 1);
    run("zk_knowledge", 1, [&](uint8_t* o) {
        return ufsecp_gpu_zk_knowledge_verify_batch(ctx, p64.get(), u65.get(), m32.get(), count, o);
    });
    run("zk_dleq", 1, [&](uint8_t* o) {
        return ufsecp_gpu_zk_dleq_verify_batch(ctx, p64.get(), u65.get(), u65.get(), u65.get(), u65.get(), count, o);
    });
    run("bulletproof", 1, [&](uint8_t* o) {
        return ufsecp_gpu_bulletproof_verify_batch(ctx, p324.get(), u65.get(), u65.get(), count, o);
    });
    {
        Buf o(count, 0xEE);
        Buf k02(count * 65, 2);
        check(ufsecp_gpu_zk_knowledge_verify_batch(ctx, p64.get(), k02.get(), m32.get(), count, o.get()) ==
                      UFSECP_ERR_BAD_PUBKEY && o.all(0),
              "zk_knowledge with a compressed prefix", __LINE__);
    }
    run("zk_ecdsa_snark", UFSECP_ECDSA_SNARK_WITNESS_BYTES, [&](uint8_t* o) {
        return ufsecp_gpu_zk_ecdsa_snark_witness_batch(ctx, in.get(), in.get(), in.get(), count, o);
    });
    run("zk_schnorr_snark", UFSECP_SCHNORR_SNARK_WITNESS_BYTES, [&](uint8_t* o) {
        return ufsecp_gpu_zk_schnorr_snark_witness_batch(ctx, in.get(), in.get(), in.get(), count, o);
    });
    CHECK(ufsecp_gpu_last_error(ctx) == UFSECP_ERR_GPU_UNSUPPORTED);
}

void verify_forms(ufsecp_gpu_ctx* ctx) {
    for (int kind = 0; kind < 2; ++kind) {
        const bool ecdsa = kind == 0;
        Items it(ecdsa, 3001, 100 + kind);
        Buf out(it.n, 0xEE);
        auto verify = ecdsa ? ufsecp_gpu_ecdsa_verify_batch : ufsecp_gpu_schnorr_verify_batch;
        CHECK(verify(ctx, it.msgs.get(), it.pubs.get(), it.sigs.get(), it.n, out.get()) == UFSECP_OK);
        CHECK(it.verdicts_ok(out.get()));
        CHECK(verify(ctx, it.msgs.get(), it.pubs.get(), it.sigs.get(), 0, nullptr) == UFSECP_OK);
        CHECK(verify(ctx, it.msgs.get(), it.pubs.get(), it.sigs.get(), kTooMany, out.get()) == UFSECP_ERR_BAD_INPUT);
        CHECK(verify(ctx, nullptr, it.pubs.get(), it.sigs.get(), it.n, out.get()) == UFSECP_ERR_NULL_ARG);
        // collect: valid items zero their byte, invalid ones keep it
        Buf keys(it.n, 0x5C);
        auto collect = ecdsa ? ufsecp_gpu_ecdsa_verify_collect : ufsecp_gpu_schnorr_verify_collect;
        CHECK(collect(ctx, it.msgs.get(), it.pubs.get(), it.sigs.get(), it.n, keys.get()) == UFSECP_OK);
        bool ok = true;
        for (size_t i = 0; i < it.n; ++i) ok &= keys.p[i] == (it.corrupted.p[i] ? 0x5C : 0);
        CHECK(ok);
        // columns
        Buf cout_(it.n, 0xEE);
        if (ecdsa) {
            Buf opaque(it.n * 64);
            for (size_t i = 0; i < it.n; ++i)
                for (int b = 0; b < 32; ++b) {
                    opaque.p[64 * i + b] = it.sigs.p[64 * i + 31 - b];
                    opaque.p[64 * i + 32 + b] = it.sigs.p[64 * i + 63 - b];
                }
            CHECK(ufsecp_gpu_ecdsa_verify_lbtc_columns(ctx, it.msgs.get(), it.pubs.get(), opaque.get(), it.n,
                                                       cout_.get()) == UFSECP_OK);
            CHECK(it.verdicts_ok(cout_.get()));
            // rows at an odd stride; the last row has no tail
            const size_t stride = 131;
            Buf rows((it.n - 1) * stride + 129, 0xC3);
            for (size_t i = 0; i < it.n; ++i) {
                std::memcpy(rows.get() + i * stride, it.msgs.get() + 32 * i, 32);
                std::memcpy(rows.get() + i * stride + 32, it.pubs.get() + 33 * i, 33);
                std::memcpy(rows.get() + i * stride + 65, opaque.get() + 64 * i, 64);
            }
            Buf rout(it.n, 0xEE);
            CHECK(ufsecp_gpu_ecdsa_verify_opaque_rows(ctx, rows.get(), stride, it.n, rout.get()) == UFSECP_OK);
            CHECK(it.verdicts_ok(rout.get()));
            CHECK(ufsecp_gpu_ecdsa_verify_lbtc_rows(ctx, rows.get(), 128, it.n, rout.get()) == UFSECP_ERR_BAD_INPUT);
            // key operations over the same keys
            Buf kv(it.n, 0xEE);
            CHECK(ufsecp_gpu_pubkey_validate(ctx, it.pubs.get(), it.n, kv.get()) == UFSECP_OK && kv.all(1));
            // recovery with ids 0..3 (each key is recovered by one of ids 0/1)
            std::unique_ptr<int[]> ids(new int[it.n]);
            for (size_t i = 0; i < it.n; ++i) ids[i] = (int)(i & 1);
            Buf rk(it.n * 33, 0xEE), rv(it.n, 0xEE);
            CHECK(ufsecp_gpu_ecrecover_batch(ctx, it.msgs.get(), it.sigs.get(), ids.get(), it.n, rk.get(), rv.get()) ==
                  UFSECP_OK);
            ids[it.n - 1] = 4;
            CHECK(ufsecp_gpu_ecrecover_batch(ctx, it.msgs.get(), it.sigs.get(), ids.get(), it.n, rk.get(), rv.get()) ==
                  UFSECP_ERR_BAD_INPUT);
            CHECK(rk.all(0) && rv.all(0));
        } else {
            CHECK(ufsecp_gpu_schnorr_verify_lbtc_columns(ctx, it.msgs.get(), it.pubs.get(), it.sigs.get(), it.n,
                                                         cout_.get()) == UFSECP_OK);
            CHECK(it.verdicts_ok(cout_.get()));
            Buf kv(it.n, 0xEE);
            CHECK(ufsecp_gpu_xonly_validate(ctx, it.pubs.get(), it.n, kv.get()) == UFSECP_OK && kv.all(1));
            Buf tw(it.n * 32, 0), par(it.n, 0), cv(it.n, 0xEE);
            CHECK(ufsecp_gpu_commitment_verify(ctx, it.pubs.get(), tw.get(), it.pubs.get(), par.get(), it.n,
                                               cv.get()) == UFSECP_OK);
            int all_valid = 7;
            CHECK(ufsecp_gpu_ext_schnorr_batch_check(ctx, it.msgs.get(), it.pubs.get(), it.sigs.get(), it.n,
                                                     &all_valid) == UFSECP_OK);
            bool any_corrupted = false;
            for (size_t i = 0; i < it.n; ++i) any_corrupted |= it.corrupted.p[i] != 0;
            CHECK(all_valid == (any_corrupted ? 0 : 1));
        }
    }
}

void msm_forms(ufsecp_gpu_ctx* ctx) {
    const size_t n = 5000;
    Buf sc(n * 32), pts(n * 33), p64(n * 64), ks(n * 32);
    CHECK(ufsecp_wl_msm(0x3E5, 0, n, sc.get(), pts.get(), p64.get(), ks.get(), 4) == 0);
    Buf a(33, 0xEE), b(33, 0xEE);
    CHECK(ufsecp_gpu_msm(ctx, sc.get(), pts.get(), n, a.get()) == UFSECP_OK);
    CHECK(ufsecp_gpu_ext_msm(ctx, sc.get(), pts.get(), n, b.get(), 7) == UFSECP_OK);
    CHECK(std::memcmp(a.get(), b.get(), 33) == 0);
    CHECK(ufsecp_gpu_msm(ctx, nullptr, nullptr, 0, nullptr) == UFSECP_OK);
    CHECK(ufsecp_gpu_ext_msm(ctx, sc.get(), pts.get(), n, b.get(), 21) == UFSECP_ERR_BAD_INPUT);
    Buf bad(n * 33);
    std::memcpy(bad.get(), pts.get(), n * 33);
    bad.p[33 * 17] = 0x05;
    CHECK(ufsecp_gpu_msm(ctx, sc.get(), bad.get(), n, b.get()) == UFSECP_ERR_BAD_PUBKEY && b.all(0));

    // device forms on a stream of the caller's
    void *d_s = nullptr, *d_p = nullptr, *d_p64 = nullptr, *d_o = nullptr, *d_st = nullptr, *stream = nullptr;
    CHECK(ufsecp_gpu_ext_alloc(ctx, n * 32, &d_s) == UFSECP_OK);
    CHECK(ufsecp_gpu_ext_alloc(ctx, n * 33, &d_p) == UFSECP_OK);
    CHECK(ufsecp_gpu_ext_alloc(ctx, n * 64, &d_p64) == UFSECP_OK);
    CHECK(ufsecp_gpu_ext_alloc(ctx, 64, &d_o) == UFSECP_OK);
    CHECK(ufsecp_gpu_ext_alloc(ctx, 16, &d_st) == UFSECP_OK);
    CHECK(ufsecp_gpu_ext_stream_create(ctx, &stream) == UFSECP_OK);
    CHECK(ufsecp_gpu_ext_copy_to_device(ctx, d_s, sc.get(), n * 32) == UFSECP_OK);
    CHECK(ufsecp_gpu_ext_copy_to_device(ctx, d_p, pts.get(), n * 33) == UFSECP_OK);
    CHECK(ufsecp_gpu_ext_copy_to_device(ctx, d_p64, p64.get(), n * 64) == UFSECP_OK);
    Buf o(33, 0xEE);
    int32_t status = -1;
    auto fetch = [&]() {
        CHECK(ufsecp_gpu_ext_stream_synchronize(stream) == UFSECP_OK);
        CHECK(ufsecp_gpu_ext_copy_to_host(ctx, o.get(), d_o, 33) == UFSECP_OK);
        CHECK(ufsecp_gpu_ext_copy_to_host(ctx, &status, d_st, 4) == UFSECP_OK);
        return status == 0 && std::memcmp(o.get(), a.get(), 33) == 0;
    };
    CHECK(ufsecp_gpu_ext_msm_device(ctx, (const uint8_t*)d_s, (const uint8_t*)d_p, n, (uint8_t*)d_o,
                                    (int32_t*)d_st, 0, stream) == UFSECP_OK);
    CHECK(fetch());
    CHECK(ufsecp_gpu_ext_msm_affine_device(ctx, (const uint8_t*)d_s, (const uint8_t*)d_p64, n, (uint8_t*)d_o,
                                           (int32_t*)d_st, 0, stream) == UFSECP_OK);
    CHECK(fetch());
    const uint64_t allocs = ufsecp_gpu_ext_buffer_alloc_count();
    CHECK(ufsecp_gpu_ext_msm_device_noalloc(ctx, (const uint8_t*)d_s, (const uint8_t*)d_p, n, (uint8_t*)d_o,
                                            (int32_t*)d_st, 0, stream) == UFSECP_OK);
    CHECK(fetch());
    CHECK(ufsecp_gpu_ext_buffer_alloc_count() == allocs);
    // events around a verify on the caller's stream
    void *e0 = nullptr, *e1 = nullptr;
    CHECK(ufsecp_gpu_ext_event_create(ctx, &e0) == UFSECP_OK && ufsecp_gpu_ext_event_create(ctx, &e1) == UFSECP_OK);
    CHECK(ufsecp_gpu_ext_event_record(ctx, e0, stream) == UFSECP_OK);
    CHECK(ufsecp_gpu_ext_msm_device(ctx, (const uint8_t*)d_s, (const uint8_t*)d_p, n, (uint8_t*)d_o,
                                    (int32_t*)d_st, 11, stream) == UFSECP_OK);
    CHECK(ufsecp_gpu_ext_event_record(ctx, e1, stream) == UFSECP_OK);
    CHECK(fetch());
    float ms = -1;
    CHECK(ufsecp_gpu_ext_event_elapsed(e0, e1, &ms) == UFSECP_OK && ms >= 0);
    ufsecp_gpu_ext_event_destroy(e0);
    ufsecp_gpu_ext_event_destroy(e1);
    ufsecp_gpu_ext_stream_destroy(stream);
    for (void* d : {d_s, d_p, d_p64, d_o, d_st}) CHECK(ufsecp_gpu_ext_free(ctx, d) == UFSECP_OK);
}

void multi_batch(ufsecp_gpu_ctx* ctx) {
    Items e(true, 700, 11), s(false, 1300, 12);
    std::vector<void*> dev;
    auto up = [&](const Buf& b) {
        void* d = nullptr;
        CHECK(ufsecp_gpu_ext_alloc(ctx, b.n, &d) == UFSECP_OK);
        CHECK(ufsecp_gpu_ext_copy_to_device(ctx, d, b.get(), b.n) == UFSECP_OK);
        dev.push_back(d);
        return (uint8_t*)d;
    };
    ufsecp_gpu_verify_batch_t t[2] = {};
    t[0] = {UFSECP_GPU_KIND_ECDSA, (uint32_t)e.n, up(e.msgs), up(e.pubs), up(e.sigs), nullptr, 0, 0};
    t[1] = {UFSECP_GPU_KIND_SCHNORR, (uint32_t)s.n, up(s.msgs), up(s.pubs), up(s.sigs), nullptr, 0, 0};
    Buf z(2000, 0xEE);
    t[0].out_results = up(z);
    t[1].out_results = up(z);
    for (int mode = 0; mode < 3; ++mode) {
        const ufsecp_error_t rc = mode == 2 ? ufsecp_gpu_ext_verify_each_device(ctx, t, 2, nullptr)
                                            : ufsecp_gpu_ext_verify_multi_device(ctx, t, 2, mode, nullptr);
        CHECK(rc == UFSECP_OK);
        CHECK(ufsecp_gpu_ext_synchronize(ctx) == UFSECP_OK);
        Buf oe(e.n), os(s.n);
        CHECK(ufsecp_gpu_ext_copy_to_host(ctx, oe.get(), t[0].out_results, e.n) == UFSECP_OK);
        CHECK(ufsecp_gpu_ext_copy_to_host(ctx, os.get(), t[1].out_results, s.n) == UFSECP_OK);
        CHECK(e.verdicts_ok(oe.get()) && s.verdicts_ok(os.get()));
    }
    t[1].reserved = 1;
    CHECK(ufsecp_gpu_ext_verify_multi_device(ctx, t, 2, 0, nullptr) == UFSECP_ERR_BAD_INPUT);
    for (void* d : dev) CHECK(ufsecp_gpu_ext_free(ctx, d) == UFSECP_OK);
}

void serving(ufsecp_gpu_ctx* ctx) {
    ufsecp_gpu_server_config_t cfg{};
    cfg.hbm_budget_bytes = 16u << 20;
    cfg.inflight_groups = 2;
    cfg.key_cache_log2 = 12;
    ufsecp_gpu_server* srv = nullptr;
    CHECK(ufsecp_gpu_ext_server_create(ctx, &cfg, &srv) == UFSECP_OK && srv);
    if (!srv) return;
    Items e(true, 8000, 21), s(false, 4000, 22);
    Buf oe(e.n, 0xEE), os(s.n, 0xEE);
    // four threads submit slices of both workloads of uneven sizes
    std::vector<std::thread> ths;
    for (int t = 0; t < 4; ++t)
        ths.emplace_back([&, t] {
            const Items& it = (t & 1) ? s : e;
            Buf& out = (t & 1) ? os : oe;
            const uint32_t kind = (t & 1) ? UFSECP_GPU_KIND_SCHNORR : UFSECP_GPU_KIND_ECDSA;
            const size_t ps = (t & 1) ? 32 : 33;
            const size_t half = it.n / 2, lo = (t >> 1) * half, hi = (t >> 1) ? it.n : half;
            size_t size = 1;
            for (size_t i = lo; i < hi; i += size, size = size * 3 % 1000 + 1) {
                const size_t m = std::min(size, hi - i);
                uint64_t rid = 0;
                check(ufsecp_gpu_ext_server_submit(srv, kind, it.msgs.get() + 32 * i, it.pubs.get() + ps * i,
                                                   it.sigs.get() + 64 * i, m, out.get() + i, &rid) == UFSECP_OK,
                      "server submit", __LINE__);
                if (i % 7 == 0) (void)ufsecp_gpu_ext_server_poll(srv, rid);
            }
        });
    for (auto& th : ths) th.join();
    CHECK(ufsecp_gpu_ext_server_drain(srv) == UFSECP_OK);
    CHECK(e.verdicts_ok(oe.get()) && s.verdicts_ok(os.get()));
    // mapped submission from pinned memory
    void* h = nullptr;
    CHECK(ufsecp_gpu_ext_host_alloc(ctx, s.n * 128, &h) == UFSECP_OK && h);
    uint8_t* hm = (uint8_t*)h;
    std::memcpy(hm, s.msgs.get(), s.n * 32);
    std::memcpy(hm + s.n * 32, s.pubs.get(), s.n * 32);
    std::memcpy(hm + s.n * 64, s.sigs.get(), s.n * 64);
    Buf om(s.n, 0xEE);
    uint64_t rid = 0;
    CHECK(ufsecp_gpu_ext_server_submit_mapped(srv, UFSECP_GPU_KIND_SCHNORR, hm, hm + s.n * 32, hm + s.n * 64, s.n,
                                              om.get(), &rid) == UFSECP_OK);
    CHECK(ufsecp_gpu_ext_server_wait(srv, rid) == UFSECP_OK);
    CHECK(s.verdicts_ok(om.get()));
    CHECK(ufsecp_gpu_ext_server_poll(srv, rid) == -1);
    ufsecp_gpu_server_stats_t st{};
    CHECK(ufsecp_gpu_ext_server_stats(srv, &st, 1) == UFSECP_OK && st.items_completed == e.n + s.n + s.n);
    CHECK(ufsecp_gpu_ext_server_submit(srv, 9, s.msgs.get(), s.pubs.get(), s.sigs.get(), 1, om.get(), &rid) ==
          UFSECP_ERR_BAD_INPUT);
    ufsecp_gpu_ext_server_destroy(srv);
    CHECK(ufsecp_gpu_ext_host_free(ctx, h) == UFSECP_OK);
}

// Fail the k-th library allocation from now (1-based); 0 clears the knob.
void fail_alloc_at(uint64_t k) {
    if (k == 0) {
        unsetenv("UFSECP_GPU_FAIL_ALLOC");
        return;
    }
    const std::string v = std::to_string(ufsecp_gpu_ext_buffer_alloc_count() + k);
    setenv("UFSECP_GPU_FAIL_ALLOC", v.c_str(), 1);
}

void injected_failures() {
    uint64_t before = ufsecp_gpu_ext_buffer_alloc_count();
    ufsecp_gpu_ctx* ctx = nullptr;
    CHECK(ufsecp_gpu_ctx_create(&ctx, UFSECP_GPU_BACKEND_ROCM, 0) == UFSECP_OK);
    ufsecp_gpu_ctx_destroy(ctx);
    const uint64_t ctx_allocs = ufsecp_gpu_ext_buffer_alloc_count() - before;
    for (uint64_t k = 1; k <= ctx_allocs; ++k) {
        fail_alloc_at(k);
        ctx = reinterpret_cast<ufsecp_gpu_ctx*>(0x1);
        check(ufsecp_gpu_ctx_create(&ctx, UFSECP_GPU_BACKEND_ROCM, 0) == UFSECP_ERR_GPU_MEMORY && !ctx,
              "ctx_create under injected failure", __LINE__);
        fail_alloc_at(0);
    }
    Items it(false, 2000, 31);
    const size_t n = 3000;
    Buf sc(n * 32), pts(n * 33), ks(n * 32);
    CHECK(ufsecp_wl_msm(0x3E6, 0, n, sc.get(), pts.get(), nullptr, ks.get(), 4) == 0);
    for (int op = 0; op < 3; ++op) {
        // allocations the first call of this kind makes on a fresh context
        CHECK(ufsecp_gpu_ctx_create(&ctx, UFSECP_GPU_BACKEND_ROCM, 0) == UFSECP_OK);
        Buf out(op == 1 ? 33 : it.n, 0xEE);
        ufsecp_gpu_server_config_t cfg{};
        cfg.hbm_budget_bytes = 16u << 20;
        cfg.key_cache_log2 = 10;
        ufsecp_gpu_server* srv = nullptr;
        auto call = [&]() -> ufsecp_error_t {
            if (op == 0) return ufsecp_gpu_schnorr_verify_batch(ctx, it.msgs.get(), it.pubs.get(), it.sigs.get(),
                                                                it.n, out.get());
            if (op == 1) return ufsecp_gpu_msm(ctx, sc.get(), pts.get(), n, out.get());
            const ufsecp_error_t r = ufsecp_gpu_ext_server_create(ctx, &cfg, &srv);
            if (r == UFSECP_OK) ufsecp_gpu_ext_server_destroy(srv);
            return r;
        };
        before = ufsecp_gpu_ext_buffer_alloc_count();
        CHECK(call() == UFSECP_OK);
        const uint64_t made = ufsecp_gpu_ext_buffer_alloc_count() - before;
        ufsecp_gpu_ctx_destroy(ctx);
        CHECK(made >= 1);
        for (uint64_t k = 1; k <= made; ++k) {
            CHECK(ufsecp_gpu_ctx_create(&ctx, UFSECP_GPU_BACKEND_ROCM, 0) == UFSECP_OK);
            std::memset(out.get(), 0xEE, out.n);
            fail_alloc_at(k);
            const ufsecp_error_t rc = call();
            fail_alloc_at(0);
            check(rc == UFSECP_ERR_GPU_MEMORY, "call under injected failure", __LINE__);
            if (op < 2) check(out.all(0), "outputs cleared", __LINE__);
            check(call() == UFSECP_OK, "retry", __LINE__);
            if (op == 0) check(it.verdicts_ok(out.get()), "retry verdicts", __LINE__);
            ufsecp_gpu_ctx_destroy(ctx);
        }
    }
}

}  // namespace

int main(int argc, char** argv) {
    const bool device = argc > 1 && std::strcmp(argv[1], "--device") == 0;
    host_contract();
    if (device) {
        ufsecp_gpu_ctx* ctx = nullptr;
        CHECK(ufsecp_gpu_ctx_create(&ctx, UFSECP_GPU_BACKEND_ROCM, 0) == UFSECP_OK);
        if (ctx) {
            CHECK(ufsecp_gpu_is_ready(ctx) == 1);
            unsupported_on_live_context(ctx);
            verify_forms(ctx);
            msm_forms(ctx);
            multi_batch(ctx);
            serving(ctx);
            ufsecp_gpu_ctx_destroy(ctx);
            injected_failures();
            // every context, server and buffer above has been released
            CHECK(ufsecp_gpu_ext_buffer_live_count() == 0);
        }
    }
    std::printf("abi driver: %d checks, %d failed\n", g_checks.load(), g_failed.load());
    return g_failed.load() ? 1 : 0;
}
