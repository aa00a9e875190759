/* Multi-threaded driver for per-item verification through a CPU library that
 * exports the reference CPU C ABI (ufsecp_ctx_create, ufsecp_ecdsa_verify,
 * ufsecp_schnorr_verify) -- TEST INFRASTRUCTURE ONLY (CPU baseline timing and
 * parity cross-checks).  The library is opened with dlopen so the driver has
 * no link-time dependency on it.  One ufsecp context per thread, as the
 * reference requires (include/ufsecp/ufsecp.h: contexts are single-thread). */
#include <dlfcn.h>
#include <pthread.h>
#include <stddef.h>
#include <stdint.h>
#include <stdlib.h>
#include <time.h>

typedef int (*ctx_create_fn)(void**);
typedef void (*ctx_destroy_fn)(void*);
typedef int (*verify_fn)(void*, const uint8_t*, const uint8_t*, const uint8_t*);

typedef struct {
    ctx_create_fn create;
    ctx_destroy_fn destroy;
    verify_fn verify;
    int schnorr;
    const uint8_t *msgs, *pubs, *sigs;
    uint8_t* out;
    size_t lo, hi;
    int rc;
} rjob_t;

static void* rjob_run(void* arg) {
    rjob_t* j = (rjob_t*)arg;
    void* ctx = 0;
    if (j->create(&ctx) != 0) {
        j->rc = -1;
        return 0;
    }
    const size_t ps = j->schnorr ? 32 : 33;
    for (size_t i = j->lo; i < j->hi; ++i)
        j->out[i] = (uint8_t)(j->verify(ctx, j->msgs + 32 * i, j->sigs + 64 * i, j->pubs + ps * i) == 0);
    if (j->destroy) j->destroy(ctx);
    j->rc = 0;
    return 0;
}

/* Verifies n items with `threads` threads; returns wall seconds (< 0 on error). */
__attribute__((visibility("default"))) double orc_ref_verify_batch(const char* lib_path, int schnorr,
                                                                     const uint8_t* msgs, const uint8_t* pubs,
                                                                     const uint8_t* sigs, size_t n, uint8_t* out,
                                                                     int threads) {
    void* h = dlopen(lib_path, RTLD_NOW | RTLD_LOCAL);
    if (!h) return -1.0;
    ctx_create_fn create = (ctx_create_fn)dlsym(h, "ufsecp_ctx_create");
    ctx_destroy_fn destroy = (ctx_destroy_fn)dlsym(h, "ufsecp_ctx_destroy");
    verify_fn verify = (verify_fn)dlsym(h, schnorr ? "ufsecp_schnorr_verify" : "ufsecp_ecdsa_verify");
    if (!create || !verify) return -2.0;
    if (threads < 1) threads = 1;
    if (threads > 256) threads = 256;
    pthread_t th[256];
    rjob_t jobs[256];
    const size_t chunk = (n + (size_t)threads - 1) / (size_t)threads;
    struct timespec t0, t1;
    clock_gettime(CLOCK_MONOTONIC, &t0);
    for (int t = 0; t < threads; ++t) {
        rjob_t* j = &jobs[t];
        j->create = create; j->destroy = destroy; j->verify = verify; j->schnorr = schnorr;
        j->msgs = msgs; j->pubs = pubs; j->sigs = sigs; j->out = out;
        j->lo = (size_t)t * chunk < n ? (size_t)t * chunk : n;
        j->hi = (size_t)(t + 1) * chunk < n ? (size_t)(t + 1) * chunk : n;
        j->rc = 0;
        if (pthread_create(&th[t], 0, rjob_run, j) != 0) {
            rjob_run(j);
            th[t] = 0;
        }
    }
    int rc = 0;
    for (int t = 0; t < threads; ++t) {
        if (th[t]) pthread_join(th[t], 0);
        rc |= jobs[t].rc;
    }
    clock_gettime(CLOCK_MONOTONIC, &t1);
    if (rc) return -3.0;
    return (double)(t1.tv_sec - t0.tv_sec) + 1e-9 * (double)(t1.tv_nsec - t0.tv_nsec);
}

typedef int (*rows_fn)(void*, const uint8_t*, size_t, size_t, uint8_t*);
typedef int (*cols_fn)(void*, const uint8_t*, const uint8_t*, const uint8_t*, size_t, uint8_t*);
typedef int (*conv_fn)(void*, const uint8_t*, uint8_t*);

/* Opaque-signature ECDSA through the reference CPU C ABI, single-threaded:
 * rows != NULL -> ufsecp_ecdsa_verify_opaque_rows(rows, stride, count),
 * else ufsecp_ecdsa_verify_opaque_batch(msgs, pubs, sigs, count).  Returns the
 * ABI code, or a negative value if the library or symbol is missing. */
__attribute__((visibility("default"))) int orc_ref_verify_opaque(const char* lib_path, const uint8_t* rows,
                                                                 size_t stride, const uint8_t* msgs,
                                                                 const uint8_t* pubs, const uint8_t* sigs,
                                                                 size_t count, uint8_t* out) {
    void* h = dlopen(lib_path, RTLD_NOW | RTLD_LOCAL);
    if (!h) return -1;
    ctx_create_fn create = (ctx_create_fn)dlsym(h, "ufsecp_ctx_create");
    ctx_destroy_fn destroy = (ctx_destroy_fn)dlsym(h, "ufsecp_ctx_destroy");
    rows_fn vrows = (rows_fn)dlsym(h, "ufsecp_ecdsa_verify_opaque_rows");
    cols_fn vcols = (cols_fn)dlsym(h, "ufsecp_ecdsa_verify_opaque_batch");
    if (!create || !vrows || !vcols) return -2;
    void* ctx = 0;
    if (create(&ctx) != 0) return -3;
    const int rc = rows ? vrows(ctx, rows, stride, count, out) : vcols(ctx, msgs, pubs, sigs, count, out);
    if (destroy) destroy(ctx);
    return rc;
}

/* The reference's compact (big-endian r||s) -> opaque conversion,
 * ufsecp_ecdsa_sig_compact_to_opaque.  Returns the ABI code (< 0: missing). */
__attribute__((visibility("default"))) int orc_ref_compact_to_opaque(const char* lib_path, const uint8_t* sig64,
                                                                     uint8_t* opaque64) {
    void* h = dlopen(lib_path, RTLD_NOW | RTLD_LOCAL);
    if (!h) return -1;
    ctx_create_fn create = (ctx_create_fn)dlsym(h, "ufsecp_ctx_create");
    ctx_destroy_fn destroy = (ctx_destroy_fn)dlsym(h, "ufsecp_ctx_destroy");
    conv_fn conv = (conv_fn)dlsym(h, "ufsecp_ecdsa_sig_compact_to_opaque");
    if (!create || !conv) return -2;
    void* ctx = 0;
    if (create(&ctx) != 0) return -3;
    const int rc = conv(ctx, sig64, opaque64);
    if (destroy) destroy(ctx);
    return rc;
}

typedef int (*msm_fn)(void*, const uint8_t*, const uint8_t*, size_t, uint8_t*);

typedef struct {
    ctx_create_fn create;
    ctx_destroy_fn destroy;
    msm_fn msm;
    const uint8_t *scalars, *points;
    size_t lo, hi;
    uint8_t part[33];
    int rc;
} mjob_t;

static void* mjob_run(void* arg) {
    mjob_t* j = (mjob_t*)arg;
    void* ctx = 0;
    if (j->create(&ctx) != 0) {
        j->rc = -1;
        return 0;
    }
    j->rc = j->msm(ctx, j->scalars + 32 * j->lo, j->points + 33 * j->lo, j->hi - j->lo, j->part);
    if (j->destroy) j->destroy(ctx);
    return 0;
}

/* Multi-scalar multiplication through the reference CPU C ABI
 * (ufsecp_multi_scalar_mul) on `threads` threads: the n terms are cut into
 * contiguous chunks, one call per chunk and thread, and the partial sums
 * (33-byte compressed points; a chunk whose sum is the point at infinity,
 * ERR_ARITH, is left out) are added by one more call with unit scalars.
 * *rc_out gets the first failing chunk's code, else the final call's.
 * Returns wall seconds (< 0 if the library or a symbol is missing). */
__attribute__((visibility("default"))) double orc_ref_msm(const char* lib_path, const uint8_t* scalars,
                                                            const uint8_t* points, size_t n, uint8_t* out33,
                                                            int threads, int* rc_out) {
    void* h = dlopen(lib_path, RTLD_NOW | RTLD_LOCAL);
    if (!h) return -1.0;
    ctx_create_fn create = (ctx_create_fn)dlsym(h, "ufsecp_ctx_create");
    ctx_destroy_fn destroy = (ctx_destroy_fn)dlsym(h, "ufsecp_ctx_destroy");
    msm_fn msm = (msm_fn)dlsym(h, "ufsecp_multi_scalar_mul");
    if (!create || !msm) return -2.0;
    if (threads < 1) threads = 1;
    if (threads > 256) threads = 256;
    if ((size_t)threads > n) threads = n ? (int)n : 1;
    pthread_t th[256];
    mjob_t jobs[256];
    const size_t chunk = (n + (size_t)threads - 1) / (size_t)threads;
    struct timespec t0, t1;
    clock_gettime(CLOCK_MONOTONIC, &t0);
    for (int t = 0; t < threads; ++t) {
        mjob_t* j = &jobs[t];
        j->create = create; j->destroy = destroy; j->msm = msm;
        j->scalars = scalars; j->points = points;
        j->lo = (size_t)t * chunk < n ? (size_t)t * chunk : n;
        j->hi = (size_t)(t + 1) * chunk < n ? (size_t)(t + 1) * chunk : n;
        j->rc = 0;
        if (pthread_create(&th[t], 0, mjob_run, j) != 0) {
            mjob_run(j);
            th[t] = 0;
        }
    }
    for (int t = 0; t < threads; ++t)
        if (th[t]) pthread_join(th[t], 0);
    uint8_t one[32 * 256] = {0}, parts[33 * 256];
    size_t k = 0;
    int rc = 0;
    for (int t = 0; t < threads && rc == 0; ++t) {
        if (jobs[t].rc == 7) continue;        /* this chunk sums to infinity */
        if (jobs[t].rc != 0) { rc = jobs[t].rc; break; }
        one[32 * k + 31] = 1;
        for (int b = 0; b < 33; ++b) parts[33 * k + b] = jobs[t].part[b];
        ++k;
    }
    if (rc == 0) {
        void* ctx = 0;
        if (create(&ctx) != 0) return -3.0;
        rc = k ? msm(ctx, one, parts, k, out33) : 7;
        if (destroy) destroy(ctx);
    }
    clock_gettime(CLOCK_MONOTONIC, &t1);
    *rc_out = rc;
    return (double)(t1.tv_sec - t0.tv_sec) + 1e-9 * (double)(t1.tv_nsec - t0.tv_nsec);
}
