/* CPU oracle for the batch-verify / MSM path -- TEST INFRASTRUCTURE ONLY.
 *
 * A plain-C restatement of the reference CPU semantics that the GPU kernels
 * must match bit-exactly.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py use it; the product library never links it.
 *
 * Functions and the reference code they restate:
 *   orc_ecdsa_verify   src/cpu/src/impl/ufsecp_ecdsa.cpp:244-275 (strict r,s,
 *                      low-S), ecdsa.cpp:891-927 (strict compressed key),
 *                      ecdsa.cpp:929-952 (u1 = z/s, u2 = r/s), ecdsa.cpp:764-850
 *                      (x(R) == r including the r + n case)
 *   orc_ecdsa_verify_opaque  src/cpu/src/impl/ufsecp_ecdsa.cpp:105-117,
 *                      128-140, 180-189 (opaque little-endian r,s strict in
 *                      [1,n), s normalised: high-S accepted), :412-452
 *                      (per-row verdict); then the orc_ecdsa_verify core
 *   orc_ecdsa_recover  src/cpu/src/impl/ufsecp_ecdsa.cpp:710-739 (strict
 *                      r, s; recid 0..3), recovery.cpp:158-235 (r + n only
 *                      when r < p - n, lift with parity, Q finite)
 *   orc_pubkey_validate, orc_xonly_validate
 *                      ufsecp_pubkey_parse of 02/03||x (reference
 *                      compat/libbitcoin_bridge/src/ufsecp_libbitcoin.cpp:
 *                      1423-1427, 1531-1534; ecdsa.cpp:891-927)
 *   orc_commitment_verify  the same file :1138-1149 (1*P + t*G by
 *                      ufsecp_multi_scalar_mul, compressed compare)
 *   orc_schnorr_verify src/cpu/src/impl/ufsecp_ecdsa.cpp:992-1019,
 *                      schnorr.cpp:355-371 (r < p, 0 < s < n), :275-300
 *                      (lift_x, x < p), :745-806 (R = sG - eP, even y, x == r)
 *   orc_msm            src/cpu/src/impl/ufsecp_taproot.cpp:957-991 (strict
 *                      scalar < n then point per term, infinity -> ARITH);
 *                      multiscalar.cpp:77-312 computes the same exact sum
 *   orc_ecdsa_sign     src/cpu/src/ct_sign.cpp:40-70 + ecdsa.cpp:335-430
 *                      (RFC 6979, h1 = raw digest, low-S)
 *   orc_schnorr_sign   BIP-340 signing with aux_rand (reference
 *                      impl/ufsecp_ecdsa.cpp:745-783)
 *
 * Pinned by tests/test_oracle.py against the reference's golden vectors
 * (BIP-340 vectors 0-14, RFC 6979 vectors, edge cases and workload digests
 * recorded from the reference CPU library) and against closed forms.
 *
 * Arithmetic: 4 x 64-bit limbs, unsigned __int128 products, fully reduced
 * values; scalar multiplication by 4-bit windows (Shamir's trick for the
 * double multiplication). */
#include <pthread.h>
#include <stdint.h>
#include <string.h>

typedef unsigned __int128 u128;
typedef uint64_t u64;
typedef uint32_t u32;
typedef uint8_t u8;

#define ORC_API __attribute__((visibility("default")))

typedef struct { u64 v[4]; } u256;

static const u256 P_ = {{0xFFFFFFFEFFFFFC2FULL, 0xFFFFFFFFFFFFFFFFULL, 0xFFFFFFFFFFFFFFFFULL, 0xFFFFFFFFFFFFFFFFULL}};
static const u256 N_ = {{0xBFD25E8CD0364141ULL, 0xBAAEDCE6AF48A03BULL, 0xFFFFFFFFFFFFFFFEULL, 0xFFFFFFFFFFFFFFFFULL}};
static const u256 NH_ = {{0xDFE92F46681B20A0ULL, 0x5D576E7357A4501DULL, 0xFFFFFFFFFFFFFFFFULL, 0x7FFFFFFFFFFFFFFFULL}};
static const u256 GX_ = {{0x59F2815B16F81798ULL, 0x029BFCDB2DCE28D9ULL, 0x55A06295CE870B07ULL, 0x79BE667EF9DCBBACULL}};
static const u256 GY_ = {{0x9C47D08FFB10D4B8ULL, 0xFD17B448A6855419ULL, 0x5DA4FBFC0E1108A8ULL, 0x483ADA7726A3C465ULL}};

/* ---- 256-bit integers --------------------------------------------------- */
static u256 be_load(const u8* b) {
    u256 r;
    for (int i = 0; i < 4; ++i) {
        u64 x = 0;
        for (int j = 0; j < 8; ++j) x = (x << 8) | b[(3 - i) * 8 + j];
        r.v[i] = x;
    }
    return r;
}
static void be_store(u8* b, u256 a) {
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 8; ++j) b[(3 - i) * 8 + j] = (u8)(a.v[i] >> (56 - 8 * j));
}
static int u_zero(u256 a) { return (a.v[0] | a.v[1] | a.v[2] | a.v[3]) == 0; }
static int u_cmp(u256 a, u256 b) {
    for (int i = 3; i >= 0; --i) {
        if (a.v[i] != b.v[i]) return a.v[i] < b.v[i] ? -1 : 1;
    }
    return 0;
}
static u64 u_add(u256* r, u256 a, u256 b) {
    u128 c = 0;
    for (int i = 0; i < 4; ++i) { c += (u128)a.v[i] + b.v[i]; r->v[i] = (u64)c; c >>= 64; }
    return (u64)c;
}
static u64 u_sub(u256* r, u256 a, u256 b) {
    u64 bw = 0;
    for (int i = 0; i < 4; ++i) {
        u128 d = (u128)a.v[i] - b.v[i] - bw;
        r->v[i] = (u64)d;
        bw = (u64)(d >> 127) & 1u;
    }
    return bw;
}
static void u_mul(u64 t[8], u256 a, u256 b) {
    memset(t, 0, 8 * sizeof(u64));
    for (int i = 0; i < 4; ++i) {
        u128 c = 0;
        for (int j = 0; j < 4; ++j) {
            c += (u128)a.v[i] * b.v[j] + t[i + j];
            t[i + j] = (u64)c;
            c >>= 64;
        }
        t[i + 4] = (u64)c;
    }
}

/* generic reduction of a 512-bit value modulo m = 2^256 - c (c small) */
static u256 u_reduce(const u64 t_in[8], u256 m, u256 c) {
    u64 t[8];
    memcpy(t, t_in, sizeof(t));
    for (int round = 0; round < 4; ++round) {
        u256 hi = {{t[4], t[5], t[6], t[7]}};
        if (u_zero(hi)) break;
        u64 h[8];
        u_mul(h, hi, c);
        u128 cy = 0;
        for (int i = 0; i < 8; ++i) {
            cy += (u128)h[i] + (i < 4 ? t[i] : 0);
            t[i] = (u64)cy;
            cy >>= 64;
        }
    }
    u256 r = {{t[0], t[1], t[2], t[3]}};
    while (u_cmp(r, m) >= 0) u_sub(&r, r, m);
    return r;
}

/* ---- field mod p -------------------------------------------------------- */
static const u256 PC_ = {{0x1000003D1ULL, 0, 0, 0}};
static u256 f_mul(u256 a, u256 b) { u64 t[8]; u_mul(t, a, b); return u_reduce(t, P_, PC_); }
static u256 f_sqr(u256 a) { return f_mul(a, a); }
static u256 f_add(u256 a, u256 b) {
    u256 r;
    if (u_add(&r, a, b) || u_cmp(r, P_) >= 0) u_sub(&r, r, P_);
    return r;
}
static u256 f_sub(u256 a, u256 b) {
    u256 r;
    if (u_sub(&r, a, b)) u_add(&r, r, P_);
    return r;
}
static u256 f_small(u64 x) { u256 r = {{x, 0, 0, 0}}; return r; }
static u256 f_pow(u256 a, u256 e) {
    u256 r = f_small(1);
    for (int i = 255; i >= 0; --i) {
        r = f_sqr(r);
        if ((e.v[i >> 6] >> (i & 63)) & 1) r = f_mul(r, a);
    }
    return r;
}
static u256 f_inv(u256 a) { u256 e; u_sub(&e, P_, f_small(2)); return f_pow(a, e); }
/* sqrt exponent (p+1)/4 */
static int f_sqrt(u256* r, u256 a) {
    u256 e = P_;
    u_add(&e, e, f_small(1));      /* p + 1 fits: p < 2^256 - 1 */
    for (int i = 0; i < 3; ++i) e.v[i] = (e.v[i] >> 2) | (e.v[i + 1] << 62);
    e.v[3] >>= 2;
    *r = f_pow(a, e);
    return u_cmp(f_sqr(*r), a) == 0;
}

/* ---- scalars mod n ------------------------------------------------------ */
static const u256 NC_ = {{0x402DA1732FC9BEBFULL, 0x4551231950B75FC4ULL, 1, 0}};
static u256 s_mul(u256 a, u256 b) { u64 t[8]; u_mul(t, a, b); return u_reduce(t, N_, NC_); }
static u256 s_add(u256 a, u256 b) {
    u256 r;
    if (u_add(&r, a, b) || u_cmp(r, N_) >= 0) u_sub(&r, r, N_);
    return r;
}
static u256 s_neg(u256 a) {
    if (u_zero(a)) return a;
    u256 r;
    u_sub(&r, N_, a);
    return r;
}
static u256 s_inv(u256 a) {
    u256 e, r = f_small(1);
    u_sub(&e, N_, f_small(2));
    for (int i = 255; i >= 0; --i) {
        r = s_mul(r, r);
        if ((e.v[i >> 6] >> (i & 63)) & 1) r = s_mul(r, a);
    }
    return r;
}
static u256 s_from_bytes(const u8* b) {   /* one conditional subtraction (scalar.cpp:158-200) */
    u256 r = be_load(b);
    if (u_cmp(r, N_) >= 0) u_sub(&r, r, N_);
    return r;
}

/* ---- group law (Jacobian, a = 0) ---------------------------------------- */
typedef struct { u256 x, y, z; int inf; } jac;
typedef struct { u256 x, y; } aff;

static jac j_inf(void) { jac r; memset(&r, 0, sizeof(r)); r.inf = 1; return r; }
static jac j_from(aff a) { jac r; r.x = a.x; r.y = a.y; r.z = f_small(1); r.inf = 0; return r; }

static jac j_dbl(jac a) {
    if (a.inf || u_zero(a.y)) return j_inf();
    u256 yy = f_sqr(a.y);
    u256 s = f_mul(a.x, yy); s = f_add(s, s); s = f_add(s, s);          /* 4xy^2 */
    u256 m = f_sqr(a.x); m = f_add(f_add(m, m), m);                      /* 3x^2  */
    jac r;
    r.x = f_sub(f_sqr(m), f_add(s, s));
    u256 y4 = f_sqr(yy); y4 = f_add(y4, y4); y4 = f_add(y4, y4); y4 = f_add(y4, y4);
    r.y = f_sub(f_mul(m, f_sub(s, r.x)), y4);
    r.z = f_mul(a.y, a.z); r.z = f_add(r.z, r.z);
    r.inf = 0;
    return r;
}

static jac j_add(jac a, jac b) {
    if (a.inf) return b;
    if (b.inf) return a;
    u256 z1 = f_sqr(a.z), z2 = f_sqr(b.z);
    u256 u1 = f_mul(a.x, z2), u2 = f_mul(b.x, z1);
    u256 s1 = f_mul(a.y, f_mul(z2, b.z)), s2 = f_mul(b.y, f_mul(z1, a.z));
    if (u_cmp(u1, u2) == 0) {
        if (u_cmp(s1, s2) == 0) return j_dbl(a);
        return j_inf();
    }
    u256 h = f_sub(u2, u1), r = f_sub(s2, s1);
    u256 h2 = f_sqr(h), h3 = f_mul(h2, h), v = f_mul(u1, h2);
    jac o;
    o.x = f_sub(f_sub(f_sqr(r), h3), f_add(v, v));
    o.y = f_sub(f_mul(r, f_sub(v, o.x)), f_mul(s1, h3));
    o.z = f_mul(f_mul(a.z, b.z), h);
    o.inf = 0;
    return o;
}

static int j_to_aff(aff* r, jac a) {
    if (a.inf) return 0;
    u256 zi = f_inv(a.z), zi2 = f_sqr(zi);
    r->x = f_mul(a.x, zi2);
    r->y = f_mul(a.y, f_mul(zi2, zi));
    return 1;
}

/* a*A + b*B with 4-bit windows sharing the doublings (Shamir). */
static jac j_mul2(u256 a, jac A, u256 b, jac B) {
    jac ta[16], tb[16];
    ta[0] = j_inf(); tb[0] = j_inf();
    for (int i = 1; i < 16; ++i) { ta[i] = j_add(ta[i - 1], A); tb[i] = j_add(tb[i - 1], B); }
    jac acc = j_inf();
    for (int w = 63; w >= 0; --w) {
        for (int d = 0; d < 4; ++d) acc = j_dbl(acc);
        unsigned da = (unsigned)(a.v[w >> 4] >> (4 * (w & 15))) & 15u;
        unsigned db = (unsigned)(b.v[w >> 4] >> (4 * (w & 15))) & 15u;
        if (da) acc = j_add(acc, ta[da]);
        if (db) acc = j_add(acc, tb[db]);
    }
    return acc;
}

static jac G_(void) { aff g = {GX_, GY_}; return j_from(g); }

/* strict x decode + lift with parity; 0 on failure */
static int lift(aff* out, const u8* x32, int odd) {
    u256 x = be_load(x32);
    if (u_cmp(x, P_) >= 0) return 0;
    u256 y2 = f_add(f_mul(f_sqr(x), x), f_small(7)), y;
    if (!f_sqrt(&y, y2)) return 0;
    if ((int)(y.v[0] & 1) != odd) y = f_sub(f_small(0), y);
    out->x = x;
    out->y = y;
    return 1;
}

static int pub33_parse(aff* out, const u8* pub) {
    if (pub[0] != 2 && pub[0] != 3) return 0;
    return lift(out, pub + 1, pub[0] == 3);
}

/* ---- SHA-256 ------------------------------------------------------------ */
static const u32 K_[64] = {
    0x428a2f98, 0x71374491, 0xb5c0fbcf, 0xe9b5dba5, 0x3956c25b, 0x59f111f1, 0x923f82a4, 0xab1c5ed5,
    0xd807aa98, 0x12835b01, 0x243185be, 0x550c7dc3, 0x72be5d74, 0x80deb1fe, 0x9bdc06a7, 0xc19bf174,
    0xe49b69c1, 0xefbe4786, 0x0fc19dc6, 0x240ca1cc, 0x2de92c6f, 0x4a7484aa, 0x5cb0a9dc, 0x76f988da,
    0x983e5152, 0xa831c66d, 0xb00327c8, 0xbf597fc7, 0xc6e00bf3, 0xd5a79147, 0x06ca6351, 0x14292967,
    0x27b70a85, 0x2e1b2138, 0x4d2c6dfc, 0x53380d13, 0x650a7354, 0x766a0abb, 0x81c2c92e, 0x92722c85,
    0xa2bfe8a1, 0xa81a664b, 0xc24b8b70, 0xc76c51a3, 0xd192e819, 0xd6990624, 0xf40e3585, 0x106aa070,
    0x19a4c116, 0x1e376c08, 0x2748774c, 0x34b0bcb5, 0x391c0cb3, 0x4ed8aa4a, 0x5b9cca4f, 0x682e6ff3,
    0x748f82ee, 0x78a5636f, 0x84c87814, 0x8cc70208, 0x90befffa, 0xa4506ceb, 0xbef9a3f7, 0xc67178f2};
#define ROR(x, n) (((x) >> (n)) | ((x) << (32 - (n))))
static void sha_block(u32 h[8], const u8* p) {
    u32 w[64];
    for (int i = 0; i < 16; ++i) w[i] = (u32)p[4 * i] << 24 | (u32)p[4 * i + 1] << 16 | (u32)p[4 * i + 2] << 8 | p[4 * i + 3];
    for (int i = 16; i < 64; ++i)
        w[i] = w[i - 16] + (ROR(w[i - 15], 7) ^ ROR(w[i - 15], 18) ^ (w[i - 15] >> 3)) + w[i - 7] +
               (ROR(w[i - 2], 17) ^ ROR(w[i - 2], 19) ^ (w[i - 2] >> 10));
    u32 a = h[0], b = h[1], c = h[2], d = h[3], e = h[4], f = h[5], g = h[6], k = h[7];
    for (int i = 0; i < 64; ++i) {
        u32 t1 = k + (ROR(e, 6) ^ ROR(e, 11) ^ ROR(e, 25)) + ((e & f) ^ (~e & g)) + K_[i] + w[i];
        u32 t2 = (ROR(a, 2) ^ ROR(a, 13) ^ ROR(a, 22)) + ((a & b) ^ (a & c) ^ (b & c));
        k = g; g = f; f = e; e = d + t1; d = c; c = b; b = a; a = t1 + t2;
    }
    h[0] += a; h[1] += b; h[2] += c; h[3] += d; h[4] += e; h[5] += f; h[6] += g; h[7] += k;
}
/* SHA-256 of the concatenation of up to four pieces (total < 2^16 bytes) */
static void sha256v(u8 out[32], const u8* p[], const size_t n[], int cnt) {
    u32 h[8] = {0x6a09e667, 0xbb67ae85, 0x3c6ef372, 0xa54ff53a, 0x510e527f, 0x9b05688c, 0x1f83d9ab, 0x5be0cd19};
    u8 buf[64];
    size_t fill = 0, total = 0;
    for (int k = 0; k < cnt; ++k)
        for (size_t i = 0; i < n[k]; ++i) {
            buf[fill++] = p[k][i];
            ++total;
            if (fill == 64) { sha_block(h, buf); fill = 0; }
        }
    buf[fill++] = 0x80;
    if (fill > 56) { while (fill < 64) buf[fill++] = 0; sha_block(h, buf); fill = 0; }
    while (fill < 56) buf[fill++] = 0;
    const u64 bits = (u64)total * 8;
    for (int i = 0; i < 8; ++i) buf[56 + i] = (u8)(bits >> (56 - 8 * i));
    sha_block(h, buf);
    for (int i = 0; i < 8; ++i) {
        out[4 * i] = (u8)(h[i] >> 24); out[4 * i + 1] = (u8)(h[i] >> 16);
        out[4 * i + 2] = (u8)(h[i] >> 8); out[4 * i + 3] = (u8)h[i];
    }
}
static void tagged(u8 out[32], const char* tag, const u8* a, size_t na, const u8* b, size_t nb, const u8* c, size_t nc) {
    u8 th[32];
    const u8* p1[1] = {(const u8*)tag};
    size_t n1[1] = {strlen(tag)};
    sha256v(th, p1, n1, 1);
    const u8* p[5] = {th, th, a, b, c};
    size_t n[5] = {32, 32, na, nb, nc};
    sha256v(out, p, n, 5);
}
static void hmac(u8 out[32], const u8 key[32], const u8* m, size_t nm) {
    u8 ip[64], op[64], in[32];
    for (int i = 0; i < 64; ++i) { u8 k = i < 32 ? key[i] : 0; ip[i] = k ^ 0x36; op[i] = k ^ 0x5c; }
    const u8* p1[2] = {ip, m};
    size_t n1[2] = {64, nm};
    sha256v(in, p1, n1, 2);
    const u8* p2[2] = {op, in};
    size_t n2[2] = {64, 32};
    sha256v(out, p2, n2, 2);
}

/* ---- verification ------------------------------------------------------- */
static int ecdsa_core(const u8* msg32, u256 r, u256 s, const u8* pub33, int require_low_s) {
    if (u_zero(r) || u_cmp(r, N_) >= 0 || u_zero(s) || u_cmp(s, N_) >= 0) return 0;
    if (require_low_s && u_cmp(s, NH_) > 0) return 0;
    aff Q;
    if (!pub33_parse(&Q, pub33)) return 0;
    u256 z = s_from_bytes(msg32), w = s_inv(s);
    jac R = j_mul2(s_mul(z, w), G_(), s_mul(r, w), j_from(Q));
    if (R.inf) return 0;
    /* x(R) mod n == r  <=>  X == r Z^2 or (r + n < p and X == (r+n) Z^2) */
    u256 z2 = f_sqr(R.z);
    if (u_cmp(f_mul(r, z2), R.x) == 0) return 1;
    u256 rn;
    if (!u_add(&rn, r, N_) && u_cmp(rn, P_) < 0 && u_cmp(f_mul(rn, z2), R.x) == 0) return 1;
    return 0;
}

ORC_API int orc_ecdsa_verify(const u8* msg32, const u8* sig64, const u8* pub33) {
    return ecdsa_core(msg32, be_load(sig64), be_load(sig64 + 32), pub33, 1);
}

/* 256-bit little-endian integer (the opaque signature scalar layout). */
static u256 le_load(const u8* b) {
    u256 r;
    for (int i = 0; i < 4; ++i) {
        u64 v = 0;
        for (int j = 7; j >= 0; --j) v = (v << 8) | b[8 * i + j];
        r.v[i] = v;
    }
    return r;
}

/* Opaque-signature ECDSA: r, s little-endian, strict in [1,n); high-S is
 * accepted (the reference normalises s, which keeps x(R)). */
ORC_API int orc_ecdsa_verify_opaque(const u8* msg32, const u8* opaque64, const u8* pub33) {
    return ecdsa_core(msg32, le_load(opaque64), le_load(opaque64 + 32), pub33, 0);
}

ORC_API int orc_schnorr_verify(const u8* msg32, const u8* sig64, const u8* px32) {
    u256 r = be_load(sig64), s = be_load(sig64 + 32);
    if (u_cmp(r, P_) >= 0 || u_zero(s) || u_cmp(s, N_) >= 0) return 0;
    aff Pk;
    if (!lift(&Pk, px32, 0)) return 0;
    u8 eh[32];
    tagged(eh, "BIP0340/challenge", sig64, 32, px32, 32, msg32, 32);
    u256 e = s_from_bytes(eh);
    jac R = j_mul2(s, G_(), s_neg(e), j_from(Pk));
    aff A;
    if (!j_to_aff(&A, R)) return 0;
    return u_cmp(A.x, r) == 0 && (A.y.v[0] & 1) == 0;
}

/* ---- public-key recovery ------------------------------------------------- */
/* Q = (-z/r) G + (s/r) R, R = lift(r [+ n for ids 2, 3], parity = id & 1);
 * 1 and the compressed key on success, 0 and 33 zero bytes on failure. */
ORC_API int orc_ecdsa_recover(const u8* msg32, const u8* sig64, int recid, u8* out33) {
    memset(out33, 0, 33);
    if (recid < 0 || recid > 3) return 0;
    u256 r = be_load(sig64), s = be_load(sig64 + 32);
    if (u_zero(r) || u_cmp(r, N_) >= 0 || u_zero(s) || u_cmp(s, N_) >= 0) return 0;
    u256 rx = r;
    if (recid & 2) {
        u256 pmn;
        u_sub(&pmn, P_, N_);
        if (u_cmp(r, pmn) >= 0) return 0;
        u_add(&rx, r, N_);
    }
    u8 xb[32];
    be_store(xb, rx);
    aff R;
    if (!lift(&R, xb, recid & 1)) return 0;
    u256 ri = s_inv(r);
    u256 u1 = s_mul(s_neg(s_from_bytes(msg32)), ri), u2 = s_mul(s, ri);
    jac Q = j_mul2(u1, G_(), u2, j_from(R));
    aff A;
    if (!j_to_aff(&A, Q)) return 0;
    out33[0] = (u8)(2 | (A.y.v[0] & 1));
    be_store(out33 + 1, A.x);
    return 1;
}

/* ---- public-data key operations ----------------------------------------- */
ORC_API int orc_pubkey_validate(const u8* pub33) {
    aff Q;
    return pub33_parse(&Q, pub33);
}

ORC_API int orc_xonly_validate(const u8* x32) {
    aff Q;
    return lift(&Q, x32, 0);
}

/* BIP-341 tweak-add check as the reference bridge computes it: the two-term
 * MSM 1*lift_x(internal, even y) + t*G (t >= n, a bad point or infinity
 * fail), then the compressed result against (parity ? 03 : 02) || tweaked_x. */
ORC_API int orc_commitment_verify(const u8* internal_x32, const u8* tweak32, const u8* tweaked_x32, int parity) {
    u256 t = be_load(tweak32);
    if (u_cmp(t, N_) >= 0) return 0;
    aff P;
    if (!lift(&P, internal_x32, 0)) return 0;
    jac Q = j_mul2(t, G_(), f_small(1), j_from(P));
    aff A;
    if (!j_to_aff(&A, Q)) return 0;
    return u_cmp(A.x, be_load(tweaked_x32)) == 0 && (int)(A.y.v[0] & 1) == (parity != 0);
}

/* reference CPU ufsecp_multi_scalar_mul codes: 0 OK, 5 BAD_INPUT, 3 BAD_PUBKEY, 7 ARITH */
ORC_API int orc_msm(const u8* scalars32, const u8* points33, size_t n, u8* out33) {
    memset(out33, 0, 33);
    if (n == 0) return 5;
    jac acc = j_inf();
    for (size_t i = 0; i < n; ++i) {
        u256 s = be_load(scalars32 + 32 * i);
        if (u_cmp(s, N_) >= 0) return 5;
        aff Pt;
        if (!pub33_parse(&Pt, points33 + 33 * i)) return 3;
        acc = j_add(acc, j_mul2(s, j_from(Pt), f_small(0), j_inf()));
    }
    aff A;
    if (!j_to_aff(&A, acc)) return 7;
    out33[0] = (u8)(2 | (A.y.v[0] & 1));
    be_store(out33 + 1, A.x);
    return 0;
}

/* ---- signing (test-input generation only) ------------------------------- */
static aff g_mul_aff(u256 k) {
    aff A;
    j_to_aff(&A, j_mul2(k, G_(), f_small(0), j_inf()));
    return A;
}

ORC_API int orc_pubkey33(const u8* d32, u8* out33) {
    u256 d = be_load(d32);
    if (u_zero(d) || u_cmp(d, N_) >= 0) return 0;
    aff A = g_mul_aff(d);
    out33[0] = (u8)(2 | (A.y.v[0] & 1));
    be_store(out33 + 1, A.x);
    return 1;
}

static u256 rfc6979_k(u256 d, const u8* msg) {
    u8 x[32], V[32], K[32], buf[97], t33[33];
    be_store(x, d);
    memset(V, 1, 32);
    memset(K, 0, 32);
    memcpy(buf, V, 32); buf[32] = 0; memcpy(buf + 33, x, 32); memcpy(buf + 65, msg, 32);
    hmac(K, K, buf, 97);
    hmac(V, K, V, 32);
    memcpy(buf, V, 32); buf[32] = 1; memcpy(buf + 33, x, 32); memcpy(buf + 65, msg, 32);
    hmac(K, K, buf, 97);
    hmac(V, K, V, 32);
    hmac(V, K, V, 32);
    u256 c1 = be_load(V);
    int ok1 = !u_zero(c1) && u_cmp(c1, N_) < 0;
    memcpy(t33, V, 32); t33[32] = 0;
    hmac(K, K, t33, 33);
    hmac(V, K, V, 32);
    hmac(V, K, V, 32);
    return ok1 ? c1 : be_load(V);
}

ORC_API int orc_ecdsa_sign(const u8* msg32, const u8* d32, u8* sig64, u8* nonce32_out) {
    u256 d = be_load(d32);
    if (u_zero(d) || u_cmp(d, N_) >= 0) return 0;
    u256 k = rfc6979_k(d, msg32);
    if (nonce32_out) be_store(nonce32_out, k);
    aff R = g_mul_aff(k);
    u256 r = R.x;
    if (u_cmp(r, N_) >= 0) u_sub(&r, r, N_);
    u256 s = s_mul(s_inv(k), s_add(s_from_bytes(msg32), s_mul(r, d)));
    if (u_cmp(s, NH_) > 0) s = s_neg(s);
    be_store(sig64, r);
    be_store(sig64 + 32, s);
    return 1;
}

ORC_API int orc_schnorr_sign(const u8* msg32, const u8* d32, const u8* aux32, u8* sig64, u8* px_out) {
    u256 d = be_load(d32);
    if (u_zero(d) || u_cmp(d, N_) >= 0) return 0;
    aff Pk = g_mul_aff(d);
    if (Pk.y.v[0] & 1) d = s_neg(d);
    u8 px[32], db[32], t[32], ah[32], rnd[32], rx[32], eh[32];
    be_store(px, Pk.x);
    be_store(db, d);
    tagged(ah, "BIP0340/aux", aux32, 32, 0, 0, 0, 0);
    for (int i = 0; i < 32; ++i) t[i] = db[i] ^ ah[i];
    tagged(rnd, "BIP0340/nonce", t, 32, px, 32, msg32, 32);
    u256 k = s_from_bytes(rnd);
    if (u_zero(k)) return 0;
    aff R = g_mul_aff(k);
    if (R.y.v[0] & 1) k = s_neg(k);
    be_store(rx, R.x);
    tagged(eh, "BIP0340/challenge", rx, 32, px, 32, msg32, 32);
    u256 s = s_add(k, s_mul(s_from_bytes(eh), d));
    memcpy(sig64, rx, 32);
    be_store(sig64 + 32, s);
    if (px_out) memcpy(px_out, px, 32);
    return 1;
}

/* ---- multi-threaded batch verify (CPU baseline / large parity checks) ---- */
typedef struct {
    int kind;   /* 0 ECDSA, 1 Schnorr, 2 opaque ECDSA */
    const u8 *msgs, *pubs, *sigs;
    u8* out;
    size_t lo, hi;
} job_t;

static void* run_job(void* arg) {
    job_t* j = (job_t*)arg;
    const size_t ps = j->kind == 1 ? 32 : 33;
    for (size_t i = j->lo; i < j->hi; ++i) {
        const u8 *m = j->msgs + 32 * i, *sg = j->sigs + 64 * i, *pk = j->pubs + ps * i;
        j->out[i] = (u8)(j->kind == 1   ? orc_schnorr_verify(m, sg, pk)
                         : j->kind == 2 ? orc_ecdsa_verify_opaque(m, sg, pk)
                                        : orc_ecdsa_verify(m, sg, pk));
    }
    return 0;
}

ORC_API int orc_verify_batch(int kind, const u8* msgs, const u8* pubs, const u8* sigs, size_t n, u8* out,
                             int threads) {
    if (threads < 1) threads = 1;
    if (threads > 256) threads = 256;
    pthread_t th[256];
    job_t jobs[256];
    const size_t chunk = (n + (size_t)threads - 1) / (size_t)threads;
    int started = 0;
    for (int t = 0; t < threads; ++t) {
        jobs[t].kind = kind; jobs[t].msgs = msgs; jobs[t].pubs = pubs; jobs[t].sigs = sigs; jobs[t].out = out;
        jobs[t].lo = (size_t)t * chunk < n ? (size_t)t * chunk : n;
        jobs[t].hi = (size_t)(t + 1) * chunk < n ? (size_t)(t + 1) * chunk : n;
        if (pthread_create(&th[t], 0, run_job, &jobs[t]) == 0) {
            ++started;
        } else {
            run_job(&jobs[t]);
            th[t] = 0;
        }
    }
    for (int t = 0; t < threads; ++t)
        if (th[t]) pthread_join(th[t], 0);
    return started;
}
