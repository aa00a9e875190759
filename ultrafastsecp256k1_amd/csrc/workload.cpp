// Synthetic workload generator (CPU, multi-threaded): keys, RFC6979 ECDSA and
// BIP-340 signatures and MSM terms following the input specification in
// ultrafastsecp256k1_amd/workload.py.  Only public test data is produced for
// the GPU; the synthetic secret keys never leave this host-side generator.
//
// Signing semantics reproduce the reference CPU signers byte for byte:
//   ECDSA  : src/cpu/src/ct_sign.cpp:40-70 + ecdsa.cpp:335-430 (RFC 6979 with
//            h1 = the raw 32-byte digest, first valid of two candidates,
//            s normalised to low-S);
//   BIP-340: the published BIP-340 signing algorithm with aux_rand.
// Arithmetic: 4 x 64-bit limbs with unsigned __int128; k*G by an 8-bit comb
// over a 512 KiB affine table built once; Z inversions batched per chunk.
#include <pthread.h>

#include <algorithm>
#include <array>
#include <atomic>
#include <cstdint>
#include <cstring>
#include <mutex>
#include <thread>
#include <vector>

#define WL_API extern "C" __attribute__((visibility("default")))

namespace {

typedef unsigned __int128 u128;
typedef uint64_t u64;
typedef uint32_t u32;
typedef uint8_t u8;

// ---------------------------------------------------------------------------
// SHA-256
// ---------------------------------------------------------------------------
const u32 K256[64] = {
    0x428a2f98, 0x71374491, 0xb5c0fbcf, 0xe9b5dba5, 0x3956c25b, 0x59f111f1, 0x923f82a4, 0xab1c5ed5,
    0xd807aa98, 0x12835b01, 0x243185be, 0x550c7dc3, 0x72be5d74, 0x80deb1fe, 0x9bdc06a7, 0xc19bf174,
    0xe49b69c1, 0xefbe4786, 0x0fc19dc6, 0x240ca1cc, 0x2de92c6f, 0x4a7484aa, 0x5cb0a9dc, 0x76f988da,
    0x983e5152, 0xa831c66d, 0xb00327c8, 0xbf597fc7, 0xc6e00bf3, 0xd5a79147, 0x06ca6351, 0x14292967,
    0x27b70a85, 0x2e1b2138, 0x4d2c6dfc, 0x53380d13, 0x650a7354, 0x766a0abb, 0x81c2c92e, 0x92722c85,
    0xa2bfe8a1, 0xa81a664b, 0xc24b8b70, 0xc76c51a3, 0xd192e819, 0xd6990624, 0xf40e3585, 0x106aa070,
    0x19a4c116, 0x1e376c08, 0x2748774c, 0x34b0bcb5, 0x391c0cb3, 0x4ed8aa4a, 0x5b9cca4f, 0x682e6ff3,
    0x748f82ee, 0x78a5636f, 0x84c87814, 0x8cc70208, 0x90befffa, 0xa4506ceb, 0xbef9a3f7, 0xc67178f2};

inline u32 rotr(u32 x, int n) { return (x >> n) | (x << (32 - n)); }

void sha256_block(u32 st[8], const u8* p) {
    u32 w[64];
    for (int i = 0; i < 16; ++i) w[i] = (u32)p[4 * i] << 24 | (u32)p[4 * i + 1] << 16 | (u32)p[4 * i + 2] << 8 | p[4 * i + 3];
    for (int i = 16; i < 64; ++i) {
        u32 s0 = rotr(w[i - 15], 7) ^ rotr(w[i - 15], 18) ^ (w[i - 15] >> 3);
        u32 s1 = rotr(w[i - 2], 17) ^ rotr(w[i - 2], 19) ^ (w[i - 2] >> 10);
        w[i] = w[i - 16] + s0 + w[i - 7] + s1;
    }
    u32 a = st[0], b = st[1], c = st[2], d = st[3], e = st[4], f = st[5], g = st[6], h = st[7];
    for (int i = 0; i < 64; ++i) {
        u32 t1 = h + (rotr(e, 6) ^ rotr(e, 11) ^ rotr(e, 25)) + ((e & f) ^ (~e & g)) + K256[i] + w[i];
        u32 t2 = (rotr(a, 2) ^ rotr(a, 13) ^ rotr(a, 22)) + ((a & b) ^ (a & c) ^ (b & c));
        h = g; g = f; f = e; e = d + t1; d = c; c = b; b = a; a = t1 + t2;
    }
    st[0] += a; st[1] += b; st[2] += c; st[3] += d; st[4] += e; st[5] += f; st[6] += g; st[7] += h;
}

struct Sha256 {
    u32 st[8] = {0x6a09e667, 0xbb67ae85, 0x3c6ef372, 0xa54ff53a, 0x510e527f, 0x9b05688c, 0x1f83d9ab, 0x5be0cd19};
    u8 buf[64];
    size_t len = 0;
    void update(const u8* p, size_t n) {
        while (n > 0) {
            size_t take = std::min(n, 64 - (len & 63));
            std::memcpy(buf + (len & 63), p, take);
            len += take;
            p += take;
            n -= take;
            if ((len & 63) == 0) sha256_block(st, buf);
        }
    }
    void final(u8 out[32]) {
        const u64 bits = (u64)len * 8;
        u8 pad = 0x80;
        update(&pad, 1);
        u8 z = 0;
        while ((len & 63) != 56) update(&z, 1);
        u8 lb[8];
        for (int i = 0; i < 8; ++i) lb[i] = (u8)(bits >> (56 - 8 * i));
        update(lb, 8);
        for (int i = 0; i < 8; ++i) {
            out[4 * i] = (u8)(st[i] >> 24); out[4 * i + 1] = (u8)(st[i] >> 16);
            out[4 * i + 2] = (u8)(st[i] >> 8); out[4 * i + 3] = (u8)st[i];
        }
    }
};

void sha256(const u8* p, size_t n, u8 out[32]) {
    Sha256 h;
    h.update(p, n);
    h.final(out);
}

void hmac_sha256(const u8 key[32], const u8* msg, size_t n, u8 out[32]) {
    u8 k[64] = {0}, ipad[64], opad[64], inner[32];
    std::memcpy(k, key, 32);
    for (int i = 0; i < 64; ++i) { ipad[i] = k[i] ^ 0x36; opad[i] = k[i] ^ 0x5c; }
    Sha256 hi; hi.update(ipad, 64); hi.update(msg, n); hi.final(inner);
    Sha256 ho; ho.update(opad, 64); ho.update(inner, 32); ho.final(out);
}

void tagged_hash(const char* tag, const u8* a, size_t na, const u8* b, size_t nb, const u8* c, size_t nc,
                 u8 out[32]) {
    u8 th[32];
    sha256((const u8*)tag, std::strlen(tag), th);
    Sha256 h;
    h.update(th, 32); h.update(th, 32);
    if (na) h.update(a, na);
    if (nb) h.update(b, nb);
    if (nc) h.update(c, nc);
    h.final(out);
}

// ---------------------------------------------------------------------------
// 256-bit helpers (little-endian 64-bit limbs)
// ---------------------------------------------------------------------------
struct U256 { u64 v[4]; };

const U256 P = {{0xFFFFFFFEFFFFFC2FULL, 0xFFFFFFFFFFFFFFFFULL, 0xFFFFFFFFFFFFFFFFULL, 0xFFFFFFFFFFFFFFFFULL}};
const U256 N = {{0xBFD25E8CD0364141ULL, 0xBAAEDCE6AF48A03BULL, 0xFFFFFFFFFFFFFFFEULL, 0xFFFFFFFFFFFFFFFFULL}};
const U256 NHALF = {{0xDFE92F46681B20A0ULL, 0x5D576E7357A4501DULL, 0xFFFFFFFFFFFFFFFFULL, 0x7FFFFFFFFFFFFFFFULL}};
const U256 NC = {{0x402DA1732FC9BEBFULL, 0x4551231950B75FC4ULL, 1, 0}};   // 2^256 - n
const U256 GX = {{0x59F2815B16F81798ULL, 0x029BFCDB2DCE28D9ULL, 0x55A06295CE870B07ULL, 0x79BE667EF9DCBBACULL}};
const U256 GY = {{0x9C47D08FFB10D4B8ULL, 0xFD17B448A6855419ULL, 0x5DA4FBFC0E1108A8ULL, 0x483ADA7726A3C465ULL}};

U256 from_be(const u8* b) {
    U256 r;
    for (int i = 0; i < 4; ++i) {
        u64 x = 0;
        for (int j = 0; j < 8; ++j) x = x << 8 | b[(3 - i) * 8 + j];
        r.v[i] = x;
    }
    return r;
}

void to_be(const U256& a, u8* b) {
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 8; ++j) b[(3 - i) * 8 + j] = (u8)(a.v[i] >> (56 - 8 * j));
}

bool is_zero(const U256& a) { return (a.v[0] | a.v[1] | a.v[2] | a.v[3]) == 0; }

int cmp(const U256& a, const U256& b) {
    for (int i = 3; i >= 0; --i) {
        if (a.v[i] < b.v[i]) return -1;
        if (a.v[i] > b.v[i]) return 1;
    }
    return 0;
}

u64 add_to(U256& r, const U256& a, const U256& b) {
    u128 c = 0;
    for (int i = 0; i < 4; ++i) { c += (u128)a.v[i] + b.v[i]; r.v[i] = (u64)c; c >>= 64; }
    return (u64)c;
}

u64 sub_to(U256& r, const U256& a, const U256& b) {
    u64 bw = 0;
    for (int i = 0; i < 4; ++i) {
        u128 d = (u128)a.v[i] - b.v[i] - bw;
        r.v[i] = (u64)d;
        bw = (u64)(d >> 127) & 1;
    }
    return bw;
}

void mul_wide(u64 t[8], const U256& a, const U256& b) {
    for (int i = 0; i < 8; ++i) t[i] = 0;
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

// ---- field mod p (fully reduced) -------------------------------------------
U256 fe_reduce_wide(const u64 t[8]) {
    // t = lo + hi*2^256, 2^256 = 0x1000003D1 (mod p)
    const u64 K = 0x1000003D1ULL;
    U256 r;
    u128 c = 0;
    for (int i = 0; i < 4; ++i) {
        c += (u128)t[4 + i] * K + t[i];
        r.v[i] = (u64)c;
        c >>= 64;
    }
    // c < 2^34: fold again
    u128 c2 = c * K;
    for (int i = 0; i < 4; ++i) {
        c2 += r.v[i];
        r.v[i] = (u64)c2;
        c2 >>= 64;
    }
    if (c2) {   // one more wrap: add K
        u128 c3 = (u128)r.v[0] + K;
        r.v[0] = (u64)c3;
        c3 >>= 64;
        for (int i = 1; i < 4 && c3; ++i) { c3 += r.v[i]; r.v[i] = (u64)c3; c3 >>= 64; }
    }
    if (cmp(r, P) >= 0) sub_to(r, r, P);
    return r;
}

U256 fmul(const U256& a, const U256& b) { u64 t[8]; mul_wide(t, a, b); return fe_reduce_wide(t); }
U256 fsqr(const U256& a) { return fmul(a, a); }
U256 fadd(const U256& a, const U256& b) {
    U256 r;
    u64 c = add_to(r, a, b);
    if (c || cmp(r, P) >= 0) sub_to(r, r, P);
    return r;
}
U256 fsub(const U256& a, const U256& b) {
    U256 r;
    if (sub_to(r, a, b)) add_to(r, r, P);
    return r;
}
U256 fneg(const U256& a) { U256 z = {{0, 0, 0, 0}}; return fsub(z, a); }

U256 fpow(const U256& a, const U256& e) {
    U256 r = {{1, 0, 0, 0}};
    for (int i = 255; i >= 0; --i) {
        r = fsqr(r);
        if ((e.v[i >> 6] >> (i & 63)) & 1) r = fmul(r, a);
    }
    return r;
}

U256 finv(const U256& a) {
    U256 e;
    U256 two = {{2, 0, 0, 0}};
    sub_to(e, P, two);
    return fpow(a, e);
}

// ---- scalar mod n ----------------------------------------------------------
U256 sc_reduce_wide(const u64 t[8]) {
    // repeatedly fold hi * NC (NC < 2^129)
    u64 w[8];
    std::memcpy(w, t, sizeof(w));
    for (int round = 0; round < 3; ++round) {
        U256 hi = {{w[4], w[5], w[6], w[7]}};
        if (is_zero(hi)) break;
        u64 m[8];
        mul_wide(m, hi, NC);
        u128 c = 0;
        for (int i = 0; i < 8; ++i) {
            c += (u128)m[i] + (i < 4 ? w[i] : 0);
            w[i] = (u64)c;
            c >>= 64;
        }
    }
    U256 r = {{w[0], w[1], w[2], w[3]}};
    while (cmp(r, N) >= 0) sub_to(r, r, N);
    return r;
}

U256 smul(const U256& a, const U256& b) { u64 t[8]; mul_wide(t, a, b); return sc_reduce_wide(t); }
U256 sadd(const U256& a, const U256& b) {
    U256 r;
    u64 c = add_to(r, a, b);
    if (c || cmp(r, N) >= 0) sub_to(r, r, N);
    return r;
}
U256 sneg(const U256& a) {
    if (is_zero(a)) return a;
    U256 r;
    sub_to(r, N, a);
    return r;
}
U256 sinv(const U256& a) {
    U256 e, two = {{2, 0, 0, 0}};
    sub_to(e, N, two);
    U256 r = {{1, 0, 0, 0}};
    for (int i = 255; i >= 0; --i) {
        r = smul(r, r);
        if ((e.v[i >> 6] >> (i & 63)) & 1) r = smul(r, a);
    }
    return r;
}
U256 sc_from_bytes_reduced(const u8* b) {
    U256 r = from_be(b);
    if (cmp(r, N) >= 0) sub_to(r, r, N);
    return r;
}

// ---- points ----------------------------------------------------------------
struct Aff { U256 x, y; };
struct Jac { U256 x, y, z; bool inf; };

Jac jdbl(const Jac& a) {
    if (a.inf) return a;
    U256 A = fsqr(a.x), B = fsqr(a.y), C = fsqr(B);
    U256 D = fsub(fsub(fsqr(fadd(a.x, B)), A), C);
    D = fadd(D, D);
    U256 E = fadd(fadd(A, A), A), F = fsqr(E);
    Jac r;
    r.x = fsub(F, fadd(D, D));
    U256 C8 = fadd(C, C); C8 = fadd(C8, C8); C8 = fadd(C8, C8);
    r.y = fsub(fmul(E, fsub(D, r.x)), C8);
    r.z = fmul(a.y, a.z); r.z = fadd(r.z, r.z);
    r.inf = false;
    return r;
}

Jac jadd_aff(const Jac& a, const Aff& b) {
    if (a.inf) return Jac{b.x, b.y, {{1, 0, 0, 0}}, false};
    U256 z2 = fsqr(a.z);
    U256 u2 = fmul(b.x, z2), s2 = fmul(b.y, fmul(z2, a.z));
    U256 h = fsub(u2, a.x), r = fsub(s2, a.y);
    if (is_zero(h)) {
        if (is_zero(r)) return jdbl(a);
        Jac inf{}; inf.inf = true; return inf;
    }
    U256 h2 = fsqr(h), h3 = fmul(h2, h), v = fmul(a.x, h2);
    Jac o;
    o.x = fsub(fsub(fsqr(r), h3), fadd(v, v));
    o.y = fsub(fmul(r, fsub(v, o.x)), fmul(a.y, h3));
    o.z = fmul(a.z, h);
    o.inf = false;
    return o;
}

// comb table: T[i][j] = j * 2^(8i) * G for j in 1..255 (index j-1)
std::vector<Aff> g_comb;
std::once_flag g_comb_once;

void batch_to_affine(const std::vector<Jac>& in, std::vector<Aff>& out) {
    const size_t n = in.size();
    out.resize(n);
    std::vector<U256> pre(n);
    U256 acc = {{1, 0, 0, 0}};
    for (size_t i = 0; i < n; ++i) { pre[i] = acc; acc = fmul(acc, in[i].z); }
    U256 inv = finv(acc);
    for (size_t i = n; i-- > 0;) {
        U256 zi = fmul(inv, pre[i]);
        inv = fmul(inv, in[i].z);
        U256 zi2 = fsqr(zi);
        out[i].x = fmul(in[i].x, zi2);
        out[i].y = fmul(in[i].y, fmul(zi2, zi));
    }
}

void build_comb() {
    std::vector<Jac> jt(32 * 255);
    Jac base{GX, GY, {{1, 0, 0, 0}}, false};
    for (int i = 0; i < 32; ++i) {
        std::vector<Aff> b1;
        batch_to_affine(std::vector<Jac>{base}, b1);
        Jac acc{b1[0].x, b1[0].y, {{1, 0, 0, 0}}, false};
        jt[i * 255 + 0] = acc;
        for (int j = 2; j <= 255; ++j) {
            acc = jadd_aff(acc, b1[0]);
            jt[i * 255 + j - 1] = acc;
        }
        for (int d = 0; d < 8; ++d) base = jdbl(base);
    }
    batch_to_affine(jt, g_comb);
}

Jac gmul(const U256& k) {
    std::call_once(g_comb_once, build_comb);
    Jac acc{};
    acc.inf = true;
    for (int i = 0; i < 32; ++i) {
        const u32 byte = (u32)(k.v[i >> 3] >> (8 * (i & 7))) & 0xFF;
        if (byte) acc = jadd_aff(acc, g_comb[i * 255 + byte - 1]);
    }
    return acc;
}

// RFC 6979 nonce as in the reference CPU signer (h1 = raw digest, first valid
// of two candidates).
U256 rfc6979(const U256& d, const u8 msg[32]) {
    u8 x[32], V[32], K[32], buf[97], t33[33];
    to_be(d, x);
    std::memset(V, 0x01, 32);
    std::memset(K, 0x00, 32);
    std::memcpy(buf, V, 32); buf[32] = 0x00; std::memcpy(buf + 33, x, 32); std::memcpy(buf + 65, msg, 32);
    hmac_sha256(K, buf, 97, K);
    hmac_sha256(K, V, 32, V);
    std::memcpy(buf, V, 32); buf[32] = 0x01; std::memcpy(buf + 33, x, 32); std::memcpy(buf + 65, msg, 32);
    hmac_sha256(K, buf, 97, K);
    hmac_sha256(K, V, 32, V);
    hmac_sha256(K, V, 32, V);
    U256 c1 = from_be(V);
    const bool ok1 = !is_zero(c1) && cmp(c1, N) < 0;
    std::memcpy(t33, V, 32); t33[32] = 0x00;
    hmac_sha256(K, t33, 33, K);
    hmac_sha256(K, V, 32, V);
    hmac_sha256(K, V, 32, V);
    U256 c2 = from_be(V);
    return ok1 ? c1 : c2;
}

// ---- SplitMix64 item streams ----------------------------------------------
struct Rng {
    u64 s;
    Rng(u64 seed, u64 item) : s(seed * 0x9E3779B97F4A7C15ULL + (item + 1) * 0xD1B54A32D192ED03ULL) {}
    u64 next() {
        s += 0x9E3779B97F4A7C15ULL;
        u64 z = s;
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
        return z ^ (z >> 31);
    }
    void next256(u8 out[32]) {
        for (int i = 0; i < 4; ++i) {
            u64 v = next();
            for (int j = 0; j < 8; ++j) out[8 * i + j] = (u8)(v >> (56 - 8 * j));
        }
    }
    U256 key() {
        u8 b[32];
        for (;;) {
            next256(b);
            U256 v = from_be(b);
            if (!is_zero(v) && cmp(v, N) < 0) return v;
        }
    }
};

void corrupt(u64 c, u8* msg, u8* sig, u8* flag) {
    *flag = (c % 100 == 0) ? 1 : 0;
    if (!*flag) return;
    const u32 field = (u32)((c >> 8) % 3), bit = (u32)((c >> 16) % 256);
    u8* tgt = field == 2 ? msg : (sig + (field == 0 ? 0 : 32));
    tgt[bit >> 3] ^= (u8)(1u << (bit & 7));
}

template <class F>
void parallel_chunks(u64 n, int threads, F&& f) {
    if (threads <= 0) threads = (int)std::max(1u, std::thread::hardware_concurrency());
    threads = std::min<int>(threads, 64);
    const u64 chunk = 256;
    std::atomic<u64> next{0};
    std::vector<std::thread> ts;
    for (int t = 0; t < threads; ++t)
        ts.emplace_back([&] {
            for (;;) {
                const u64 b = next.fetch_add(chunk);
                if (b >= n) break;
                f(b, std::min(n, b + chunk));
            }
        });
    for (auto& t : ts) t.join();
}

// Key of an item.  Plain workloads: the item stream's first key() draw.
// Serving workloads (pool_size > 0): a first draw u picks, with probability
// pool_permille / 1000, key number (u >> 20) % pool_size of the pool (pool key
// j = the first key() draw of stream (pool_seed, j)); otherwise the item
// stream's next key() draw.
struct KeyPool {
    u64 seed = 0;
    u64 size = 0;
    u32 permille = 0;
    U256 key(Rng& r) const {
        if (size == 0) return r.key();
        const u64 u = r.next();
        if (u % 1000 < permille) {
            Rng p(seed, (u >> 20) % size);
            return p.key();
        }
        return r.key();
    }
};

int wl_ecdsa(const KeyPool& kp, u64 seed, u64 first, u64 n, u8* msgs, u8* pubs, u8* sigs, u8* corrupted,
             int threads);
int wl_schnorr(const KeyPool& kp, u64 seed, u64 first, u64 n, u8* msgs, u8* pubs, u8* sigs, u8* corrupted,
               int threads);

}  // namespace

// ECDSA items [first, first+n): msgs n*32, pubs n*33, sigs n*64, corrupted n.
WL_API int ufsecp_wl_ecdsa(u64 seed, u64 first, u64 n, u8* msgs, u8* pubs, u8* sigs, u8* corrupted,
                           int threads) {
    return wl_ecdsa(KeyPool{}, seed, first, n, msgs, pubs, sigs, corrupted, threads);
}

// BIP-340 items: msgs n*32, pubs (x-only) n*32, sigs n*64, corrupted n.
WL_API int ufsecp_wl_schnorr(u64 seed, u64 first, u64 n, u8* msgs, u8* pubs, u8* sigs, u8* corrupted,
                             int threads) {
    return wl_schnorr(KeyPool{}, seed, first, n, msgs, pubs, sigs, corrupted, threads);
}

// Serving workloads: as above, with a share of the keys drawn from a pool.
WL_API int ufsecp_wl_ecdsa_pool(u64 seed, u64 first, u64 n, u64 pool_seed, u64 pool_size, u32 pool_permille,
                                u8* msgs, u8* pubs, u8* sigs, u8* corrupted, int threads) {
    return wl_ecdsa(KeyPool{pool_seed, pool_size, pool_permille}, seed, first, n, msgs, pubs, sigs, corrupted,
                    threads);
}

WL_API int ufsecp_wl_schnorr_pool(u64 seed, u64 first, u64 n, u64 pool_seed, u64 pool_size, u32 pool_permille,
                                  u8* msgs, u8* pubs, u8* sigs, u8* corrupted, int threads) {
    return wl_schnorr(KeyPool{pool_seed, pool_size, pool_permille}, seed, first, n, msgs, pubs, sigs, corrupted,
                      threads);
}

namespace {

int wl_ecdsa(const KeyPool& kp, u64 seed, u64 first, u64 n, u8* msgs, u8* pubs, u8* sigs, u8* corrupted,
             int threads) {
    if (n == 0) return 0;
    parallel_chunks(n, threads, [&](u64 b, u64 e) {
        const size_t m = e - b;
        std::vector<U256> d(m), k(m), z(m);
        std::vector<u64> cw(m);
        std::vector<Jac> Q(m), R(m);
        for (size_t j = 0; j < m; ++j) {
            Rng r(seed, first + b + j);
            d[j] = kp.key(r);
            u8* msg = msgs + 32 * (b + j);
            r.next256(msg);
            cw[j] = r.next();
            z[j] = sc_from_bytes_reduced(msg);
            k[j] = rfc6979(d[j], msg);
            Q[j] = gmul(d[j]);
            R[j] = gmul(k[j]);
        }
        std::vector<Aff> Qa, Ra;
        batch_to_affine(Q, Qa);
        batch_to_affine(R, Ra);
        // batch inverse of k
        std::vector<U256> pre(m);
        U256 acc = {{1, 0, 0, 0}};
        for (size_t j = 0; j < m; ++j) { pre[j] = acc; acc = smul(acc, k[j]); }
        U256 inv = sinv(acc);
        std::vector<U256> kinv(m);
        for (size_t j = m; j-- > 0;) { kinv[j] = smul(inv, pre[j]); inv = smul(inv, k[j]); }
        for (size_t j = 0; j < m; ++j) {
            const size_t i = b + j;
            u8* pub = pubs + 33 * i;
            pub[0] = (u8)(0x02 | (Qa[j].y.v[0] & 1));
            to_be(Qa[j].x, pub + 1);
            U256 r = Ra[j].x;
            if (cmp(r, N) >= 0) sub_to(r, r, N);
            U256 s = smul(kinv[j], sadd(z[j], smul(r, d[j])));
            if (cmp(s, NHALF) > 0) s = sneg(s);
            u8* sig = sigs + 64 * i;
            to_be(r, sig);
            to_be(s, sig + 32);
            corrupt(cw[j], msgs + 32 * i, sig, corrupted + i);
        }
    });
    return 0;
}

int wl_schnorr(const KeyPool& kp, u64 seed, u64 first, u64 n, u8* msgs, u8* pubs, u8* sigs, u8* corrupted,
               int threads) {
    if (n == 0) return 0;
    parallel_chunks(n, threads, [&](u64 b, u64 e) {
        const size_t m = e - b;
        std::vector<U256> d(m);
        std::vector<u64> cw(m);
        std::vector<Jac> Pj(m), Rj(m);
        std::vector<std::array<u8, 32>> aux(m);
        for (size_t j = 0; j < m; ++j) {
            Rng r(seed, first + b + j);
            d[j] = kp.key(r);
            r.next256(msgs + 32 * (b + j));
            r.next256(aux[j].data());
            cw[j] = r.next();
            Pj[j] = gmul(d[j]);
        }
        std::vector<Aff> Pa;
        batch_to_affine(Pj, Pa);
        std::vector<U256> kk(m), dd(m);
        for (size_t j = 0; j < m; ++j) {
            const size_t i = b + j;
            dd[j] = (Pa[j].y.v[0] & 1) ? sneg(d[j]) : d[j];
            u8 px[32], db[32], t[32], ah[32], rnd[32];
            to_be(Pa[j].x, px);
            std::memcpy(pubs + 32 * i, px, 32);
            to_be(dd[j], db);
            tagged_hash("BIP0340/aux", aux[j].data(), 32, nullptr, 0, nullptr, 0, ah);
            for (int q = 0; q < 32; ++q) t[q] = db[q] ^ ah[q];
            tagged_hash("BIP0340/nonce", t, 32, px, 32, msgs + 32 * i, 32, rnd);
            kk[j] = sc_from_bytes_reduced(rnd);
            Rj[j] = gmul(kk[j]);
        }
        std::vector<Aff> Ra;
        batch_to_affine(Rj, Ra);
        for (size_t j = 0; j < m; ++j) {
            const size_t i = b + j;
            U256 k = (Ra[j].y.v[0] & 1) ? sneg(kk[j]) : kk[j];
            u8 rx[32], ev[32];
            to_be(Ra[j].x, rx);
            tagged_hash("BIP0340/challenge", rx, 32, pubs + 32 * i, 32, msgs + 32 * i, 32, ev);
            U256 ee = sc_from_bytes_reduced(ev);
            U256 s = sadd(k, smul(ee, dd[j]));
            u8* sig = sigs + 64 * i;
            std::memcpy(sig, rx, 32);
            to_be(s, sig + 32);
            corrupt(cw[j], msgs + 32 * i, sig, corrupted + i);
        }
    });
    return 0;
}

}  // namespace

// MSM terms: scalars n*32 (uniform 256-bit reduced mod n), points n*33 = k_i*G
// compressed, ks n*32 (the k_i, for the closed form).
// points: n x 33 compressed; points64 (optional, may be null): the same
// bases as n x 64 big-endian affine x || y.
WL_API int ufsecp_wl_msm(u64 seed, u64 first, u64 n, u8* scalars, u8* points, u8* points64, u8* ks, int threads) {
    if (n == 0) return 0;
    parallel_chunks(n, threads, [&](u64 b, u64 e) {
        const size_t m = e - b;
        std::vector<Jac> K(m);
        for (size_t j = 0; j < m; ++j) {
            Rng r(seed, first + b + j);
            U256 k = r.key();
            u8 sb[32];
            r.next256(sb);
            U256 s = from_be(sb);
            if (cmp(s, N) >= 0) sub_to(s, s, N);
            to_be(s, scalars + 32 * (b + j));
            to_be(k, ks + 32 * (b + j));
            K[j] = gmul(k);
        }
        std::vector<Aff> Ka;
        batch_to_affine(K, Ka);
        for (size_t j = 0; j < m; ++j) {
            u8* p = points + 33 * (b + j);
            p[0] = (u8)(0x02 | (Ka[j].y.v[0] & 1));
            to_be(Ka[j].x, p + 1);
            if (points64) {
                to_be(Ka[j].x, points64 + 64 * (b + j));
                to_be(Ka[j].y, points64 + 64 * (b + j) + 32);
            }
        }
    });
    return 0;
}
