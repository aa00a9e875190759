// SHA-256 compression on the device, used for the BIP-340 challenge
// e = SHA256(SHA256(tag) || SHA256(tag) || r || P.x || m) mod n.
// The 64-byte tag prefix is absorbed once into a midstate at context set-up,
// so each signature costs exactly two compressions (reference:
// src/cuda/include/schnorr.cuh:55-82 tagged_hash_fast, ecdsa.cuh:57-86).
#pragma once
#include "secp_field.h"

namespace ufs {

__constant__ const u32 kSha256K[64] = {
    0x428a2f98, 0x71374491, 0xb5c0fbcf, 0xe9b5dba5, 0x3956c25b, 0x59f111f1, 0x923f82a4, 0xab1c5ed5,
    0xd807aa98, 0x12835b01, 0x243185be, 0x550c7dc3, 0x72be5d74, 0x80deb1fe, 0x9bdc06a7, 0xc19bf174,
    0xe49b69c1, 0xefbe4786, 0x0fc19dc6, 0x240ca1cc, 0x2de92c6f, 0x4a7484aa, 0x5cb0a9dc, 0x76f988da,
    0x983e5152, 0xa831c66d, 0xb00327c8, 0xbf597fc7, 0xc6e00bf3, 0xd5a79147, 0x06ca6351, 0x14292967,
    0x27b70a85, 0x2e1b2138, 0x4d2c6dfc, 0x53380d13, 0x650a7354, 0x766a0abb, 0x81c2c92e, 0x92722c85,
    0xa2bfe8a1, 0xa81a664b, 0xc24b8b70, 0xc76c51a3, 0xd192e819, 0xd6990624, 0xf40e3585, 0x106aa070,
    0x19a4c116, 0x1e376c08, 0x2748774c, 0x34b0bcb5, 0x391c0cb3, 0x4ed8aa4a, 0x5b9cca4f, 0x682e6ff3,
    0x748f82ee, 0x78a5636f, 0x84c87814, 0x8cc70208, 0x90befffa, 0xa4506ceb, 0xbef9a3f7, 0xc67178f2};

#define UFS_SHA256_IV 0x6a09e667U, 0xbb67ae85U, 0x3c6ef372U, 0xa54ff53aU, 0x510e527fU, 0x9b05688cU, 0x1f83d9abU, 0x5be0cd19U

UFS_DEV u32 rotr32(u32 x, u32 n) { return __builtin_amdgcn_alignbit(x, x, n); }

// st <- compress(st, w[0..15]) where w are big-endian message words.
UFS_DEV void sha256_compress(u32 st[8], const u32 win[16]) {
    u32 w[16];
#pragma unroll
    for (int i = 0; i < 16; ++i) w[i] = win[i];
    u32 a = st[0], b = st[1], c = st[2], d = st[3], e = st[4], f = st[5], g = st[6], h = st[7];
#pragma unroll
    for (int i = 0; i < 64; ++i) {
        u32 wi;
        if (i < 16) {
            wi = w[i];
        } else {
            const u32 x15 = w[(i - 15) & 15], x2 = w[(i - 2) & 15];
            const u32 s0 = rotr32(x15, 7) ^ rotr32(x15, 18) ^ (x15 >> 3);
            const u32 s1 = rotr32(x2, 17) ^ rotr32(x2, 19) ^ (x2 >> 10);
            wi = w[i & 15] + s0 + w[(i - 7) & 15] + s1;
            w[i & 15] = wi;
        }
        const u32 S1 = rotr32(e, 6) ^ rotr32(e, 11) ^ rotr32(e, 25);
        const u32 ch = (e & f) ^ (~e & g);
        const u32 t1 = h + S1 + ch + kSha256K[i] + wi;
        const u32 S0 = rotr32(a, 2) ^ rotr32(a, 13) ^ rotr32(a, 22);
        const u32 mj = (a & b) ^ (a & c) ^ (b & c);
        const u32 t2 = S0 + mj;
        h = g; g = f; f = e; e = d + t1; d = c; c = b; b = a; a = t1 + t2;
    }
    st[0] += a; st[1] += b; st[2] += c; st[3] += d;
    st[4] += e; st[5] += f; st[6] += g; st[7] += h;
}

// BIP-340 challenge hash from the tag midstate: returns the 32-byte digest as
// eight big-endian words.  r, px, m are eight big-endian words each.
UFS_DEV void bip340_challenge(u32 out[8], const u32 mid[8], const u32 r[8], const u32 px[8], const u32 m[8]) {
    u32 st[8], blk[16];
#pragma unroll
    for (int i = 0; i < 8; ++i) st[i] = mid[i];
#pragma unroll
    for (int i = 0; i < 8; ++i) { blk[i] = r[i]; blk[8 + i] = px[i]; }
    sha256_compress(st, blk);
#pragma unroll
    for (int i = 0; i < 8; ++i) blk[i] = m[i];
    blk[8] = 0x80000000U;
#pragma unroll
    for (int i = 9; i < 15; ++i) blk[i] = 0;
    blk[15] = (64 + 96) * 8;   // message bits after the 64-byte tag prefix
    sha256_compress(st, blk);
#pragma unroll
    for (int i = 0; i < 8; ++i) out[i] = st[i];
}

UFS_DEV u32 bswap32(u32 x) { return __builtin_bswap32(x); }

}  // namespace ufs
