// SHA-256 (FIPS 180-4), usable on the host and in device code.  The round
// loop is fully unrolled so the round constants become immediates on the GPU.
#pragma once

#include <stddef.h>
#include <stdint.h>

#include "fp64.hpp"  // ZHD

namespace zoda {

struct Sha256 {
  uint32_t h[8];
  uint8_t buf[64];
  uint32_t nbuf;
  uint64_t bits;

  static ZHD uint32_t rotr(uint32_t x, int n) { return (x >> n) | (x << (32 - n)); }

  ZHD void init() {
    h[0] = 0x6a09e667u; h[1] = 0xbb67ae85u; h[2] = 0x3c6ef372u; h[3] = 0xa54ff53au;
    h[4] = 0x510e527fu; h[5] = 0x9b05688cu; h[6] = 0x1f83d9abu; h[7] = 0x5be0cd19u;
    nbuf = 0;
    bits = 0;
  }

  ZHD void block(const uint8_t* p) {
    uint32_t w[16];
#pragma unroll
    for (int i = 0; i < 16; i++)
      w[i] = ((uint32_t)p[4 * i] << 24) | ((uint32_t)p[4 * i + 1] << 16) | ((uint32_t)p[4 * i + 2] << 8) |
             (uint32_t)p[4 * i + 3];
    compress(w);
  }

  // One block given as its 16 big-endian message words (overwritten: the
  // message schedule runs in place).
  ZHD void compress(uint32_t (&w)[16]) {
    constexpr uint32_t K[64] = {
        0x428a2f98u, 0x71374491u, 0xb5c0fbcfu, 0xe9b5dba5u, 0x3956c25bu, 0x59f111f1u, 0x923f82a4u,
        0xab1c5ed5u, 0xd807aa98u, 0x12835b01u, 0x243185beu, 0x550c7dc3u, 0x72be5d74u, 0x80deb1feu,
        0x9bdc06a7u, 0xc19bf174u, 0xe49b69c1u, 0xefbe4786u, 0x0fc19dc6u, 0x240ca1ccu, 0x2de92c6fu,
        0x4a7484aau, 0x5cb0a9dcu, 0x76f988dau, 0x983e5152u, 0xa831c66du, 0xb00327c8u, 0xbf597fc7u,
        0xc6e00bf3u, 0xd5a79147u, 0x06ca6351u, 0x14292967u, 0x27b70a85u, 0x2e1b2138u, 0x4d2c6dfcu,
        0x53380d13u, 0x650a7354u, 0x766a0abbu, 0x81c2c92eu, 0x92722c85u, 0xa2bfe8a1u, 0xa81a664bu,
        0xc24b8b70u, 0xc76c51a3u, 0xd192e819u, 0xd6990624u, 0xf40e3585u, 0x106aa070u, 0x19a4c116u,
        0x1e376c08u, 0x2748774cu, 0x34b0bcb5u, 0x391c0cb3u, 0x4ed8aa4au, 0x5b9cca4fu, 0x682e6ff3u,
        0x748f82eeu, 0x78a5636fu, 0x84c87814u, 0x8cc70208u, 0x90befffau, 0xa4506cebu, 0xbef9a3f7u,
        0xc67178f2u};
    uint32_t a = h[0], b = h[1], c = h[2], d = h[3], e = h[4], f = h[5], g = h[6], hh = h[7];
#pragma unroll
    for (int i = 0; i < 64; i++) {
      uint32_t wi;
      if (i < 16) {
        wi = w[i];
      } else {
        const uint32_t w15 = w[(i + 1) & 15], w2 = w[(i + 14) & 15];
        const uint32_t s0 = rotr(w15, 7) ^ rotr(w15, 18) ^ (w15 >> 3);
        const uint32_t s1 = rotr(w2, 17) ^ rotr(w2, 19) ^ (w2 >> 10);
        wi = w[i & 15] = w[i & 15] + s0 + w[(i + 9) & 15] + s1;
      }
      const uint32_t S1 = rotr(e, 6) ^ rotr(e, 11) ^ rotr(e, 25);
      const uint32_t ch = (e & f) ^ (~e & g);
      const uint32_t t1 = hh + S1 + ch + K[i] + wi;
      const uint32_t S0 = rotr(a, 2) ^ rotr(a, 13) ^ rotr(a, 22);
      const uint32_t mj = (a & b) ^ (a & c) ^ (b & c);
      const uint32_t t2 = S0 + mj;
      hh = g;
      g = f;
      f = e;
      e = d + t1;
      d = c;
      c = b;
      b = a;
      a = t1 + t2;
    }
    h[0] += a; h[1] += b; h[2] += c; h[3] += d; h[4] += e; h[5] += f; h[6] += g; h[7] += hh;
  }

  ZHD void update(const uint8_t* p, size_t n) {
    bits += (uint64_t)n * 8;
    while (n > 0) {
      if (nbuf == 0 && n >= 64) {
        block(p);
        p += 64;
        n -= 64;
        continue;
      }
      const size_t take = (64 - nbuf) < n ? (64 - nbuf) : n;
      for (size_t i = 0; i < take; i++) buf[nbuf + i] = p[i];
      nbuf += (uint32_t)take;
      p += take;
      n -= take;
      if (nbuf == 64) {
        block(buf);
        nbuf = 0;
      }
    }
  }

  ZHD void final(uint8_t out[32]) {
    const uint64_t total = bits;
    buf[nbuf++] = 0x80;
    if (nbuf > 56) {
      while (nbuf < 64) buf[nbuf++] = 0;
      block(buf);
      nbuf = 0;
    }
    while (nbuf < 56) buf[nbuf++] = 0;
    for (int i = 0; i < 8; i++) buf[56 + i] = (uint8_t)(total >> (56 - 8 * i));
    block(buf);
    for (int i = 0; i < 8; i++) {
      out[4 * i] = (uint8_t)(h[i] >> 24);
      out[4 * i + 1] = (uint8_t)(h[i] >> 16);
      out[4 * i + 2] = (uint8_t)(h[i] >> 8);
      out[4 * i + 3] = (uint8_t)h[i];
    }
  }
};

// The four limbs from a finished SHA-256 state: limb l is the big-endian u64
// in digest bytes [8l, 8l + 8) reduced mod p.
ZHD void rlc_limbs(const uint32_t h[8], uint32_t out[4]) {
  for (int l = 0; l < 4; l++) out[l] = (uint32_t)((((uint64_t)h[2 * l] << 32) | h[2 * l + 1]) % 2013265921ull);
}

// RLC coefficient of column `col` for a data root whose length is a multiple
// of 64 bytes, from `mid`, the SHA-256 state after the root's blocks: the one
// block left holds the index's 8 little-endian bytes, the 0x80 terminator and
// the message length in bits (total_bits = root bytes * 8 + 64).
ZHD void rlc_coefficient_mid(const uint32_t mid[8], uint64_t total_bits, uint64_t col, uint32_t out[4]) {
  Sha256 s;
  for (int i = 0; i < 8; i++) s.h[i] = mid[i];
  const uint32_t lo = (uint32_t)col, hi = (uint32_t)(col >> 32);
  uint32_t w[16] = {0};
  w[0] = (lo << 24) | ((lo << 8) & 0xff0000u) | ((lo >> 8) & 0xff00u) | (lo >> 24);  // bytes in order
  w[1] = (hi << 24) | ((hi << 8) & 0xff0000u) | ((hi >> 8) & 0xff00u) | (hi >> 24);
  w[2] = 0x80000000u;
  w[14] = (uint32_t)(total_bits >> 32);
  w[15] = (uint32_t)total_bits;
  s.compress(w);
  rlc_limbs(s.h, out);
}

// RLC coefficient of column `col`: SHA-256 of the data root's text followed
// by the column index as 8 little-endian bytes; limb l is the big-endian u64
// in digest bytes [8l, 8l + 8) reduced mod p (src/zoda/rlc.rs:48-66).
ZHD void rlc_coefficient(const uint8_t* root, uint32_t root_len, uint64_t col, uint32_t out[4]) {
  Sha256 s;
  s.init();
  s.update(root, root_len);
  uint8_t le[8];
  for (int i = 0; i < 8; i++) le[i] = (uint8_t)(col >> (8 * i));
  s.update(le, 8);
  uint8_t d[32];
  s.final(d);  // leaves the digest words in s.h
  rlc_limbs(s.h, out);
}

}  // namespace zoda
