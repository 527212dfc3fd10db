// BLAKE2b (RFC 7693), 64-byte digests, optional key, for the prover's Fiat-Shamir
// transcript (blake2b_simd State keyed "Domain separator for transcript",
// proofs/src/transcript/implementors.rs:13-29) and the verifying-key hash
// (personalisation "Halo2-Verify-Key", proofs/src/plonk/mod.rs:249-283).
// digest() does not consume the state: hashing may continue afterwards, as the
// transcript's squeeze requires.
#pragma once
#include <stdint.h>
#include <string.h>

namespace mzk {
namespace host {

class Blake2b {
 public:
  // key: up to 64 bytes; personal: exactly 16 bytes or null
  Blake2b(const uint8_t* key, size_t keylen, const uint8_t* personal = nullptr) {
    static const uint64_t IV[8] = {0x6a09e667f3bcc908ull, 0xbb67ae8584caa73bull, 0x3c6ef372fe94f82bull,
                                   0xa54ff53a5f1d36f1ull, 0x510e527fade682d1ull, 0x9b05688c2b3e6c1full,
                                   0x1f83d9abfb41bd6bull, 0x5be0cd19137e2179ull};
    for (int i = 0; i < 8; i++) h_[i] = IV[i];
    h_[0] ^= 0x01010000ull ^ ((uint64_t)keylen << 8) ^ 64ull;   // digest 64, fanout 1, depth 1
    if (personal) {
      uint64_t p0, p1;
      memcpy(&p0, personal, 8);
      memcpy(&p1, personal + 8, 8);
      h_[6] ^= p0;
      h_[7] ^= p1;
    }
    t_ = 0;
    buflen_ = 0;
    if (keylen) {
      uint8_t block[128] = {0};
      memcpy(block, key, keylen);
      update(block, 128);
    }
  }

  void update(const uint8_t* in, size_t len) {
    while (len > 0) {
      if (buflen_ == 128) {   // a full block is compressed only once more input follows
        t_ += 128;
        compress(buf_, false);
        buflen_ = 0;
      }
      const size_t take = len < 128 - buflen_ ? len : 128 - buflen_;
      memcpy(buf_ + buflen_, in, take);
      buflen_ += take;
      in += take;
      len -= take;
    }
  }

  void digest(uint8_t out[64]) const {
    Blake2b s = *this;
    s.t_ += s.buflen_;
    memset(s.buf_ + s.buflen_, 0, 128 - s.buflen_);
    s.compress(s.buf_, true);
    memcpy(out, s.h_, 64);
  }

 private:
  uint64_t h_[8];
  uint64_t t_;   // bytes hashed (messages here stay far below 2^64)
  uint8_t buf_[128];
  size_t buflen_;

  static uint64_t rotr(uint64_t x, int n) { return (x >> n) | (x << (64 - n)); }

  void compress(const uint8_t block[128], bool last) {
    static const uint8_t SIGMA[12][16] = {
        {0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13, 14, 15}, {14, 10, 4, 8, 9, 15, 13, 6, 1, 12, 0, 2, 11, 7, 5, 3},
        {11, 8, 12, 0, 5, 2, 15, 13, 10, 14, 3, 6, 7, 1, 9, 4}, {7, 9, 3, 1, 13, 12, 11, 14, 2, 6, 5, 10, 4, 0, 15, 8},
        {9, 0, 5, 7, 2, 4, 10, 15, 14, 1, 11, 12, 6, 8, 3, 13}, {2, 12, 6, 10, 0, 11, 8, 3, 4, 13, 7, 5, 15, 14, 1, 9},
        {12, 5, 1, 15, 14, 13, 4, 10, 0, 7, 6, 3, 9, 2, 8, 11}, {13, 11, 7, 14, 12, 1, 3, 9, 5, 0, 15, 4, 8, 6, 2, 10},
        {6, 15, 14, 9, 11, 3, 0, 8, 12, 2, 13, 7, 1, 4, 10, 5}, {10, 2, 8, 4, 7, 6, 1, 5, 15, 11, 9, 14, 3, 12, 13, 0},
        {0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13, 14, 15}, {14, 10, 4, 8, 9, 15, 13, 6, 1, 12, 0, 2, 11, 7, 5, 3}};
    static const uint64_t IV[8] = {0x6a09e667f3bcc908ull, 0xbb67ae8584caa73bull, 0x3c6ef372fe94f82bull,
                                   0xa54ff53a5f1d36f1ull, 0x510e527fade682d1ull, 0x9b05688c2b3e6c1full,
                                   0x1f83d9abfb41bd6bull, 0x5be0cd19137e2179ull};
    uint64_t m[16], v[16];
    for (int i = 0; i < 16; i++) memcpy(&m[i], block + 8 * i, 8);
    for (int i = 0; i < 8; i++) {
      v[i] = h_[i];
      v[i + 8] = IV[i];
    }
    v[12] ^= t_;
    if (last) v[14] = ~v[14];
    for (int r = 0; r < 12; r++) {
      const uint8_t* s = SIGMA[r];
      auto G = [&](int a, int b, int c, int d, uint64_t x, uint64_t y) {
        v[a] = v[a] + v[b] + x;
        v[d] = rotr(v[d] ^ v[a], 32);
        v[c] = v[c] + v[d];
        v[b] = rotr(v[b] ^ v[c], 24);
        v[a] = v[a] + v[b] + y;
        v[d] = rotr(v[d] ^ v[a], 16);
        v[c] = v[c] + v[d];
        v[b] = rotr(v[b] ^ v[c], 63);
      };
      G(0, 4, 8, 12, m[s[0]], m[s[1]]);
      G(1, 5, 9, 13, m[s[2]], m[s[3]]);
      G(2, 6, 10, 14, m[s[4]], m[s[5]]);
      G(3, 7, 11, 15, m[s[6]], m[s[7]]);
      G(0, 5, 10, 15, m[s[8]], m[s[9]]);
      G(1, 6, 11, 12, m[s[10]], m[s[11]]);
      G(2, 7, 8, 13, m[s[12]], m[s[13]]);
      G(3, 4, 9, 14, m[s[14]], m[s[15]]);
    }
    for (int i = 0; i < 8; i++) h_[i] ^= v[i] ^ v[i + 8];
  }
};

}  // namespace host
}  // namespace mzk
