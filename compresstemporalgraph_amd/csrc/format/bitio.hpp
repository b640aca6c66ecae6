// Host-side bit I/O and instantaneous codes for the ChronoGraph / WebGraph
// on-disk format.
//
// Bit order: MSB-first inside each byte, the convention of dsiutils'
// OutputBitStream that the reference GPU reader assumes when it scans
// `CHECK_BIT(byte, 7 - i % 8)` (gpu_cuda/shared/bfs.cu:166-175).
//
// Codes (dsiutils definitions, as decoded by gpu_cuda/shared/bfs.cu:166-175,
// 202-256, 590-665):
//   unary(x)  : x zeros then a one
//   gamma(x)  : v = x + 1, m = floor(log2 v); m zeros, then v in m+1 bits
//   zeta_k(x) : v = x + 1, h = floor(floor(log2 v) / k); unary(h), then the
//               minimal binary code of v - 2^(hk) over [0, 2^(hk+k) - 2^(hk))
//   int2nat   : v >= 0 -> 2v, v < 0 -> -2v - 1
#pragma once
#include <cstdint>
#include <cstring>
#include <stdexcept>
#include <vector>

namespace tgaf {

inline int msb64(uint64_t v) { return 63 - __builtin_clzll(v); }  // v > 0

inline uint64_t int2nat(int64_t x) { return x >= 0 ? (uint64_t)x << 1 : ((uint64_t)(-(x + 1)) << 1) | 1u; }
inline int64_t nat2int(uint64_t n) { return (n & 1) ? -(int64_t)(n >> 1) - 1 : (int64_t)(n >> 1); }

inline int gamma_len(uint64_t x) { return 2 * msb64(x + 1) + 1; }
inline int unary_len(uint64_t x) { return (int)x + 1; }
inline int zeta_len(uint64_t x, int k) {
  const uint64_t v = x + 1;
  const int h = msb64(v) / k;
  const uint64_t left = 1ull << (h * k);
  return h + 1 + ((v - left < left) ? h * k + k - 1 : h * k + k);
}

// Append-only MSB-first bit writer.  Bits are staged in a 64-bit accumulator
// and flushed byte by byte.
class BitWriter {
 public:
  uint64_t bits() const { return nbits_; }

  void put(uint64_t value, int len) {  // low `len` bits of value, MSB first; len <= 64
    if (len < 64) value &= (1ull << len) - 1;
    if (len > 32) {
      put32(value >> 32, len - 32);
      put32(value & 0xffffffffu, 32);
    } else if (len > 0) {
      put32(value & ((len == 32) ? 0xffffffffu : ((1u << len) - 1)), len);
    }
  }
  void unary(uint64_t x) {
    while (x >= 32) { put32(0, 32); x -= 32; }
    put32(1, (int)x + 1);
  }
  void gamma(uint64_t x) {
    const uint64_t v = x + 1;
    const int m = msb64(v);
    unary(m);
    put(v, m);  // low m bits (the leading one is implicit in the unary part)
  }
  void zeta(uint64_t x, int k) {
    const uint64_t v = x + 1;
    const int h = msb64(v) / k;
    unary(h);
    const uint64_t left = 1ull << (h * k);
    if (v - left < left) put(v - left, h * k + k - 1);
    else put(v, h * k + k);
  }
  // Flushes the partial byte (zero padded) and returns the byte buffer.
  std::vector<uint8_t>& finish() {
    if (fill_ > 0) {
      out_.push_back((uint8_t)(acc_ << (8 - fill_)));
      fill_ = 0;
      acc_ = 0;
    }
    return out_;
  }
  // Appends `nbits` bits of another MSB-first stream.
  void append_stream(const std::vector<uint8_t>& src, uint64_t nbits) {
    uint64_t i = 0;
    for (; i + 8 <= nbits; i += 8) put32(src[i >> 3], 8);
    for (; i < nbits; ++i) put32((src[i >> 3] >> (7 - (i & 7))) & 1, 1);
  }

 private:
  void put32(uint64_t value, int len) {  // len <= 32, value < 2^len
    acc_ = (acc_ << len) | value;
    fill_ += len;
    nbits_ += (uint64_t)len;
    while (fill_ >= 8) {
      fill_ -= 8;
      out_.push_back((uint8_t)(acc_ >> fill_));
    }
    acc_ &= (fill_ == 0) ? 0 : ((1ull << fill_) - 1);
  }
  std::vector<uint8_t> out_;
  uint64_t acc_ = 0;
  int fill_ = 0;
  uint64_t nbits_ = 0;
};

// MSB-first bit reader over a byte buffer (host only; used by the index
// builders and the self-checks of the encoder).
class BitReader {
 public:
  BitReader(const uint8_t* data, uint64_t nbytes) : d_(data), n_(nbytes) {}
  void seek(uint64_t bitpos) { pos_ = bitpos; }
  uint64_t pos() const { return pos_; }
  int bit() {
    if ((pos_ >> 3) >= n_) throw std::out_of_range("bit stream exhausted");
    const int b = (d_[pos_ >> 3] >> (7 - (pos_ & 7))) & 1;
    ++pos_;
    return b;
  }
  uint64_t get(int len) {
    uint64_t v = 0;
    for (int i = 0; i < len; ++i) v = (v << 1) | (uint64_t)bit();
    return v;
  }
  uint64_t unary() {
    uint64_t z = 0;
    while (bit() == 0) ++z;
    return z;
  }
  uint64_t gamma() {
    const int m = (int)unary();
    return ((1ull << m) | get(m)) - 1;
  }
  uint64_t zeta(int k) {
    const int h = (int)unary();
    const uint64_t left = 1ull << (h * k);
    const uint64_t m = get(h * k + k - 1);
    if (m < left) return m + left - 1;
    return (m << 1) + (uint64_t)bit() - 1;
  }

 private:
  const uint8_t* d_;
  uint64_t n_;
  uint64_t pos_ = 0;
};

}  // namespace tgaf
