// Host side of loading a dataset: the stream files as device word layouts
// (decode.hpp) and the sampled select directory over the Elias-Fano upper
// bits.  Shared by the engine's tga_open and the CPU decode harness
// (tests/host_decode), so both read the same layout.
#pragma once
#include <stdint.h>

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include "decode.hpp"

namespace tga {

inline bool read_file(const std::string& path, std::vector<uint8_t>& out) {
  FILE* f = std::fopen(path.c_str(), "rb");
  if (!f) return false;
  std::fseek(f, 0, SEEK_END);
  const long len = std::ftell(f);
  std::rewind(f);
  if (len < 0) {
    std::fclose(f);
    return false;
  }
  out.resize((size_t)len);
  const bool ok = std::fread(out.data(), 1, out.size(), f) == out.size();
  std::fclose(f);
  return ok;
}

// Bytes -> 64-bit words whose most significant byte is the first stream byte,
// plus kPadWords zero words (bit buffers read up to five words from any
// position inside the stream).
constexpr size_t kPadWords = 5;
// Large streams are converted by up to 16 threads, one slice of words each
// (a little-endian load of 8 bytes, byte-swapped, puts the first byte on top).
inline std::vector<uint64_t> be_words(const std::vector<uint8_t>& b) {
  const size_t nw = (b.size() + 7) / 8;
  std::vector<uint64_t> w(nw + kPadWords, 0);
  auto slice = [&](size_t lo, size_t hi) {
    const size_t bytes = std::min(b.size(), hi * 8) - lo * 8;
    std::memcpy(w.data() + lo, b.data() + lo * 8, bytes);
    for (size_t i = lo; i < hi; ++i) w[i] = __builtin_bswap64(w[i]);
  };
  const size_t nth = nw >= ((size_t)1 << 22) ? std::min<size_t>(16, std::max(1u, std::thread::hardware_concurrency())) : 1;
  if (nth == 1) {
    slice(0, nw);
  } else {
    std::vector<std::thread> pool;
    for (size_t k = 0; k < nth; ++k) pool.emplace_back(slice, nw * k / nth, nw * (k + 1) / nth);
    for (auto& th : pool) th.join();
  }
  return w;
}

inline std::vector<uint64_t> le_words(const std::vector<uint8_t>& b) {
  std::vector<uint64_t> w((b.size() + 7) / 8 + 2, 0);
  std::memcpy(w.data(), b.data(), b.size());
  return w;
}

inline int msb64_host(uint64_t v) { return 63 - __builtin_clzll(v); }

// Sampled select directory over the upper bits (see decode.hpp).
inline void build_select(const std::vector<uint64_t>& up, size_t words, uint64_t& ones, std::vector<uint64_t>& samp,
                  std::vector<uint64_t>& spill) {
  std::vector<uint64_t> block;
  ones = 0;
  samp.clear();
  spill.clear();
  auto flush = [&]() {
    if (block.empty()) return;
    if (block.back() - block.front() > kSpanBits) {
      samp.push_back((1ull << 63) | spill.size());
      for (size_t i = 0; i < 64; ++i) spill.push_back(i < block.size() ? block[i] : block.back());
    } else {
      samp.push_back(block.front());
    }
    block.clear();
  };
  for (size_t w = 0; w < words; ++w) {
    uint64_t word = up[w];
    while (word) {
      const int b = __builtin_ctzll(word);
      word &= word - 1;
      block.push_back(w * 64 + (uint64_t)b);
      ++ones;
      if (block.size() == 64) flush();
    }
  }
  flush();
  samp.push_back(0);
  spill.push_back(0);
}

}  // namespace tga
