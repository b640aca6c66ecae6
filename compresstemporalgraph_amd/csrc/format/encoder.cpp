// Format kit: BV multigraph encoder, timestamp stream writer, GPU Elias-Fano
// index writer, synthetic generators and raw .gz I/O.
//
// The record layout written here is the one the reference's Java writer emits
// (BVMultiGraph.java:1640-1852 diffComp, :1855-1990 CompressionThread.call)
// and the one its GPU reader walks (gpu_cuda/shared/bfs.cu:1485-1633):
//
//   gamma(d)                                      d = contacts of x
//   if d > 0:
//     unary(r)                                    0 <= r <= 7, reference x - r
//     if r > 0: gamma(bc), gamma(b0), gamma(b_i - 1) for i >= 1
//     if extras + multiples > 0:
//       gamma(mc), multiples (first gamma(int2nat(v0 - x)), then gamma(gap - 1)),
//                  each followed by gamma(len - 2)
//       gamma(ic), intervals (first gamma(int2nat(l0 - x)), then
//                  gamma(l_i - (l_{i-1} + len_{i-1}) - 1)), each gamma(len - 4)
//       residuals: zeta_3(int2nat(r0 - x)), then zeta_3(r_i - r_{i-1} - 1)
//
// The code below is written from that layout; helper names follow the
// Java writer where the step is the same (intervalize, reference search).
#include <zlib.h>

#include <algorithm>
#include <atomic>
#include <thread>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <numeric>
#include <random>
#include <string>
#include <vector>

#include "../../../include/tga_format.h"
#include "bitio.hpp"

namespace tgaf {

static thread_local std::string g_err;
static int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}

constexpr int kWindow = 7;        // BVMultiGraph.DEFAULT_WINDOW_SIZE
constexpr int kMaxRefCount = 3;   // BVMultiGraph.DEFAULT_MAX_REF_COUNT
constexpr int kMinInterval = 4;   // BVMultiGraph.DEFAULT_MIN_INTERVAL_LENGTH
constexpr int kMinMultiple = 2;   // BVMultiGraph.DEFAULT_MIN_MULTIPLE_LENGTH
constexpr int kResidualZeta = 3;  // BVMultiGraph.DEFAULT_ZETA_K
constexpr int kTimestampZeta = 2; // BaselineRunner.ZETA_K

// ---------------------------------------------------------------------------
// One successor record, planned against a candidate reference list.
struct Plan {
  std::vector<int> blocks;
  std::vector<int> extras;
  std::vector<int> ileft, ilen, resid;
};

struct Encoder {
  // Successor multisets in CSR form (sorted by destination).
  const int32_t* adj;
  const int64_t* start;  // n + 1
  int64_t n;

  // Per-vertex split of the multiset into multiples runs and singles.
  std::vector<int> mleft, mlen, singles;

  // Splits a sorted multiset into runs of equal values (length >= 2) and
  // singles (BVMultiGraph.intervalizeMultiples, BVMultiGraph.java:1400-1430).
  void split_multiples(const int32_t* v, int64_t len) {
    mleft.clear();
    mlen.clear();
    singles.clear();
    int64_t i = 0;
    while (i < len) {
      int64_t j = i + 1;
      while (j < len && v[j] == v[i]) ++j;
      if (j - i > 1) {
        mleft.push_back(v[i]);
        mlen.push_back((int)(j - i));
      } else {
        singles.push_back(v[i]);
      }
      i = j;
    }
  }

  // Maximal runs of consecutive integers of length >= 4 become intervals,
  // everything else residuals (BVMultiGraph.intervalize, :1375-1398).
  static void intervalize(const std::vector<int>& x, Plan& p) {
    p.ileft.clear();
    p.ilen.clear();
    p.resid.clear();
    const size_t vl = x.size();
    for (size_t i = 0; i < vl; ++i) {
      size_t j = 0;
      if (i + 1 < vl && x[i] + 1 == x[i + 1]) {
        do ++j; while (i + j + 1 < vl && x[i + j] + 1 == x[i + j + 1]);
        ++j;
        if (j >= (size_t)kMinInterval) {
          p.ileft.push_back(x[i]);
          p.ilen.push_back((int)j);
          i += j - 1;
        }
      }
      if (j < (size_t)kMinInterval) p.resid.push_back(x[i]);
    }
  }

  // Copy/skip blocks of the singles against the reference multiset; whatever
  // is not copied goes to the extras (the block loop of diffComp,
  // BVMultiGraph.java:1666-1720).
  void plan_blocks(const int32_t* ref, int64_t reflen, Plan& p) const {
    p.blocks.clear();
    p.extras.clear();
    const int64_t cl = (int64_t)singles.size();
    int64_t j = 0, k = 0;
    int cur = 0;
    bool copying = true;
    while (j < cl && k < reflen) {
      const int cj = singles[j];
      const int rk = ref[k];
      if (copying) {
        if (cj > rk) {
          p.blocks.push_back(cur);
          copying = false;
          cur = 0;
        } else if (cj < rk) {
          p.extras.push_back(cj);
          ++j;
        } else {
          ++j;
          ++k;
          ++cur;
        }
      } else {
        if (cj < rk) {
          p.extras.push_back(cj);
          ++j;
        } else if (cj > rk) {
          ++k;
          ++cur;
        } else {
          p.blocks.push_back(cur);
          copying = true;
          cur = 0;
        }
      }
    }
    if (copying && k < reflen) p.blocks.push_back(cur);
    while (j < cl) p.extras.push_back(singles[j++]);
  }

  // Bits the record body takes for reference distance r (outdegree excluded).
  uint64_t cost(int64_t x, int r, const Plan& p) const {
    uint64_t bits = unary_len(r);
    if (r != 0) {
      bits += gamma_len(p.blocks.size());
      for (size_t i = 0; i < p.blocks.size(); ++i) bits += gamma_len(i == 0 ? p.blocks[0] : p.blocks[i] - 1);
    }
    if (!p.extras.empty() || !mleft.empty()) {
      bits += gamma_len(mleft.size());
      int64_t prev = 0;
      for (size_t i = 0; i < mleft.size(); ++i) {
        bits += (i == 0) ? gamma_len(int2nat((int64_t)mleft[0] - x)) : gamma_len(mleft[i] - prev - 1);
        prev = mleft[i];
        bits += gamma_len(mlen[i] - kMinMultiple);
      }
      bits += gamma_len(p.ileft.size());
      for (size_t i = 0; i < p.ileft.size(); ++i) {
        bits += (i == 0) ? gamma_len(int2nat((int64_t)p.ileft[0] - x)) : gamma_len(p.ileft[i] - prev - 1);
        prev = (int64_t)p.ileft[i] + p.ilen[i];
        bits += gamma_len(p.ilen[i] - kMinInterval);
      }
      if (!p.resid.empty()) {
        bits += zeta_len(int2nat((int64_t)p.resid[0] - x), kResidualZeta);
        for (size_t i = 1; i < p.resid.size(); ++i) bits += zeta_len(p.resid[i] - p.resid[i - 1] - 1, kResidualZeta);
      }
    }
    return bits;
  }

  void write(BitWriter& w, int64_t x, int r, const Plan& p) const {
    w.unary(r);
    if (r != 0) {
      w.gamma(p.blocks.size());
      for (size_t i = 0; i < p.blocks.size(); ++i) w.gamma(i == 0 ? p.blocks[0] : p.blocks[i] - 1);
    }
    if (!p.extras.empty() || !mleft.empty()) {
      w.gamma(mleft.size());
      int64_t prev = 0;
      for (size_t i = 0; i < mleft.size(); ++i) {
        w.gamma(i == 0 ? int2nat((int64_t)mleft[0] - x) : (uint64_t)(mleft[i] - prev - 1));
        prev = mleft[i];
        w.gamma(mlen[i] - kMinMultiple);
      }
      w.gamma(p.ileft.size());
      for (size_t i = 0; i < p.ileft.size(); ++i) {
        w.gamma(i == 0 ? int2nat((int64_t)p.ileft[0] - x) : (uint64_t)(p.ileft[i] - prev - 1));
        prev = (int64_t)p.ileft[i] + p.ilen[i];
        w.gamma(p.ilen[i] - kMinInterval);
      }
      if (!p.resid.empty()) {
        w.zeta(int2nat((int64_t)p.resid[0] - x), kResidualZeta);
        for (size_t i = 1; i < p.resid.size(); ++i) w.zeta(p.resid[i] - p.resid[i - 1] - 1, kResidualZeta);
      }
    }
  }
};

static bool write_file(const std::string& path, const void* data, size_t len) {
  FILE* f = std::fopen(path.c_str(), "wb");
  if (!f) return false;
  const bool ok = std::fwrite(data, 1, len, f) == len;
  return std::fclose(f) == 0 && ok;
}

static bool read_file(const std::string& path, std::vector<uint8_t>& out) {
  FILE* f = std::fopen(path.c_str(), "rb");
  if (!f) return false;
  std::fseek(f, 0, SEEK_END);
  const long len = std::ftell(f);
  std::rewind(f);
  out.resize((size_t)std::max(0L, len));
  const bool ok = len >= 0 && std::fread(out.data(), 1, out.size(), f) == out.size();
  std::fclose(f);
  return ok;
}

// Elias-Fano lower-bit width the reference computes at load time from the
// stream's file size: (int)log2(bytes * 8 / MAX_LABEL) (bfs.cu:1785-1786),
// evaluated exactly as the most significant bit of the integer quotient.
static int ef_low_width(uint64_t file_bytes, int64_t max_label, std::string& why) {
  if (max_label < 1) {
    why = "MAX_LABEL must be >= 1 (a graph needs at least two vertex ids)";
    return -1;
  }
  const uint64_t q = file_bytes * 8 / (uint64_t)max_label;
  if (q == 0) {
    why = "stream has fewer bits than MAX_LABEL: log2(0) in the index width";
    return -1;
  }
  return msb64(q);
}

// Writes value_i = offsets[i] (i = 0..count-1) as the GPU Elias-Fano index:
// lower = low l bits of each value, MSB-first, back to back; upper = unary
// high parts, bit (value_i >> l) + i set, LSB-first within bytes.  Both files
// are zero padded to a multiple of 8 bytes (the reader loads 64-bit words).
static bool write_ef_index(const std::string& lower_path, const std::string& upper_path,
                           const std::vector<uint64_t>& offsets, int l) {
  BitWriter lw;
  for (uint64_t v : offsets) lw.put(v, l);
  std::vector<uint8_t> lower = lw.finish();
  lower.resize((lower.size() + 7) / 8 * 8 + 8, 0);
  const uint64_t count = offsets.size();
  const uint64_t upper_bits = (offsets.empty() ? 0 : (offsets.back() >> l)) + count;
  std::vector<uint8_t> upper((upper_bits + 7) / 8, 0);
  upper.resize((upper.size() + 7) / 8 * 8 + 8, 0);
  for (uint64_t i = 0; i < count; ++i) {
    const uint64_t p = (offsets[i] >> l) + i;
    upper[p >> 3] |= (uint8_t)(1u << (p & 7));
  }
  return write_file(lower_path, lower.data(), lower.size()) && write_file(upper_path, upper.data(), upper.size());
}

static std::string join(const char* dir, const std::string& file) {
  std::string d(dir);
  while (!d.empty() && d.back() == '/') d.pop_back();
  return d + "/" + file;
}

static int write_indexes(const char* dir, const std::vector<uint64_t>& goff, uint64_t graph_bytes,
                         const std::vector<uint64_t>& toff, uint64_t ts_bytes, int64_t n, tgaf_stats* st) {
  std::string why;
  const int lg = ef_low_width(graph_bytes, n - 1, why);
  if (lg < 0) return fail(TGAF_ERANGE, "graph index: " + why);
  const int lt = ef_low_width(ts_bytes, n - 1, why);
  if (lt < 0) return fail(TGAF_ERANGE, "timestamp index: " + why);
  if (!write_ef_index(join(dir, "graph_lower_bits.bin"), join(dir, "graph_upper_bits.bin"), goff, lg) ||
      !write_ef_index(join(dir, "timestamp_lower_bits.bin"), join(dir, "timestamp_upper_bits.bin"), toff, lt))
    return fail(TGAF_EIO, std::string("cannot write index files in ") + dir);
  if (st) {
    st->l_graph = lg;
    st->l_timestamps = lt;
  }
  return TGAF_OK;
}

}  // namespace tgaf

using namespace tgaf;

// Progress lines on stderr for long generations and encodings.
static bool progress_on() {
  const char* e = std::getenv("TGAF_PROGRESS");
  return e && *e && *e != '0';
}

extern "C" const char* tgaf_last_error(void) { return g_err.c_str(); }

static int write_properties(const char* dir, const std::string& base, int64_t n, int64_t m, uint64_t graph_bits,
                            const tgaf_stats& s, tgaf_stats* st);

// The encoder behind tgaf_encode and tgaf_encode_graph: with t, every file;
// without (t == nullptr), the graph stream and .properties only, the record
// bit offsets going to goff_out (n + 1 entries).
static int encode_impl(const int32_t* src, const int32_t* dst, const int64_t* t, int64_t m, int64_t n_min,
                       const char* dir, const char* name, tgaf_stats* st, uint64_t* goff_out) {
  if (!src || !dst || m <= 0 || !dir || !name || (!t && !goff_out)) return fail(TGAF_EINVAL, "tgaf_encode: bad arguments");
  int64_t n = n_min;
  int64_t tmin = INT64_MAX;
  for (int64_t i = 0; i < m; ++i) {
    if (src[i] < 0 || dst[i] < 0) return fail(TGAF_EFORMAT, "negative node identifier at contact " + std::to_string(i));
    if (i > 0 && (src[i] < src[i - 1] || (src[i] == src[i - 1] && dst[i] < dst[i - 1])))
      return fail(TGAF_EFORMAT, "contacts must be sorted by (src, dst); violated at contact " + std::to_string(i));
    if (t && (t[i] < INT32_MIN || t[i] > INT32_MAX))
      return fail(TGAF_ERANGE, "timestamp outside int32 at contact " + std::to_string(i));
    n = std::max<int64_t>(n, (int64_t)std::max(src[i], dst[i]) + 1);
    if (t) tmin = std::min(tmin, t[i]);
  }
  if (n < 2) return fail(TGAF_ERANGE, "a graph needs at least two vertex ids");

  std::vector<int64_t> start((size_t)n + 1, 0);
  for (int64_t i = 0; i < m; ++i) start[(size_t)src[i] + 1]++;
  for (int64_t x = 0; x < n; ++x) start[(size_t)x + 1] += start[(size_t)x];

  // ---- .graph and .offsets (BVMultiGraph.CompressionThread.call) ----------
  Encoder enc{dst, start.data(), n};
  BitWriter gw, ow;
  std::vector<uint64_t> goff((size_t)n + 1);
  std::vector<int> refcount((size_t)n, 0), depth((size_t)n, 0);
  Plan cand, best;
  tgaf_stats s{};
  uint64_t prev_bits = 0;
  const bool talk = progress_on();
  for (int64_t x = 0; x < n; ++x) {
    if (talk && (x & ((1 << 22) - 1)) == 0 && x > 0)
      std::fprintf(stderr, "[encode] graph: %lld/%lld vertices\n", (long long)x, (long long)n);
    goff[(size_t)x] = gw.bits();
    ow.gamma(gw.bits() - prev_bits);
    prev_bits = gw.bits();
    const int64_t d = start[(size_t)x + 1] - start[(size_t)x];
    gw.gamma((uint64_t)d);
    if (d == 0) continue;
    const int32_t* list = dst + start[(size_t)x];
    enc.split_multiples(list, d);
    uint64_t best_bits = UINT64_MAX;
    int best_r = 0;
    for (int r = 0; r <= kWindow; ++r) {
      const int64_t y = x - r;
      if (r > 0) {
        if (y < 0) break;
        const int64_t dy = start[(size_t)y + 1] - start[(size_t)y];
        if (dy == 0 || refcount[(size_t)y] >= kMaxRefCount) continue;
        enc.plan_blocks(dst + start[(size_t)y], dy, cand);
      } else {
        enc.plan_blocks(nullptr, 0, cand);
      }
      Encoder::intervalize(cand.extras, cand);
      const uint64_t bits = enc.cost(x, r, cand);
      if (bits < best_bits) {
        best_bits = bits;
        best_r = r;
        std::swap(best, cand);
      }
    }
    refcount[(size_t)x] = best_r == 0 ? 0 : refcount[(size_t)(x - best_r)] + 1;
    depth[(size_t)x] = refcount[(size_t)x];
    enc.write(gw, x, best_r, best);
    // statistics
    int64_t copied = 0;
    for (int64_t i : enc.mlen) s.multiple_arcs += i;
    for (int i : best.ilen) s.intervalised_arcs += i;
    s.residual_arcs += (int64_t)best.resid.size();
    copied = (int64_t)enc.singles.size() - (int64_t)best.extras.size();
    s.copied_arcs += copied;
    if (best_r > 0) s.records_with_ref++;
    s.max_ref_depth = std::max(s.max_ref_depth, refcount[(size_t)x]);
  }
  goff[(size_t)n] = gw.bits();
  ow.gamma(gw.bits() - prev_bits);
  const uint64_t graph_bits = gw.bits();
  std::vector<uint8_t>& graph = gw.finish();
  std::vector<uint8_t>& offs = ow.finish();

  const std::string base(name);
  s.nodes = n;
  s.arcs = m;
  s.graph_bits = graph_bits;
  if (!t) {
    if (!write_file(join(dir, base + ".graph"), graph.data(), graph.size()))
      return fail(TGAF_EIO, std::string("cannot write dataset files in ") + dir);
    std::copy(goff.begin(), goff.end(), goff_out);
    return write_properties(dir, base, n, m, graph_bits, s, st);
  }

  // ---- .timestamps (EvolvingMultiGraph.storeTimestampsAndIndex) -----------
  BitWriter tw;
  tw.put((uint64_t)tmin, 64);
  std::vector<uint64_t> toff((size_t)n + 1);
  for (int64_t x = 0; x < n; ++x) {
    if (talk && (x & ((1 << 22) - 1)) == 0 && x > 0)
      std::fprintf(stderr, "[encode] timestamps: %lld/%lld vertices\n", (long long)x, (long long)n);
    toff[(size_t)x] = tw.bits();
    int64_t prev = tmin;
    for (int64_t i = start[(size_t)x]; i < start[(size_t)x + 1]; ++i) {
      tw.zeta(int2nat(t[i] - prev), kTimestampZeta);
      prev = t[i];
    }
  }
  toff[(size_t)n] = tw.bits();
  const uint64_t ts_bits = tw.bits();
  std::vector<uint8_t>& ts = tw.finish();

  if (!write_file(join(dir, base + ".graph"), graph.data(), graph.size()) ||
      !write_file(join(dir, base + ".offsets"), offs.data(), offs.size()) ||
      !write_file(join(dir, base + ".timestamps"), ts.data(), ts.size()))
    return fail(TGAF_EIO, std::string("cannot write dataset files in ") + dir);

  s.min_timestamp = tmin;
  s.timestamps_bits = ts_bits;
  const int rc = write_indexes(dir, goff, graph.size(), toff, ts.size(), n, &s);
  if (rc != TGAF_OK) return rc;
  return write_properties(dir, base, n, m, graph_bits, s, st);
}

// .properties (BVMultiGraph.java:2113-2150; keys the loader checks), then the
// statistics to the caller.
static int write_properties(const char* dir, const std::string& base, int64_t n, int64_t m, uint64_t graph_bits,
                            const tgaf_stats& s, tgaf_stats* st) {
  {
    char buf[2048];
    std::snprintf(buf, sizeof buf,
                  "#BVMultiGraph properties\n"
                  "nodes=%lld\narcs=%lld\nwindowsize=%d\nmaxrefcount=%d\nminintervallength=%d\nzetak=%d\n"
                  "compressionflags=\ncopiedarcs=%lld\nintervalisedarcs=%lld\nresidualarcs=%lld\n"
                  "bitsperlink=%.3f\nbitspernode=%.3f\n"
                  "graphclass=gr.uoa.di.networkanalysis.BVMultiGraph\nversion=0\n",
                  (long long)n, (long long)m, kWindow, kMaxRefCount, kMinInterval, kResidualZeta,
                  (long long)s.copied_arcs, (long long)s.intervalised_arcs, (long long)s.residual_arcs,
                  (double)graph_bits / (double)m, (double)graph_bits / (double)n);
    const std::string props(buf);
    if (!write_file(join(dir, base + ".properties"), props.data(), props.size()))
      return fail(TGAF_EIO, "cannot write .properties");
  }
  if (st) *st = s;
  return TGAF_OK;
}

extern "C" int tgaf_encode(const int32_t* src, const int32_t* dst, const int64_t* t, int64_t m, int64_t n_min,
                           const char* dir, const char* name, tgaf_stats* st) {
  if (!t) return fail(TGAF_EINVAL, "tgaf_encode: bad arguments");
  return encode_impl(src, dst, t, m, n_min, dir, name, st, nullptr);
}

extern "C" int tgaf_encode_graph(const int32_t* src, const int32_t* dst, int64_t m, int64_t n_min, const char* dir,
                                 const char* name, uint64_t* goff, tgaf_stats* st) {
  return encode_impl(src, dst, nullptr, m, n_min, dir, name, st, goff);
}

extern "C" int tgaf_build_gpu_index(const char* dir, const char* name, tgaf_stats* st) {
  if (!dir || !name) return fail(TGAF_EINVAL, "tgaf_build_gpu_index: bad arguments");
  const std::string base(name);
  std::vector<uint8_t> graph, offs, ts;
  if (!read_file(join(dir, base + ".graph"), graph) || !read_file(join(dir, base + ".offsets"), offs) ||
      !read_file(join(dir, base + ".timestamps"), ts))
    return fail(TGAF_EIO, std::string("cannot read .graph/.offsets/.timestamps in ") + dir);
  // Node count from .properties.
  int64_t n = -1;
  {
    std::ifstream in(join(dir, base + ".properties"));
    std::string line;
    while (std::getline(in, line))
      if (line.rfind("nodes=", 0) == 0) n = std::stoll(line.substr(6));
  }
  if (n < 2) return fail(TGAF_EFORMAT, "missing or invalid nodes= in .properties");
  std::vector<uint64_t> goff((size_t)n + 1), toff((size_t)n + 1);
  try {
    BitReader orr(offs.data(), offs.size());
    uint64_t pos = 0;
    for (int64_t x = 0; x <= n; ++x) {
      pos += orr.gamma();
      goff[(size_t)x] = pos;
    }
    // Timestamp offsets: each record holds outdegree-many zeta_2 codes.
    BitReader gr(graph.data(), graph.size()), tr(ts.data(), ts.size());
    tr.seek(64);
    for (int64_t x = 0; x < n; ++x) {
      toff[(size_t)x] = tr.pos();
      gr.seek(goff[(size_t)x]);
      const uint64_t d = gr.gamma();
      for (uint64_t i = 0; i < d; ++i) tr.zeta(kTimestampZeta);
    }
    toff[(size_t)n] = tr.pos();
  } catch (const std::exception& e) {
    return fail(TGAF_EFORMAT, std::string("truncated stream while indexing: ") + e.what());
  }
  tgaf_stats s{};
  s.nodes = n;
  const int rc = write_indexes(dir, goff, graph.size(), toff, ts.size(), n, &s);
  if (rc == TGAF_OK && st) *st = s;
  return rc;
}

// ---------------------------------------------------------------------------
// Generators

// Sorts contacts by (src, dst, t): counting sort on src, then each source's
// bucket by (dst, t), buckets spread over threads.
static void sort_contacts(int64_t m, int32_t* src, int32_t* dst, int64_t* t) {
  int32_t maxs = 0;
  for (int64_t i = 0; i < m; ++i) maxs = std::max(maxs, src[i]);
  std::vector<int64_t> start((size_t)maxs + 2, 0);
  for (int64_t i = 0; i < m; ++i) start[(size_t)src[i] + 1]++;
  for (size_t x = 0; x + 1 < start.size(); ++x) start[x + 1] += start[x];
  struct DT {
    int32_t d;
    int64_t t;
  };
  std::vector<DT> buf((size_t)m);
  {
    std::vector<int64_t> fill(start.begin(), start.end() - 1);
    for (int64_t i = 0; i < m; ++i) buf[(size_t)fill[(size_t)src[i]]++] = DT{dst[i], t[i]};
  }
  const unsigned nth = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
  std::atomic<int64_t> next{0};
  auto work = [&]() {
    for (;;) {
      const int64_t x0 = next.fetch_add(4096);
      if (x0 > maxs) break;
      const int64_t x1 = std::min<int64_t>(x0 + 4096, (int64_t)maxs + 1);
      for (int64_t x = x0; x < x1; ++x)
        std::sort(buf.begin() + start[(size_t)x], buf.begin() + start[(size_t)x + 1],
                  [](const DT& p, const DT& q) { return p.d != q.d ? p.d < q.d : p.t < q.t; });
    }
  };
  std::vector<std::thread> pool;
  for (unsigned i = 0; i < nth; ++i) pool.emplace_back(work);
  for (auto& th : pool) th.join();
  for (int64_t x = 0; x <= maxs; ++x)
    for (int64_t i = start[(size_t)x]; i < start[(size_t)x + 1]; ++i) {
      src[i] = (int32_t)x;
      dst[i] = buf[(size_t)i].d;
      t[i] = buf[(size_t)i].t;
    }
}

static uint64_t splitmix64(uint64_t x) {
  x += 0x9E3779B97F4A7C15ULL;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ULL;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBULL;
  return x ^ (x >> 31);
}

extern "C" int tgaf_gen_rmat(int32_t scale, double a, double b, double c, double p_geo, int64_t t_span,
                             uint64_t seed, int64_t m, int32_t* src, int32_t* dst, int64_t* t) {
  if (scale < 1 || scale > 30 || m <= 0 || !src || !dst || !t || p_geo <= 0 || p_geo > 1 || t_span <= 0 ||
      a < 0 || b < 0 || c < 0 || a + b + c > 1)
    return fail(TGAF_EINVAL, "tgaf_gen_rmat: bad arguments");
  std::mt19937_64 rng(seed);
  std::uniform_real_distribution<double> U(0.0, 1.0);
  std::geometric_distribution<int64_t> G(p_geo);  // failures before success: +1 gives support 1..
  std::uniform_int_distribution<int64_t> T(0, t_span - 1);
  int64_t k = 0;
  while (k < m) {
    int64_t u = 0, v = 0;
    for (int lvl = 0; lvl < scale; ++lvl) {
      const double r = U(rng);
      const int qu = (r >= a + b) ? 1 : 0;                          // lower half for c, d
      const int qv = (r >= a && r < a + b) || (r >= a + b + c) ? 1 : 0;  // right half for b, d
      u = (u << 1) | qu;
      v = (v << 1) | qv;
    }
    int64_t cnt = G(rng) + 1;
    if (cnt > m - k) cnt = m - k;
    for (int64_t i = 0; i < cnt; ++i, ++k) {
      src[k] = (int32_t)u;
      dst[k] = (int32_t)v;
      t[k] = T(rng);
    }
  }
  sort_contacts(m, src, dst, t);
  return TGAF_OK;
}

// Alias table (Vose) over the power-law weights w_i = (i+1)^(-1/(exponent-1)).
static void power_law_alias(int64_t n, double exponent, std::vector<double>& prob, std::vector<int32_t>& alias) {
  const double beta = 1.0 / (exponent - 1.0);
  prob.assign((size_t)n, 0.0);
  alias.assign((size_t)n, 0);
  {
    double total = 0.0;
    for (int64_t i = 0; i < n; ++i) total += (prob[(size_t)i] = std::pow((double)(i + 1), -beta));
    std::vector<int32_t> small, large;
    for (int64_t i = 0; i < n; ++i) {
      prob[(size_t)i] *= (double)n / total;
      (prob[(size_t)i] < 1.0 ? small : large).push_back((int32_t)i);
    }
    while (!small.empty() && !large.empty()) {
      const int32_t s = small.back(), l = large.back();
      small.pop_back();
      alias[(size_t)s] = l;
      prob[(size_t)l] -= 1.0 - prob[(size_t)s];
      if (prob[(size_t)l] < 1.0) {
        large.pop_back();
        small.push_back(l);
      }
    }
    for (int32_t i : large) prob[(size_t)i] = 1.0;
    for (int32_t i : small) prob[(size_t)i] = 1.0;
  }
}

extern "C" int tgaf_gen_chung_lu(int64_t n, double exponent, int64_t t_span, int64_t period, uint64_t seed,
                                 int64_t m, int32_t* src, int32_t* dst, int64_t* t) {
  if (n < 2 || n > INT32_MAX || exponent <= 1.0 || t_span <= 0 || period <= 0 || m <= 0 || !src || !dst || !t)
    return fail(TGAF_EINVAL, "tgaf_gen_chung_lu: bad arguments");
  std::vector<double> prob;
  std::vector<int32_t> alias;
  power_law_alias(n, exponent, prob, alias);
  // Fixed chunking with one generator per chunk: the output does not depend
  // on the number of threads.
  constexpr int64_t kChunks = 64;
  const double two_pi = 6.283185307179586;
  std::atomic<int64_t> next{0};
  auto work = [&]() {
    for (;;) {
      const int64_t c = next.fetch_add(1);
      if (c >= kChunks) break;
      std::mt19937_64 rng(splitmix64(seed * 0x100000001B3ULL + (uint64_t)c));
      std::uniform_real_distribution<double> U(0.0, 1.0);
      auto draw = [&]() -> int32_t {
        const double u = U(rng) * (double)n;
        int64_t i = (int64_t)u;
        if (i >= n) i = n - 1;
        return (U(rng) < prob[(size_t)i]) ? (int32_t)i : alias[(size_t)i];
      };
      const int64_t k0 = m * c / kChunks, k1 = m * (c + 1) / kChunks;
      for (int64_t k = k0; k < k1; ++k) {
        src[k] = draw();
        dst[k] = draw();
        // Rejection sampling of the diurnal intensity 1 + 0.5 sin(2 pi t / P).
        for (;;) {
          const int64_t cand = std::min<int64_t>((int64_t)(U(rng) * (double)t_span), t_span - 1);
          const double lam = 1.0 + 0.5 * std::sin(two_pi * (double)cand / (double)period);
          if (U(rng) * 1.5 < lam) {
            t[k] = cand;
            break;
          }
        }
      }
    }
  };
  const unsigned nth = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
  std::vector<std::thread> pool;
  for (unsigned i = 0; i < nth; ++i) pool.emplace_back(work);
  for (auto& th : pool) th.join();
  sort_contacts(m, src, dst, t);
  return TGAF_OK;
}

extern "C" int tgaf_gen_netflow(int64_t n, double exponent, double count_min, double gap_min, int64_t t_span,
                                uint64_t seed, int64_t m, int32_t* src, int32_t* dst, int64_t* t) {
  if (n < 2 || n > INT32_MAX || exponent <= 1.0 || count_min < 1.0 || gap_min <= 0.0 || t_span <= 0 ||
      t_span > ((int64_t)1 << 31) || m <= 0 || !src || !dst || !t)
    return fail(TGAF_EINVAL, "tgaf_gen_netflow: bad arguments");
  const bool talk = progress_on();
  std::vector<double> prob;
  std::vector<int32_t> alias;
  power_law_alias(n, exponent, prob, alias);
  // Pairs, drawn serially until their contact counts reach m (the last one
  // cut to fit): both ends from the power-law weights, the count a Pareto
  // (alpha 2) variate round(count_min * U^(-1/2)), mean about 2 count_min.
  struct Pair {
    uint64_t key;  // src << 32 | dst
    int64_t cnt;
  };
  std::vector<Pair> pairs;
  pairs.reserve((size_t)(m / (2.0 * count_min) * 1.1) + 16);
  {
    std::mt19937_64 rng(splitmix64(seed));
    std::uniform_real_distribution<double> U(0.0, 1.0);
    auto draw = [&]() -> uint32_t {
      const double u = U(rng) * (double)n;
      int64_t i = (int64_t)u;
      if (i >= n) i = n - 1;
      return (uint32_t)((U(rng) < prob[(size_t)i]) ? (int32_t)i : alias[(size_t)i]);
    };
    int64_t total = 0;
    while (total < m) {
      const uint32_t a = draw(), b = draw();
      int64_t c = std::max<int64_t>(1, std::llround(count_min / std::sqrt(1.0 - U(rng))));
      if (c > m - total) c = m - total;
      pairs.push_back(Pair{(uint64_t)a << 32 | b, c});
      total += c;
    }
  }
  std::sort(pairs.begin(), pairs.end(), [](const Pair& x, const Pair& y) { return x.key < y.key; });
  size_t np = 0;  // merge repeated pairs
  for (size_t i = 0; i < pairs.size(); ++i) {
    if (np > 0 && pairs[np - 1].key == pairs[i].key)
      pairs[np - 1].cnt += pairs[i].cnt;
    else
      pairs[np++] = pairs[i];
  }
  pairs.resize(np);
  std::vector<int64_t> off(np + 1, 0);
  for (size_t i = 0; i < np; ++i) off[i + 1] = off[i] + pairs[i].cnt;
  if (talk) std::fprintf(stderr, "[gen] %zu distinct pairs, %lld contacts\n", np, (long long)off[np]);
  // Contacts of a pair: a burst from a uniform start with Pareto (alpha 1.5)
  // gaps of at least gap_min, folded into [0, t_span) and sorted.  Fixed
  // chunks with one generator each, so the output does not depend on the
  // number of threads.
  constexpr int64_t kChunks = 256;
  std::atomic<int64_t> next{0};
  auto work = [&]() {
    for (;;) {
      const int64_t c = next.fetch_add(1);
      if (c >= kChunks) break;
      std::mt19937_64 rng(splitmix64(seed * 0x100000001B3ULL + 0x5bd1e995ULL + (uint64_t)c));
      std::uniform_real_distribution<double> U(0.0, 1.0);
      const size_t p0 = np * (size_t)c / kChunks, p1 = np * (size_t)(c + 1) / kChunks;
      for (size_t p = p0; p < p1; ++p) {
        const int32_t a = (int32_t)(pairs[p].key >> 32), b = (int32_t)(pairs[p].key & 0xffffffffu);
        int64_t tt = std::min<int64_t>((int64_t)(U(rng) * (double)t_span), t_span - 1);
        bool wrapped = false;
        for (int64_t k = off[p]; k < off[p + 1]; ++k) {
          src[k] = a;
          dst[k] = b;
          t[k] = tt;
          const double gap = gap_min * std::pow(1.0 - U(rng), -1.0 / 1.5);
          tt += (int64_t)std::min(gap, (double)t_span);
          if (tt >= t_span) {
            tt %= t_span;
            wrapped = true;
          }
        }
        if (wrapped) std::sort(t + off[p], t + off[p + 1]);
      }
      if (talk && (c & 31) == 31) std::fprintf(stderr, "[gen] contact chunks %lld/%lld\n", (long long)c + 1, (long long)kChunks);
    }
  };
  const unsigned nth = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
  std::vector<std::thread> pool;
  for (unsigned i = 0; i < nth; ++i) pool.emplace_back(work);
  for (auto& th : pool) th.join();
  return TGAF_OK;
}

// The neighbor-access checksum of a raw contact list: the wrapping 64-bit sum
// over its contacts of h = (t ^ r * 0x9E3779B1 ^ x * 0x85EBCA77) * 0x7FEB352D
// (uint32 arithmetic; the sum the decoding pass returns).
extern "C" int tgaf_contact_checksum(const int32_t* src, const int32_t* dst, const int64_t* t, int64_t m,
                                     uint64_t* out) {
  if (m < 0 || (m > 0 && (!src || !dst || !t)) || !out) return fail(TGAF_EINVAL, "tgaf_contact_checksum: bad arguments");
  constexpr int64_t kChunks = 64;
  std::vector<uint64_t> part(kChunks, 0);
  std::atomic<int64_t> next{0};
  auto term = [](uint32_t x, uint32_t r, uint32_t t) -> uint64_t {
    return (uint64_t)((t ^ (r * 0x9E3779B1u) ^ (x * 0x85EBCA77u)) * 0x7FEB352Du);
  };
  auto work = [&]() {
    for (;;) {
      const int64_t c = next.fetch_add(1);
      if (c >= kChunks) break;
      uint64_t sum = 0;
      for (int64_t i = m * c / kChunks; i < m * (c + 1) / kChunks; ++i)
        sum += term((uint32_t)src[i], (uint32_t)dst[i], (uint32_t)(int32_t)t[i]);
      part[(size_t)c] = sum;
    }
  };
  const unsigned nth = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
  std::vector<std::thread> pool;
  for (unsigned i = 0; i < nth; ++i) pool.emplace_back(work);
  for (auto& th : pool) th.join();
  uint64_t sum = 0;
  for (uint64_t v : part) sum += v;
  *out = sum;
  return TGAF_OK;
}

// ---------------------------------------------------------------------------
// Raw text I/O

extern "C" int tgaf_write_raw_gz(const char* path, const int32_t* src, const int32_t* dst, const int64_t* t,
                                 int64_t m) {
  if (!path || m < 0 || (m > 0 && (!src || !dst || !t))) return fail(TGAF_EINVAL, "tgaf_write_raw_gz: bad arguments");
  gzFile f = gzopen(path, "wb6");
  if (!f) return fail(TGAF_EIO, std::string("cannot open ") + path);
  std::string chunk;
  chunk.reserve(1 << 20);
  char line[96];
  for (int64_t i = 0; i < m; ++i) {
    const int len = std::snprintf(line, sizeof line, "%d %d 1 %lld\n", src[i], dst[i], (long long)t[i]);
    chunk.append(line, (size_t)len);
    if (chunk.size() > (1u << 20) - 128) {
      if (gzwrite(f, chunk.data(), (unsigned)chunk.size()) != (int)chunk.size()) {
        gzclose(f);
        return fail(TGAF_EIO, "gzwrite failed");
      }
      chunk.clear();
    }
  }
  if (!chunk.empty() && gzwrite(f, chunk.data(), (unsigned)chunk.size()) != (int)chunk.size()) {
    gzclose(f);
    return fail(TGAF_EIO, "gzwrite failed");
  }
  return gzclose(f) == Z_OK ? TGAF_OK : fail(TGAF_EIO, "gzclose failed");
}

extern "C" int tgaf_read_raw_gz(const char* path, int headers, int64_t* m_out, int32_t* src, int32_t* dst,
                                int64_t* t) {
  if (!path || !m_out) return fail(TGAF_EINVAL, "tgaf_read_raw_gz: bad arguments");
  gzFile f = gzopen(path, "rb");
  if (!f) return fail(TGAF_EIO, std::string("cannot open ") + path);
  std::vector<char> buf(1 << 16);
  int64_t k = 0, lineno = 0;
  int32_t last_src = -1;
  bool skip_header = headers != 0;
  while (gzgets(f, buf.data(), (int)buf.size())) {
    ++lineno;
    if (skip_header) {
      skip_header = false;
      continue;
    }
    // Tokenise on whitespace; blank lines are skipped (EvolvingMultiGraph.java:236-244).
    long long tok[4];
    int nt = 0;
    char* p = buf.data();
    bool bad = false;
    while (*p && nt < 4) {
      while (*p == ' ' || *p == '\t' || *p == '\r' || *p == '\n') ++p;
      if (!*p) break;
      char* end = nullptr;
      tok[nt] = std::strtoll(p, &end, 10);
      if (end == p || (*end && *end != ' ' && *end != '\t' && *end != '\r' && *end != '\n')) {
        bad = true;
        break;
      }
      ++nt;
      p = end;
    }
    if (nt == 0 && !bad) continue;
    if (bad || nt < 4) {
      gzclose(f);
      return fail(TGAF_EFORMAT, "expected at least four numeric columns at input line " + std::to_string(lineno));
    }
    if (tok[0] < 0 || tok[1] < 0 || tok[0] > INT32_MAX || tok[1] > INT32_MAX) {
      gzclose(f);
      return fail(TGAF_EFORMAT, "node identifier out of range at input line " + std::to_string(lineno));
    }
    if ((int32_t)tok[0] < last_src) {
      gzclose(f);
      return fail(TGAF_EFORMAT, "source nodes must be sorted at input line " + std::to_string(lineno));
    }
    last_src = (int32_t)tok[0];
    if (src) {
      if (k >= *m_out) {
        gzclose(f);
        return fail(TGAF_EINVAL, "output arrays too small");
      }
      src[k] = (int32_t)tok[0];
      dst[k] = (int32_t)tok[1];
      t[k] = tok[3];
    }
    ++k;
  }
  gzclose(f);
  *m_out = k;
  return TGAF_OK;
}

// ---------------------------------------------------------------------------
// Decode tables and the Java priority scheme

extern "C" int tgaf_decode_table(int kind, int32_t* out) {
  if (!out || (kind != 0 && kind != 1)) return fail(TGAF_EINVAL, "tgaf_decode_table: bad arguments");
  for (uint32_t w = 0; w < 65536; ++w) {
    out[w] = 0;
    if (w == 0) continue;
    const int z = __builtin_clz(w) - 16;  // leading zeros inside the 16-bit window
    if (kind == 0) {
      const int len = 2 * z + 1;
      if (len > 16) continue;
      const uint32_t v = w >> (16 - len);  // z zeros then the m+1 value bits
      out[w] = (int32_t)(((uint32_t)len << 16) | (v - 1));
    } else {
      const int k = 3, h = z;
      const int nbits = h * k + k - 1;
      int len = h + 1 + nbits;
      if (len > 16) continue;
      const uint32_t mval = (w >> (16 - len)) & ((1u << nbits) - 1);
      const uint32_t left = 1u << (h * k);
      uint32_t val;
      if (mval < left) {
        val = mval + left - 1;
      } else {
        if (len + 1 > 16) continue;
        const uint32_t bit = (w >> (16 - len - 1)) & 1u;
        val = (mval << 1) + bit - 1;
        len += 1;
      }
      out[w] = (int32_t)(((uint32_t)len << 16) | val);
    }
  }
  return TGAF_OK;
}

extern "C" int tgaf_java_random_ints(int64_t seed, int32_t bound, int64_t count, int32_t* out) {
  if (bound <= 0 || count < 0 || (count > 0 && !out)) return fail(TGAF_EINVAL, "tgaf_java_random_ints: bad arguments");
  // java.util.Random: 48-bit linear congruential generator (Java SE API, class Random).
  const uint64_t mult = 0x5DEECE66DULL, addend = 0xBULL, mask = (1ULL << 48) - 1;
  uint64_t s = ((uint64_t)seed ^ mult) & mask;
  auto next31 = [&]() -> int32_t {
    s = (s * mult + addend) & mask;
    return (int32_t)(s >> (48 - 31));
  };
  for (int64_t i = 0; i < count; ++i) {
    int32_t r = next31();
    const int32_t mm = bound - 1;
    if ((bound & mm) == 0) {
      r = (int32_t)(((int64_t)bound * (int64_t)r) >> 31);
    } else {
      for (int32_t u = r; u - (r = u % bound) + mm < 0; u = next31()) {
      }
    }
    out[i] = r;
  }
  return TGAF_OK;
}

// ---------------------------------------------------------------------------
// Code writer for tests and tools: kind 0 unary, 1 gamma, k >= 2 zeta_k
// (k < 64), 64 a 64-bit big-endian integer.

extern "C" int tgaf_encode_codes(const int32_t* kinds, const uint64_t* values, int64_t n, uint8_t* out,
                                 int64_t cap, uint64_t* nbits) {
  if (n < 0 || (n > 0 && (!kinds || !values)) || !out || !nbits) return fail(TGAF_EINVAL, "tgaf_encode_codes: bad arguments");
  BitWriter w;
  for (int64_t i = 0; i < n; ++i) {
    const int k = kinds[i];
    if (k == 0) {
      if (values[i] > (1u << 20)) return fail(TGAF_ERANGE, "unary value too large");
      w.unary(values[i]);
    } else if (k == 1) {
      if (values[i] == UINT64_MAX) return fail(TGAF_ERANGE, "gamma value too large");
      w.gamma(values[i]);
    } else if (k >= 2 && k < 64) {
      if (values[i] >= (1ull << 62)) return fail(TGAF_ERANGE, "zeta value too large");
      w.zeta(values[i], k);
    } else if (k == 64) {
      w.put(values[i], 64);
    } else {
      return fail(TGAF_EINVAL, "unknown code kind");
    }
  }
  *nbits = w.bits();
  std::vector<uint8_t>& b = w.finish();
  if ((int64_t)b.size() > cap) return fail(TGAF_ERANGE, "output buffer too small");
  std::memcpy(out, b.data(), b.size());
  return TGAF_OK;
}
