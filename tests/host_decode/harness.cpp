// CPU build of the engine's device decoder (test infrastructure).
//
// Loads a dataset with the engine's own host loader (loader.hpp), then runs
// the very code the kernels run -- record headers and Elias-Fano offsets
// (k_headers), the checkpoint build of long records (ckpt_build_record), and
// every work item (run_item over whole records and checkpointed chunks, with
// the decoder depth the engine's launch groups use) -- sequentially on the
// host.  Built under address / undefined-behaviour sanitizers with pattern-
// initialised locals, and optimised with zero-initialised locals, a read of
// an uninitialised decoder field, an out-of-range private index or a read
// past a stream aborts or changes the output, instead of being masked by
// whatever a register held.
//
// Usage: harness <dataset dir> <basename> <out file> [flat chunk stride [burst [col]]]
// (col = 1: timestamps from the column copy the engine builds at open)
// Writes, per vertex, the decoded contact count (int64) and the sum of the
// neighbor-access checksum terms of its contacts (uint64), then the error
// word (int64).  Exit status 0 when the walk finished; the caller compares
// the per-vertex sums with the raw contact list.
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "decode.hpp"
#include "loader.hpp"

using namespace tga;

namespace {

int die(const char* msg, long long v = 0) {
  std::fprintf(stderr, "harness: %s %lld\n", msg, v);
  return 2;
}

// Per-item act: accumulates the checksum terms per vertex and checks what a
// walk promises -- contacts of an item in order, neighbors non-decreasing
// within a record, and the split-run links of chunks (a chunk that starts
// inside a run starts with that run's neighbor; one that ends inside a run
// ends with it).
struct Act {
  std::vector<int64_t>* cnt;
  std::vector<uint64_t>* sum;
  int64_t* bad;
  Split sp{-1, -1, -1, -1};
  int32_t k_expect = -1, prev_r = -1, last_r = -1, first_r = -1;
  bool chunked = false;
  bool begin(int32_t, bool ch, Split s) {
    sp = s;
    chunked = ch;
    k_expect = -1;
    prev_r = last_r = first_r = -1;
    return true;
  }
  bool contact(int32_t x, int32_t r, int32_t t, int32_t k) {
    if (k_expect >= 0 && k != k_expect) ++*bad;
    k_expect = k + 1;
    if (r < prev_r) ++*bad;
    if (first_r < 0) first_r = r;
    prev_r = last_r = r;
    (*cnt)[(size_t)x] += 1;
    (*sum)[(size_t)x] += contact_hash(x, r, t);
    return true;
  }
  void end(int32_t, bool) {
    if (!chunked) return;
    if (sp.cont_slot >= 0 && first_r != sp.cont_r) ++*bad;
    if (sp.tail_slot >= 0 && last_r != sp.tail_r) ++*bad;
  }
};

}  // namespace

int main(int argc, char** argv) {
  if (argc < 4 || argc > 7) return die("usage: harness <dir> <basename> <out> [flat stride [burst [col]]]");
  const std::string dir = argv[1], name = argv[2], base = dir + "/" + name;
  std::vector<uint8_t> graph, ts, glo, gup, tlo, tup;
  const std::pair<std::string, std::vector<uint8_t>*> files[] = {
      {base + ".graph", &graph},
      {base + ".timestamps", &ts},
      {dir + "/graph_lower_bits.bin", &glo},
      {dir + "/graph_upper_bits.bin", &gup},
      {dir + "/timestamp_lower_bits.bin", &tlo},
      {dir + "/timestamp_upper_bits.bin", &tup}};
  for (const auto& f : files)
    if (!read_file(f.first, *f.second)) return die("cannot read a dataset file");
  int64_t max_label = -1;
  {
    std::vector<uint8_t> props;
    if (!read_file(base + ".properties", props)) return die("cannot read .properties");
    const std::string p(props.begin(), props.end());
    const size_t at = p.find("\nnodes=") != std::string::npos ? p.find("\nnodes=") + 1 : p.find("nodes=");
    if (at == std::string::npos) return die("no nodes= in .properties");
    max_label = std::atoll(p.c_str() + at + 6) - 1;
  }
  if (max_label < 1) return die("bad MAX_LABEL", max_label);
  const int64_t n = max_label + 1;
  const uint64_t qg = (uint64_t)graph.size() * 8 / (uint64_t)max_label;
  const uint64_t qt = (uint64_t)ts.size() * 8 / (uint64_t)max_label;
  if (qg == 0 || qt == 0) return die("log2(0) index width");
  int64_t min_ts = 0;
  for (int i = 0; i < 8; ++i) min_ts = (int64_t)(((uint64_t)min_ts << 8) | ts[(size_t)i]);

  const std::vector<uint64_t> gw = be_words(graph), tw = be_words(ts), glw = be_words(glo), tlw = be_words(tlo);
  const std::vector<uint64_t> guw = le_words(gup), tuw = le_words(tup);
  uint64_t gones = 0, tones = 0;
  std::vector<uint64_t> gsamp, gspill, tsamp, tspill;
  build_select(guw, (gup.size() + 7) / 8, gones, gsamp, gspill);
  build_select(tuw, (tup.size() + 7) / 8, tones, tsamp, tspill);
  if (gones < (uint64_t)n || tones < (uint64_t)n) return die("index shorter than MAX_LABEL + 1");

  int32_t err = 0;
  Streams S{};
  S.g = gw.data();
  S.t = tw.data();
  S.g_bits = (uint64_t)graph.size() * 8;
  S.t_bits = (uint64_t)ts.size() * 8;
  S.g_last = (uint32_t)(gw.size() - 1);
  S.t_last = (uint32_t)(tw.size() - 1);
  S.gi = EFDir{glw.data(), guw.data(), gsamp.data(), gspill.data(), (gup.size() + 7) / 8, msb64_host(qg)};
  S.ti = EFDir{tlw.data(), tuw.data(), tsamp.data(), tspill.data(), (tup.size() + 7) / 8, msb64_host(qt)};
  S.n = (int32_t)n;
  S.min_ts = (int32_t)min_ts;
  S.err = &err;

  // Headers and the record start table (k_headers).
  std::vector<uint64_t> vpos(2 * (size_t)n);
  std::vector<int32_t> deg((size_t)n);
  std::vector<int8_t> depth((size_t)n, 0);
  const Rd g{S.g, S.g_bits, 0};
  for (int64_t x = 0; x < n; ++x) {
    vpos[2 * (size_t)x] = ef_get(S, S.gi, (uint64_t)x);
    vpos[2 * (size_t)x + 1] = ef_get(S, S.ti, (uint64_t)x);
    uint64_t pos = vpos[2 * (size_t)x];
    deg[(size_t)x] = (int32_t)rd_gamma(S, g, pos);
    const int r = deg[(size_t)x] > 0 ? (int)rd_unary(S, g, pos) : 0;
    if (r > 0) {
      if (r > kWindow || r > x) return die("reference outside the window at", x);
      const int dp = depth[(size_t)(x - r)] + 1;
      if (dp > kMaxDepth) return die("chain deeper than 3 at", x);
      depth[(size_t)x] = (int8_t)dp;
    }
  }
  S.vpos = vpos.data();
  if (err) return die("decode error in the headers", err);
  int64_t contacts = 0;
  for (int64_t x = 0; x < n; ++x) contacts += deg[(size_t)x];
  S.flat_stride = argc >= 5 && std::atoi(argv[4]) > 0 ? std::atoi(argv[4]) : flat_stride_for(contacts);
  S.burst = argc >= 6 ? std::atoi(argv[5]) : kBurstDefault;
  const bool col = argc >= 7 && std::atoi(argv[6]) != 0;

  // Checkpoints of long records, class by class (a record's references have
  // smaller depth, so their checkpoints exist when it is built).
  std::vector<int32_t> vinfo((size_t)n);
  std::vector<int64_t> slot((size_t)n, -1);
  int64_t slots = 0;
  for (int64_t x = 0; x < n; ++x) {
    vinfo[(size_t)x] = deg[(size_t)x] == 0 ? -1 - (kMaxDepth + 1) : -1 - depth[(size_t)x];
    const int c = depth[(size_t)x];
    if (deg[(size_t)x] <= chunk_stride(c, S.flat_stride)) continue;
    slot[(size_t)x] = slots;
    vinfo[(size_t)x] = (int32_t)slots;
    slots += (deg[(size_t)x] + chunk_stride(c, S.flat_stride) - 1) / chunk_stride(c, S.flat_stride);
  }
  std::vector<Ckpt> ck((size_t)std::max<int64_t>(slots, 1));
  std::memset((void*)ck.data(), 0, sizeof(Ckpt) * ck.size());  // as the engine clears the table
  S.ck = ck.data();
  S.vinfo = vinfo.data();
  for (int c = 0; c <= kMaxDepth; ++c)
    for (int64_t x = 0; x < n; ++x) {
      if (slot[(size_t)x] < 0 || depth[(size_t)x] != c) continue;
      int32_t nch = 0;
      Ckpt* out = ck.data() + slot[(size_t)x];
      if (c == 0) nch = ckpt_build_record<0>(S, (int32_t)x, slot[(size_t)x], out);
      if (c == 1) nch = ckpt_build_record<1>(S, (int32_t)x, slot[(size_t)x], out);
      if (c == 2) nch = ckpt_build_record<2>(S, (int32_t)x, slot[(size_t)x], out);
      if (c == 3) nch = ckpt_build_record<3>(S, (int32_t)x, slot[(size_t)x], out);
      if (nch != (deg[(size_t)x] + chunk_stride(c, S.flat_stride) - 1) / chunk_stride(c, S.flat_stride)) return die("checkpoint build at", x);
    }
  if (err) return die("decode error in the checkpoint build", err);

  // Every work item, with the decoder depth of its launch group.
  std::vector<int64_t> cnt((size_t)n, 0);
  std::vector<uint64_t> sum((size_t)n, 0);
  int64_t bad = 0;
  std::vector<int2> ring_store(kRing);
  std::vector<uint32_t> stage_store((kStageDw + 2) * 8);  // slot pitch 8 (the walk's minimum)
  const RunRing ring{ring_store.data(), 1};
  const TsStage stage{stage_store.data(), 8, 0, 0, 0, 0, 0};
  std::vector<int32_t> list, lgroup;  // items in vertex order (chunks in order) and their launch group
  for (int64_t x = 0; x < n; ++x) {
    if (deg[(size_t)x] == 0) continue;
    const int c = depth[(size_t)x];
    if (slot[(size_t)x] < 0) {
      list.push_back((int32_t)x);
      lgroup.push_back(c);
    } else {
      const int32_t nch = ck[(size_t)slot[(size_t)x]].nchunk;
      for (int32_t j = 0; j < nch; ++j) {
        list.push_back(-(int32_t)(slot[(size_t)x] + j) - 1);
        lgroup.push_back(c);
      }
    }
  }
  // With `col`, the timestamps come from a column copy of the list built as
  // the engine builds it at open (k_tcol_len, k_tcol_fill: groups of 64
  // positions, unit u of position 64q + l at goff[q] + 64u + l); units past
  // an item's codes hold all-ones words, so a walk that read them would
  // decode wrong times.
  std::vector<uint64_t> cw, goff, gcw, ggoff;
  int64_t start_words = 0, start_bad = 0;
  std::vector<uint32_t> nbits(list.size()), gbits(list.size(), 0);
  if (col) {
    // Graph columns of the whole records without references (k_gcol_len,
    // k_gcol_fill): ceil(bits / 64) + 5 words each, word j of position
    // 64q + l at ggoff[q] + 64j + l; other words all ones.
    const size_t gng = (list.size() + 63) / 64;
    ggoff.resize(gng + 1);
    uint64_t gat = 0;
    for (size_t i = 0; i < list.size(); ++i)
      if (list[i] >= 0 && lgroup[i] == 0) {
        const int32_t x = list[i];
        const uint64_t e = x + 1 < n ? vpos[2 * (size_t)x + 2] : S.g_bits;
        gbits[i] = (uint32_t)(e - vpos[2 * (size_t)x]);
      }
    for (size_t q = 0; q < gng; ++q) {
      uint32_t rows = 0;
      for (size_t i = 64 * q; i < std::min(list.size(), 64 * q + 64); ++i)
        if (gbits[i]) rows = std::max(rows, (gbits[i] + 63) / 64 + 6);
      ggoff[q] = gat;
      gat += 64ull * rows;
    }
    ggoff[gng] = gat;
    gcw.assign((size_t)gat, ~0ull);
    for (size_t i = 0; i < list.size(); ++i) {
      if (!gbits[i]) continue;
      const uint64_t b = vpos[2 * (size_t)list[i]];
      const size_t at = (size_t)(ggoff[i / 64] + (i % 64));
      gcw[at] = flat_start_word(S, b);
      if (gcw[at] != kNoStart) ++start_words;
      for (uint64_t j = 0; j < (gbits[i] + 63) / 64 + 5; ++j) {
        const uint64_t p = b + 64 * j, wi = p >> 6;
        const uint32_t sh = (uint32_t)p & 63u;
        const uint64_t a = wi <= S.g_last ? S.g[wi] : 0, c = wi + 1 <= S.g_last ? S.g[wi + 1] : 0;
        gcw[at + 64 * (j + 1)] = sh ? (a << sh) | (c >> (64u - sh)) : a;
      }
      // The start word must give the level state dec_init reaches.
      if (gcw[at] != kNoStart) {
        Dec<0> A, B;
        std::memset((void*)&A, 0, sizeof A);
        std::memset((void*)&B, 0, sizeof B);
        int32_t depth = 0;
        const int32_t da = dec_init<0, 0>(S, A, list[i], b, depth);
        Streams SC = S;
        SC.g = gcw.data() + at + 64;
        SC.g_sh = 6;
        SC.g_last = (gbits[i] + 63) / 64 + 4;
        SC.g_bits = 64ull * SC.g_last;
        const int32_t db = dec_init_flat<0>(SC, B, list[i], gcw[at]);
        const Level& la = A.lv[0];
        const Level& lb = B.lv[0];
        const bool same = da == db && la.hm == lb.hm && la.mrem == lb.mrem && la.mleft == lb.mleft &&
                          la.hi == lb.hi && la.irem == lb.irem && la.ileft == lb.ileft &&
                          (la.ileft == 0 || la.ipos - b == lb.ipos) && la.hr == lb.hr && la.rleft == lb.rleft &&
                          (la.rleft == 0 || A.rb0.pos() - b == B.rb0.pos()) &&
                          (la.mleft == 0 || A.mb0.pos() - b == B.mb0.pos());
        if (!same) ++start_bad;
      }
    }
    std::vector<uint64_t> tstart(list.size());
    const int64_t nck = (int64_t)slots;
    for (size_t i = 0; i < list.size(); ++i) {
      const int32_t code = list[i];
      uint64_t b, e;
      if (code >= 0) {
        b = vpos[2 * (size_t)code + 1];
        e = rec_tend(S, code);
      } else {
        const int64_t j = -(int64_t)code - 1;
        b = ck[(size_t)j].tpos;
        e = (j + 1 < nck && ck[(size_t)j + 1].x == ck[(size_t)j].x && ck[(size_t)j + 1].k0 == ck[(size_t)j].k1)
                ? ck[(size_t)j + 1].tpos
                : rec_tend(S, ck[(size_t)j].x);
      }
      tstart[i] = b;
      nbits[i] = (uint32_t)(e - b);
    }
    const size_t ng = (list.size() + 63) / 64;
    goff.resize(ng + 1);
    uint64_t at = 0;
    for (size_t q = 0; q < ng; ++q) {
      uint32_t rows = 0;
      for (size_t i = 64 * q; i < std::min(list.size(), 64 * q + 64); ++i) rows = std::max(rows, (nbits[i] + 127) / 128);
      goff[q] = at;
      at += 64ull * rows;
    }
    goff[ng] = at;
    cw.assign(2 * (size_t)at, ~0ull);
    for (size_t i = 0; i < list.size(); ++i)
      for (uint64_t u = 0; 128 * u < nbits[i]; ++u) {
        const size_t k = 2 * (size_t)(goff[i / 64] + 64 * u + (i % 64));
        cw[k] = peek64(S.t, tstart[i] + 128 * u);
        cw[k + 1] = peek64(S.t, tstart[i] + 128 * u + 64);
      }
  }
  for (size_t i = 0; i < list.size(); ++i) {
    Act a{&cnt, &sum, &bad};
    TsStage st = stage;
    if (col) {
      st.uw = cw.data() + 2 * (goff[i / 64] + (i % 64));
      st.ustride = 64;
      st.nbits = nbits[i];
      if (gbits[i]) {
        st.gw = gcw.data() + ggoff[i / 64] + (i % 64);
        st.gbits = gbits[i];
      }
    }
    const int c = lgroup[i];
    if (c == 0) run_item<0>(S, ck.data(), list[i], a, ring, st);
    else if (c == 1) run_item<1>(S, ck.data(), list[i], a, ring, st);
    else run_item<kMaxDepth>(S, ck.data(), list[i], a, ring, st);
  }
  if (bad) return die("walk order / split-run links broken in items:", bad);
  if (start_bad) return die("start words that do not give dec_init's state:", start_bad);
  if (col && start_words == 0) return die("no record got a start word");

  FILE* f = std::fopen(argv[3], "wb");
  if (!f) return die("cannot write the output");
  const int64_t e64 = err;
  bool ok = std::fwrite(cnt.data(), 8, cnt.size(), f) == cnt.size() &&
            std::fwrite(sum.data(), 8, sum.size(), f) == sum.size() && std::fwrite(&e64, 8, 1, f) == 1;
  ok = std::fclose(f) == 0 && ok;
  return ok ? 0 : die("write failed");
}
