// MI355X engine: loader, decode schedule, algorithm kernels and the C ABI of
// include/tga.h.
//
// Work decomposition.  Every algorithm is one or more passes over "work
// items".  An item is either a whole successor record (small vertices) or a
// chunk of a long record that starts from a saved decoder state (Ckpt,
// decode.hpp).  At open the loader
//   1. reads every record header on the GPU (outdegree d, reference r),
//   2. derives each record's reference-chain depth on the host,
//   3. decodes every record longer than its chunk stride once, sequentially,
//      saving the decoder state every stride contacts (runs of one neighbor
//      that cross a boundary are tracked as split runs, see Ckpt), and
//   4. builds three item lists: flat records (no reference: a one-level
//      decoder with few registers), depth-1 records (two levels) and deeper
//      records (four levels), each ordered by decreasing cost so waves get
//      items of similar length.
// Full passes (RA, DC, PR, HITS) launch over the three lists concurrently.
// Frontier algorithms (BFS, CC, GC, earliest arrival) mark vertices in
// per-vertex arrays (level, change stamp, color) and select each round's
// vertices by scanning the active range with a predicate (k_expand), which
// writes the round's item queues without per-vertex atomics; the host reads
// the few 4-byte queue counts to size the launches (the reference copies its
// queue size back per level too, bfs.cu:2028).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <memory>
#include <mutex>
#include <numeric>
#include <string>
#include <thread>
#include <utility>
#include <vector>

#include "../../../include/tga.h"
#include "decode.hpp"
#include "loader.hpp"

namespace tga {
// sort.hip: stable radix sort of (key, value) pairs by key bits [0, end_bit).
hipError_t sort_pairs_u32_u64(void* tmp, size_t& tmp_bytes, const uint32_t* keys_in, uint32_t* keys_out,
                              const uint64_t* vals_in, uint64_t* vals_out, size_t n, int end_bit,
                              hipStream_t stream);
// sort.hip: per-row sums over pairs [begin, begin + count) (deterministic
// reduce-by-key); mode 0 PageRank in-rows, 1 HITS authority in-rows, 2 HITS
// hub out-rows.
hipError_t reduce_rows(int mode, void* tmp, size_t& tmp_bytes, const uint32_t* sk, const uint64_t* val,
                       const uint32_t* pk, const float* vec, int32_t n, size_t begin, size_t count, uint32_t* ukeys,
                       double* sums, uint32_t* runs, hipStream_t stream);
}  // namespace tga

using namespace tga;

namespace {

thread_local std::string g_err;

int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}

#define HIP_TRY(call)                                                                   \
  do {                                                                                  \
    hipError_t e_ = (call);                                                             \
    if (e_ != hipSuccess)                                                               \
      return fail(TGA_EDEVICE, std::string(#call " failed: ") + hipGetErrorString(e_)); \
  } while (0)

constexpr int kBlock = 256;

// Records fall into classes by their reference-chain depth (0: no reference
// ... 3); the class fixes the chunk stride of long records.  A pass launches
// three groups of work items, each with a decoder of the levels it needs:
// flat records (class 0, one level), depth-1 records (two levels) and deeper
// ones (four levels, few records).  Three concurrent launches plus the
// handle's stream stay within the hardware queues a process gets by default
// (4), so no launch of a pass waits behind another on a shared queue.
constexpr int kClasses = kMaxDepth + 1;
constexpr int kGroups = 3;
__host__ __device__ constexpr int group_of(int depth) { return depth < 2 ? depth : 2; }
// vinfo[x]: whole record of depth d -> -1 - d; no contacts -> kEmpty;
// chunked record -> index of its first checkpoint.
constexpr int32_t kEmpty = -1 - kClasses;

struct Queues {
  int32_t* q[kGroups];
};

struct Window {
  int64_t ws, we;
  int32_t ab, ae;  // active range [ab, ae)
  __device__ __forceinline__ bool in(int32_t t) const { return (int64_t)t >= ws && (int64_t)t <= we; }
  __device__ __forceinline__ bool active(int32_t v) const { return v >= ab && v < ae; }
};

template <class T>
__device__ __forceinline__ T wave_sum(T v) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}


// Accumulator pairs of the neighbor-access checksum and count.
constexpr int kSumSlots = 256;

// Column copy of the timestamp codes of a static item list (built at open,
// tcol_build): list positions 64q .. 64q + 63 -- the items one wave walks
// together, as FOR_ITEMS hands them out with whole waves of consecutive
// positions -- form group q, whose region starts at unit goff[q] and holds
// rows of 64 16-byte units: unit u of position 64q + l at goff[q] + 64u + l,
// the item's code bits from bit 0 of its unit 0 on.  A group has as many rows
// as its longest item needs.
struct TsCols {
  const uint64_t* w = nullptr;     // units as word pairs (stream word layout)
  const uint64_t* goff = nullptr;  // per group: first unit
  const uint32_t* nbits = nullptr; // per list position: the item's code bits
  // The flat list's whole records also have their graph bits in columns of
  // words: word j of position 64q + l at gw[ggoff[q] + 64j + l]; word 0 is
  // the record's start word (flat_start_word), then the record's bits from
  // bit 0 of word 1 on, ceil(gbits / 64) + 5 words (gbits 0: a chunk, read
  // from the stream).
  const uint64_t* gw = nullptr;
  const uint64_t* ggoff = nullptr;  // per group: first word
  const uint32_t* gbits = nullptr;  // per list position
};

// An item list and its length; static lists carry their column copy.
struct Items {
  const int32_t* code;
  uint32_t count;
  TsCols tc{};
  // The lane's timestamp window for list position i: from the column copy
  // where the list has one, else from the linear stream.
  __device__ __forceinline__ TsStage stage(TsStage st, uint32_t i) const {
    if (tc.w) {
      st.uw = tc.w + 2 * (tc.goff[i >> 6] + (i & 63));
      st.ustride = 64;
      st.nbits = tc.nbits[i];
    }
    if (tc.gw) {
      st.gbits = tc.gbits[i];
      st.gw = st.gbits ? tc.gw + tc.ggoff[i >> 6] + (i & 63) : nullptr;
    }
    return st;
  }
};

// The calling lane's ring of decoded runs and its timestamp window
// (decode.hpp walk) in LDS, for blocks of at most kBlock threads, interleaved
// by thread.
#define LANE_RING()                                                       \
  __shared__ int2 ring_lds_[kRing * kBlock];                              \
  __shared__ uint32_t stage_lds_[(kStageDw + 2) * kBlock];                \
  const RunRing ring_{ring_lds_ + threadIdx.x, kBlock};                   \
  const TsStage stage_ {                                                  \
    stage_lds_ + threadIdx.x, kBlock, 0, 0, 0, 0, 0                       \
  }

#define FOR_ITEMS(I, i) \
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < (I).count; i += gridDim.x * blockDim.x)

// Per-item arrays (list counts and offsets) index whole records by vertex
// and checkpoint chunk j as n + j.
__device__ __forceinline__ int32_t item_index(int32_t code, int32_t n) { return code >= 0 ? code : n + (-code - 1); }

// ------------------------------------------------------- schedule kernels ---

// Stream words in place: the file's bytes were copied into the word array,
// and each word's bytes are swapped so that its most significant byte is the
// first stream byte (the layout of decode.hpp).
__global__ void __launch_bounds__(kBlock) k_bswap_words(uint64_t* w, int64_t nw) {
  for (int64_t i = blockIdx.x * (int64_t)kBlock + threadIdx.x; i < nw; i += (int64_t)gridDim.x * kBlock)
    w[i] = __builtin_bswap64(w[i]);
}

// First timestamp bit and code bits of every position of a static item list
// (a whole record: its record's codes; a chunk: from its checkpoint's cursor
// to the next chunk's of the same record, or to the record's end).
__global__ void __launch_bounds__(kBlock) k_tcol_len(Streams S, const Ckpt* ck, int64_t nck, Items I, uint64_t* tstart,
                                                     uint32_t* nbits) {
  FOR_ITEMS(I, i) {
    const int32_t code = I.code[i];
    uint64_t b, e;
    if (code >= 0) {
      b = S.vpos[2 * (size_t)code + 1];
      e = rec_tend(S, code);
    } else {
      const int64_t j = -(int64_t)code - 1;
      const Ckpt& c = ck[j];
      b = c.tpos;
      e = (j + 1 < nck && ck[j + 1].x == c.x && ck[j + 1].k0 == c.k1) ? ck[j + 1].tpos : rec_tend(S, c.x);
    }
    tstart[i] = b;
    nbits[i] = (uint32_t)(e - b);
  }
}

// Copies every item's codes into its column (TsCols): unit u of position i
// holds the 128 stream bits from tstart[i] + 128u on, as two stream-layout
// words.  Units past an item's last code are not written (never read).
__global__ void __launch_bounds__(kBlock) k_tcol_fill(Streams S, Items I, const uint64_t* tstart, uint64_t* w) {
  FOR_ITEMS(I, i) {
    const uint64_t b = tstart[i];
    const uint32_t nb = I.tc.nbits[i];
    uint64_t* out = w + 2 * (I.tc.goff[i >> 6] + (i & 63));
    for (uint32_t u = 0; 128ull * u < nb; ++u) {
      out[128ull * u] = peek64(S.t, b + 128ull * u);
      out[128ull * u + 1] = peek64(S.t, b + 128ull * u + 64);
    }
  }
}

// Graph bits of every whole record of a flat list (0 for a chunk), and where
// they start.
__global__ void __launch_bounds__(kBlock) k_gcol_len(Streams S, Items I, uint64_t* gstart, uint32_t* gbits) {
  FOR_ITEMS(I, i) {
    const int32_t x = I.code[i];
    uint64_t b = 0, e = 0;
    if (x >= 0) {
      b = S.vpos[2 * (size_t)x];
      e = x + 1 < S.n ? S.vpos[2 * (size_t)x + 2] : S.g_bits;
    }
    gstart[i] = b;
    gbits[i] = (uint32_t)(e - b);
  }
}

// Copies every whole record's graph bits into its column (TsCols::gw): the
// record's start word (flat_start_word), then ceil(gbits / 64) + 5 words of
// the stream from the record's first bit on (words past the stream's
// padding read as 0).
__global__ void __launch_bounds__(kBlock) k_gcol_fill(Streams S, Items I, const uint64_t* gstart, uint64_t* w) {
  FOR_ITEMS(I, i) {
    const uint32_t nb = I.tc.gbits[i];
    if (nb == 0) continue;
    const uint64_t b = gstart[i];
    uint64_t* out = w + I.tc.ggoff[i >> 6] + (i & 63);
    out[0] = flat_start_word(S, b);
    const uint32_t words = (nb + 63) / 64 + 5;
    for (uint32_t j = 0; j < words; ++j) {
      const uint64_t p = b + 64ull * j;
      const uint64_t wi = p >> 6;
      const uint32_t sh = (uint32_t)p & 63u;
      const uint64_t a = wi <= S.g_last ? S.g[wi] : 0;
      const uint64_t c = wi + 1 <= S.g_last ? S.g[wi + 1] : 0;
      out[64ull * (j + 1)] = sh ? (a << sh) | (c >> (64u - sh)) : a;
    }
  }
}

// Record header of every vertex: outdegree and reference distance.
// Also fills the record start table (Streams::vpos) from the two
// Elias-Fano indexes.
__global__ void __launch_bounds__(kBlock) k_headers(Streams S, int32_t* deg, int8_t* ref, uint64_t* vpos) {
  const Rd g{S.g, S.g_bits, 0};
  for (int64_t x = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; x < S.n; x += (int64_t)gridDim.x * blockDim.x) {
    const uint64_t gpos = ef_get(S, S.gi, (uint64_t)x);
    const uint64_t tpos = ef_get(S, S.ti, (uint64_t)x);
    uint64_t pos = gpos;
    const int32_t d = (int32_t)rd_gamma(S, g, pos);
    int32_t r = 0;
    if (d > 0) r = (int32_t)rd_unary(S, g, pos);
    deg[x] = d;
    ref[x] = (int8_t)min(r, 127);
    vpos[2 * x] = gpos;
    vpos[2 * x + 1] = tpos;
  }
}

// Decodes each long record once and saves a checkpoint every chunk_stride
// contacts.  A run of one neighbor crossing a boundary gets a split-run slot:
// the index of the first chunk it continues into (see Ckpt).
template <int MAXD>
__global__ void __launch_bounds__(64) k_ckpt_build(Streams S, const int32_t* bigv, const int64_t* slot, int32_t nbig,
                                                   Ckpt* ck, int32_t* nchunk_out) {
  const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nbig) return;
  nchunk_out[i] = ckpt_build_record<MAXD>(S, bigv[i], slot[i], ck + slot[i]);
}

// Selection of the vertices a round expands (evaluated on device arrays).
enum : int32_t {
  kSelAll = 0,        // every vertex of the range or list
  kSelEq = 1,         // a[v] == val (BFS level, round stamp of a changed vertex)
  kSelUncolored = 2,  // a[v] == kInf (graph coloring's worklist)
  kSelMask = 3,       // m[v] != 0 (batched BFS: some query has v in its frontier)
  kSelGc = 4,         // a[v] == kInf and not held by its cached blocker: b[v] < 0 or a[b[v]] != kInf
};
struct Sel {
  int32_t kind;
  const int32_t* a;
  int32_t val;
  const uint64_t* m = nullptr;
  const int32_t* b = nullptr;
  __device__ __forceinline__ bool operator()(int32_t v) const {
    if (kind == kSelAll) return true;
    if (kind == kSelMask) return m[v] != 0;
    const int32_t x = a[v];
    if (kind == kSelGc) {
      if (x != kInf) return false;
      const int32_t u = b[v];
      return u < 0 || a[u] != kInf;
    }
    return kind == kSelEq ? x == val : x == kInf;
  }
};

constexpr int kExpandBlocks = 1024;  // grid of k_expand (one queue reservation per group per block)
constexpr uint32_t kExpandCoop = 16;  // chunk lists longer than this are written by a whole wave

// Inclusive prefix sum over the 64 lanes of a wave.
__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t v) {
  const int lane = threadIdx.x & 63;
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    const uint32_t u = __shfl_up(v, off, 64);
    if (lane >= off) v += u;
  }
  return v;
}

// Exclusive prefix of v over the block's threads (in thread order) for
// kGroups + 1 counters at once; tot receives the block totals.  All threads
// of the block must call it.
__device__ __forceinline__ void block_excl_scan(const uint32_t (&v)[kGroups + 1], uint32_t (&pre)[kGroups + 1],
                                                uint32_t (&tot)[kGroups + 1]) {
  __shared__ uint32_t wsum[kBlock / 64][kGroups + 1];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  uint32_t inc[kGroups + 1];
#pragma unroll
  for (int g = 0; g <= kGroups; ++g) {
    inc[g] = wave_incl_scan(v[g]);
    if (lane == 63) wsum[wave][g] = inc[g];
  }
  __syncthreads();
#pragma unroll
  for (int g = 0; g <= kGroups; ++g) {
    uint32_t before = 0, all = 0;
    for (int w = 0; w < kBlock / 64; ++w) {
      if (w < wave) before += wsum[w][g];
      all += wsum[w][g];
    }
    pre[g] = before + inc[g] - v[g];
    tot[g] = all;
  }
  __syncthreads();
}

// Expands the selected vertices of an explicit list, or of the range
// [base, base + count), into per-group item queues: cnt[g] receives the count
// of group g, and, when bigq is given, chunked records are also listed in bigq
// (counted in cnt[kGroups]).  keep_empty keeps records without contacts
// (graph coloring colors them) as group-0 items.  Each block takes one
// contiguous slice: it counts its items, reserves queue space with one atomic
// per group, then writes them in vertex order.
__global__ void __launch_bounds__(kBlock) k_expand(const int32_t* front, int32_t base, uint32_t count, Sel sel,
                                                   const int32_t* vinfo, const Ckpt* ck, Queues Q, uint32_t* cnt,
                                                   int32_t* bigq, int keep_empty) {
  __shared__ uint32_t at[kGroups + 1];
  const uint32_t per = (count + gridDim.x - 1) / gridDim.x;
  const uint32_t slice = (per + kBlock - 1) / kBlock * kBlock;
  const uint32_t lo = min(count, blockIdx.x * slice), hi = min(count, lo + slice);
  // (vertex, group, items) of element i of the slice; items 0: not expanded.
  auto classify = [&](uint32_t i, int32_t& x, int32_t& info, int& grp, uint32_t& items) {
    x = front ? front[i] : base + (int32_t)i;
    items = 0;
    grp = 0;
    info = kEmpty;
    if (!sel(x)) return;
    info = vinfo[x];
    if (info == kEmpty) {
      items = keep_empty ? 1u : 0u;
    } else if (info < 0) {
      grp = group_of(-1 - info);
      items = 1;
    } else {
      grp = group_of(ck[info].depth);
      items = (uint32_t)ck[info].nchunk;
    }
  };
  uint32_t own[kGroups + 1] = {};
  for (uint32_t i = lo + threadIdx.x; i < hi; i += kBlock) {
    int32_t x, info;
    int grp;
    uint32_t items;
    classify(i, x, info, grp, items);
#pragma unroll
    for (int g = 0; g < kGroups; ++g) own[g] += g == grp ? items : 0u;
    if (items && info >= 0) ++own[kGroups];
  }
  uint32_t pre[kGroups + 1], tot[kGroups + 1];
  block_excl_scan(own, pre, tot);
  if (threadIdx.x == 0) {
#pragma unroll
    for (int g = 0; g < kGroups; ++g) at[g] = tot[g] ? atomicAdd(&cnt[g], tot[g]) : 0;
    at[kGroups] = (bigq && tot[kGroups]) ? atomicAdd(&cnt[kGroups], tot[kGroups]) : 0;
  }
  __syncthreads();
  uint32_t run[kGroups + 1];
#pragma unroll
  for (int g = 0; g <= kGroups; ++g) run[g] = at[g];
  for (uint32_t t0 = lo; t0 < hi; t0 += kBlock) {
    const uint32_t i = t0 + threadIdx.x;
    int32_t x = 0, info = kEmpty;
    int grp = 0;
    uint32_t items = 0;
    if (i < hi) classify(i, x, info, grp, items);
    uint32_t v[kGroups + 1];
#pragma unroll
    for (int g = 0; g < kGroups; ++g) v[g] = g == grp ? items : 0u;
    v[kGroups] = (items && info >= 0) ? 1u : 0u;
    block_excl_scan(v, pre, tot);
    // Queue offset of this thread's items (static indices only: a dynamically
    // indexed register array spills).
    const uint32_t qoff = grp == 0 ? run[0] + pre[0] : (grp == 1 ? run[1] + pre[1] : run[2] + pre[2]);
    bool coop = false;
    if (items) {
      int32_t* q = (grp == 0 ? Q.q[0] : (grp == 1 ? Q.q[1] : Q.q[2])) + qoff;
      if (info < 0) {
        q[0] = x;
      } else {
        if (items <= kExpandCoop)
          for (uint32_t j = 0; j < items; ++j) q[j] = -(info + (int32_t)j) - 1;
        else
          coop = true;
        if (bigq) bigq[run[kGroups] + pre[kGroups]] = x;
      }
    }
    // Long chunk lists are written by the whole wave, one record at a time,
    // as coalesced stores (one lane writing thousands of entries would hold
    // the block back).
    const int lane = threadIdx.x & 63;
    for (uint64_t todo = __ballot(coop); todo; todo &= todo - 1) {
      const int l = __ffsll((unsigned long long)todo) - 1;
      const int lg = __shfl(grp, l, 64);
      const uint32_t lo_ = (uint32_t)__shfl((int32_t)qoff, l, 64);
      const int32_t li = __shfl(info, l, 64);
      const uint32_t ln = (uint32_t)__shfl((int32_t)items, l, 64);
      int32_t* q = (lg == 0 ? Q.q[0] : (lg == 1 ? Q.q[1] : Q.q[2])) + lo_;
      for (uint32_t j = (uint32_t)lane; j < ln; j += 64) q[j] = -(li + (int32_t)j) - 1;
    }
#pragma unroll
    for (int g = 0; g <= kGroups; ++g) run[g] += tot[g];
  }
}

// -------------------------------------------------------------- actions ---

struct RaAct {
  uint64_t cs = 0;
  uint32_t cnt = 0;  // contacts of this lane (far below 2^32 per pass)
  uint32_t kx = 0;
  __device__ bool begin(int32_t x, bool, Split) {
    kx = contact_key_x(x);
    return true;
  }
  // The walk hands contact() the neighbor's checksum key, computed once per run.
  __device__ int32_t ring_val(int32_t r) const { return (int32_t)contact_key_r(kx, r); }
  __device__ bool contact(int32_t, int32_t kr, int32_t t, int32_t) {
    cs += contact_mix((uint32_t)kr, t);
    ++cnt;
    return true;
  }
  __device__ void end(int32_t, bool) {}
};

template <int MAXD>
__global__ void __launch_bounds__(kBlock) k_ra(Streams S, const Ckpt* ck, Items I, unsigned long long* out) {
  RaAct a;
  LANE_RING();
  FOR_ITEMS(I, i) run_item<MAXD>(S, ck, I.code[i], a, ring_, I.stage(stage_, i));
  const unsigned long long cs = wave_sum((unsigned long long)a.cs);
  const unsigned long long cnt = wave_sum((unsigned long long)a.cnt);
  // Spread over kSumSlots accumulator pairs: tens of thousands of waves
  // adding into one address would serialize on it.
  if ((threadIdx.x & 63) == 0 && cnt) {
    unsigned long long* o = out + 2 * (blockIdx.x & (kSumSlots - 1));
    atomicAdd(&o[0], cs);
    atomicAdd(&o[1], cnt);
  }
}

// Contact-time cache (window-independent, built once per handle by
// k_ctimes): the time t of contact k of active x at ct[coff[x] + k], or
// t + kInactiveBias when its neighbor is inactive.  Biased values lie above
// the int32 range, so a window clamped to int32 (clamp_lo / clamp_hi; t is
// int32, so nothing changes) never holds them.  The degree centrality of long
// records and the batched degree centrality and PageRank read it instead of
// decoding.
constexpr int64_t kNoTime = INT64_MIN;  // below every window (padding)
constexpr int64_t kInactiveBias = (int64_t)1 << 40;
__device__ __forceinline__ int64_t ct_time(int64_t v) { return v >= (kInactiveBias >> 1) ? v - kInactiveBias : v; }
__host__ __device__ __forceinline__ int64_t clamp_lo(int64_t ws) { return ws < INT32_MIN ? (int64_t)INT32_MIN : ws; }
__host__ __device__ __forceinline__ int64_t clamp_hi(int64_t we) { return we > INT32_MAX ? (int64_t)INT32_MAX : we; }

struct CtAct {
  Window W;  // active range
  int64_t* ct;
  const uint64_t* coff;
  int64_t* tb = nullptr;
  __device__ bool begin(int32_t x, bool, Split) {
    tb = ct + coff[x];
    return W.active(x);
  }
  __device__ bool contact(int32_t, int32_t r, int32_t t, int32_t k) {
    tb[k] = W.active(r) ? (int64_t)t : (int64_t)t + kInactiveBias;
    return true;
  }
  __device__ void end(int32_t, bool) {}
};

template <int MAXD>
__global__ void __launch_bounds__(kBlock) k_ctimes(Streams S, const Ckpt* ck, Items I, Window W, int64_t* ct,
                                                   const uint64_t* coff) {
  CtAct a{W, ct, coff};
  LANE_RING();
  FOR_ITEMS(I, i) run_item<MAXD>(S, ck, I.code[i], a, ring_, I.stage(stage_, i));
}

// Degree count and Java running mean of one record (whole records also get
// the centrality formula).
struct DcAct {
  Window W;
  int32_t* count;
  int32_t* centr;
  int32_t deg = 0;
  int64_t avg = 0;
  __device__ bool begin(int32_t x, bool, Split) {
    deg = 0;
    avg = 0;
    return W.active(x);
  }
  __device__ bool contact(int32_t, int32_t r, int32_t t, int32_t) {
    if (W.in(t) && W.active(r)) {
      ++deg;
      avg += ((int64_t)t - avg) / deg;  // truncating, as Java's long division
    }
    return true;
  }
  __device__ void end(int32_t x, bool chunked);
};

#pragma clang fp contract(off)
// Java's (int)(0.85 * c + 0.15 * c * (mean - ws) / (we - ws) + 1), evaluated
// left to right in double without fused multiply-adds (BaselineRunner.java:325-331).
__device__ int32_t dc_formula(int32_t deg, int64_t avg, int64_t ws, int64_t we) {
  if (ws == we) return deg;
  const double c = 0.85 * (double)deg + 0.15 * (double)deg * (double)(avg - ws) / (double)(we - ws) + 1.0;
  return (int32_t)c;
}
#pragma clang fp contract(on)

// Writes only the outputs it was given (k_dc_long passes none and reads
// deg / avg itself).
__device__ void DcAct::end(int32_t x, bool chunked) {
  if (!count) return;
  if (chunked) {
    if (deg) atomicAdd(&count[x], deg);
  } else {
    count[x] = deg;
    if (centr) centr[x] = dc_formula(deg, avg, W.ws, W.we);
  }
}

template <int MAXD>
__global__ void __launch_bounds__(kBlock) k_dc(Streams S, const Ckpt* ck, Items I, Window W, int32_t* count,
                                               int32_t* centr) {
  DcAct a{W, count, centr};
  LANE_RING();
  FOR_ITEMS(I, i) run_item<MAXD>(S, ck, I.code[i], a, ring_, I.stage(stage_, i));
}

// Java's long division a / k (truncating toward zero) for k >= 1 and
// |a| < 2^40, given rk = 1.0 / k in fp64: a * rk is within 2^-12 of a / k,
// so its truncation is within one of the quotient; one correction step.
__device__ __forceinline__ int64_t div_trunc(int64_t a, int64_t k, double rk) {
  int64_t q = (int64_t)((double)a * rk);
  const int64_t r = a - q * k;
  if (a >= 0) {
    if (r >= k) ++q;
    else if (r < 0) --q;
  } else {
    if (r <= -k) --q;
    else if (r > 0) ++q;
  }
  return q;
}

// Lane moves of a 64-bit value: DPP (row shifts and the GFX9 row broadcasts,
// no LDS round trip) and readlane with a wave-uniform lane.
template <int CTRL>
__device__ __forceinline__ int64_t dpp64(int64_t v) {
  const int lo = __builtin_amdgcn_mov_dpp((int)(uint32_t)v, CTRL, 0xf, 0xf, false);
  const int hi = __builtin_amdgcn_mov_dpp((int)(uint32_t)((uint64_t)v >> 32), CTRL, 0xf, 0xf, false);
  return (int64_t)(((uint64_t)(uint32_t)hi << 32) | (uint32_t)lo);
}

__device__ __forceinline__ int64_t readlane64(int64_t v, int lane) {
  const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)v, lane);
  const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)((uint64_t)v >> 32), lane);
  return (int64_t)(((uint64_t)hi << 32) | lo);
}

// Inclusive prefix sum over the wave: row_shr 1, 2, 4, 8 within rows of 16
// lanes, then row_bcast:15 into rows 1 and 3 and row_bcast:31 into rows 2
// and 3.  Every lane must call it.
__device__ __forceinline__ int64_t wave_incl_scan64(int64_t v) {
  const int lane = threadIdx.x & 63, rl = lane & 15;
  int64_t t = dpp64<0x111>(v);
  if (rl >= 1) v += t;
  t = dpp64<0x112>(v);
  if (rl >= 2) v += t;
  t = dpp64<0x114>(v);
  if (rl >= 4) v += t;
  t = dpp64<0x118>(v);
  if (rl >= 8) v += t;
  t = dpp64<0x142>(v);
  if ((lane & 31) >= 16) v += t;
  t = dpp64<0x143>(v);
  if (lane >= 32) v += t;
  return v;
}

// Count and Java running mean of the cached contact times tb[0, len) that lie
// in [ws, we], by one wave, 64 contacts a step.  The mean is sequential
// (avg_k = avg_k-1 + (t_k - avg_k-1) / k, truncating), so each step is
// speculative: every lane divides with the step's starting mean, a prefix sum
// of those quotients gives each lane its starting mean, and the lanes
// recompute; from the first lane whose quotient changed (the lanes before it
// are exact, and so is its new quotient) the step restarts.  Each round fixes
// at least one lane.  Every lane of the wave must call it.
__device__ void wave_mean(const int64_t* tb, int32_t len, int64_t ws, int64_t we, int32_t& cnt_out,
                          int64_t& avg_out) {
  const int lane = threadIdx.x & 63;
  ws = clamp_lo(ws);
  we = clamp_hi(we);
  int64_t avg = 0;
  int32_t cnt = 0;
  // Times are loaded kAhead steps of 64 at a time, the next group while the
  // current one is processed.
  constexpr int kAhead = 4;
  int64_t cur[kAhead], nxt[kAhead];
#pragma unroll
  for (int j = 0; j < kAhead; ++j) {
    const int32_t i = j * 64 + lane;
    cur[j] = i < len ? tb[i] : kNoTime;
  }
  for (int32_t g0 = 0; g0 < len; g0 += kAhead * 64) {
#pragma unroll
    for (int j = 0; j < kAhead; ++j) {
      const int32_t i = g0 + (kAhead + j) * 64 + lane;
      nxt[j] = i < len ? tb[i] : kNoTime;
    }
#pragma unroll
    for (int j = 0; j < kAhead; ++j) {
      const int64_t t = cur[j];
      const bool valid = t != kNoTime && t >= ws && t <= we;
      const uint64_t vm = __ballot(valid);
      if (!vm) continue;
      const int64_t k = cnt + __popcll(vm & (lane == 63 ? ~0ull : ((2ull << lane) - 1)));
      const double rk = 1.0 / (double)k;  // off the sequential chain
      int first = 0;
      int64_t base = avg;
      int64_t q = valid ? div_trunc(t - base, k, rk) : 0;
      while (true) {
        const int64_t inc = wave_incl_scan64(q);
        const int64_t before = base + inc - q;
        const int64_t q2 = (valid && lane >= first) ? div_trunc(t - before, k, rk) : 0;
        const uint64_t mm = __ballot(q2 != q);
        if (!mm) {
          avg = base + readlane64(inc, 63);
          break;
        }
        const int f = __ffsll((unsigned long long)mm) - 1;
        base = readlane64(before + q2, f);
        first = f + 1;
        q = (valid && lane >= first) ? div_trunc(t - base, k, rk) : 0;
      }
      cnt += __popcll(vm);
    }
#pragma unroll
    for (int j = 0; j < kAhead; ++j) cur[j] = nxt[j];
  }
  cnt_out = cnt;
  avg_out = avg;
}

// Centrality of each long record from the contact-time cache, one wave per
// record.
__global__ void __launch_bounds__(kBlock) k_dc_rec(const int32_t* bigv, int32_t nbig, const int32_t* vdeg,
                                                   const uint64_t* coff, const int64_t* ct, Window W,
                                                   int32_t* centr) {
  const int lane = threadIdx.x & 63;
  const int32_t wid = blockIdx.x * (kBlock / 64) + (threadIdx.x >> 6);
  if (wid >= nbig) return;
  const int32_t x = bigv[wid];
  int32_t cnt = 0;
  int64_t avg = 0;
  if (W.active(x)) wave_mean(ct + coff[x], vdeg[x], W.ws, W.we, cnt, avg);
  if (lane == 0) centr[x] = dc_formula(cnt, avg, W.ws, W.we);
}

// Batched degree centrality, up to 64 windows (wins: ws, we per query):
// count / centr row q receives what tga_degree_centrality gives for window q.
// Records that are not chunked: one wave per vertex, lane q running query q's
// count and mean over the cached times (the loads are the same for all
// lanes).  Long records are left to k_dcb_long.
__global__ void __launch_bounds__(kBlock) k_dcb(Window A, int32_t n, const int32_t* vinfo, const int32_t* vdeg,
                                                const uint64_t* coff, const int64_t* ct, const int64_t* wins,
                                                int32_t nq, int32_t* count, int32_t* centr) {
  const int lane = threadIdx.x & 63;
  const bool on = lane < nq;
  const int64_t ws = on ? wins[2 * lane] : 0, we = on ? wins[2 * lane + 1] : 0;
  const int64_t cws = clamp_lo(ws), cwe = clamp_hi(we);
  const int64_t nw = (int64_t)gridDim.x * (kBlock / 64);
  for (int64_t v = blockIdx.x * (int64_t)(kBlock / 64) + (threadIdx.x >> 6); v < n; v += nw) {
    const bool act = A.active((int32_t)v);
    if (act && vinfo[v] >= 0) continue;  // chunked: k_dcb_long
    int32_t cnt = 0;
    int64_t avg = 0;
    if (act) {
      const int64_t* tb = ct + coff[v];
      const int32_t len = vdeg[v];
      for (int32_t i = 0; i < len; ++i) {
        const int64_t t = tb[i];
        if (t >= cws && t <= cwe) {
          ++cnt;
          avg += div_trunc(t - avg, cnt, 1.0 / (double)cnt);
        }
      }
    }
    if (on) {
      count[(int64_t)lane * n + v] = cnt;
      if (centr) centr[(int64_t)lane * n + v] = dc_formula(cnt, avg, ws, we);
    }
  }
}

// Batched degree centrality of long records: one wave per (record, query).
__global__ void __launch_bounds__(kBlock) k_dcb_long(const int32_t* bigv, int32_t nbig, Window A, int32_t n,
                                                     const int32_t* vdeg, const uint64_t* coff, const int64_t* ct,
                                                     const int64_t* wins, int32_t nq, int32_t* count,
                                                     int32_t* centr) {
  const int lane = threadIdx.x & 63;
  const int64_t wid = blockIdx.x * (int64_t)(kBlock / 64) + (threadIdx.x >> 6);
  if (wid >= (int64_t)nbig * nq) return;
  const int32_t x = bigv[wid / nq], q = (int32_t)(wid % nq);
  const int64_t ws = wins[2 * q], we = wins[2 * q + 1];
  int32_t cnt = 0;
  int64_t avg = 0;
  if (A.active(x)) wave_mean(ct + coff[x], vdeg[x], ws, we, cnt, avg);
  if (lane == 0) {
    count[(int64_t)q * n + x] = cnt;
    if (centr) centr[(int64_t)q * n + x] = dc_formula(cnt, avg, ws, we);
  }
}

// ------------------------------------------------- batched PageRank ---
//
// Up to 64 PageRank queries (windows) advance together as FP32 columns: rank,
// coefficient and window degree are [vertex][64] arrays, so lane q of a wave
// holds query q and a vertex's 64 values are one coalesced row.  The pair
// structure is window-independent and built once per handle: every run of
// equal neighbors in every record, (x, r, contacts [k0, k1)), from a counting
// and a filling decode pass (a chunk skips the run that continues into it from
// the previous chunk, so each run is listed once, where it starts, and ends
// where the next run of its record starts), radix-sorted by r (runs to
// inactive r last) and cut into segments of at most kPrbSeg runs.  Query q's
// weight of a run is 1 - exp(-0.1 t) at the run's first contact inside its
// window, read from the contact-time cache.  An iteration is one wave per
// segment: lane q sums (double)(coef[x][q] * w) over the segment's runs in
// order; rows of several segments are summed in segment order by k_prb_fix.
constexpr int kPrbSeg = 512;
constexpr int kPrbLongDeg = 1024;  // window degrees of longer records: one wave per (record, query)
constexpr int kPrbRedBlocks = 512;
constexpr int kCcbGroup = 8;  // batched CC rounds per replayed graph

struct RunAct {
  Window A;             // active range
  uint32_t* cnt;        // counting pass: per item
  const uint64_t* off;  // filling pass: per item
  int32_t *rx, *rr, *rk;
  int32_t n;
  uint64_t pos = 0;
  uint32_t c = 0;
  int32_t idx = 0, x = 0, last = -1, cont_r = -1;
  bool cont = false;
  __device__ void item(int32_t code) { idx = item_index(code, n); }
  __device__ bool begin(int32_t v, bool, Split s) {
    if (!A.active(v)) return false;
    x = v;
    c = 0;
    last = -1;
    cont = s.cont_slot >= 0;
    cont_r = s.cont_r;
    if (off) pos = off[idx];
    return true;
  }
  __device__ bool contact(int32_t, int32_t r, int32_t, int32_t k) {
    if (cont) {
      if (r == cont_r) return true;  // the run that started in the previous chunk
      cont = false;
    }
    if (r == last) return true;
    last = r;
    if (off) {
      rx[pos] = x;
      rr[pos] = r;
      rk[pos] = k;
      ++pos;
    } else {
      ++c;
    }
    return true;
  }
  __device__ void end(int32_t, bool) {
    if (!off) cnt[idx] = c;
  }
};

template <int MAXD>
__global__ void __launch_bounds__(kBlock) k_runs(Streams S, const Ckpt* ck, Items I, Window A, uint32_t* cnt,
                                                 const uint64_t* off, int32_t* rx, int32_t* rr, int32_t* rk) {
  RunAct a{A, cnt, off, rx, rr, rk, S.n};
  LANE_RING();
  FOR_ITEMS(I, i) run_item<MAXD>(S, ck, I.code[i], a, ring_, I.stage(stage_, i));
}

// Run ends (the next run of the record, or its outdegree) and sort keys
// (r, or n for an inactive r, which sorts last); vals = run index.
__global__ void __launch_bounds__(kBlock) k_run_keys(int64_t nr, const int32_t* rx, const int32_t* rr,
                                                     const int32_t* rk, const int32_t* vdeg, Window A, int32_t n,
                                                     int32_t* k1, uint32_t* keys, uint64_t* vals) {
  for (int64_t i = blockIdx.x * (int64_t)kBlock + threadIdx.x; i < nr; i += (int64_t)gridDim.x * kBlock) {
    k1[i] = (i + 1 < nr && rx[i + 1] == rx[i]) ? rk[i + 1] : vdeg[rx[i]];
    keys[i] = A.active(rr[i]) ? (uint32_t)rr[i] : (uint32_t)n;
    vals[i] = (uint64_t)i;
  }
}

// in_off[v] = first position of key v in the sorted keys, v in [0, n].
__global__ void __launch_bounds__(kBlock) k_key_bounds(const uint32_t* sk, uint64_t count, int32_t n,
                                                       uint64_t* in_off) {
  for (int64_t v = blockIdx.x * (int64_t)kBlock + threadIdx.x; v <= n; v += (int64_t)gridDim.x * kBlock) {
    uint64_t lo = 0, hi = count;
    while (lo < hi) {
      const uint64_t mid = (lo + hi) >> 1;
      if (sk[mid] < (uint32_t)v)
        lo = mid + 1;
      else
        hi = mid;
    }
    in_off[v] = lo;
  }
}

// Sorted runs as (x, k0, k1, t0) with t0 the time of the run's first contact,
// and f0 = 1 - exp(-0.1 t0), its weight for every query whose window holds t0.
__global__ void __launch_bounds__(kBlock) k_run_gather(uint64_t nv, const uint64_t* sv, const int32_t* rx,
                                                       const int32_t* rk, const int32_t* k1, const uint64_t* coff,
                                                       const int64_t* ct, int4* runs, float* f0) {
  for (int64_t j = blockIdx.x * (int64_t)kBlock + threadIdx.x; j < (int64_t)nv; j += (int64_t)gridDim.x * kBlock) {
    const uint64_t i = sv[j];
    const int32_t t0 = (int32_t)ct[coff[rx[i]] + rk[i]];  // r is active: the time is stored as is
    runs[j] = int4{rx[i], rk[i], k1[i], t0};
    f0[j] = 1.0f - expf(-0.1f * (float)t0);
  }
}

// Runs with more than one contact: flag (for a scan), then their weight slot
// midx (-1 for a single-contact run) and the list of them.
__global__ void __launch_bounds__(kBlock) k_multi_flag(uint64_t nv, const int4* runs, uint32_t* flag) {
  for (int64_t j = blockIdx.x * (int64_t)kBlock + threadIdx.x; j < (int64_t)nv; j += (int64_t)gridDim.x * kBlock)
    flag[j] = runs[j].z - runs[j].y > 1 ? 1u : 0u;
}

__global__ void __launch_bounds__(kBlock) k_multi_list(uint64_t nv, const int4* runs, const uint64_t* off,
                                                       int32_t* midx, int32_t* mlist) {
  for (int64_t j = blockIdx.x * (int64_t)kBlock + threadIdx.x; j < (int64_t)nv; j += (int64_t)gridDim.x * kBlock) {
    const bool multi = runs[j].z - runs[j].y > 1;
    midx[j] = multi ? (int32_t)off[j] : -1;
    if (multi) mlist[off[j]] = (int32_t)j;
  }
}

// Per batch: query q's weight of every multi-contact run, 1 - exp(-0.1 t) at
// its first contact inside window q (0 when none), one wave per run with the
// lanes over 64 contacts at a time and one ballot per query still open.
__global__ void __launch_bounds__(kBlock) k_prb_wmulti(const int32_t* mlist, int32_t nm, const int4* runs,
                                                       const uint64_t* coff, const int64_t* ct, const int64_t* wins,
                                                       int32_t nq, float* wm, uint64_t only) {
  const int lane = threadIdx.x & 63;
  const bool on = lane < nq;
  const int64_t ws = on ? clamp_lo(wins[2 * lane]) : 1, we = on ? clamp_hi(wins[2 * lane + 1]) : 0;
  const uint64_t all = (nq >= 64 ? ~0ull : ((1ull << nq) - 1)) & only;
  const int64_t nw = (int64_t)gridDim.x * (kBlock / 64);
  for (int64_t i = blockIdx.x * (int64_t)(kBlock / 64) + (threadIdx.x >> 6); i < nm; i += nw) {
    const int4 run = runs[mlist[i]];
    const int64_t* tb = ct + coff[run.x];
    uint64_t found = 0;
    float w = 0.0f;
    for (int32_t b0 = run.y; b0 < run.z && found != all; b0 += 64) {
      const int32_t k = b0 + lane;
      const int64_t t = k < run.z ? tb[k] : kNoTime;  // r is active: times as stored
      for (int q = 0; q < nq; ++q) {
        if ((found | ~all) >> q & 1) continue;
        const int64_t qs = readlane64(ws, q), qe = readlane64(we, q);
        const uint64_t m = __ballot(t >= qs && t <= qe);
        if (!m) continue;
        const int64_t tf = readlane64(t, __ffsll((unsigned long long)m) - 1);
        if (lane == q) w = 1.0f - expf(-0.1f * (float)tf);
        found |= 1ull << q;
      }
    }
    if (only >> lane & 1) wm[(int64_t)i * 64 + lane] = w;
  }
}

__global__ void __launch_bounds__(kBlock) k_seg_count(int32_t n, const uint64_t* in_off, uint32_t* nseg) {
  for (int64_t r = blockIdx.x * (int64_t)kBlock + threadIdx.x; r < n; r += (int64_t)gridDim.x * kBlock) {
    const uint64_t c = in_off[r + 1] - in_off[r];
    nseg[r] = (uint32_t)((c + kPrbSeg - 1) / kPrbSeg);
  }
}

// Segment s of row r: runs [beg, beg + len); rows of several segments are
// listed in mrows (their count in *nm).
__global__ void __launch_bounds__(kBlock) k_seg_fill(int32_t n, const uint64_t* in_off, const uint64_t* seg_off,
                                                     int32_t* seg_row, uint64_t* seg_beg, int32_t* seg_len,
                                                     int32_t* mrows, uint32_t* nm) {
  for (int64_t r = blockIdx.x * (int64_t)kBlock + threadIdx.x; r < n; r += (int64_t)gridDim.x * kBlock) {
    const uint64_t b = in_off[r], e = in_off[r + 1], s0 = seg_off[r], s1 = seg_off[r + 1];
    for (uint64_t s = s0; s < s1; ++s) {
      const uint64_t at = b + (s - s0) * kPrbSeg;
      seg_row[s] = (int32_t)r;
      seg_beg[s] = at;
      seg_len[s] = (int32_t)min((uint64_t)kPrbSeg, e - at);
    }
    if (s1 - s0 > 1) mrows[atomicAdd(nm, 1u)] = (int32_t)r;
  }
}

// Window degree of every vertex for every query (all window contacts, any
// neighbor): one wave per vertex, lane q counting query q; records longer
// than kPrbLongDeg are left to k_prb_du_long.
// Batch-state kernels take `only`: the lanes whose columns they write (the
// lanes a query server admits; ~0 for a whole batch).
__global__ void __launch_bounds__(kBlock) k_prb_du(Window A, int32_t n, const int32_t* vdeg, const uint64_t* coff,
                                                   const int64_t* ct, const int64_t* wins, int32_t nq, int32_t* du,
                                                   uint64_t only) {
  const int lane = threadIdx.x & 63;
  const bool on = lane < nq;
  const bool mine = only >> lane & 1;
  const int64_t ws = on ? clamp_lo(wins[2 * lane]) : 1, we = on ? clamp_hi(wins[2 * lane + 1]) : 0;
  const int64_t nw = (int64_t)gridDim.x * (kBlock / 64);
  for (int64_t v = blockIdx.x * (int64_t)(kBlock / 64) + (threadIdx.x >> 6); v < n; v += nw) {
    const bool act = A.active((int32_t)v);
    const int32_t len = act ? vdeg[v] : 0;
    if (len > kPrbLongDeg) {
      if (mine) du[v * 64 + lane] = 0;  // counted by k_prb_du_long
      continue;
    }
    int32_t c = 0;
    const int64_t* tb = ct + (act ? coff[v] : 0);
    for (int32_t i = 0; i < len; ++i) {
      const int64_t t = ct_time(tb[i]);
      c += (t >= ws && t <= we) ? 1 : 0;
    }
    if (mine) du[v * 64 + lane] = c;
  }
}

// Window degrees of long records: one wave per (record, kPrbDuChunk contacts)
// item, lane q counting query q over the item's contacts, added with one
// integer atomic per query.
constexpr int kPrbDuChunk = 4096;

__global__ void __launch_bounds__(kBlock) k_prb_du_long(const int2* items, int32_t ni, const uint64_t* coff,
                                                        const int32_t* vdeg, const int64_t* ct, const int64_t* wins,
                                                        int32_t nq, int32_t* du, uint64_t only) {
  const int lane = threadIdx.x & 63;
  const bool on = lane < nq && (only >> lane & 1);
  const int64_t ws = on ? clamp_lo(wins[2 * lane]) : 1, we = on ? clamp_hi(wins[2 * lane + 1]) : 0;
  const int64_t nw = (int64_t)gridDim.x * (kBlock / 64);
  for (int64_t i = blockIdx.x * (int64_t)(kBlock / 64) + (threadIdx.x >> 6); i < ni; i += nw) {
    const int2 it = items[i];
    const int32_t k1 = min(vdeg[it.x], it.y + kPrbDuChunk);
    const int64_t* tb = ct + coff[it.x];
    int32_t c = 0;
    for (int32_t k = it.y; k < k1; ++k) {
      const int64_t t = ct_time(tb[k]);
      c += (t >= ws && t <= we) ? 1 : 0;
    }
    if (on && c) atomicAdd(&du[(int64_t)it.x * 64 + lane], c);
  }
}

// The (record, first contact) items of the long records.
__global__ void __launch_bounds__(kBlock) k_prb_du_items(const int32_t* lv, int32_t nl, const int32_t* vdeg,
                                                         int2* items, uint32_t* cnt) {
  for (int32_t i = blockIdx.x * kBlock + threadIdx.x; i < nl; i += gridDim.x * kBlock) {
    const int32_t x = lv[i], nch = (vdeg[x] + kPrbDuChunk - 1) / kPrbDuChunk;
    const uint32_t at = atomicAdd(cnt, (uint32_t)nch);
    for (int32_t j = 0; j < nch; ++j) items[at + j] = int2{x, j * kPrbDuChunk};
  }
}

// rank = 1/N for active vertices (all 64 columns), coef and base per iteration.
__global__ void __launch_bounds__(kBlock) k_prb_init(Window A, int32_t n, float inv, float* rank, uint64_t only) {
  for (int64_t i = blockIdx.x * (int64_t)kBlock + threadIdx.x; i < (int64_t)n * 64; i += (int64_t)gridDim.x * kBlock)
    if (only >> (i & 63) & 1) rank[i] = A.active((int32_t)(i >> 6)) ? inv : 0.0f;
}
// One iteration's coefficients for lanes at different iterations (a query
// server): lanes in `fresh` start from their initial rank, the others from
// the last iteration's sums.
__global__ void __launch_bounds__(kBlock) k_prb_coef_step(Window A, int32_t n, const float* rank, const double* sums,
                                                          const int32_t* du, float damping, float* coef, double base,
                                                          double* acc, uint64_t fresh) {
  for (int64_t i = blockIdx.x * (int64_t)kBlock + threadIdx.x; i < (int64_t)n * 64; i += (int64_t)gridDim.x * kBlock) {
    const bool act = A.active((int32_t)(i >> 6));
    float r;
    if (fresh >> (i & 63) & 1) {
      r = rank[i];
    } else {
      const double s = sums[i & 63];
      r = (float)(s > 0 ? acc[i] / s : acc[i]);
    }
    coef[i] = (act && du[i] > 0) ? damping * r / (float)du[i] : 0.0f;
    acc[i] = act ? base : 0.0;
  }
}
// Column q of [vertex][64] values as a row (query servers' results).
template <class T>
__global__ void __launch_bounds__(kBlock) k_col_out(int32_t n, int32_t q, const T* cols, T* out) {
  for (int64_t v = blockIdx.x * (int64_t)kBlock + threadIdx.x; v < n; v += (int64_t)gridDim.x * kBlock)
    out[v] = cols[v * 64 + q];
}
__global__ void __launch_bounds__(kBlock) k_prb_col_out(int32_t n, int32_t q, const double* acc, const double* sums,
                                                        float* out) {
  const double s = sums[q];
  for (int64_t v = blockIdx.x * (int64_t)kBlock + threadIdx.x; v < n; v += (int64_t)gridDim.x * kBlock)
    out[v] = (float)(s > 0 ? acc[v * 64 + q] / s : acc[v * 64 + q]);
}

__global__ void __launch_bounds__(kBlock) k_prb_coef(Window A, int32_t n, const float* rank, const int32_t* du,
                                                     float damping, float* coef, double base, double* acc) {
  for (int64_t i = blockIdx.x * (int64_t)kBlock + threadIdx.x; i < (int64_t)n * 64; i += (int64_t)gridDim.x * kBlock) {
    const bool act = A.active((int32_t)(i >> 6));
    coef[i] = (act && du[i] > 0) ? damping * rank[i] / (float)du[i] : 0.0f;
    acc[i] = act ? base : 0.0;
  }
}

// One wave per segment: lane q sums query q's terms over the segment's runs,
// in run order.  The runs' descriptors are read 64 at a time (one per lane)
// and handed out by readlane; the 64-wide coefficient rows are requested
// kPrbAhead runs at a time so their loads overlap.  A single-contact run's
// weight is f0 where its time lies in the query's window; a multi-contact
// run's comes from this batch's k_prb_wmulti.  A row of one segment gets
// base + sum in acc, else the sum goes to partial.
constexpr int kPrbAhead = 8;

__device__ __forceinline__ int32_t rl32(int32_t v, int l) { return __builtin_amdgcn_readlane(v, l); }

__global__ void __launch_bounds__(kBlock) k_prb_pull(int64_t nseg, const int32_t* seg_row, const uint64_t* seg_beg,
                                                     const int32_t* seg_len, const uint64_t* seg_off, const int4* runs,
                                                     const float* f0, const int32_t* midx, const float* wm,
                                                     const float* coef, const int64_t* wins, int32_t nq, double base,
                                                     double* acc, double* partial) {
  const int lane = threadIdx.x & 63;
  const bool on = lane < nq;
  const int64_t ws = on ? clamp_lo(wins[2 * lane]) : 1, we = on ? clamp_hi(wins[2 * lane + 1]) : 0;
  const int64_t nw = (int64_t)gridDim.x * (kBlock / 64);
  for (int64_t s = blockIdx.x * (int64_t)(kBlock / 64) + (threadIdx.x >> 6); s < nseg; s += nw) {
    const int32_t r = seg_row[s];
    const uint64_t sb = seg_beg[s];
    const int32_t len = seg_len[s];
    double sum = 0;
    for (int32_t b0 = 0; b0 < len; b0 += 64) {
      const int32_t j = b0 + lane;
      const int4 my = j < len ? runs[sb + j] : int4{0, 0, 0, 0};
      const float myf = j < len ? f0[sb + j] : 0.0f;
      const int32_t mym = j < len ? midx[sb + j] : -1;
      const int32_t cnt = min(64, len - b0);
      for (int32_t u0 = 0; u0 < cnt; u0 += kPrbAhead) {
        float c[kPrbAhead];
#pragma unroll
        for (int u = 0; u < kPrbAhead; ++u)
          c[u] = (u0 + u < cnt) ? coef[(int64_t)rl32(my.x, u0 + u) * 64 + lane] : 0.0f;
#pragma unroll
        for (int u = 0; u < kPrbAhead; ++u) {
          if (u0 + u >= cnt) break;
          const int l = u0 + u;
          const int32_t mi = rl32(mym, l);
          float w;
          bool hit;
          if (mi >= 0) {  // several contacts: this batch's weight, per query
            w = wm[(int64_t)mi * 64 + lane];
            hit = w != 0.0f;
          } else {
            const int32_t t0 = rl32(my.w, l);
            hit = t0 >= ws && t0 <= we;
            w = hit ? __int_as_float(rl32(__float_as_int(myf), l)) : 0.0f;
          }
          if (hit) sum += (double)(c[u] * w);
        }
      }
    }
    if (seg_off[r + 1] - seg_off[r] > 1)
      partial[s * 64 + lane] = sum;
    else
      acc[(int64_t)r * 64 + lane] = base + sum;
  }
}

// Iterations after the first: rank = float(acc / sum) of the previous
// iteration (the normalisation), then coef and base as k_prb_coef.
__global__ void __launch_bounds__(kBlock) k_prb_step(Window A, int32_t n, const double* sums, const int32_t* du,
                                                     float damping, float* coef, double base, double* acc) {
  for (int64_t i = blockIdx.x * (int64_t)kBlock + threadIdx.x; i < (int64_t)n * 64; i += (int64_t)gridDim.x * kBlock) {
    const bool act = A.active((int32_t)(i >> 6));
    const double s = sums[i & 63];
    const float rank = (float)(s > 0 ? acc[i] / s : acc[i]);
    coef[i] = (act && du[i] > 0) ? damping * rank / (float)du[i] : 0.0f;
    acc[i] = act ? base : 0.0;
  }
}

__global__ void __launch_bounds__(kBlock) k_prb_fix(const int32_t* mrows, const uint32_t* nm, const uint64_t* seg_off,
                                                    const double* partial, double base, double* acc) {
  const int lane = threadIdx.x & 63;
  const uint32_t cnt = *nm;
  const uint32_t nw = gridDim.x * (kBlock / 64);
  for (uint32_t i = blockIdx.x * (kBlock / 64) + (threadIdx.x >> 6); i < cnt; i += nw) {
    const int32_t r = mrows[i];
    double sum = 0;
    for (uint64_t s = seg_off[r]; s < seg_off[r + 1]; ++s) sum += partial[s * 64 + lane];
    acc[(int64_t)r * 64 + lane] = base + sum;
  }
}

// Column sums of acc (fixed grid and order): per-block partials, then one wave.
__global__ void __launch_bounds__(kBlock) k_prb_colsum(int32_t n, const double* acc, double* part) {
  __shared__ double sh[kBlock / 64][64];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  double s = 0;
  for (int64_t v = blockIdx.x * (int64_t)(kBlock / 64) + wave; v < n; v += (int64_t)gridDim.x * (kBlock / 64))
    s += acc[v * 64 + lane];
  sh[wave][lane] = s;
  __syncthreads();
  if (wave == 0) {
    double t = 0;
    for (int w = 0; w < kBlock / 64; ++w) t += sh[w][lane];
    part[(int64_t)blockIdx.x * 64 + lane] = t;
  }
}

__global__ void __launch_bounds__(64) k_prb_colsum2(const double* part, int32_t nb, double* sums) {
  double t = 0;
  for (int b = 0; b < nb; ++b) t += part[(int64_t)b * 64 + threadIdx.x];
  sums[threadIdx.x] = t;
}

// --------------------------------------------- batched connected components ---
//
// Up to 64 min-label queries (windows) over the same run structure, with
// int32 label columns label[v][q] (active v starts at v, inactive at -1).  A
// run x -> r links x to r in query q when one of its contacts lies in window
// q: for a single-contact run its time t0, for a multi-contact run this
// batch's bit mask (k_ccb_mask).  A round is one wave per segment: lane q
// takes the least label[x][q] over the linked runs and lowers label[r][q] to
// it.  Labels are updated in place -- they only fall, and every value read is
// the label of a vertex that reaches r, so the fixpoint (the least active id
// reaching each vertex) is the one of the plain rounds.  Rows of several
// segments go through partial minima and k_ccb_fix.  Rounds repeat until none
// lowers a label.
__global__ void __launch_bounds__(kBlock) k_ccb_mask(const int32_t* mlist, int32_t nm, const int4* runs,
                                                     const uint64_t* coff, const int64_t* ct, const int64_t* wins,
                                                     int32_t nq, uint64_t* mask, uint64_t only) {
  const int lane = threadIdx.x & 63;
  const bool on = lane < nq;
  const int64_t ws = on ? clamp_lo(wins[2 * lane]) : 1, we = on ? clamp_hi(wins[2 * lane + 1]) : 0;
  const uint64_t all = (nq >= 64 ? ~0ull : ((1ull << nq) - 1)) & only;
  const int64_t nw = (int64_t)gridDim.x * (kBlock / 64);
  for (int64_t i = blockIdx.x * (int64_t)(kBlock / 64) + (threadIdx.x >> 6); i < nm; i += nw) {
    const int4 run = runs[mlist[i]];
    const int64_t* tb = ct + coff[run.x];
    uint64_t bits = 0;
    for (int32_t b0 = run.y; b0 < run.z && bits != all; b0 += 64) {
      const int32_t k = b0 + lane;
      const int64_t t = k < run.z ? tb[k] : kNoTime;
      for (int q = 0; q < nq; ++q) {
        if ((bits | ~all) >> q & 1) continue;
        const int64_t qs = readlane64(ws, q), qe = readlane64(we, q);
        if (__ballot(t >= qs && t <= qe)) bits |= 1ull << q;
      }
    }
    if (lane == 0) mask[i] = only == ~0ull ? bits : ((mask[i] & ~only) | bits);
  }
}

__global__ void __launch_bounds__(kBlock) k_ccb_init(Window A, int32_t n, int32_t* label, uint64_t only) {
  for (int64_t i = blockIdx.x * (int64_t)kBlock + threadIdx.x; i < (int64_t)n * 64; i += (int64_t)gridDim.x * kBlock) {
    const int32_t v = (int32_t)(i >> 6);
    if (only >> (i & 63) & 1) label[i] = A.active(v) ? v : -1;
  }
}

__global__ void __launch_bounds__(kBlock) k_ccb_pull(int64_t nseg, const int32_t* seg_row, const uint64_t* seg_beg,
                                                     const int32_t* seg_len, const uint64_t* seg_off, const int4* runs,
                                                     const int32_t* midx, const uint64_t* mask, const int64_t* wins,
                                                     int32_t nq, int32_t* label, int32_t* partial, int32_t* changed,
                                                     const int32_t* run_if, unsigned long long* lane_changed) {
  if (run_if && *run_if == 0) return;  // an earlier round of the group lowered nothing: converged
  const int lane = threadIdx.x & 63;
  const bool on = lane < nq;
  const int64_t ws = on ? clamp_lo(wins[2 * lane]) : 1, we = on ? clamp_hi(wins[2 * lane + 1]) : 0;
  const int64_t nw = (int64_t)gridDim.x * (kBlock / 64);
  for (int64_t s = blockIdx.x * (int64_t)(kBlock / 64) + (threadIdx.x >> 6); s < nseg; s += nw) {
    const int32_t r = seg_row[s];
    const uint64_t sb = seg_beg[s];
    const int32_t len = seg_len[s];
    int32_t cur = INT32_MAX;
    for (int32_t b0 = 0; b0 < len; b0 += 64) {
      const int32_t j = b0 + lane;
      const int4 my = j < len ? runs[sb + j] : int4{0, 0, 0, 0};
      const int32_t mym = j < len ? midx[sb + j] : -1;
      const int32_t cnt = min(64, len - b0);
      for (int32_t u0 = 0; u0 < cnt; u0 += kPrbAhead) {
        int32_t lab[kPrbAhead];
#pragma unroll
        for (int u = 0; u < kPrbAhead; ++u)
          lab[u] = (u0 + u < cnt) ? label[(int64_t)rl32(my.x, u0 + u) * 64 + lane] : INT32_MAX;
#pragma unroll
        for (int u = 0; u < kPrbAhead; ++u) {
          if (u0 + u >= cnt) break;
          const int l = u0 + u;
          const int32_t mi = rl32(mym, l);
          bool hit;
          if (mi >= 0) {
            hit = mask[mi] >> lane & 1;
          } else {
            const int32_t t0 = rl32(my.w, l);
            hit = t0 >= ws && t0 <= we;
          }
          if (hit) cur = min(cur, lab[u]);
        }
      }
    }
    bool lowered = false;
    if (seg_off[r + 1] - seg_off[r] > 1) {
      partial[s * 64 + lane] = cur;
    } else if (on) {
      int32_t* p = label + (int64_t)r * 64 + lane;
      if (cur < *p) {
        *p = cur;
        lowered = true;
      }
    }
    const uint64_t low = __ballot(lowered);
    if (low && lane == 0) {
      *changed = 1;
      if (lane_changed) atomicOr(lane_changed, (unsigned long long)low);
    }
  }
}

__global__ void __launch_bounds__(kBlock) k_ccb_fix(const int32_t* mrows, const uint32_t* nm, const uint64_t* seg_off,
                                                    const int32_t* partial, int32_t nq, int32_t* label,
                                                    int32_t* changed, const int32_t* run_if,
                                                    unsigned long long* lane_changed) {
  if (run_if && *run_if == 0) return;
  const int lane = threadIdx.x & 63;
  const uint32_t cnt = *nm;
  const uint32_t nw = gridDim.x * (kBlock / 64);
  for (uint32_t i = blockIdx.x * (kBlock / 64) + (threadIdx.x >> 6); i < cnt; i += nw) {
    const int32_t r = mrows[i];
    int32_t cur = INT32_MAX;
    for (uint64_t s = seg_off[r]; s < seg_off[r + 1]; ++s) cur = min(cur, partial[s * 64 + lane]);
    bool lowered = false;
    if (lane < nq) {
      int32_t* p = label + (int64_t)r * 64 + lane;
      if (cur < *p) {
        *p = cur;
        lowered = true;
      }
    }
    const uint64_t low = __ballot(lowered);
    if (low && lane == 0) {
      *changed = 1;
      if (lane_changed) atomicOr(lane_changed, (unsigned long long)low);
    }
  }
}

// out[q][v] = label[v][q] for q < nq.
__global__ void __launch_bounds__(kBlock) k_ccb_out(int32_t n, int32_t nq, const int32_t* label, int32_t* out) {
  for (int64_t i = blockIdx.x * (int64_t)kBlock + threadIdx.x; i < (int64_t)n * 64; i += (int64_t)gridDim.x * kBlock) {
    const int32_t q = (int32_t)(i & 63);
    if (q < nq) out[(int64_t)q * n + (i >> 6)] = label[i];
  }
}

// ------------------------------------------------------- batched HITS ---
//
// Up to 64 HITS queries as FP32 hub and authority columns.  The authority
// step pulls over the destination-sorted runs (in-runs of r), the hub step
// over the same runs grouped by source in fill order (out-runs of x: per
// run its destination, or -1 when inactive, and its index among the sorted
// runs, which keys the single-contact time and the multi-contact link mask of
// batched CC).  Rows are cut into segments like the in-rows; both steps are
// followed by fixed-order L2 column norms.
__global__ void __launch_bounds__(kBlock) k_run_inv(uint64_t nv, const uint64_t* sv, int32_t* inv) {
  for (int64_t j = blockIdx.x * (int64_t)kBlock + threadIdx.x; j < (int64_t)nv; j += (int64_t)gridDim.x * kBlock)
    inv[sv[j]] = (int32_t)j;
}

// Out-run i (fill order): (r or -1, sorted index or -1, t0, 0).
__global__ void __launch_bounds__(kBlock) k_out_runs(int64_t nr, const int32_t* rr, const int32_t* inv,
                                                     const int4* runs, int4* out) {
  for (int64_t i = blockIdx.x * (int64_t)kBlock + threadIdx.x; i < nr; i += (int64_t)gridDim.x * kBlock) {
    const int32_t j = inv[i];
    out[i] = j >= 0 ? int4{rr[i], j, runs[j].w, 0} : int4{-1, -1, 0, 0};
  }
}

// Out-row of every active x: its items' runs (one range: its chunks are
// consecutive items), and its segment count.
__global__ void __launch_bounds__(kBlock) k_out_rows(Window A, int32_t n, const int32_t* vinfo, const Ckpt* ck,
                                                     const uint64_t* off, uint64_t* obeg, int32_t* olen,
                                                     uint32_t* nseg) {
  for (int64_t v = blockIdx.x * (int64_t)kBlock + threadIdx.x; v < n; v += (int64_t)gridDim.x * kBlock) {
    uint64_t b = 0, e = 0;
    if (A.active((int32_t)v)) {
      const int32_t info = vinfo[v];
      if (info >= 0) {
        b = off[n + info];
        e = off[n + info + ck[info].nchunk];
      } else {
        b = off[v];
        e = off[v + 1];
      }
    }
    obeg[v] = b;
    olen[v] = (int32_t)(e - b);
    nseg[v] = (uint32_t)((e - b + kPrbSeg - 1) / kPrbSeg);
  }
}

__global__ void __launch_bounds__(kBlock) k_oseg_fill(int32_t n, const uint64_t* obeg, const int32_t* olen,
                                                      const uint64_t* seg_off, int32_t* seg_row, uint64_t* seg_beg,
                                                      int32_t* seg_len, int32_t* mrows, uint32_t* nm) {
  for (int64_t r = blockIdx.x * (int64_t)kBlock + threadIdx.x; r < n; r += (int64_t)gridDim.x * kBlock) {
    const uint64_t s0 = seg_off[r], s1 = seg_off[r + 1];
    for (uint64_t s = s0; s < s1; ++s) {
      const int64_t j = (int64_t)(s - s0) * kPrbSeg;
      seg_row[s] = (int32_t)r;
      seg_beg[s] = obeg[r] + (uint64_t)j;
      seg_len[s] = (int32_t)min((int64_t)kPrbSeg, (int64_t)olen[r] - j);
    }
    if (s1 - s0 > 1) mrows[atomicAdd(nm, 1u)] = (int32_t)r;
  }
}

// One wave per segment: lane q sums (double)vec[y][q] over the segment's
// runs linked in query q, y the run's other end.  OUT: out-runs (y = r, -1
// for an inactive r); else in-runs (y = x).  Single-segment rows store the
// sum in acc, others in partial (k_prb_fix adds them with base 0).
template <int OUT>
__global__ void __launch_bounds__(kBlock) k_hitsb_pull(int64_t nseg, const int32_t* seg_row, const uint64_t* seg_beg,
                                                       const int32_t* seg_len, const uint64_t* seg_off,
                                                       const int4* rr, const int32_t* midx, const uint64_t* mask,
                                                       const int64_t* wins, int32_t nq, const float* vec,
                                                       double* acc, double* partial) {
  const int lane = threadIdx.x & 63;
  const bool on = lane < nq;
  const int64_t ws = on ? clamp_lo(wins[2 * lane]) : 1, we = on ? clamp_hi(wins[2 * lane + 1]) : 0;
  const int64_t nw = (int64_t)gridDim.x * (kBlock / 64);
  for (int64_t s = blockIdx.x * (int64_t)(kBlock / 64) + (threadIdx.x >> 6); s < nseg; s += nw) {
    const int32_t row = seg_row[s];
    const uint64_t sb = seg_beg[s];
    const int32_t len = seg_len[s];
    double sum = 0;
    for (int32_t b0 = 0; b0 < len; b0 += 64) {
      const int32_t j = b0 + lane;
      const int4 my = j < len ? rr[sb + j] : int4{-1, -1, 0, 0};
      const int32_t y = my.x;
      const int32_t t0 = OUT ? my.z : my.w;
      const int32_t si = OUT ? my.y : (j < len ? (int32_t)(sb + j) : -1);
      const int32_t mym = si >= 0 ? midx[si] : -1;
      const int32_t cnt = min(64, len - b0);
      for (int32_t u0 = 0; u0 < cnt; u0 += kPrbAhead) {
        float vv[kPrbAhead];
#pragma unroll
        for (int u = 0; u < kPrbAhead; ++u) {
          const int32_t yy = (u0 + u < cnt) ? rl32(y, u0 + u) : -1;
          vv[u] = yy >= 0 ? vec[(int64_t)yy * 64 + lane] : 0.0f;
        }
#pragma unroll
        for (int u = 0; u < kPrbAhead; ++u) {
          if (u0 + u >= cnt) break;
          const int l = u0 + u;
          if (rl32(y, l) < 0) continue;
          const int32_t mi = rl32(mym, l);
          bool hit;
          if (mi >= 0) {
            hit = mask[mi] >> lane & 1;
          } else {
            const int32_t tt = rl32(t0, l);
            hit = tt >= ws && tt <= we;
          }
          if (hit) sum += (double)vv[u];
        }
      }
    }
    if (seg_off[row + 1] - seg_off[row] > 1)
      partial[s * 64 + lane] = sum;
    else
      acc[(int64_t)row * 64 + lane] = sum;
  }
}

constexpr int kGcbPieceBits = 3;
constexpr int32_t kGcbPiece = 64;  // runs per piece
static_assert((kGcbPiece << kGcbPieceBits) >= kPrbSeg, "a segment must fit its pieces");
// Batched graph coloring.  The round rule (an uncolored active vertex takes
// color c in round c unless an outranking active window out-neighbor is
// still uncolored) gives every vertex one more than the largest color of its
// outranking window out-neighbors, and 0 when it has none: the longest path
// in the outranking order, a DAG.  So the colors of 64 windows come from one
// sweep over priority levels, highest first: a level's vertices only need the
// final colors of higher levels, and passes within a level settle the
// (rare) outranking neighbors inside it.  One wave per out-segment of the
// level's vertices, lane q for window q: the largest color + 1 over the
// segment's runs linked in q whose destination outranks the row, merged into
// the row's column with atomicMax.  `changed` records a raised color; a pass
// whose `run_if` flag is 0 returns at once (the level had settled).
__global__ void __launch_bounds__(kBlock) k_gcb_level(const int32_t* segs, int32_t nsegs, const int32_t* seg_row,
                                                      const uint64_t* seg_beg, const int32_t* seg_len,
                                                      const int4* rr, const int32_t* midx, const uint64_t* mask,
                                                      const int64_t* wins, int32_t nq, const int32_t* prio,
                                                      int32_t* col, int32_t* changed, const int32_t* run_if) {
  if (run_if && *run_if == 0) return;
  const int lane = threadIdx.x & 63;
  const bool on = lane < nq;
  const int64_t ws = on ? clamp_lo(wins[2 * lane]) : 1, we = on ? clamp_hi(wins[2 * lane + 1]) : 0;
  const int32_t nw = gridDim.x * (kBlock / 64);
  for (int32_t i = blockIdx.x * (kBlock / 64) + (threadIdx.x >> 6); i < nsegs; i += nw) {
    // a piece of kGcbPiece runs of out-segment s (short pieces keep the
    // dependent gathers of one wave few, and levels run one after another)
    const int32_t s = segs[i] >> kGcbPieceBits, piece = segs[i] & ((1 << kGcbPieceBits) - 1);
    const int32_t row = seg_row[s];
    const int32_t pv = prio[row];
    const uint64_t sb = seg_beg[s] + (uint64_t)piece * kGcbPiece;
    const int32_t len = min(kGcbPiece, seg_len[s] - piece * kGcbPiece);
    int32_t cur = 0;
    for (int32_t b0 = 0; b0 < len; b0 += 64) {
      const int32_t j = b0 + lane;
      const int4 my = j < len ? rr[sb + j] : int4{-1, -1, 0, 0};
      // the run's destination outranks the row (priority, then the smaller
      // id); other runs are dropped from the wave's list
      const int32_t y = (my.x >= 0 && (prio[my.x] > pv || (prio[my.x] == pv && my.x < row))) ? my.x : -1;
      const uint64_t upm = __ballot(y >= 0);
      const int32_t mym = (y >= 0 && my.y >= 0) ? midx[my.y] : -1;
      const int32_t nup = __popcll(upm);
      // the outranking runs, compacted in lane order: the k-th is lane kth[k]
      const int32_t rank = __popcll(upm & ((1ull << lane) - 1));
      __shared__ int32_t kth_lds[kBlock];
      int32_t* kth = kth_lds + (threadIdx.x & ~63);
      if (y >= 0) kth[rank] = lane;
      __builtin_amdgcn_wave_barrier();
      for (int32_t u0 = 0; u0 < nup; u0 += kPrbAhead) {
        int32_t c[kPrbAhead];
#pragma unroll
        for (int u = 0; u < kPrbAhead; ++u) {
          const int l = __builtin_amdgcn_readfirstlane(u0 + u < nup ? kth[u0 + u] : 0);  // same in every lane
          const int32_t yy = rl32(y, l);
          c[u] = (u0 + u < nup && on) ? col[(int64_t)yy * 64 + lane] : -1;
        }
#pragma unroll
        for (int u = 0; u < kPrbAhead; ++u) {
          if (u0 + u >= nup) break;
          const int l = __builtin_amdgcn_readfirstlane(kth[u0 + u]);
          const int32_t mi = rl32(mym, l);
          bool hit;
          if (mi >= 0) {
            hit = mask[mi] >> lane & 1;
          } else {
            const int32_t tt = rl32(my.z, l);
            hit = tt >= ws && tt <= we;
          }
          if (hit) cur = max(cur, c[u] + 1);
        }
      }
      __builtin_amdgcn_wave_barrier();
    }
    bool raised = false;
    if (on && cur > 0) raised = atomicMax(&col[(int64_t)row * 64 + lane], cur) < cur;
    if (__ballot(raised) && lane == 0) *changed = 1;
  }
}

// Out-segments whose row has an outranking run inside its own level (the
// only ones a second pass of the level can change): flag[s] = 1.
__global__ void __launch_bounds__(kBlock) k_gcb_inner(int64_t nseg, const int32_t* seg_row, const uint64_t* seg_beg,
                                                      const int32_t* seg_len, const int4* rr, const int32_t* prio,
                                                      const int32_t* vlevel, int32_t* flag) {
  const int lane = threadIdx.x & 63;
  const int64_t nw = (int64_t)gridDim.x * (kBlock / 64);
  for (int64_t s = blockIdx.x * (int64_t)(kBlock / 64) + (threadIdx.x >> 6); s < nseg; s += nw) {
    const int32_t row = seg_row[s];
    const int32_t pv = prio[row], lv = vlevel[row];
    const uint64_t sb = seg_beg[s];
    const int32_t len = seg_len[s];
    bool in = false;
    for (int32_t j = lane; j < len; j += 64) {
      const int32_t y = rr[sb + j].x;
      if (y >= 0 && vlevel[y] == lv && (prio[y] > pv || (prio[y] == pv && y < row))) in = true;
    }
    const bool any = __ballot(in) != 0;  // every lane of the wave votes
    if (lane == 0) flag[s] = any;
  }
}

// Colors past the round limit stay uncolored (kInf), as do inactive vertices.
__global__ void __launch_bounds__(kBlock) k_gcb_out(Window A, int32_t n, int32_t nq, int32_t rounds,
                                                    const int32_t* col, int32_t* out) {
  for (int64_t i = blockIdx.x * (int64_t)kBlock + threadIdx.x; i < (int64_t)n * nq;
       i += (int64_t)gridDim.x * kBlock) {
    const int32_t q = (int32_t)(i / n), v = (int32_t)(i % n);
    const int32_t c = col[(int64_t)v * 64 + q];
    out[i] = (A.active(v) && c < rounds) ? c : kInf;
  }
}

// Column sums of squares (k_prb_colsum's fixed order) and their square roots.
__global__ void __launch_bounds__(kBlock) k_colsq(int32_t n, const double* acc, double* part) {
  __shared__ double sh[kBlock / 64][64];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  double s = 0;
  for (int64_t v = blockIdx.x * (int64_t)(kBlock / 64) + wave; v < n; v += (int64_t)gridDim.x * (kBlock / 64)) {
    const double x = acc[v * 64 + lane];
    s += x * x;
  }
  sh[wave][lane] = s;
  __syncthreads();
  if (wave == 0) {
    double t = 0;
    for (int w = 0; w < kBlock / 64; ++w) t += sh[w][lane];
    part[(int64_t)blockIdx.x * 64 + lane] = t;
  }
}

__global__ void __launch_bounds__(64) k_colsq2(const double* part, int32_t nb, double* norms) {
  double t = 0;
  for (int b = 0; b < nb; ++b) t += part[(int64_t)b * 64 + threadIdx.x];
  norms[threadIdx.x] = sqrt(t);
}

__global__ void __launch_bounds__(kBlock) k_colscale(int32_t n, const double* acc, const double* norms, float* out) {
  for (int64_t i = blockIdx.x * (int64_t)kBlock + threadIdx.x; i < (int64_t)n * 64; i += (int64_t)gridDim.x * kBlock) {
    const double s = norms[i & 63];
    out[i] = (float)(s > 0 ? acc[i] / s : acc[i]);
  }
}

__global__ void __launch_bounds__(kBlock) k_cols_init(Window A, int32_t n, float v, float* x, float* y) {
  for (int64_t i = blockIdx.x * (int64_t)kBlock + threadIdx.x; i < (int64_t)n * 64; i += (int64_t)gridDim.x * kBlock) {
    const float a = A.active((int32_t)(i >> 6)) ? v : 0.0f;
    x[i] = a;
    y[i] = a;
  }
}

// out[q][v] = col[v][q] for q < nq.
__global__ void __launch_bounds__(kBlock) k_cols_out(int32_t n, int32_t nq, const float* col, float* out) {
  for (int64_t i = blockIdx.x * (int64_t)kBlock + threadIdx.x; i < (int64_t)n * 64; i += (int64_t)gridDim.x * kBlock) {
    const int32_t q = (int32_t)(i & 63);
    if (q < nq) out[(int64_t)q * n + (i >> 6)] = col[i];
  }
}

// Active vertices with more than kPrbLongDeg contacts (unordered).
__global__ void __launch_bounds__(kBlock) k_long_list(Window A, int32_t n, const int32_t* vdeg, int32_t* list,
                                                      uint32_t* cnt) {
  for (int64_t v = blockIdx.x * (int64_t)kBlock + threadIdx.x; v < n; v += (int64_t)gridDim.x * kBlock)
    if (A.active((int32_t)v) && vdeg[v] > kPrbLongDeg) list[atomicAdd(cnt, 1u)] = (int32_t)v;
}

// out[q][v] = rank of v for query q < nq: float(acc / sum) after the last
// iteration (sums non-null), else the initial rank.
__global__ void __launch_bounds__(kBlock) k_prb_out(int32_t n, int32_t nq, const double* acc, const double* sums,
                                                    const float* rank, float* out) {
  for (int64_t i = blockIdx.x * (int64_t)kBlock + threadIdx.x; i < (int64_t)n * 64; i += (int64_t)gridDim.x * kBlock) {
    const int32_t q = (int32_t)(i & 63);
    if (q >= nq) continue;
    float r = rank[i];
    if (sums) {
      const double s = sums[q];
      r = (float)(s > 0 ? acc[i] / s : acc[i]);
    }
    out[(int64_t)q * n + (i >> 6)] = r;
  }
}

// The running mean of long records is sequential: one thread per record.
template <int MAXD>
__global__ void __launch_bounds__(64) k_dc_long(Streams S, const int32_t* bigv, int32_t nbig, Window W,
                                                int32_t* centr) {
  LANE_RING();
  const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nbig) return;
  DcAct a{W, nullptr, nullptr};
  const int32_t x = bigv[i];
  run_item<MAXD>(S, nullptr, x, a, ring_, stage_);
  centr[x] = W.active(x) ? dc_formula(a.deg, a.avg, W.ws, W.we) : dc_formula(0, 0, W.ws, W.we);
}

// Adds the wave's counts into *out (one atomic per wave).
__device__ __forceinline__ void add_wave_count(uint32_t c, uint32_t* out) {
  for (int off = 32; off > 0; off >>= 1) c += __shfl_xor(c, off, 64);
  if ((threadIdx.x & 63) == 0 && c) atomicAdd(out, c);
}

// BFS level step: every unreached active neighbor over a window contact gets
// level lvl + 1 (all writers of a pass store the same value, so plain stores
// suffice); the next level is selected by its level value.
struct BfsAct {
  Window W;
  int32_t* level;
  int32_t lvl;
  __device__ bool begin(int32_t x, bool, Split) { return W.active(x); }
  __device__ bool contact(int32_t, int32_t r, int32_t t, int32_t) {
    if (W.in(t) && W.active(r) && level[r] == kInf) level[r] = lvl + 1;
    return true;
  }
  __device__ void end(int32_t, bool) {}
};

template <int MAXD>
__global__ void __launch_bounds__(kBlock) k_bfs(Streams S, const Ckpt* ck, Items I, Window W, int32_t* level,
                                                int32_t lvl) {
  BfsAct a{W, level, lvl};
  LANE_RING();
  FOR_ITEMS(I, i) run_item<MAXD>(S, ck, I.code[i], a, ring_, I.stage(stage_, i));
}

// The same step for a small level that also lists the next one: a level is
// claimed with a compare-and-swap so that exactly one lane reaches each
// vertex, and that lane appends it to `next` (wave-aggregated: one counter
// atomic per wave and contact step), so the next level expands the list
// instead of scanning every vertex.  The swap's returned value is waited for,
// which a large level would pay for at every discovery.
struct BfsListAct {
  Window W;
  int32_t* level;
  int32_t lvl;
  int32_t* next;    // next level's vertices (at most one entry per vertex)
  uint32_t* nnext;  // their count
  __device__ bool begin(int32_t x, bool, Split) { return W.active(x); }
  __device__ bool contact(int32_t, int32_t r, int32_t t, int32_t) {
    const bool won = W.in(t) && W.active(r) && level[r] == kInf && atomicCAS(&level[r], kInf, lvl + 1) == kInf;
    const uint64_t m = __ballot(won);
    if (m) {
      const int lane = (int)(threadIdx.x & 63);
      const int first = __ffsll((unsigned long long)m) - 1;
      uint32_t at = 0;
      if (lane == first) at = atomicAdd(nnext, (uint32_t)__popcll(m));
      at = (uint32_t)__shfl((int32_t)at, first, 64);
      if (won) next[at + (uint32_t)__popcll(m & ((1ull << lane) - 1))] = r;
    }
    return true;
  }
  __device__ void end(int32_t, bool) {}
};

template <int MAXD>
__global__ void __launch_bounds__(kBlock) k_bfsl(Streams S, const Ckpt* ck, Items I, Window W, int32_t* level,
                                                 int32_t lvl, int32_t* next, uint32_t* nnext) {
  BfsListAct a{W, level, lvl, next, nnext};
  LANE_RING();
  FOR_ITEMS(I, i) run_item<MAXD>(S, ck, I.code[i], a, ring_, I.stage(stage_, i));
}

// Min-label propagation: a vertex whose label drops in round r is stamped r
// and expanded in round r + 1.  A round runs either over a queue of the
// stamped vertices' items (k_expand) or, when many are stamped, over the
// static lists with `sel` set: an item then runs when its vertex's stamp is
// r - 1 as the item begins.  A vertex stamped again during the round before
// its item begins waits for the next round; labels only fall and every value
// is the label of a vertex reaching the one it is stored at, so the fixpoint
// (the least active id reaching each vertex) is the same.  The kernel counts
// the vertices it stamps (nx; a race may count one twice), which the host
// uses to pick the next round's form and to stop.
template <bool LIST>
struct CcActT {
  Window W;
  int32_t* label;
  int32_t* stamp;
  int32_t round;
  const int32_t* sel;  // stamp, in a round over the static lists
  int32_t* next;       // LIST: the stamped vertices, appended (at most cap)
  uint32_t* nnext;     // LIST: their count
  uint32_t cap;
  uint32_t nx = 0;     // vertices stamped (not LIST: summed per wave into *nnext)
  int32_t lx = 0;
  __device__ bool begin(int32_t x, bool, Split) {
    if (!W.active(x) || (sel && sel[x] != round - 1)) return false;
    lx = label[x];
    return true;
  }
  __device__ bool contact(int32_t, int32_t r, int32_t t, int32_t) {
    bool st = false;
    if (W.in(t) && W.active(r) && lx < label[r]) {
      const int32_t old = atomicMin(&label[r], lx);
      if (lx < old && stamp[r] != round) {
        stamp[r] = round;
        st = true;
      }
    }
    if constexpr (LIST) {
      const uint64_t m = __ballot(st);
      if (m) {
        const int lane = (int)(threadIdx.x & 63);
        const int first = __ffsll((unsigned long long)m) - 1;
        uint32_t at = 0;
        if (lane == first) at = atomicAdd(nnext, (uint32_t)__popcll(m));
        at = (uint32_t)__shfl((int32_t)at, first, 64) + (uint32_t)__popcll(m & ((1ull << lane) - 1));
        if (st && at < cap) next[at] = r;
      }
    } else {
      nx += st;
    }
    return true;
  }
  __device__ void end(int32_t, bool) {}
};
using CcAct = CcActT<false>;

template <int MAXD>
__global__ void __launch_bounds__(kBlock) k_cc(Streams S, const Ckpt* ck, Items I, Window W, int32_t* label,
                                               int32_t* stamp, int32_t round, const int32_t* sel,
                                               uint32_t* next_count) {
  CcAct a{W, label, stamp, round, sel, nullptr, nullptr, 0};
  LANE_RING();
  FOR_ITEMS(I, i) run_item<MAXD>(S, ck, I.code[i], a, ring_, I.stage(stage_, i));
  add_wave_count(a.nx, next_count);
}

// A sparse round that also lists the vertices it stamps (a vertex stamped by
// two lanes at once may be listed twice, which only repeats its items), so
// that the next round expands the list instead of scanning every vertex.
template <int MAXD>
__global__ void __launch_bounds__(kBlock) k_ccl(Streams S, const Ckpt* ck, Items I, Window W, int32_t* label,
                                                int32_t* stamp, int32_t round, int32_t* next, uint32_t* nnext,
                                                uint32_t cap) {
  CcActT<true> a{W, label, stamp, round, nullptr, next, nnext, cap};
  LANE_RING();
  FOR_ITEMS(I, i) run_item<MAXD>(S, ck, I.code[i], a, ring_, I.stage(stage_, i));
}

// Graph coloring round: a neighbor blocks when it has higher priority and is
// uncolored at the start of the round (its color is kInf, or `round` if it was
// colored earlier in this same round).  Chunks of a long record report to
// blocked[x]; k_gc_finish decides those vertices.  A blocked vertex caches
// its blocker in blocker[x] and is not decoded again while that blocker
// stays uncolored (one uncolored higher-priority neighbor is enough to block,
// and priorities, windows and activity do not change between rounds).
struct GcAct {
  Window W;
  const int32_t* prio;
  int32_t* color;
  int32_t* blocked;
  int32_t* blocker;
  int32_t round;
  int32_t v = 0, pv = 0, bu = -1;
  bool blk = false;
  __device__ bool begin(int32_t x, bool chunked, Split) {
    if (!W.active(x) || color[x] != kInf) return false;
    if (chunked && blocked[x]) return false;  // another chunk already found a blocker
    v = x;
    pv = prio[x];
    blk = false;
    return true;
  }
  __device__ bool contact(int32_t, int32_t u, int32_t t, int32_t) {
    if (W.in(t) && W.active(u)) {
      const int32_t pu = prio[u];
      if ((pu > pv || (pu == pv && u < v)) && color[u] >= round) {
        blk = true;
        bu = u;
        return false;
      }
    }
    return true;
  }
  __device__ void end(int32_t x, bool chunked) {
    if (blk) blocker[x] = bu;  // any chunk's blocker will do
    if (chunked) {
      if (blk) atomicOr(&blocked[x], 1);
    } else if (!blk) {
      color[x] = round;
    }
  }
};

template <int MAXD>
__global__ void __launch_bounds__(kBlock) k_gc(Streams S, const Ckpt* ck, Items I, Window W, const int32_t* prio,
                                               int32_t* color, int32_t* blocked, int32_t* blocker, int32_t round) {
  GcAct a{W, prio, color, blocked, blocker, round};
  LANE_RING();
  FOR_ITEMS(I, i) run_item<MAXD>(S, ck, I.code[i], a, ring_, I.stage(stage_, i));
}

__global__ void __launch_bounds__(kBlock) k_gc_finish(const int32_t* bigq, uint32_t nbig, Window W, int32_t* color,
                                                      int32_t* blocked, int32_t round) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < nbig; i += gridDim.x * blockDim.x) {
    const int32_t x = bigq[i];
    if (!W.active(x) || color[x] != kInf) continue;
    if (blocked[x])
      blocked[x] = 0;
    else
      color[x] = round;
  }
}

// ------------------------------------------- graph coloring over lists ---
//
// The rounds above decode every candidate again.  When memory allows, the
// records are decoded twice in all instead: a counting pass and a filling
// pass write, for each active vertex x, the list of window contacts to an
// active u that outranks x (priority[u] > priority[x], ties to the smaller
// id).  x is blocked in round c iff some u on its list has color[u] >= c.
// Colors never change once set, so an entry with color[u] < c never blocks
// again: each vertex keeps a cursor that only moves forward, and a round is
// one scan step per candidate (a whole wave per candidate for long lists).
//
// List layout: item index i = code for a whole record, n + j for checkpoint
// chunk j; off[] is the exclusive scan of the per-item counts, so the list
// of a chunked vertex (its chunks are consecutive checkpoints) is one range.

struct GcListAct {
  Window W;
  const int32_t* prio;
  uint32_t* cnt;       // counting pass: per item
  const uint64_t* off;  // filling pass: per item
  int32_t* hp;
  int32_t n;
  uint64_t pos = 0;
  uint32_t c = 0;
  int32_t idx = 0, v = 0, pv = 0;
  __device__ void item(int32_t code) { idx = item_index(code, n); }
  __device__ bool begin(int32_t x, bool, Split) {
    if (!W.active(x)) return false;
    v = x;
    pv = prio[x];
    c = 0;
    if (off) pos = off[idx];
    return true;
  }
  __device__ bool contact(int32_t, int32_t u, int32_t t, int32_t) {
    if (W.in(t) && W.active(u)) {
      const int32_t pu = prio[u];
      if (pu > pv || (pu == pv && u < v)) {
        if (off)
          hp[pos++] = u;
        else
          ++c;
      }
    }
    return true;
  }
  __device__ void end(int32_t, bool) {
    if (!off) cnt[idx] = c;
  }
};

template <int MAXD>
__global__ void __launch_bounds__(kBlock) k_gc_list(Streams S, const Ckpt* ck, Items I, Window W, const int32_t* prio,
                                                    uint32_t* cnt, const uint64_t* off, int32_t* hp) {
  GcListAct a{W, prio, cnt, off, hp, S.n};
  LANE_RING();
  FOR_ITEMS(I, i) run_item<MAXD>(S, ck, I.code[i], a, ring_, I.stage(stage_, i));
}

// Exclusive prefix of v over the block (thread order); total receives the sum.
__device__ __forceinline__ uint64_t block_excl_scan64(uint64_t v, uint64_t& total) {
  __shared__ uint64_t ws[kBlock / 64];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  uint64_t inc = v;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const uint64_t u = __shfl_up((unsigned long long)inc, d, 64);
    if (lane >= d) inc += u;
  }
  if (lane == 63) ws[wave] = inc;
  __syncthreads();
  uint64_t before = 0;
  total = 0;
  for (int w = 0; w < kBlock / 64; ++w) {
    if (w < wave) before += ws[w];
    total += ws[w];
  }
  __syncthreads();
  return before + inc - v;
}

// Exclusive scan of m 32-bit counts into m + 1 64-bit offsets, in tiles of
// kScanTile: tile sums, a one-block scan of the tile sums, then each tile.
constexpr int kScanPer = 8;
constexpr int kScanTile = kBlock * kScanPer;

__device__ __forceinline__ uint64_t scan_thread_sum(const uint32_t* c, int64_t m, int64_t b) {
  uint64_t s = 0;
#pragma unroll
  for (int j = 0; j < kScanPer; ++j) s += b + j < m ? c[b + j] : 0u;
  return s;
}

__global__ void __launch_bounds__(kBlock) k_scan_tiles(const uint32_t* c, int64_t m, uint64_t* tsum) {
  const int64_t b = (int64_t)blockIdx.x * kScanTile + (int64_t)threadIdx.x * kScanPer;
  uint64_t total;
  block_excl_scan64(scan_thread_sum(c, m, b), total);
  if (threadIdx.x == 0) tsum[blockIdx.x] = total;
}

__global__ void __launch_bounds__(kBlock) k_scan_top(uint64_t* tsum, int64_t tiles) {
  uint64_t carry = 0;
  for (int64_t b = 0; b < tiles; b += kBlock) {
    const int64_t i = b + threadIdx.x;
    const uint64_t v = i < tiles ? tsum[i] : 0;
    uint64_t total;
    const uint64_t pre = block_excl_scan64(v, total);
    if (i < tiles) tsum[i] = carry + pre;
    carry += total;
  }
}

__global__ void __launch_bounds__(kBlock) k_scan_apply(const uint32_t* c, int64_t m, const uint64_t* tpre,
                                                       uint64_t* off) {
  const int64_t b = (int64_t)blockIdx.x * kScanTile + (int64_t)threadIdx.x * kScanPer;
  uint64_t total;
  uint64_t run = tpre[blockIdx.x] + block_excl_scan64(scan_thread_sum(c, m, b), total);
#pragma unroll
  for (int j = 0; j < kScanPer; ++j) {
    if (b + j <= m) off[b + j] = run;
    if (b + j < m) run += c[b + j];
  }
}

// Candidates are (vertex, cursor, blocker, -): the blocker is the list entry
// at the cursor (-1 before the first scan); vertex -1 marks a candidate
// colored since the last compaction.  A round reads the blocker's color
// first; only when it got colored does the list scan resume, and only then
// is the candidate written back (rounds take no atomics: k_gc_compact drops
// the colored candidates every few rounds).  Scans with more than kGcSerial
// entries left are done by the whole wave, 64 entries a step, one such
// candidate after another; shorter ones by their own lane.
constexpr int32_t kGcSerial = 64;

// List range of every active vertex, and the first candidates: every active
// vertex, blocked by its first entry if it has one, in vertex order.
__global__ void __launch_bounds__(kBlock) k_gc_ranges(int32_t ab, int32_t ae, int32_t n, const int32_t* vinfo,
                                                      const Ckpt* ck, const uint64_t* off, const int32_t* hp,
                                                      uint64_t* lbeg, int32_t* llen, int4* list, uint32_t* count) {
  if (blockIdx.x == 0 && threadIdx.x == 0) *count = (uint32_t)(ae - ab);
  for (int64_t v = ab + (int64_t)blockIdx.x * kBlock + threadIdx.x; v < ae; v += (int64_t)gridDim.x * kBlock) {
    const int32_t info = vinfo[v];
    uint64_t b, end;
    if (info >= 0) {
      b = off[n + info];
      end = off[n + info + ck[info].nchunk];
    } else {
      b = off[v];
      end = off[v + 1];
    }
    lbeg[v] = b;
    llen[v] = (int32_t)(end - b);
    list[v - ab] = int4{(int32_t)v, 0, end > b ? hp[b] : -1, 0};
  }
}

// One coloring round over list[0, *cnt), in place.
__global__ void __launch_bounds__(kBlock) k_gc_round(int4* list, const uint32_t* cnt_p, const int32_t* hp,
                                                     const uint64_t* lbeg, const int32_t* llen, int32_t* color,
                                                     int32_t round) {
  const int lane = threadIdx.x & 63;
  const uint32_t cnt = *cnt_p;
  for (uint32_t i0 = blockIdx.x * kBlock + (threadIdx.x & ~63u); i0 < cnt; i0 += gridDim.x * kBlock) {
    const uint32_t i = i0 + lane;
    int4 e{-1, 0, -1, 0};
    bool coop = false, dirty = false;
    int32_t len = 0, start = 0;
    uint64_t lb = 0;
    if (i < cnt) e = list[i];
    if (e.x >= 0 && !(e.z >= 0 && color[e.z] >= round)) {
      // the blocker got colored (or there was none yet): resume the scan
      dirty = true;
      len = llen[e.x];
      lb = lbeg[e.x];
      start = e.z >= 0 ? e.y + 1 : e.y;
      if (len - start > kGcSerial) {
        coop = true;
      } else {
        bool blocked = false;
        for (int32_t c = start; c < len; ++c) {
          const int32_t u = hp[lb + c];
          if (color[u] >= round) {
            blocked = true;
            e.y = c;
            e.z = u;
            break;
          }
        }
        if (!blocked) {
          color[e.x] = round;
          e.x = -1;
        }
      }
    }
    uint64_t todo = __ballot(coop);
    while (todo) {
      const int l = __ffsll((unsigned long long)todo) - 1;
      todo &= todo - 1;
      const int32_t s = __shfl(start, l, 64), ln = __shfl(len, l, 64);
      const uint64_t b = __shfl((unsigned long long)lb, l, 64);
      int32_t found = -1, fu = -1;
      for (int32_t c = s; c < ln; c += 64) {
        const int32_t j = c + lane;
        const int32_t u = j < ln ? hp[b + j] : -1;
        const uint64_t m = __ballot(u >= 0 && color[u] >= round);
        if (m) {
          const int f = __ffsll((unsigned long long)m) - 1;
          found = c + f;
          fu = __shfl(u, f, 64);
          break;
        }
      }
      if (lane == l) {
        if (found >= 0) {
          e.y = found;
          e.z = fu;
        } else {
          color[e.x] = round;
          e.x = -1;
        }
      }
    }
    if (dirty) list[i] = e;
  }
}

// Copies the live candidates of in[0, *cin) to out, appending at *cout (one
// atomic per block; each block keeps its slice's order).
__global__ void __launch_bounds__(kBlock) k_gc_compact(const int4* in, const uint32_t* cin, int4* out,
                                                       uint32_t* cout) {
  __shared__ uint32_t at;
  const uint32_t count = *cin;
  const uint32_t per = (count + gridDim.x - 1) / gridDim.x;
  const uint32_t slice = (per + kBlock - 1) / kBlock * kBlock;
  const uint32_t lo = min(count, blockIdx.x * slice), hi = min(count, lo + slice);
  uint64_t own = 0, tot;
  for (uint32_t i = lo + threadIdx.x; i < hi; i += kBlock) own += in[i].x >= 0;
  block_excl_scan64(own, tot);
  if (threadIdx.x == 0) at = tot ? atomicAdd(cout, (uint32_t)tot) : 0u;
  __syncthreads();
  uint32_t run = at;
  for (uint32_t t0 = lo; t0 < hi; t0 += kBlock) {
    const uint32_t i = t0 + threadIdx.x;
    int4 e{-1, 0, -1, 0};
    if (i < hi) e = in[i];
    uint64_t t;
    const uint32_t pre = (uint32_t)block_excl_scan64(e.x >= 0 ? 1u : 0u, t);
    if (e.x >= 0) out[run + pre] = e;
    run += (uint32_t)t;
  }
}

// ---------------------------------------------------------- batched BFS ---
//
// Up to 64 BFS queries, each with its own source and window, share one
// decode pass per level as bit lanes: bit q of front[v] / vis[v] / next[v]
// belongs to query q.  The queries containing a timestamp t are found by a
// binary search over the sorted window boundaries: boundary list b (at most
// kMaxBounds, ascending), and mask[i] = the queries whose window contains
// [b[i-1], b[i]) (mask[0] below b[0], mask[nb] from b[nb-1] on).
constexpr int kMaxQueries = 64;
constexpr int kMaxBounds = 2 * kMaxQueries;

struct WinTable {
  const int64_t* bound;
  const uint64_t* mask;
  int32_t nb;
};

struct BfsbAct {
  Window W;
  const uint64_t* front;
  const uint64_t* vis;
  uint64_t* next;
  const int64_t* bound;  // LDS copies
  const uint64_t* mask;
  int32_t nb;
  uint64_t fx = 0;
  __device__ bool begin(int32_t x, bool, Split) {
    if (!W.active(x)) return false;
    fx = front[x];
    return fx != 0;
  }
  __device__ __forceinline__ uint64_t winmask(int32_t t) const {
    int lo = 0, hi = nb;  // first boundary > t
    while (lo < hi) {
      const int mid = (lo + hi) >> 1;
      if (bound[mid] <= (int64_t)t) lo = mid + 1; else hi = mid;
    }
    return mask[lo];
  }
  __device__ bool contact(int32_t, int32_t r, int32_t t, int32_t) {
    if (!W.active(r)) return true;
    const uint64_t m = fx & winmask(t);
    if (!m) return true;
    const uint64_t nv = m & ~vis[r];
    if (nv && (nv & ~next[r])) atomicOr((unsigned long long*)&next[r], (unsigned long long)nv);
    return true;
  }
  __device__ void end(int32_t, bool) {}
};

template <int MAXD>
__global__ void __launch_bounds__(kBlock) k_bfsb(Streams S, const Ckpt* ck, Items I, Window W, const uint64_t* front,
                                                 const uint64_t* vis, uint64_t* next, WinTable tab) {
  __shared__ int64_t bound[kMaxBounds];
  __shared__ uint64_t mask[kMaxBounds + 1];
  for (int i = threadIdx.x; i <= tab.nb; i += blockDim.x) {
    if (i < tab.nb) bound[i] = tab.bound[i];
    mask[i] = tab.mask[i];
  }
  __syncthreads();
  BfsbAct a{W, front, vis, next, bound, mask, tab.nb};
  LANE_RING();
  FOR_ITEMS(I, i) run_item<MAXD>(S, ck, I.code[i], a, ring_, I.stage(stage_, i));
}

// Ends step `lvl` of a batched BFS: the bits newly set in next[v] become the
// frontier, join vis[v], and give their queries level lvl + 1 - start[q] at v
// (levels is rows of n; start[q] is the step at which query q began, or
// null for 0).  alive (when given) receives the OR of the new frontier.
__global__ void __launch_bounds__(kBlock) k_bfsb_finish(int32_t n, uint64_t* front, uint64_t* vis, uint64_t* next,
                                                        int32_t* levels, int32_t lvl, const int32_t* start,
                                                        unsigned long long* alive) {
  uint64_t any = 0;
  for (int64_t v = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; v < n; v += (int64_t)gridDim.x * blockDim.x) {
    const uint64_t nx = next[v] & ~vis[v];
    next[v] = 0;
    front[v] = nx;
    if (!nx) continue;
    any |= nx;
    vis[v] |= nx;
    for (uint64_t b = nx; b; b &= b - 1) {
      const int64_t q = __ffsll((unsigned long long)b) - 1;
      levels[q * n + v] = lvl + 1 - (start ? start[q] : 0);
    }
  }
  if (alive) {
    for (int off = 32; off > 0; off >>= 1) any |= (uint64_t)__shfl_xor((unsigned long long)any, off, 64);
    if ((threadIdx.x & 63) == 0 && any) atomicOr(alive, (unsigned long long)any);
  }
}

// Admits the lanes of `lanes` (a bit mask) into a running batched BFS: clears
// their bits from vis over all vertices and their level rows, then starts
// lane q at src[q].
__global__ void __launch_bounds__(kBlock) k_bfsb_admit(int32_t n, uint64_t lanes, const int32_t* src, uint64_t* vis,
                                                       uint64_t* front, int32_t* levels) {
  for (int64_t v = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; v < n; v += (int64_t)gridDim.x * blockDim.x) {
    vis[v] &= ~lanes;
    for (uint64_t b = lanes; b; b &= b - 1) levels[(int64_t)(__ffsll((unsigned long long)b) - 1) * n + v] = kInf;
  }
}

__global__ void k_bfsb_admit_src(int32_t n, uint64_t lanes, const int32_t* src, uint64_t* vis, uint64_t* front,
                                 int32_t* levels) {
  if (threadIdx.x != 0) return;
  for (uint64_t b = lanes; b; b &= b - 1) {
    const int q = __ffsll((unsigned long long)b) - 1;
    const int32_t s = src[q];
    front[s] |= 1ull << q;
    vis[s] |= 1ull << q;
    levels[(int64_t)q * n + s] = 0;
  }
}

// Starts a batched BFS: query q's source gets level 0 and joins the frontier.
__global__ void k_bfsb_start(int32_t n, const int32_t* src, int32_t nq, uint64_t* front, uint64_t* vis,
                             int32_t* levels) {
  for (int q = threadIdx.x; q < nq; q += blockDim.x) {
    const int32_t s = src[q];
    atomicOr((unsigned long long*)&front[s], 1ull << q);
    atomicOr((unsigned long long*)&vis[s], 1ull << q);
    levels[(int64_t)q * n + s] = 0;
  }
}

// Rules applied once per distinct neighbor use the pair's first window
// contact.  Contacts of a run split across chunks are deferred: each chunk
// posts its first window contact of the run to the run's slot as
// (contact index << 32 | t) with an atomic minimum, and a fix-up kernel
// applies the rule once per slot after the pass.
__device__ __forceinline__ bool defer(const Split& sp, uint64_t* runkey, int32_t r, int32_t t, int32_t k,
                                      const Window& W) {
  const int32_t s = sp.slot_of(r);
  if (s < 0) return false;
  if (W.in(t)) atomicMin((unsigned long long*)&runkey[s], ((uint64_t)(uint32_t)k << 32) | (uint32_t)t);
  return true;
}

// Earliest arrival (bfs.cu:1452-1471 rule): first window contact per pair.
struct EaAct {
  Window W;
  int32_t* arr;
  int32_t* stamp;
  int32_t round;
  uint64_t* runkey;
  Split sp{-1, -1, -1, -1};
  int32_t ax = 0, last = -1;
  __device__ bool begin(int32_t x, bool, Split s) {
    if (!W.active(x)) return false;
    ax = arr[x];
    last = -1;
    sp = s;
    return true;
  }
  __device__ void relax(int32_t r, int32_t t) {
    if (W.active(r) && ax <= t && t < arr[r]) {
      const int32_t old = atomicMin(&arr[r], t);
      if (t < old) stamp[r] = round;
    }
  }
  __device__ bool contact(int32_t, int32_t r, int32_t t, int32_t k) {
    if (defer(sp, runkey, r, t, k, W)) return true;
    if (!W.in(t) || r == last) return true;
    last = r;
    relax(r, t);
    return true;
  }
  __device__ void end(int32_t, bool) {}
};

template <int MAXD>
__global__ void __launch_bounds__(kBlock) k_ea(Streams S, const Ckpt* ck, Items I, Window W, int32_t* arr,
                                               int32_t* stamp, int32_t round, uint64_t* runkey) {
  EaAct a{W, arr, stamp, round, runkey};
  LANE_RING();
  FOR_ITEMS(I, i) run_item<MAXD>(S, ck, I.code[i], a, ring_, I.stage(stage_, i));
}

// Resolved split runs: the slot's pair (x, r) and its first window contact t.
struct SplitRun {
  int32_t x, r, t;
};

__device__ __forceinline__ bool split_run(const Ckpt* ck, uint64_t* runkey, int32_t s, SplitRun& out) {
  const uint64_t key = runkey[s];
  if (key == ~0ull) return false;
  runkey[s] = ~0ull;  // ready for the next pass
  out = SplitRun{ck[s].x, ck[s].cont_r, (int32_t)(uint32_t)key};
  return true;
}

__global__ void __launch_bounds__(kBlock) k_ea_fix(const Ckpt* ck, int32_t nck, uint64_t* runkey, Window W,
                                                   int32_t* arr, int32_t* stamp, int32_t round) {
  for (int32_t s = blockIdx.x * blockDim.x + threadIdx.x; s < nck; s += gridDim.x * blockDim.x) {
    SplitRun q;
    if (!split_run(ck, runkey, s, q)) continue;
    EaAct a{W, arr, stamp, round, runkey};
    a.ax = arr[q.x];
    a.relax(q.r, q.t);
  }
}

// Latest departure to a target (temporal_ligra LatestDeparture-Compute.C:27-51
// rule over the transposed graph, with instantaneous contacts): a window
// contact (x, r, t) between active vertices with r reached gives x the
// departure t when t <= ld[r].  Every item pulls the best such t over its
// contacts; a vertex split into chunks merges them with an atomic maximum.
// Values only rise and each is met by a real path, so the rounds reach the
// same fixpoint as the reference's frontier loop whatever order the waves run
// in.  Unreached vertices keep the reference's -(INT_MAX / 2) (:116-119).
constexpr int32_t kLdUnreached = -(INT32_MAX / 2);
struct LdAct {
  Window W;
  int32_t* ld;
  uint32_t* changed;
  int32_t best = INT32_MIN;
  __device__ bool begin(int32_t x, bool, Split) {
    if (!W.active(x)) return false;
    best = ld[x];
    return true;
  }
  __device__ bool contact(int32_t, int32_t r, int32_t t, int32_t) {
    if (t > best && W.in(t) && W.active(r)) {
      const int32_t lr = ld[r];
      if (lr != kLdUnreached && t <= lr) best = t;
    }
    return true;
  }
  __device__ void end(int32_t x, bool chunked) {
    if (chunked) {
      if (best > ld[x] && atomicMax(&ld[x], best) < best) *changed = 1u;
    } else if (best > ld[x]) {
      ld[x] = best;
      *changed = 1u;
    }
  }
};

template <int MAXD>
__global__ void __launch_bounds__(kBlock) k_ld(Streams S, const Ckpt* ck, Items I, Window W, int32_t* ld,
                                               uint32_t* changed) {
  LdAct a{W, ld, changed};
  LANE_RING();
  FOR_ITEMS(I, i) run_item<MAXD>(S, ck, I.code[i], a, ring_, I.stage(stage_, i));
}

// Window contact count (PageRank divisor).
struct WdegAct {
  Window W;
  int32_t* du;
  int32_t c = 0;
  __device__ bool begin(int32_t x, bool, Split) {
    c = 0;
    return W.active(x);
  }
  __device__ bool contact(int32_t, int32_t, int32_t t, int32_t) {
    c += W.in(t) ? 1 : 0;
    return true;
  }
  __device__ void end(int32_t x, bool chunked) {
    if (chunked) {
      if (c) atomicAdd(&du[x], c);
    } else {
      du[x] = c;
    }
  }
};

template <int MAXD>
__global__ void __launch_bounds__(kBlock) k_wdeg(Streams S, const Ckpt* ck, Items I, Window W, int32_t* du) {
  WdegAct a{W, du};
  LANE_RING();
  FOR_ITEMS(I, i) run_item<MAXD>(S, ck, I.code[i], a, ring_, I.stage(stage_, i));
}

__device__ __forceinline__ float pr_coef(float damping, const float* rank, const int32_t* du, int32_t x) {
  return damping * rank[x] / (float)du[x];
}

__device__ __forceinline__ double pr_term(float coef, int32_t t) {
  const float w = 1.0f - expf(-0.1f * (float)t);
  return (double)(coef * w);
}

struct PrAct {
  Window W;
  const float* rank;
  const int32_t* du;
  float damping;
  double* acc;
  uint64_t* runkey;
  Split sp{-1, -1, -1, -1};
  float coef = 0;
  int32_t last = -1;
  __device__ bool begin(int32_t x, bool, Split s) {
    if (!W.active(x) || du[x] == 0) return false;
    coef = pr_coef(damping, rank, du, x);
    last = -1;
    sp = s;
    return true;
  }
  __device__ bool contact(int32_t, int32_t r, int32_t t, int32_t k) {
    if (defer(sp, runkey, r, t, k, W)) return true;
    if (!W.in(t) || r == last) return true;
    last = r;
    if (W.active(r)) atomicAdd(&acc[r], pr_term(coef, t));
    return true;
  }
  __device__ void end(int32_t, bool) {}
};

template <int MAXD>
__global__ void __launch_bounds__(kBlock) k_pr_push(Streams S, const Ckpt* ck, Items I, Window W, const float* rank,
                                                    const int32_t* du, float damping, double* acc, uint64_t* runkey) {
  PrAct a{W, rank, du, damping, acc, runkey};
  LANE_RING();
  FOR_ITEMS(I, i) run_item<MAXD>(S, ck, I.code[i], a, ring_, I.stage(stage_, i));
}

__global__ void __launch_bounds__(kBlock) k_pr_fix(const Ckpt* ck, int32_t nck, uint64_t* runkey, Window W,
                                                   const float* rank, const int32_t* du, float damping, double* acc) {
  for (int32_t s = blockIdx.x * blockDim.x + threadIdx.x; s < nck; s += gridDim.x * blockDim.x) {
    SplitRun q;
    if (!split_run(ck, runkey, s, q)) continue;
    if (W.active(q.r)) atomicAdd(&acc[q.r], pr_term(pr_coef(damping, rank, du, q.x), q.t));
  }
}

struct HubAct {
  Window W;
  const float* auth;
  double* acc;
  uint64_t* runkey;
  Split sp{-1, -1, -1, -1};
  double sum = 0;
  int32_t last = -1;
  __device__ bool begin(int32_t x, bool, Split s) {
    sum = 0;
    last = -1;
    sp = s;
    return W.active(x);
  }
  __device__ bool contact(int32_t, int32_t r, int32_t t, int32_t k) {
    if (defer(sp, runkey, r, t, k, W)) return true;
    if (!W.in(t) || r == last) return true;
    last = r;
    if (W.active(r)) sum += (double)auth[r];
    return true;
  }
  __device__ void end(int32_t x, bool chunked) {
    if (chunked)
      atomicAdd(&acc[x], sum);
    else
      acc[x] = sum;
  }
};

template <int MAXD>
__global__ void __launch_bounds__(kBlock) k_hits_pull(Streams S, const Ckpt* ck, Items I, Window W, const float* auth,
                                                      double* acc, uint64_t* runkey) {
  HubAct a{W, auth, acc, runkey};
  LANE_RING();
  FOR_ITEMS(I, i) run_item<MAXD>(S, ck, I.code[i], a, ring_, I.stage(stage_, i));
}

struct AuthAct {
  Window W;
  const float* hub;
  double* acc;
  uint64_t* runkey;
  Split sp{-1, -1, -1, -1};
  double h = 0;
  int32_t last = -1;
  __device__ bool begin(int32_t x, bool, Split s) {
    if (!W.active(x)) return false;
    h = (double)hub[x];
    last = -1;
    sp = s;
    return true;
  }
  __device__ bool contact(int32_t, int32_t r, int32_t t, int32_t k) {
    if (defer(sp, runkey, r, t, k, W)) return true;
    if (!W.in(t) || r == last) return true;
    last = r;
    if (W.active(r)) atomicAdd(&acc[r], h);
    return true;
  }
  __device__ void end(int32_t, bool) {}
};

template <int MAXD>
__global__ void __launch_bounds__(kBlock) k_hits_push(Streams S, const Ckpt* ck, Items I, Window W, const float* hub,
                                                      double* acc, uint64_t* runkey) {
  AuthAct a{W, hub, acc, runkey};
  LANE_RING();
  FOR_ITEMS(I, i) run_item<MAXD>(S, ck, I.code[i], a, ring_, I.stage(stage_, i));
}

// mode 0: hub[x] += auth[r]; mode 1: auth[r] += hub[x] (per resolved split run).
__global__ void __launch_bounds__(kBlock) k_hits_fix(const Ckpt* ck, int32_t nck, uint64_t* runkey, Window W,
                                                     const float* other, double* acc, int mode) {
  for (int32_t s = blockIdx.x * blockDim.x + threadIdx.x; s < nck; s += gridDim.x * blockDim.x) {
    SplitRun q;
    if (!split_run(ck, runkey, s, q)) continue;
    if (!W.active(q.r) || !W.active(q.x)) continue;
    if (mode == 0)
      atomicAdd(&acc[q.x], (double)other[q.r]);
    else
      atomicAdd(&acc[q.r], (double)other[q.x]);
  }
}

// ------------------------------------------- PageRank / HITS over pairs ---
//
// Both iterate over the distinct (x, r) neighbor pairs with a window contact
// between active vertices (PageRank weights a pair by its first window
// contact t).  Decoded every iteration, as above, the sums are device-wide
// FP64 atomics, and their throughput bounds the pass.  When memory allows,
// the pairs are decoded once per call instead: a counting pass, a scan and a
// filling pass (item order, as for the coloring lists; split runs go to one
// extra slot per checkpoint after the items' pairs), then a radix sort by r.
// Each iteration is then a pull with no atomics and a fixed summation order:
// a row sum over r's in-pairs, or over x's out-pairs in fill order (the HITS
// hub step), computed by a load-balanced reduce-by-key (rows run from one
// pair to a hub's hundreds of thousands).  A pair is key r and value (x << 32 | float bits of
// w = 1 - exp(-0.1 t)).
__device__ __forceinline__ uint64_t pair_val(int32_t x, int32_t t) {
  const float w = 1.0f - expf(-0.1f * (float)t);
  return ((uint64_t)(uint32_t)x << 32) | __float_as_uint(w);
}

struct PairAct {
  Window W;
  uint32_t* cnt;        // counting pass: per item
  const uint64_t* off;  // filling pass: per item
  uint32_t* key;
  uint64_t* val;
  uint64_t* runkey;
  int32_t n;
  Split sp{-1, -1, -1, -1};
  uint64_t pos = 0;
  uint32_t c = 0;
  int32_t idx = 0, x = 0, last = -1;
  __device__ void item(int32_t code) { idx = item_index(code, n); }
  __device__ bool begin(int32_t v, bool, Split s) {
    if (!W.active(v)) return false;
    x = v;
    sp = s;
    last = -1;
    c = 0;
    if (off) pos = off[idx];
    return true;
  }
  __device__ bool contact(int32_t, int32_t r, int32_t t, int32_t k) {
    if (off) {
      if (defer(sp, runkey, r, t, k, W)) return true;
    } else if (sp.slot_of(r) >= 0) {
      return true;  // split run: one extra slot, filled by k_pair_fix
    }
    if (!W.in(t) || r == last) return true;
    last = r;
    if (!W.active(r)) return true;
    if (off) {
      key[pos] = (uint32_t)r;
      val[pos] = pair_val(x, t);
      ++pos;
    } else {
      ++c;
    }
    return true;
  }
  __device__ void end(int32_t, bool) {
    if (!off) cnt[idx] = c;
  }
};

template <int MAXD>
__global__ void __launch_bounds__(kBlock) k_pairs(Streams S, const Ckpt* ck, Items I, Window W, uint32_t* cnt,
                                                  const uint64_t* off, uint32_t* key, uint64_t* val,
                                                  uint64_t* runkey) {
  PairAct a{W, cnt, off, key, val, runkey, S.n};
  LANE_RING();
  FOR_ITEMS(I, i) run_item<MAXD>(S, ck, I.code[i], a, ring_, I.stage(stage_, i));
}

// Split-run pairs into key/val[0, nck) (one slot per checkpoint).  An empty
// slot gets r = n (sorted after every vertex, never summed) and w = 0, with
// its chunk's vertex as x: the slots of one vertex stay one run of equal x
// (k_scatter_rows adds each run's sum without atomics).
__global__ void __launch_bounds__(kBlock) k_pair_fix(const Ckpt* ck, int32_t nck, uint64_t* runkey, Window W,
                                                     int32_t n, uint32_t* key, uint64_t* val) {
  for (int32_t s = blockIdx.x * blockDim.x + threadIdx.x; s < nck; s += gridDim.x * blockDim.x) {
    SplitRun q;
    if (split_run(ck, runkey, s, q) && W.active(q.r) && W.active(q.x)) {
      key[s] = (uint32_t)q.r;
      val[s] = pair_val(q.x, q.t);
    } else {
      key[s] = (uint32_t)n;
      val[s] = (uint64_t)(uint32_t)ck[s].x << 32;
    }
  }
}

// PageRank coefficient of every vertex: d * rank[x] / du[x] (0 without
// window contacts or when inactive).
__global__ void __launch_bounds__(kBlock) k_pr_coef(Window W, int32_t n, const float* rank, const int32_t* du,
                                                    float damping, float* coef) {
  for (int64_t v = blockIdx.x * (int64_t)kBlock + threadIdx.x; v < n; v += (int64_t)gridDim.x * kBlock)
    coef[v] = (W.active((int32_t)v) && du[v] > 0) ? pr_coef(damping, rank, du, (int32_t)v) : 0.0f;
}

// End of a PageRank iteration over the pair lists, fused: rank = acc / sum
// (as k_scale_store), the next iteration's coefficients from it (as
// k_pr_coef) and acc reset to base (as k_fill_active): one pass over the
// vertices instead of three, with the same arithmetic.
__global__ void __launch_bounds__(kBlock) k_pr_next(Window W, int32_t n, double* acc, const double* scale,
                                                    float* rank, const int32_t* du, float damping, float* coef,
                                                    double base) {
  const double s = scale[0];
  for (int64_t v = blockIdx.x * (int64_t)kBlock + threadIdx.x; v < n; v += (int64_t)gridDim.x * kBlock) {
    const double a = acc[v];
    const float r = (float)(s > 0 ? a / s : a);
    rank[v] = r;
    const bool act = W.active((int32_t)v);
    const int32_t d = du[v];
    coef[v] = (act && d > 0) ? damping * r / (float)d : 0.0f;
    acc[v] = act ? base : 0.0;
  }
}

// acc[k] = base + sums[i] (or acc[k] += sums[i] when add) for every row run
// i with key k < n.  Keys must be distinct among the runs of one call.
__global__ void __launch_bounds__(kBlock) k_scatter_rows(const uint32_t* ukeys, const double* sums,
                                                         const uint32_t* runs, int32_t n, double base, int add,
                                                         double* acc) {
  const uint32_t cnt = *runs;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < cnt; i += gridDim.x * blockDim.x) {
    const uint32_t k = ukeys[i];
    if (k < (uint32_t)n) acc[k] = add ? acc[k] + sums[i] : base + sums[i];
  }
}

// ------------------------------------------------------ elementwise, sums ---

template <class T>
__global__ void __launch_bounds__(kBlock) k_fill(T* a, int64_t n, T val) {
  for (int64_t v = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; v < n; v += (int64_t)gridDim.x * blockDim.x)
    a[v] = val;
}

template <class T>
__global__ void __launch_bounds__(kBlock) k_fill_active(Window W, int32_t n, T* a, T inside, T outside) {
  for (int64_t v = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; v < n; v += (int64_t)gridDim.x * blockDim.x)
    a[v] = W.active((int32_t)v) ? inside : outside;
}

__global__ void k_init_labels(Window W, int32_t n, int32_t* label, int32_t* stamp) {
  for (int64_t v = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; v < n; v += (int64_t)gridDim.x * blockDim.x) {
    label[v] = W.active((int32_t)v) ? (int32_t)v : -1;
    stamp[v] = -1;
  }
}

__global__ void k_copy_deg(Window W, int32_t n, const int32_t* deg, int32_t* du) {
  for (int64_t v = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; v < n; v += (int64_t)gridDim.x * blockDim.x)
    du[v] = W.active((int32_t)v) ? deg[v] : 0;
}

// Deterministic two-stage reductions (fixed grid, fixed tree order).
constexpr int kRedBlocks = 512;

template <int MODE>  // 0: sum x, 1: sum x^2, 2: sum |float(x / s) - old|
__global__ void __launch_bounds__(kBlock) k_reduce1(const double* x, const float* old, const double* scale, int64_t n,
                                                   double* partial) {
  __shared__ double sh[kBlock];
  double acc = 0;
  const double s = (MODE == 2) ? scale[0] : 0.0;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const double v = x[i];
    if (MODE == 0) acc += v;
    if (MODE == 1) acc += v * v;
    if (MODE == 2) acc += fabs((double)(float)(s > 0 ? v / s : v) - (double)old[i]);
  }
  sh[threadIdx.x] = acc;
  __syncthreads();
  for (int w = kBlock / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[blockIdx.x] = sh[0];
}

__global__ void __launch_bounds__(kBlock) k_reduce2(const double* partial, int nparts, double* out, int sqrt_out) {
  __shared__ double sh[kBlock];
  double acc = 0;
  for (int i = threadIdx.x; i < nparts; i += blockDim.x) acc += partial[i];
  sh[threadIdx.x] = acc;
  __syncthreads();
  for (int w = kBlock / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[0] = sqrt_out ? sqrt(sh[0]) : sh[0];
}

// out[i] = float(x[i] / s) (s = 0 leaves x unscaled).
__global__ void __launch_bounds__(kBlock) k_scale_store(const double* x, const double* scale, int64_t n, float* out) {
  const double s = scale[0];
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    out[i] = (float)(s > 0 ? x[i] / s : x[i]);
}

// Lanes that finished at the end of a PR server step: rank[v][q] = float(acc
// / sum) as k_prb_out, kept in the (then unused) initial-rank column.
__global__ void __launch_bounds__(kBlock) k_prb_fin(int32_t n, uint64_t lanes, const double* acc, const double* sums,
                                                    float* rank) {
  for (int64_t v = blockIdx.x * (int64_t)kBlock + threadIdx.x; v < n; v += (int64_t)gridDim.x * kBlock)
    for (uint64_t b = lanes; b; b &= b - 1) {
      const int q = __ffsll((unsigned long long)b) - 1;
      const double s = sums[q];
      const int64_t i = v * 64 + q;
      rank[i] = (float)(s > 0 ? acc[i] / s : acc[i]);
    }
}

}  // namespace

struct tga_graph;
// State shared by the query servers (BFS, CC, PR): 64 lanes, each holding
// one query from admission to the collection of its result.  Their device
// state lives in allocations of the handle (so the handle's device byte
// count includes it); closing the handle first orphans a server (g reset):
// its calls then fail, and closing it frees nothing more.
struct LaneServer {
  tga_graph* g = nullptr;
  uint64_t running = 0, pending = 0, done = 0;
  tga_window h_win[64] = {};
  std::vector<void*> mem;
  int64_t* h_wtab = nullptr;       // pinned window table being uploaded
  hipEvent_t ev_wtab = nullptr;    // its last upload
  explicit LaneServer(tga_graph* g_);
  virtual ~LaneServer();
  template <class T>
  int alloc(T*& p, size_t count);
  // A free lane for a new query, or -1.
  int free_lane() const {
    const uint64_t used = running | pending | done;
    return ~used ? __builtin_ctzll(~used) : -1;
  }
  LaneServer(const LaneServer&) = delete;
  LaneServer& operator=(const LaneServer&) = delete;
};

struct tga_graph {
  int device = 0;
  hipStream_t stream = nullptr;
  // The groups of one pass run concurrently on their own streams (default
  // priority), forked from and joined back into `stream` with events.
  hipStream_t cstream[kGroups] = {};
  hipEvent_t ev_fork = nullptr;
  hipEvent_t ev_join[kGroups] = {};
  int64_t n = 0;
  int32_t ab = 0, ae = 0;
  int32_t min_ts = 0;
  tga_graph_info info{};
  Streams S{};
  std::vector<void*> allocs;
  std::vector<size_t> alloc_bytes;
  std::vector<void**> alloc_owner;
  int64_t fail_alloc_at = 0;  // tga_debug_fail_alloc: this allocation from now fails (0: off)
  int32_t* d_err = nullptr;
  // schedule
  Ckpt* ck = nullptr;
  int64_t nck = 0;
  uint64_t* runkey = nullptr;  // per checkpoint: split-run first window contact (~0: none)
  int32_t* vinfo = nullptr;  // per vertex item code (see kEmpty)
  int32_t* vdeg = nullptr;   // per vertex outdegree
  int32_t* items[kGroups] = {};  // static item lists by group
  uint32_t n_items[kGroups] = {};
  TsCols tcols[kGroups];  // their timestamp codes in column order (TGA_TCOL=0: none)
  int32_t* big[kClasses] = {};  // chunked records by class
  int32_t n_big[kClasses] = {};
  // scratch (n-sized unless stated)
  int32_t *q0 = nullptr, *i0 = nullptr, *i1 = nullptr, *i2 = nullptr;
  Queues xq{};                // expanded item queues by group
  int32_t* xq_big = nullptr;  // chunked records of the expanded set
  uint32_t* cnt = nullptr;  // 8 counters
  unsigned long long* u64s = nullptr;  // 2 * kSumSlots values
  double* acc = nullptr;
  double* partial = nullptr;  // kRedBlocks
  double* scal = nullptr;     // 4 scalars
  float *f0 = nullptr, *f1 = nullptr;
  int32_t* prio = nullptr;  // default GC priorities, built on first use
  // per-item list offsets (graph coloring, PageRank, HITS), on first use
  uint32_t* lcnt = nullptr;   // per item: n + nck
  uint64_t* loff = nullptr;   // n + nck + 1
  uint64_t* ltsum = nullptr;  // scan tiles
  // graph coloring lists, allocated on first use (k_gc_round)
  int32_t* gc_hp = nullptr;     // one entry per contact at most
  uint64_t* gc_lbeg = nullptr;
  int32_t* gc_llen = nullptr;
  int4* gc_list[2] = {};        // candidate lists (ping-pong), n each
  uint32_t* gc_ctr = nullptr;   // candidate counts of the two lists
  // neighbor pairs of PageRank / HITS, allocated on first use (k_pairs)
  uint32_t *pk = nullptr, *sk = nullptr;  // r: fill order, sorted by r
  uint64_t *pv = nullptr, *sv = nullptr;  // (x, w): fill order, sorted by r
  int64_t* ct = nullptr;     // contact-time cache (k_ctimes), on first use
  uint64_t* coff = nullptr;  // n + 1: first contact of every record
  // batched PageRank: run structure (once per handle) and batch state
  int4* prb_runs = nullptr;  // sorted runs (x, k0, k1, t0)
  float* prb_f0 = nullptr;   // weight of each run's first contact
  int32_t *prb_midx = nullptr, *prb_mlist = nullptr;  // multi-contact runs: weight slot, list
  int32_t prb_nmulti = 0;
  float* prb_wm = nullptr;   // per batch: 64 weights per multi-contact run
  uint64_t *prb_in_off = nullptr, *prb_seg_off = nullptr, *prb_seg_beg = nullptr;
  int32_t *prb_seg_row = nullptr, *prb_seg_len = nullptr, *prb_mrows = nullptr, *prb_long = nullptr;
  uint32_t* prb_nm = nullptr;
  int64_t prb_nseg = 0;
  int32_t prb_nmrows = 0, prb_nlong = 0;
  int2* prb_du_items = nullptr;  // (long record, first contact) items of k_prb_du_long
  int32_t prb_ndu_items = 0;
  int32_t* prb_du = nullptr;
  float *prb_rank = nullptr, *prb_coef = nullptr, *prb_out = nullptr;
  double *prb_acc = nullptr, *prb_partial = nullptr, *prb_part = nullptr, *prb_sums = nullptr;
  int64_t* prb_win = nullptr;
  // batched HITS: out-runs (built with the run structure) and batch state
  int4* prb_out_runs = nullptr;
  uint64_t *prb_obeg = nullptr, *prb_oseg_off = nullptr, *prb_oseg_beg = nullptr;
  int32_t *prb_olen = nullptr, *prb_oseg_row = nullptr, *prb_oseg_len = nullptr, *prb_omrows = nullptr;
  uint32_t* prb_onm = nullptr;
  int64_t prb_noseg = 0;
  int32_t prb_nomrows = 0;
  float *hb_hub = nullptr, *hb_auth = nullptr, *hb_out = nullptr;
  double *hb_partial = nullptr, *hb_norm = nullptr;
  uint64_t* hb_mask = nullptr;
  int32_t hb_out_rows = 0;
  // batched CC state (first use)
  int32_t *ccb_label = nullptr, *ccb_partial = nullptr, *ccb_changed = nullptr, *ccb_out = nullptr;
  uint64_t* ccb_mask = nullptr;
  int32_t ccb_out_rows = 0;
  int32_t prb_out_rows = 0;
  int64_t* dcb_win = nullptr;  // batched DC windows (2 x 64)
  int32_t* dcb_out = nullptr;  // batched DC rows when the caller's output is on the host
  int32_t dcb_rows = 0;
  uint32_t* ukeys = nullptr;  // n + 2: row keys of a reduce_rows call
  double* usums = nullptr;    // n + 2: their sums
  uint32_t* nruns = nullptr;
  void* sort_tmp = nullptr;  // radix sort and reduce_rows temporary storage
  size_t sort_tmp_bytes = 0;
  // batched BFS state, allocated on first use
  uint64_t *bq_front = nullptr, *bq_vis = nullptr, *bq_next = nullptr, *bq_mask = nullptr;
  int64_t* bq_bound = nullptr;
  int32_t* bq_src = nullptr;
  int32_t* bq_levels = nullptr;  // nq rows of n when the caller's output is on the host
  int32_t bq_levels_rows = 0;
  int64_t device_bytes = 0;
  // Captured iteration loops of the batched calls (replay).
  struct Graph {
    hipGraphExec_t exec = nullptr;
    uint64_t key = 0;
  };
  Graph gr_prb, gr_hb, gr_ccb, gr_gcb;
  // batched coloring: out-segments in priority-level order, per-level flags
  int32_t* gcb_segs = nullptr;
  int32_t* gcb_flags = nullptr;
  std::vector<int32_t> h_oseg_row;         // row of every out-segment (host copy)
  std::vector<int64_t> gcb_level_off;      // level boundaries in gcb_segs
  bool gcb_default_levels = false;         // gcb_segs holds the default priorities' levels
  bool use_graphs = true;  // TGA_GRAPHS=0: launch the loops one kernel at a time
  int32_t dense_div = 5;   // TGA_DENSE_DIV (dense)
  std::vector<LaneServer*> servers;  // open query servers (orphaned when the handle closes)
  std::vector<int32_t> h_items;  // items each vertex expands to (0 empty, 1 whole record, else its chunks)
  // kernel timing (tga_profile)
  bool prof = false;
  double prof_ms = 0;
  int64_t prof_launches = 0;
  std::vector<hipEvent_t> ev;  // start/stop pairs
  size_t ev_used = 0;

  void kbegin() {
    if (!prof) return;
    if (ev_used + 2 > ev.size()) {
      hipEvent_t a, b;
      if (hipEventCreate(&a) != hipSuccess || hipEventCreate(&b) != hipSuccess) return;
      ev.push_back(a);
      ev.push_back(b);
    }
    (void)hipEventRecord(ev[ev_used], stream);
  }
  void kend() {
    if (!prof || ev_used + 2 > ev.size()) return;
    (void)hipEventRecord(ev[ev_used + 1], stream);
    ev_used += 2;
  }
  void harvest() {
    if (!ev_used) return;
    (void)hipStreamSynchronize(stream);
    for (size_t i = 0; i + 1 < ev_used; i += 2) {
      float ms = 0;
      if (hipEventElapsedTime(&ms, ev[i], ev[i + 1]) == hipSuccess) {
        prof_ms += ms;
        ++prof_launches;
      }
    }
    ev_used = 0;
  }

  // Device allocations of the handle.  Each records its size and, when the
  // pointer it was stored into is a member of this handle, that member, so
  // that a failed group of first-use allocations can be undone (AllocTxn)
  // and a staging buffer can be released before it grows.
  template <class T>
  int alloc(T*& p, size_t count) {
    const size_t bytes = std::max<size_t>(count, 1) * sizeof(T);
    if (fail_alloc_at > 0 && --fail_alloc_at == 0)
      return fail(TGA_ENOMEM, "hipMalloc: injected failure (tga_debug_fail_alloc)");
    void* q = nullptr;
    hipError_t e = hipMalloc(&q, bytes);
    if (e != hipSuccess) return fail(TGA_ENOMEM, std::string("hipMalloc: ") + hipGetErrorString(e));
    const char* at = (const char*)&p;
    const char* self = (const char*)this;
    allocs.push_back(q);
    alloc_bytes.push_back(bytes);
    alloc_owner.push_back(at >= self && at < self + sizeof(*this) ? (void**)&p : nullptr);
    device_bytes += (int64_t)bytes;
    p = (T*)q;
    return TGA_OK;
  }
  void free_at(size_t i) {
    (void)hipFree(allocs[i]);  // waits for work still using it
    if (alloc_owner[i]) *alloc_owner[i] = nullptr;
    device_bytes -= (int64_t)alloc_bytes[i];
    allocs.erase(allocs.begin() + (std::ptrdiff_t)i);
    alloc_bytes.erase(alloc_bytes.begin() + (std::ptrdiff_t)i);
    alloc_owner.erase(alloc_owner.begin() + (std::ptrdiff_t)i);
  }
  // Frees one allocation of the handle and resets the pointer.
  template <class T>
  void release(T*& p) {
    for (size_t i = allocs.size(); p && i-- > 0;)
      if (allocs[i] == (void*)p) {
        free_at(i);
        break;
      }
    p = nullptr;
  }
  // Frees every allocation made after `mark` (members are reset to null).
  void rollback(size_t mark) {
    while (allocs.size() > mark) free_at(allocs.size() - 1);
  }
  // A host staging buffer of `rows` rows of n entries: grown by releasing
  // the old one first, so a long-running caller holds one buffer of the
  // largest width it used.
  template <class T>
  int rows_buffer(T*& p, int32_t& have, int32_t rows) {
    if (have >= rows && p) return TGA_OK;
    release(p);
    have = 0;
    int rc = alloc(p, (size_t)rows * (size_t)n);
    if (!rc) have = rows;
    return rc;
  }
  template <class T>
  int upload(T*& dst, const std::vector<T>& src) {
    int rc = alloc(dst, src.size());
    if (rc) return rc;
    if (!src.empty()) HIP_TRY(hipMemcpy(dst, src.data(), src.size() * sizeof(T), hipMemcpyHostToDevice));
    return TGA_OK;
  }
  // A stream file's bytes as device words (decode.hpp layout) plus
  // kPadWords zero words: the bytes are copied as they are and swapped on the
  // device.
  int upload_stream(const uint64_t*& dst, const std::vector<uint8_t>& bytes) {
    const size_t nw = (bytes.size() + 7) / 8;
    uint64_t* p = nullptr;
    int rc = alloc(p, nw + kPadWords);
    if (rc) return rc;
    HIP_TRY(hipMemsetAsync(p, 0, (nw + kPadWords) * 8, stream));
    if (!bytes.empty()) HIP_TRY(hipMemcpyAsync(p, bytes.data(), bytes.size(), hipMemcpyHostToDevice, stream));
    if (nw) k_bswap_words<<<grid((int64_t)nw), kBlock, 0, stream>>>(p, (int64_t)nw);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(stream));  // the copy reads pageable host memory
    dst = p;
    return TGA_OK;
  }
  template <class T>
  int upload_const(const T*& dst, const std::vector<T>& src) {
    T* p = nullptr;
    int rc = upload(p, src);
    dst = p;
    return rc;
  }
  int grid(int64_t items) const {
    const int64_t b = (items + kBlock - 1) / kBlock;
    return (int)std::max<int64_t>(1, std::min<int64_t>(b, 1 << 20));
  }
  Window window(tga_window w) const { return Window{w.start, w.end, ab, ae}; }
  // Fails with the decode error bits raised since the last clear, if any.
  int check_err() {
    harvest();
    int32_t e = 0;
    HIP_TRY(hipMemcpyAsync(&e, d_err, sizeof e, hipMemcpyDeviceToHost, stream));
    HIP_TRY(hipStreamSynchronize(stream));
    if (e) {
      HIP_TRY(hipMemsetAsync(d_err, 0, sizeof(int32_t), stream));
      char buf[160];
      std::snprintf(buf, sizeof buf, "malformed compressed stream (decode error bits 0x%x)", e);
      return fail(TGA_EFORMAT, buf);
    }
    return TGA_OK;
  }
  int out_copy(void* dst, const void* src, size_t bytes, uint32_t flags) {
    HIP_TRY(hipMemcpyAsync(dst, src, bytes,
                           (flags & TGA_OUT_DEVICE) ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost, stream));
    return TGA_OK;
  }
  int read_counts(const uint32_t* dev, uint32_t* host, int k) {
    HIP_TRY(hipMemcpyAsync(host, dev, sizeof(uint32_t) * (size_t)k, hipMemcpyDeviceToHost, stream));
    HIP_TRY(hipStreamSynchronize(stream));
    return TGA_OK;
  }
  // Whether a round of connected components whose frontier holds `nf`
  // vertices runs over the static lists (which read the column copies and
  // whose kernels skip the items of vertices outside the frontier) instead
  // of a queue from k_expand: when nf is at least 1/TGA_DENSE_DIV of the
  // active vertices (default 5, the best of 2-30 on the 2^30-contact
  // netflow-like graph; 0: never).  A pass over the static lists
  // costs about one full decode pass however many of its items run; a queue
  // costs a scan of every vertex plus its items' walks from the streams.
  bool dense(uint64_t nf) const { return dense_div > 0 && tcols[0].w && nf * (uint64_t)dense_div >= (uint64_t)(ae - ab); }
  // Zeroes the next-frontier counter before a round's kernels.
  uint32_t* next_count() {
    k_fill<uint32_t><<<1, 64, 0, stream>>>(cnt + 6, 1, 0u);
    return cnt + 6;
  }
  int read_next_count(uint32_t* nf) { return read_counts(cnt + 6, nf, 1); }
  // Expands the selected vertices of a list (front) or of the range
  // [base, base + count) into the group queues (and xq_big); hc[g] receives
  // the count of group g and hc[kGroups] the number of chunked records.
  int expand(const int32_t* front, int32_t base, uint32_t count, Sel sel, bool keep_empty, bool want_big,
             uint32_t hc[kGroups + 1]) {
    HIP_TRY(hipMemsetAsync(cnt, 0, (kGroups + 1) * sizeof(uint32_t), stream));
    return expand_into(front, base, count, sel, keep_empty, want_big ? xq_big : nullptr, xq, cnt, hc);
  }
  // The same into caller-chosen queues and counters (zeroed by the caller).
  int expand_into(const int32_t* front, int32_t base, uint32_t count, Sel sel, bool keep_empty, int32_t* bigq,
                  Queues q, uint32_t* counters, uint32_t hc[kGroups + 1]) {
    if (count > 0) {
      const int blocks = (int)std::min<int64_t>(kExpandBlocks, (count + kBlock - 1) / kBlock);
      k_expand<<<blocks, kBlock, 0, stream>>>(front, base, count, sel, vinfo, ck, q, counters, bigq,
                                              keep_empty ? 1 : 0);
      HIP_TRY(hipGetLastError());
    }
    return read_counts(counters, hc, kGroups + 1);
  }
  // Runs `enqueue` (kernel launches on `stream`) from a captured graph: the
  // first call for a key captures and instantiates it, later calls with the
  // same key replay it with one launch.  The captured work launches kernels
  // only (fills are kernels too, not memsets).
  template <class F>
  int replay(Graph& gr, uint64_t key, F&& enqueue) {
    if (!use_graphs) return enqueue();
    if (!gr.exec || gr.key != key) {
      if (gr.exec) (void)hipGraphExecDestroy(gr.exec);
      gr.exec = nullptr;
      hipGraph_t graph = nullptr;
      HIP_TRY(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
      const int rc = enqueue();
      const hipError_t e = hipStreamEndCapture(stream, &graph);
      if (rc) {
        if (e == hipSuccess && graph) (void)hipGraphDestroy(graph);
        return rc;
      }
      HIP_TRY(e);
      const hipError_t ei = hipGraphInstantiate(&gr.exec, graph, nullptr, nullptr, 0);
      (void)hipGraphDestroy(graph);
      if (ei != hipSuccess) {
        gr.exec = nullptr;
        return fail(TGA_EDEVICE, std::string("hipGraphInstantiate failed: ") + hipGetErrorString(ei));
      }
      gr.key = key;
    }
    HIP_TRY(hipGraphLaunch(gr.exec, stream));
    return TGA_OK;
  }
  int create_streams() {
    HIP_TRY(hipStreamCreate(&stream));
    HIP_TRY(hipEventCreateWithFlags(&ev_fork, hipEventDisableTiming));
    for (int c = 0; c < kGroups; ++c) {
      HIP_TRY(hipStreamCreate(&cstream[c]));
      HIP_TRY(hipEventCreateWithFlags(&ev_join[c], hipEventDisableTiming));
    }
    return TGA_OK;
  }
  // Frees one allocation of the handle by address.
  void release_ptr(void* p) {
    for (size_t i = allocs.size(); p && i-- > 0;)
      if (allocs[i] == p) {
        free_at(i);
        return;
      }
  }
  ~tga_graph() {
    for (LaneServer* s : servers) s->g = nullptr;
    if (stream) (void)hipStreamSynchronize(stream);
    for (Graph* gr : {&gr_prb, &gr_hb, &gr_ccb, &gr_gcb})
      if (gr->exec) (void)hipGraphExecDestroy(gr->exec);
    for (hipEvent_t e : ev) (void)hipEventDestroy(e);
    for (void* p : allocs) (void)hipFree(p);
    for (int c = 0; c < kGroups; ++c) {
      if (ev_join[c]) (void)hipEventDestroy(ev_join[c]);
      if (cstream[c]) (void)hipStreamDestroy(cstream[c]);
    }
    if (ev_fork) (void)hipEventDestroy(ev_fork);
    if (stream) (void)hipStreamDestroy(stream);
  }
};

// A group of first-use allocations that commits together or not at all:
// unless commit() is reached, every allocation made since construction is
// freed and its handle member reset, so a later call never finds a group
// half allocated (its first pointer set, the others null).
LaneServer::LaneServer(tga_graph* g_) : g(g_) { g->servers.push_back(this); }
LaneServer::~LaneServer() {
  if (g) {
    (void)hipSetDevice(g->device);
    (void)hipStreamSynchronize(g->stream);
    for (void* p : mem) g->release_ptr(p);
    g->servers.erase(std::find(g->servers.begin(), g->servers.end(), this));
  }
  if (ev_wtab) {
    (void)hipEventSynchronize(ev_wtab);
    (void)hipEventDestroy(ev_wtab);
  }
  if (h_wtab) (void)hipHostFree(h_wtab);
}
template <class T>
int LaneServer::alloc(T*& p, size_t count) {
  const int rc = g->alloc(p, count);
  if (!rc) mem.push_back((void*)p);
  return rc;
}

struct AllocTxn {
  tga_graph* g;
  size_t mark;
  bool done = false;
  explicit AllocTxn(tga_graph* g_) : g(g_), mark(g_->allocs.size()) {}
  void commit() { done = true; }
  ~AllocTxn() {
    if (!done) g->rollback(mark);
  }
  AllocTxn(const AllocTxn&) = delete;
  AllocTxn& operator=(const AllocTxn&) = delete;
};

extern "C" const char* tga_last_error(void) { return g_err.c_str(); }

// The error text of the other translation units' entry points (encode.hip).
namespace tga {
int set_error(int code, const std::string& msg) { return fail(code, msg); }
}  // namespace tga

struct NoAllocScratch;
static int scratch_for(int device, NoAllocScratch** out);

// Flat records longer than this many contacts have their checkpoints built
// on the host, while the device builds the others.  The build decodes a
// record sequentially, and one host core does that about ten times faster
// than one GPU lane; the longest records (hubs of ~10^6 contacts) set the
// build time.
constexpr int64_t kHostCkptMin = 1 << 16;

// Builds the work schedule: headers, chain depths, checkpoints, item lists.
// HS holds host copies of the indexes (its streams and vpos unset); graph and
// ts are the stream files' bytes.
// Column copies of the static item lists (tcol_build): 2 (default) the
// timestamp codes and the flat records' graph bits, 1 the timestamp codes
// only, 0 none (TGA_TCOL).
static int tcol_level() {
  const char* e = std::getenv("TGA_TCOL");
  return e && *e ? std::max(0, std::min(2, std::atoi(e))) : 2;
}
static bool tcol_enabled() { return tcol_level() > 0; }

static int build_schedule(tga_graph* g, Streams HS, const std::vector<uint8_t>& graph, const std::vector<uint8_t>& ts,
                          int talk) {
  const int64_t n = g->n;
  Streams& S = g->S;
  int32_t* d_deg = nullptr;
  int8_t* d_ref = nullptr;
  uint64_t* d_vpos = nullptr;
  int rc = g->alloc(d_deg, (size_t)n);
  if (!rc) rc = g->alloc(d_ref, (size_t)n);
  if (!rc) rc = g->alloc(d_vpos, 2 * (size_t)n);
  if (rc) return rc;
  k_headers<<<g->grid(n), kBlock, 0, g->stream>>>(S, d_deg, d_ref, d_vpos);
  HIP_TRY(hipGetLastError());
  S.vpos = d_vpos;
  std::vector<int32_t> deg((size_t)n);
  std::vector<int8_t> ref((size_t)n);
  HIP_TRY(hipMemcpyAsync(deg.data(), d_deg, (size_t)n * 4, hipMemcpyDeviceToHost, g->stream));
  HIP_TRY(hipMemcpyAsync(ref.data(), d_ref, (size_t)n, hipMemcpyDeviceToHost, g->stream));
  if ((rc = g->check_err())) return rc;
  g->vdeg = d_deg;

  // Chain depth of every record: r > 0 follows the reference x - r.
  std::vector<int8_t> depth((size_t)n, 0);
  int64_t contacts = 0, maxd = 0;
  for (int64_t x = 0; x < n; ++x) {
    if (deg[(size_t)x] < 0) return fail(TGA_EFORMAT, "negative outdegree at vertex " + std::to_string(x));
    contacts += deg[(size_t)x];
    maxd = std::max<int64_t>(maxd, deg[(size_t)x]);
    const int r = ref[(size_t)x];
    if (deg[(size_t)x] == 0 || r == 0) continue;
    if (r > kWindow || r > x) return fail(TGA_EFORMAT, "reference outside the window at vertex " + std::to_string(x));
    const int dp = depth[(size_t)(x - r)] + 1;
    if (dp > kMaxDepth) return fail(TGA_EFORMAT, "reference chain deeper than 3 at vertex " + std::to_string(x));
    depth[(size_t)x] = (int8_t)dp;
  }
  g->info.contacts = contacts;
  g->info.max_degree = maxd;
  S.flat_stride = flat_stride_for(contacts);
  if (const char* fs = std::getenv("TGA_FLAT_STRIDE")) {  // measurement override
    const long v = std::strtol(fs, nullptr, 10);
    if (v >= 16 && v <= 1 << 16 && (v & (v - 1)) == 0) S.flat_stride = (int32_t)v;
  }

  // Long records and their checkpoint slots (ceil(d / stride) chunks each).
  std::vector<int32_t> bigv[kClasses];
  std::vector<int64_t> slotv[kClasses];
  int64_t slots = 0;
  for (int64_t x = 0; x < n; ++x) {
    const int c = depth[(size_t)x];
    if (deg[(size_t)x] <= chunk_stride(c, S.flat_stride)) continue;
    bigv[c].push_back((int32_t)x);
    slotv[c].push_back(slots);
    slots += (deg[(size_t)x] + chunk_stride(c, S.flat_stride) - 1) / chunk_stride(c, S.flat_stride);
  }
  g->nck = slots;
  const size_t nslots = (size_t)std::max<int64_t>(slots, 1);
  if ((rc = g->alloc(g->ck, nslots))) return rc;
  HIP_TRY(hipMemsetAsync(g->ck, 0, sizeof(Ckpt) * nslots, g->stream));
  if ((rc = g->alloc(g->runkey, nslots))) return rc;
  HIP_TRY(hipMemsetAsync(g->runkey, 0xff, sizeof(uint64_t) * nslots, g->stream));
  // Item code of every vertex.  Uploaded before the checkpoints are built so
  // that a record referencing a long record can jump through the latter's
  // checkpoints (classes are built in order of depth, a record's references
  // having smaller depth).
  std::vector<int32_t> vinfo((size_t)n);
  for (int64_t x = 0; x < n; ++x) vinfo[(size_t)x] = deg[(size_t)x] == 0 ? kEmpty : -1 - depth[(size_t)x];
  for (int c = 0; c < kClasses; ++c)
    for (size_t i = 0; i < bigv[c].size(); ++i) vinfo[(size_t)bigv[c][i]] = (int32_t)slotv[c][i];
  if ((rc = g->upload(g->vinfo, vinfo))) return rc;
  g->S.ck = g->ck;
  g->S.vinfo = g->vinfo;
  size_t nbig_all = 0;
  for (int c = 0; c < kClasses; ++c) nbig_all += bigv[c].size();
  int32_t* d_nch = nullptr;
  if ((rc = g->alloc(d_nch, nbig_all))) return rc;
  size_t nch_at[kClasses];
  // Flat records above kHostCkptMin contacts: built on the host (below),
  // overlapping the device's build of the rest; the device list of class 0
  // is the others, their chunk counts mapped back through dev0.
  int64_t host_min = kHostCkptMin;
  if (const char* e = std::getenv("TGA_HOST_CKPT_MIN")) host_min = std::max<long long>(0, std::atoll(e));  // tests
  std::vector<int32_t> host_x, dev0_x;
  std::vector<int64_t> host_slot, dev0_slot;
  std::vector<size_t> host_at, dev0_at;
  for (size_t i = 0; i < bigv[0].size(); ++i) {
    const bool host = deg[(size_t)bigv[0][i]] > host_min;
    (host ? host_x : dev0_x).push_back(bigv[0][i]);
    (host ? host_slot : dev0_slot).push_back(slotv[0][i]);
    (host ? host_at : dev0_at).push_back(i);
  }
  {
    size_t at = 0;
    for (int c = 0; c < kClasses; ++c) {
      nch_at[c] = at;
      at += bigv[c].size();
      const int64_t* d_slot = nullptr;
      if ((rc = g->upload(g->big[c], bigv[c]))) return rc;
      g->n_big[c] = (int32_t)bigv[c].size();
      const std::vector<int32_t>& lx = c == 0 ? dev0_x : bigv[c];
      if (lx.empty()) continue;
      const int32_t* d_x = g->big[c];
      if (c == 0 && (rc = g->upload_const(d_x, dev0_x))) return rc;
      if ((rc = g->upload_const(d_slot, c == 0 ? dev0_slot : slotv[c]))) return rc;
      const int blocks = (int)((lx.size() + 63) / 64);
      const int32_t nb = (int32_t)lx.size();
      if (c == 0) k_ckpt_build<0><<<blocks, 64, 0, g->stream>>>(S, d_x, d_slot, nb, g->ck, d_nch + nch_at[c]);
      if (c == 1) k_ckpt_build<1><<<blocks, 64, 0, g->stream>>>(S, g->big[c], d_slot, nb, g->ck, d_nch + nch_at[c]);
      if (c == 2) k_ckpt_build<2><<<blocks, 64, 0, g->stream>>>(S, g->big[c], d_slot, nb, g->ck, d_nch + nch_at[c]);
      if (c == 3) k_ckpt_build<3><<<blocks, 64, 0, g->stream>>>(S, g->big[c], d_slot, nb, g->ck, d_nch + nch_at[c]);
      HIP_TRY(hipGetLastError());
    }
  }
  std::vector<int32_t> nch_host(nbig_all);
  if (!host_x.empty()) {
    // Host build of the longest flat records (the device works meanwhile):
    // their record starts from the host indexes, then the checkpoints into
    // one host block per record, uploaded into their slots.
    const auto t0 = std::chrono::steady_clock::now();
    int32_t herr = 0;
    HS.err = &herr;
    HS.flat_stride = S.flat_stride;
    HS.burst = S.burst;
    std::unique_ptr<uint64_t[]> hvpos(new uint64_t[2 * (size_t)n + 4]);  // only the entries used are written
    for (const int32_t x : host_x)
      for (int64_t v = x; v <= std::min<int64_t>(x + 1, n - 1); ++v) {
        hvpos[2 * (size_t)v] = ef_get(HS, HS.gi, (uint64_t)v);
        hvpos[2 * (size_t)v + 1] = ef_get(HS, HS.ti, (uint64_t)v);
      }
    HS.vpos = hvpos.get();
    // Host words of the streams, converted only over the host records (and a
    // few words past each, which bit buffers read ahead); the rest of the
    // arrays is never read.
    const size_t gwn = (size_t)S.g_last + 1, twn = (size_t)S.t_last + 1;
    std::unique_ptr<uint64_t[]> hg(new uint64_t[gwn]), ht(new uint64_t[twn]);
    auto words = [](uint64_t* w, size_t nwords, const std::vector<uint8_t>& b, uint64_t bit0, uint64_t bit1) {
      // from kPadWords words before (a buffer started near the stream's end
      // is moved back to stay inside it) to 8 words past
      const size_t lo = (size_t)(bit0 >> 6) > kPadWords ? (size_t)(bit0 >> 6) - kPadWords : 0;
      const size_t hi = std::min(nwords, (size_t)(bit1 >> 6) + 8);
      for (size_t i = lo; i < hi; ++i) {
        uint64_t v = 0;
        for (size_t k = 0; k < 8; ++k) v = (v << 8) | (i * 8 + k < b.size() ? b[i * 8 + k] : 0);
        w[i] = v;
      }
    };
    for (const int32_t x : host_x) {
      const size_t v = (size_t)x, v1 = (size_t)std::min<int64_t>(x + 1, n - 1);
      const uint64_t gend = x + 1 < n ? hvpos[2 * v1] : (uint64_t)graph.size() * 8;
      const uint64_t tend = x + 1 < n ? hvpos[2 * v1 + 1] : (uint64_t)ts.size() * 8;
      words(hg.get(), gwn, graph, hvpos[2 * v], gend);
      words(ht.get(), twn, ts, hvpos[2 * v + 1], tend);
    }
    HS.g = hg.get();
    HS.t = ht.get();
    int64_t hslots = 0;
    for (const int32_t x : host_x) hslots += (deg[(size_t)x] + S.flat_stride - 1) / S.flat_stride;
    // slots of the host records are not contiguous: build into one host block
    // ordered like host_x, then upload record by record
    std::vector<Ckpt> hck((size_t)hslots);
    std::vector<int64_t> hpos(host_x.size());
    {
      int64_t at = 0;
      for (size_t i = 0; i < host_x.size(); ++i) {
        hpos[i] = at;
        at += (deg[(size_t)host_x[i]] + S.flat_stride - 1) / S.flat_stride;
      }
    }
    std::vector<int32_t> hnch(host_x.size(), 0);
    {
      // ckpt_build_record writes split-run slot numbers relative to the
      // record's own first slot (slot0 argument) and the checkpoints at out
      std::atomic<size_t> next{0};
      const int nt = (int)std::max<unsigned>(1, std::min<unsigned>(16, std::thread::hardware_concurrency()));
      auto work = [&]() {
        for (size_t i; (i = next.fetch_add(1)) < host_x.size();)
          hnch[i] = ckpt_build_record<0>(HS, host_x[i], host_slot[i], hck.data() + hpos[i]);
      };
      std::vector<std::thread> pool;
      for (int k = 1; k < nt && k < (int)host_x.size(); ++k) pool.emplace_back(work);
      work();
      for (auto& t : pool) t.join();
    }
    if (herr) return fail(TGA_EFORMAT, "decode error in the host checkpoint build (bits " + std::to_string(herr) + ")");
    for (size_t i = 0; i < host_x.size(); ++i) {
      HIP_TRY(hipMemcpyAsync(g->ck + host_slot[i], hck.data() + hpos[i], sizeof(Ckpt) * (size_t)hnch[i],
                             hipMemcpyHostToDevice, g->stream));
      nch_host[nch_at[0] + host_at[i]] = hnch[i];
    }
    HIP_TRY(hipStreamSynchronize(g->stream));  // the uploads read hck, which ends here
    if (talk)
      std::fprintf(stderr, "[tga_open] host checkpoint build: %zu records, %lld chunks, %.3f s\n", host_x.size(),
                   (long long)hslots,
                   std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count());
  }
  if (!nch_host.empty()) {
    std::vector<int32_t> dn(nbig_all);
    HIP_TRY(hipMemcpyAsync(dn.data(), d_nch, dn.size() * 4, hipMemcpyDeviceToHost, g->stream));
    if ((rc = g->check_err())) return rc;  // synchronizes the stream (the uploads of hck included)
    for (size_t j = 0; j < dev0_at.size(); ++j) nch_host[nch_at[0] + dev0_at[j]] = dn[nch_at[0] + j];
    for (int c = 1; c < kClasses; ++c)
      for (size_t i = 0; i < bigv[c].size(); ++i) nch_host[nch_at[c] + i] = dn[nch_at[c] + i];
  }
  if ((rc = g->check_err())) return rc;

  // Static item lists.  Flat group: chunks first, then whole records longest
  // first (within a length class by id, for index locality).  Chain groups:
  // chunks and whole records together, ordered by an estimate of their decode
  // latency (contacts times a per-level factor, deeper records pulling each
  // copied value through more levels), longest first.
  std::vector<int32_t> nchunk(nslots, 0);
  std::vector<int32_t> items[kGroups];
  std::vector<std::pair<int64_t, int32_t>> chain[kGroups];  // (cost, code)
  auto chain_cost = [](int64_t contacts, int depth) { return contacts * (1 + 2 * depth); };
  for (int c = 0; c < kClasses; ++c) {
    const int32_t stride = chunk_stride(c, S.flat_stride);
    for (size_t i = 0; i < bigv[c].size(); ++i) {
      const int32_t x = bigv[c][i];
      const int32_t nch = nch_host[nch_at[c] + i];
      if (nch != (deg[(size_t)x] + stride - 1) / stride)
        return fail(TGA_EFORMAT, "checkpoint build failed at vertex " + std::to_string(x));
      nchunk[(size_t)slotv[c][i]] = nch;
      for (int32_t j = 0; j < nch; ++j) {
        const int32_t code = -(int32_t)(slotv[c][i] + j) - 1;
        if (c == 0) {
          items[0].push_back(code);
        } else {
          const int64_t len = std::min<int64_t>(stride, deg[(size_t)x] - (int64_t)j * stride);
          chain[group_of(c)].emplace_back(chain_cost(len, c), code);
        }
      }
    }
  }
  {
    std::vector<int32_t> small;
    for (int64_t x = 0; x < n; ++x) {
      const int c = depth[(size_t)x];
      if (deg[(size_t)x] <= 0 || deg[(size_t)x] > chunk_stride(c, S.flat_stride)) continue;
      if (c == 0)
        small.push_back((int32_t)x);
      else
        chain[group_of(c)].emplace_back(chain_cost(deg[(size_t)x], c), (int32_t)x);
    }
    // Longest first, by a stable counting sort.  With the timestamp codes in
    // column order (the default, tcol_build) the key is the exact degree:
    // the 64 records of a wave then have equal lengths, so no lane idles
    // while the longest finishes, and only the short graph-stream reads of
    // neighbouring ids are given up.  On the linear stream the key is the
    // length class, the msb of the degree (id order within a class keeps
    // the lanes' timestamp reads on shared cache lines).
    // TGA_FLAT_CLASS_SHIFT overrides: -1 the exact degree, k >= 0 the msb
    // shifted right by k (coarser classes).
    int cshift = tcol_enabled() ? -1 : 0;
    if (const char* e = std::getenv("TGA_FLAT_CLASS_SHIFT")) cshift = std::max(-1, std::min(6, std::atoi(e)));
    const int32_t ncls = cshift < 0 ? S.flat_stride + 1 : 64;
    auto cls = [&](int32_t x) {
      return cshift < 0 ? deg[(size_t)x] : msb64_host((uint64_t)deg[(size_t)x]) >> cshift;
    };
    std::vector<size_t> at((size_t)ncls + 1, 0);
    for (const int32_t x : small) ++at[(size_t)(ncls - cls(x))];
    for (int32_t k = 1; k <= ncls; ++k) at[(size_t)k] += at[(size_t)k - 1];
    const size_t first = items[0].size();
    items[0].resize(first + small.size());
    for (const int32_t x : small) items[0][first + at[(size_t)(ncls - 1 - cls(x))]++] = x;
    for (int c = 1; c < kGroups; ++c) {
      std::stable_sort(chain[c].begin(), chain[c].end(),
                       [](const std::pair<int64_t, int32_t>& a, const std::pair<int64_t, int32_t>& b) {
                         return a.first > b.first;
                       });
      for (const auto& e : chain[c]) items[c].push_back(e.second);
    }
  }
  int64_t total_items = 0;
  for (int c = 0; c < kGroups; ++c) {
    g->n_items[c] = (uint32_t)items[c].size();
    total_items += (int64_t)items[c].size();
    if ((rc = g->upload(g->items[c], items[c]))) return rc;
  }
  g->info.chunks = total_items;
  g->h_items.assign((size_t)n, 0);
  for (int64_t x = 0; x < n; ++x) {
    const int32_t v = vinfo[(size_t)x];
    g->h_items[(size_t)x] = v == kEmpty ? 0 : (v < 0 ? 1 : nchunk[(size_t)v]);
  }
  return TGA_OK;
}

// Builds the column copy of the timestamp codes of every static item list
// (TsCols): per-position bit ranges on the device, group rows and offsets on
// the host, then the copy; then the same for the graph bits of the flat
// list's whole records.
static int tcol_build(tga_graph* g) {
  if (!tcol_enabled()) return TGA_OK;
  for (int c = 0; c < kGroups; ++c) {
    const uint32_t ni = g->n_items[c];
    if (ni == 0) continue;
    const Items I{g->items[c], ni};
    uint64_t* tstart = nullptr;
    uint32_t* nbits = nullptr;
    int rc;
    if ((rc = g->alloc(tstart, ni)) || (rc = g->alloc(nbits, ni))) return rc;
    k_tcol_len<<<g->grid(ni), kBlock, 0, g->stream>>>(g->S, g->ck, g->nck, I, tstart, nbits);
    HIP_TRY(hipGetLastError());
    std::vector<uint32_t> hb(ni);
    HIP_TRY(hipMemcpyAsync(hb.data(), nbits, (size_t)ni * 4, hipMemcpyDeviceToHost, g->stream));
    HIP_TRY(hipStreamSynchronize(g->stream));
    const size_t ng = ((size_t)ni + 63) / 64;
    std::vector<uint64_t> goff(ng + 1);
    uint64_t at = 0;
    for (size_t q = 0; q < ng; ++q) {
      uint32_t rows = 0;
      for (size_t i = 64 * q; i < std::min<size_t>(ni, 64 * q + 64); ++i) rows = std::max(rows, (hb[i] + 127) / 128);
      goff[q] = at;
      at += 64ull * rows;
    }
    goff[ng] = at;
    uint64_t* w = nullptr;
    uint64_t* dgoff = nullptr;
    if ((rc = g->alloc(w, 2 * std::max<uint64_t>(at, 1))) || (rc = g->upload(dgoff, goff))) return rc;
    const TsCols tc{w, dgoff, nbits};
    k_tcol_fill<<<g->grid(ni), kBlock, 0, g->stream>>>(g->S, Items{g->items[c], ni, tc}, tstart, w);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(g->stream));
    g->release(tstart);
    g->tcols[c] = tc;
  }
  // Graph columns of the flat list's whole records (records without
  // references: every read of their walk stays within their own bits).
  const uint32_t ni = g->n_items[0];
  if (ni == 0 || tcol_level() < 2) return TGA_OK;
  const Items I{g->items[0], ni};
  uint64_t* gstart = nullptr;
  uint32_t* gbits = nullptr;
  int rc;
  if ((rc = g->alloc(gstart, ni)) || (rc = g->alloc(gbits, ni))) return rc;
  k_gcol_len<<<g->grid(ni), kBlock, 0, g->stream>>>(g->S, I, gstart, gbits);
  HIP_TRY(hipGetLastError());
  std::vector<uint32_t> hb(ni);
  HIP_TRY(hipMemcpyAsync(hb.data(), gbits, (size_t)ni * 4, hipMemcpyDeviceToHost, g->stream));
  HIP_TRY(hipStreamSynchronize(g->stream));
  const size_t ng = ((size_t)ni + 63) / 64;
  std::vector<uint64_t> goff(ng + 1);
  uint64_t at = 0;
  for (size_t q = 0; q < ng; ++q) {
    uint32_t rows = 0;
    for (size_t i = 64 * q; i < std::min<size_t>(ni, 64 * q + 64); ++i)
      if (hb[i]) rows = std::max(rows, (hb[i] + 63) / 64 + 6);
    goff[q] = at;
    at += 64ull * rows;
  }
  goff[ng] = at;
  uint64_t* w = nullptr;
  uint64_t* dgoff = nullptr;
  if ((rc = g->alloc(w, std::max<uint64_t>(at, 1))) || (rc = g->upload(dgoff, goff))) return rc;
  TsCols tc = g->tcols[0];
  tc.gw = w;
  tc.ggoff = dgoff;
  tc.gbits = gbits;
  k_gcol_fill<<<g->grid(ni), kBlock, 0, g->stream>>>(g->S, Items{g->items[0], ni, tc}, gstart, w);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipStreamSynchronize(g->stream));
  g->release(gstart);
  g->tcols[0] = tc;
  return TGA_OK;
}

static int open_impl(const std::string& dir, const std::string& name, const tga_dataset_desc* desc, int device,
                     tga_graph** out) {
  if (!out) return fail(TGA_EINVAL, "tga_open: out is NULL");
  *out = nullptr;
  const std::string base = dir + "/" + name;
  int64_t max_label = desc ? desc->max_label : 0;
  if (max_label <= 0) {
    std::ifstream in(base + ".properties");
    if (!in) return fail(TGA_EIO, "cannot open " + base + ".properties (and no max_label given)");
    std::string line;
    while (std::getline(in, line))
      if (line.rfind("nodes=", 0) == 0) max_label = std::stoll(line.substr(6)) - 1;
    if (max_label <= 0) return fail(TGA_EIO, base + ".properties has no usable nodes=");
  }
  if (max_label < 1 || max_label >= INT32_MAX) return fail(TGA_ERANGE, "MAX_LABEL must be in [1, 2^31 - 2]");
  const int64_t n = max_label + 1;
  // Phase times on stderr when TGA_OPEN_TIMES is set.
  const char* tenv = std::getenv("TGA_OPEN_TIMES");
  const bool talk = tenv && *tenv && *tenv != '0';
  auto tick = std::chrono::steady_clock::now();
  std::string phases;
  auto phase = [&](const char* what) {
    const auto now = std::chrono::steady_clock::now();
    char buf[64];
    std::snprintf(buf, sizeof buf, "%s%s %.3f s", phases.empty() ? "" : ", ", what,
                  std::chrono::duration<double>(now - tick).count());
    phases += buf;
    tick = now;
  };

  std::vector<uint8_t> graph, ts, glo, gup, tlo, tup;
  const std::pair<std::string, std::vector<uint8_t>*> files[] = {
      {base + ".graph", &graph},
      {base + ".timestamps", &ts},
      {dir + "/graph_lower_bits.bin", &glo},
      {dir + "/graph_upper_bits.bin", &gup},
      {dir + "/timestamp_lower_bits.bin", &tlo},
      {dir + "/timestamp_upper_bits.bin", &tup}};
  for (const auto& f : files)
    if (!read_file(f.first, *f.second)) return fail(TGA_EIO, "cannot read required file: " + f.first);
  if (ts.size() < 8) return fail(TGA_EFORMAT, "timestamps file shorter than its 64-bit header");
  phase("read");

  // Lower-bit widths, computed from the stream file sizes (bfs.cu:1785-1786).
  const uint64_t qg = (uint64_t)graph.size() * 8 / (uint64_t)max_label;
  const uint64_t qt = (uint64_t)ts.size() * 8 / (uint64_t)max_label;
  if (qg == 0 || qt == 0) return fail(TGA_ERANGE, "stream has fewer bits than MAX_LABEL: log2(0) index width");
  const int lg = msb64_host(qg), lt = msb64_host(qt);

  int64_t min_ts = 0;
  for (int i = 0; i < 8; ++i) min_ts = (int64_t)(((uint64_t)min_ts << 8) | ts[(size_t)i]);
  if (min_ts < INT32_MIN || min_ts > INT32_MAX) return fail(TGA_ERANGE, "minTimestamp outside int32");
  if (desc && desc->min_timestamp != TGA_FROM_FILE && desc->min_timestamp != min_ts)
    return fail(TGA_EINVAL, "minTimestamp in the descriptor does not match the .timestamps header");

  const std::vector<uint64_t> glw = be_words(glo), tlw = be_words(tlo);
  const std::vector<uint64_t> guw = le_words(gup), tuw = le_words(tup);
  phase("words");
  uint64_t gones = 0, tones = 0;
  std::vector<uint64_t> gsamp, gspill, tsamp, tspill;
  build_select(guw, (gup.size() + 7) / 8, gones, gsamp, gspill);
  build_select(tuw, (tup.size() + 7) / 8, tones, tsamp, tspill);
  if (gones < (uint64_t)n || tones < (uint64_t)n)
    return fail(TGA_EFORMAT, "GPU index holds fewer entries than MAX_LABEL + 1");
  phase("select");
  if ((uint64_t)n * (uint64_t)lg > (uint64_t)glo.size() * 8 || (uint64_t)n * (uint64_t)lt > (uint64_t)tlo.size() * 8)
    return fail(TGA_EFORMAT, "GPU index lower-bit file too short");

  auto* g = new tga_graph();
  int rc = TGA_OK;
  auto bail = [&](int code) {
    delete g;
    return code;
  };
  if (hipSetDevice(device) != hipSuccess) return bail(fail(TGA_EDEVICE, "hipSetDevice failed"));
  g->device = device;
  if (const char* e = std::getenv("TGA_GRAPHS")) g->use_graphs = std::atoi(e) != 0;
  if (const char* e = std::getenv("TGA_DENSE_DIV")) g->dense_div = std::atoi(e);
  {
    // the allocation-free entry point's device array, found once per device
    NoAllocScratch* sc = nullptr;
    if ((rc = scratch_for(device, &sc))) return bail(rc);
  }
  if ((rc = g->create_streams())) return bail(rc);
  g->n = n;
  g->ab = (int32_t)std::max<int64_t>(0, desc ? desc->active_begin : 0);
  g->ae = (int32_t)((desc && desc->active_end > 0) ? std::min<int64_t>(desc->active_end, n) : n);
  if (g->ab > g->ae) g->ab = g->ae;
  g->min_ts = (int32_t)min_ts;

  Streams& S = g->S;
  if ((rc = g->upload_stream(S.g, graph)) || (rc = g->upload_stream(S.t, ts)) || (rc = g->upload_const(S.gi.lower, glw)) ||
      (rc = g->upload_const(S.gi.upper, guw)) || (rc = g->upload_const(S.gi.samp, gsamp)) ||
      (rc = g->upload_const(S.gi.spill, gspill)) || (rc = g->upload_const(S.ti.lower, tlw)) ||
      (rc = g->upload_const(S.ti.upper, tuw)) || (rc = g->upload_const(S.ti.samp, tsamp)) ||
      (rc = g->upload_const(S.ti.spill, tspill)))
    return bail(rc);
  S.g_bits = (uint64_t)graph.size() * 8;
  S.t_bits = (uint64_t)ts.size() * 8;
  // Bit buffers index stream words with 32 bits (streams up to 32 GiB).
  const size_t gwords = (graph.size() + 7) / 8 + kPadWords, twords = (ts.size() + 7) / 8 + kPadWords;
  if (gwords >= 0xffffffffull || twords >= 0xffffffffull)
    return bail(fail(TGA_ERANGE, "stream longer than 2^32 words"));
  S.g_last = (uint32_t)(gwords - 1);
  S.t_last = (uint32_t)(twords - 1);
  S.gi.upper_words = (gup.size() + 7) / 8;
  S.ti.upper_words = (tup.size() + 7) / 8;
  S.gi.l = lg;
  S.ti.l = lt;
  S.n = (int32_t)n;
  S.min_ts = (int32_t)min_ts;
  S.burst = kBurstDefault;
  if ((rc = g->alloc(g->d_err, 1))) return bail(rc);
  S.err = g->d_err;
  if (hipMemset(g->d_err, 0, sizeof(int32_t)) != hipSuccess) return bail(fail(TGA_EDEVICE, "hipMemset failed"));
  phase("device+upload");

    {
    Streams HS = S;  // the same fields over host copies (streams: built where needed)
    HS.g = nullptr;
    HS.t = nullptr;
    HS.gi = EFDir{glw.data(), guw.data(), gsamp.data(), gspill.data(), S.gi.upper_words, lg};
    HS.ti = EFDir{tlw.data(), tuw.data(), tsamp.data(), tspill.data(), S.ti.upper_words, lt};
    HS.err = nullptr;
    HS.ck = nullptr;
    HS.vinfo = nullptr;
    HS.vpos = nullptr;
    if ((rc = build_schedule(g, HS, graph, ts, talk ? 1 : 0))) return bail(rc);
  }
  if (hipStreamSynchronize(g->stream) != hipSuccess) return bail(fail(TGA_EDEVICE, "hipStreamSynchronize failed"));
  phase("schedule");
  if ((rc = tcol_build(g))) return bail(rc);
  phase("timestamp columns");

  const size_t qn = (size_t)n + (size_t)g->nck;
  if ((rc = g->alloc(g->q0, (size_t)n)) || (rc = g->alloc(g->i0, (size_t)n)) ||
      (rc = g->alloc(g->i1, (size_t)n)) || (rc = g->alloc(g->i2, (size_t)n)) || (rc = g->alloc(g->xq.q[0], qn)) ||
      (rc = g->alloc(g->xq.q[1], qn)) || (rc = g->alloc(g->xq.q[2], qn)) ||
      (rc = g->alloc(g->xq_big, (size_t)n)) || (rc = g->alloc(g->cnt, 8)) ||
      (rc = g->alloc(g->u64s, 2 * kSumSlots)) || (rc = g->alloc(g->acc, (size_t)n)) || (rc = g->alloc(g->partial, kRedBlocks)) ||
      (rc = g->alloc(g->scal, 4)) || (rc = g->alloc(g->f0, (size_t)n)) || (rc = g->alloc(g->f1, (size_t)n)))
    return bail(rc);

  tga_graph_info& I = g->info;
  I.n = n;
  I.min_timestamp = min_ts;
  I.graph_bytes = graph.size();
  I.timestamps_bytes = ts.size();
  I.index_bytes = glo.size() + gup.size() + tlo.size() + tup.size();
  I.l_graph = lg;
  I.l_timestamps = lt;
  I.active_begin = g->ab;
  I.active_end = g->ae;
  I.device_bytes = g->device_bytes;
  phase("buffers");
  if (talk) std::fprintf(stderr, "[tga_open] %s: %s\n", name.c_str(), phases.c_str());
  *out = g;
  return TGA_OK;
}

extern "C" int tga_open_dir(const char* dir, const char* basename, const tga_dataset_desc* desc, int device,
                            tga_graph** out) {
  if (!dir || !basename) return fail(TGA_EINVAL, "tga_open_dir: NULL path");
  std::string d(dir);
  while (d.size() > 1 && d.back() == '/') d.pop_back();
  return open_impl(d, basename, desc, device, out);
}

extern "C" int tga_open(const char* root, const char* dataset, const tga_dataset_desc* desc, int device,
                        tga_graph** out) {
  if (!dataset) return fail(TGA_EINVAL, "tga_open: NULL dataset");
  std::string r = root ? root : "";
  if (r.empty()) {
    const char* env = std::getenv("TGA_COMPRESS_ROOT");
    r = env && *env ? env : ".";
  }
  while (r.size() > 1 && r.back() == '/') r.pop_back();
  return open_impl(r + "/data/compressed/" + dataset, dataset, desc, device, out);
}

extern "C" void tga_close(tga_graph* g) { delete g; }

extern "C" int tga_info(const tga_graph* g, tga_graph_info* info) {
  if (!g || !info) return fail(TGA_EINVAL, "tga_info: NULL argument");
  *info = g->info;
  info->device_bytes = g->device_bytes;  // what the handle holds now
  return TGA_OK;
}

static bool bad_window(tga_window w) { return w.start > w.end; }

// Launches kernel K over the items of every group (K<0> for flat records,
// K<1> for depth-1 records, K<kMaxDepth> for deeper ones), each group on its
// own stream, and joins them back into the handle's stream.  One profiled
// interval covers the pass.
#define LAUNCH_CLASSES(K, ITEMS, ...)                                                       \
  do {                                                                                      \
    g->kbegin();                                                                            \
    HIP_TRY(hipEventRecord(g->ev_fork, g->stream));                                         \
    for (int c_ = kGroups - 1; c_ >= 0; --c_) {                                             \
      const Items it_ = ITEMS(c_);                                                          \
      if (it_.count == 0) continue;                                                         \
      hipStream_t cs_ = g->cstream[c_];                                                     \
      HIP_TRY(hipStreamWaitEvent(cs_, g->ev_fork, 0));                                      \
      const int grid_ = g->grid(it_.count);                                                 \
      if (c_ == 0) K<0><<<grid_, kBlock, 0, cs_>>>(g->S, g->ck, it_, __VA_ARGS__);          \
      if (c_ == 1) K<1><<<grid_, kBlock, 0, cs_>>>(g->S, g->ck, it_, __VA_ARGS__);          \
      if (c_ == 2) K<kMaxDepth><<<grid_, kBlock, 0, cs_>>>(g->S, g->ck, it_, __VA_ARGS__);  \
      HIP_TRY(hipEventRecord(g->ev_join[c_], cs_));                                         \
      HIP_TRY(hipStreamWaitEvent(g->stream, g->ev_join[c_], 0));                            \
    }                                                                                       \
    g->kend();                                                                              \
    HIP_TRY(hipGetLastError());                                                             \
  } while (0)

#define STATIC_ITEMS(c) (Items{g->items[c], g->n_items[c], g->tcols[c]})
#define QUEUED_ITEMS(c) (Items{g->xq.q[c], hc_[c]})
#define LAUNCH_STATIC(K, ...) LAUNCH_CLASSES(K, STATIC_ITEMS, __VA_ARGS__)
#define LISTED_ITEMS(c) (Items{qs_.q[c], hc_[c]})
#define LAUNCH_LISTED(K, QS, HC, ...)                \
  do {                                               \
    const Queues qs_ = (QS);                         \
    const uint32_t* hc_ = (HC);                      \
    LAUNCH_CLASSES(K, LISTED_ITEMS, __VA_ARGS__);    \
  } while (0)
#define LAUNCH_QUEUED(K, HC, ...)         \
  do {                                    \
    const uint32_t* hc_ = (HC);           \
    LAUNCH_CLASSES(K, QUEUED_ITEMS, __VA_ARGS__); \
  } while (0)

// Buffers of one neighbor-access call: checksum slots, queue counters, the
// copied vertex list and the group item queues, with their capacities.
struct RaBufs {
  unsigned long long* sums;  // 2 * kSumSlots
  uint32_t* counters;        // kGroups + 1
  int32_t* front;            // cap_verts
  Queues q;                  // kGroups queues of cap_items
  int64_t cap_verts, cap_items;
  bool batch;                // split lists that do not fit into several launches (else reject)
};

static int ra_run(tga_graph* g, const int32_t* verts, int64_t nverts, const RaBufs& B, uint64_t* checksum,
                  int64_t* contacts, const char* who) {
  k_fill<unsigned long long><<<1, kBlock, 0, g->stream>>>(B.sums, 2 * kSumSlots, 0ull);
  HIP_TRY(hipGetLastError());
  auto zero_counters = [&]() -> int {
    k_fill<uint32_t><<<1, kBlock, 0, g->stream>>>(B.counters, kGroups + 1, 0u);
    HIP_TRY(hipGetLastError());
    return TGA_OK;
  };
  if (nverts > 0) {
    // Listed vertices: expanded into flat/chain items like a frontier, in
    // batches whose expansion fits the item queues (a list may repeat a long
    // record many times).
    for (int64_t i = 0; i < nverts; ++i)
      if (verts[i] < 0 || verts[i] >= g->n) return fail(TGA_EINVAL, std::string(who) + ": vertex out of range");
    int64_t i0 = 0;
    while (i0 < nverts) {
      int64_t i1 = i0, items = 0;
      while (i1 < nverts && i1 - i0 < B.cap_verts && items + g->h_items[(size_t)verts[i1]] <= B.cap_items)
        items += g->h_items[(size_t)verts[i1++]];
      if (i1 < nverts && !B.batch)
        return fail(TGA_ERANGE, std::string(who) + ": the list exceeds " + std::to_string(B.cap_verts) +
                                    " vertices or " + std::to_string(B.cap_items) + " work items");
      if (i1 == i0) return fail(TGA_ERANGE, std::string(who) + ": a record has more chunks than the item queues");
      HIP_TRY(hipMemcpyAsync(B.front, verts + i0, (size_t)(i1 - i0) * sizeof(int32_t), hipMemcpyHostToDevice,
                             g->stream));
      uint32_t hc[kGroups + 1];
      int rc = zero_counters();
      if (!rc) rc = g->expand_into(B.front, 0, (uint32_t)(i1 - i0), Sel{kSelAll, nullptr, 0}, false, nullptr, B.q,
                                   B.counters, hc);
      if (rc) return rc;
      LAUNCH_LISTED(k_ra, B.q, hc, B.sums);
      i0 = i1;
    }
  } else if (g->ab == 0 && g->ae == g->n) {
    LAUNCH_STATIC(k_ra, B.sums);
  } else {
    int64_t items = 0;
    for (int64_t x = g->ab; x < g->ae; ++x) items += g->h_items[(size_t)x];
    if (items > B.cap_items)
      return fail(TGA_ERANGE, std::string(who) + ": the active range expands to more than " +
                                  std::to_string(B.cap_items) + " work items");
    uint32_t hc[kGroups + 1];
    int rc = zero_counters();
    if (!rc) rc = g->expand_into(nullptr, g->ab, (uint32_t)(g->ae - g->ab), Sel{kSelAll, nullptr, 0}, false,
                                 nullptr, B.q, B.counters, hc);
    if (rc) return rc;
    LAUNCH_LISTED(k_ra, B.q, hc, B.sums);
  }
  unsigned long long hv[2 * kSumSlots];
  HIP_TRY(hipMemcpyAsync(hv, B.sums, sizeof hv, hipMemcpyDeviceToHost, g->stream));
  int rc = g->check_err();
  if (rc) return rc;
  unsigned long long cs = 0, cnt = 0;
  for (int i = 0; i < kSumSlots; ++i) {
    cs += hv[2 * i];
    cnt += hv[2 * i + 1];
  }
  if (checksum) *checksum = cs;
  if (contacts) *contacts = (int64_t)cnt;
  return TGA_OK;
}

extern "C" int tga_neighbor_access(tga_graph* g, const int32_t* verts, int64_t nverts, uint64_t* checksum,
                                   int64_t* contacts) {
  if (!g || (nverts > 0 && !verts) || nverts < 0) return fail(TGA_EINVAL, "tga_neighbor_access: bad arguments");
  HIP_TRY(hipSetDevice(g->device));
  const RaBufs B{g->u64s, g->cnt, g->q0, g->xq, g->n, g->n + g->nck, true};
  return ra_run(g, verts, nverts, B, checksum, contacts, "tga_neighbor_access");
}

// Allocation-free neighbor access: the same pass with every buffer beyond the
// caller's inputs and outputs in one file-scope device array, so it calls no
// hipMalloc (for callers that cannot allocate: graph capture, tight serving
// loops).  Each GPU holds its own copy of the array, which host code reaches
// through the per-device address from hipGetSymbolAddress (the array's name
// is not a device address in host code).  Calls that overlap on one GPU share
// the array and race.
constexpr int64_t kNoAllocVerts = 1 << 16;  // listed vertices per call
constexpr int64_t kNoAllocItems = 1 << 18;  // work items (records and chunks) per group queue
struct NoAllocScratch {
  unsigned long long sums[2 * kSumSlots];
  uint32_t counters[kGroups + 1];
  int32_t front[kNoAllocVerts];
  int32_t queue[kGroups][kNoAllocItems];
};
static __device__ __align__(256) unsigned char scratch[sizeof(NoAllocScratch)];

namespace {
std::mutex g_scratch_mu;
std::vector<NoAllocScratch*> g_scratch_ptr;  // per device, found once
}  // namespace

static int scratch_for(int device, NoAllocScratch** out) {
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if ((size_t)device >= g_scratch_ptr.size()) g_scratch_ptr.resize((size_t)device + 1, nullptr);
  if (!g_scratch_ptr[(size_t)device]) {
    void* p = nullptr;
    HIP_TRY(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));  // the current device's copy
    g_scratch_ptr[(size_t)device] = (NoAllocScratch*)p;
  }
  *out = g_scratch_ptr[(size_t)device];
  return TGA_OK;
}

extern "C" int tga_neighbor_access_noalloc(tga_graph* g, const int32_t* verts, int64_t nverts, uint64_t* checksum,
                                           int64_t* contacts) {
  if (!g || (nverts > 0 && !verts) || nverts < 0)
    return fail(TGA_EINVAL, "tga_neighbor_access_noalloc: bad arguments");
  if (nverts > kNoAllocVerts)
    return fail(TGA_ERANGE, "tga_neighbor_access_noalloc: at most " + std::to_string(kNoAllocVerts) +
                                " listed vertices per call");
  HIP_TRY(hipSetDevice(g->device));
  NoAllocScratch* sc = nullptr;
  int rc = scratch_for(g->device, &sc);
  if (rc) return rc;
  Queues q;
  for (int c = 0; c < kGroups; ++c) q.q[c] = sc->queue[c];
  const RaBufs B{sc->sums, sc->counters, sc->front, q, kNoAllocVerts, kNoAllocItems, false};
  return ra_run(g, verts, nverts, B, checksum, contacts, "tga_neighbor_access_noalloc");
}

// Autotuner: the walk's burst length (contacts decoded per refill of the
// neighbor-run ring) decides how far the refill's branches are amortised
// (long runs of one neighbor favour long bursts, single-contact neighbors
// short ones).  Each candidate is timed over full neighbor-access passes and
// the fastest is kept for every later call on the handle.
extern "C" int tga_autotune(tga_graph* g, int32_t passes, tga_tune_report* report) {
  if (!g || passes < 1 || passes > 1000) return fail(TGA_EINVAL, "tga_autotune: bad arguments");
  HIP_TRY(hipSetDevice(g->device));
  static const int32_t cand[] = {4, 8, 12, 16, 20, 24, 32};
  constexpr int nc = (int)(sizeof cand / sizeof cand[0]);
  static_assert(nc <= TGA_TUNE_MAX, "report too small");
  tga_tune_report rep{};
  const bool was = g->prof;
  g->harvest();
  double best = 0;
  int32_t pick = g->S.burst;
  for (int i = 0; i < nc; ++i) {
    g->S.burst = cand[i];
    int rc = tga_neighbor_access(g, nullptr, 0, nullptr, nullptr);  // warm
    if (rc) return rc;
    g->prof = true;
    g->prof_ms = 0;
    g->prof_launches = 0;
    for (int32_t p = 0; p < passes && !rc; ++p) rc = tga_neighbor_access(g, nullptr, 0, nullptr, nullptr);
    g->harvest();
    g->prof = was;
    if (rc) return rc;
    const double ms = g->prof_ms / passes;
    rep.setting[i] = cand[i];
    rep.pass_ms[i] = ms;
    if (i == 0 || ms < best) {
      best = ms;
      pick = cand[i];
    }
  }
  g->prof_ms = 0;
  g->prof_launches = 0;
  g->S.burst = pick;
  rep.candidates = nc;
  rep.burst = pick;
  rep.flat_stride = g->S.flat_stride;
  if (report) *report = rep;
  return TGA_OK;
}

// Timestamp decode methods compared (tga_ts_decode_bench): the codes of the
// flat list's whole records, summed as uint32 times, decoded
//  0: one lane per record through the walk's window over the record's
//     column copy (the method every decode pass uses);
//  1: one wave per record, speculatively: lane l decodes a code at bit
//     cursor + l, the true codes are then found by hopping from lane to lane
//     by code length (a lane index read out per code), and the cursor moves
//     past the last code that started in the window (the ballot/shuffle
//     chain of the reference's Yahoo/neighbor_access_supernode_warp.cu:
//     1697-1725, read from the stream with every lane's loads falling into
//     the same few words).
template <int METHOD>
__global__ void __launch_bounds__(kBlock) k_ts_bench(Streams S, Items I, const int32_t* vdeg,
                                                     unsigned long long* out) {
  unsigned long long acc = 0;
  if constexpr (METHOD == 0) {
    LANE_RING();
    (void)ring_;
    FOR_ITEMS(I, i) {
      const int32_t x = I.code[i];
      if (x < 0) continue;
      TsStage ts = I.stage(stage_, i);
      if (!ts.uw) continue;
      stage_init_col(ts, S.min_ts);
      const int32_t d = vdeg[x];
      stage_fill(S, ts);
      for (int32_t j = 0; j < d; ++j) {
        if ((j & 15) == 0) stage_fill(S, ts);
        acc += (uint32_t)stage_ts_next(S, ts);
      }
    }
  } else {
    const int lane = (int)(threadIdx.x & 63);
    const uint32_t waves = (gridDim.x * blockDim.x) >> 6;
    for (uint32_t i = (blockIdx.x * blockDim.x + threadIdx.x) >> 6; i < I.count; i += waves) {
      const int32_t x = I.code[i];
      if (x < 0) continue;
      uint64_t cur = S.vpos[2 * (size_t)x + 1];
      int32_t left = vdeg[x];
      int32_t prev = S.min_ts;
      while (left > 0) {
        const uint64_t w = peek64(S.t, cur + (uint64_t)lane);
        uint32_t len;
        const int32_t d = zeta2_delta((uint32_t)(w >> 32), (uint32_t)w, len);
        uint32_t off = 0;
        while (off < 64 && left > 0) {
          const uint32_t l = (uint32_t)__shfl((int32_t)len, (int)off, 64);
          prev = (int32_t)((uint32_t)prev + (uint32_t)__shfl(d, (int)off, 64));
          if (lane == 0) acc += (uint32_t)prev;
          off += l;
          --left;
        }
        cur += off;
      }
    }
  }
  const unsigned long long t = wave_sum(acc);
  if ((threadIdx.x & 63) == 0 && t) atomicAdd(out, t);
}

extern "C" int tga_ts_decode_bench(tga_graph* g, int32_t method, int32_t passes, double* ms, uint64_t* sum) {
  if (!g || method < 0 || method > 1 || passes < 1 || passes > 1000 || !ms || !sum)
    return fail(TGA_EINVAL, "tga_ts_decode_bench: bad arguments");
  if (!g->tcols[0].w) return fail(TGA_EINVAL, "tga_ts_decode_bench: the handle has no column copies (TGA_TCOL=0)");
  HIP_TRY(hipSetDevice(g->device));
  const Items I = STATIC_ITEMS(0);
  hipEvent_t a, b;
  HIP_TRY(hipEventCreate(&a));
  HIP_TRY(hipEventCreate(&b));
  // Resident grids (eight blocks per compute unit, the items handed out by
  // grid stride): one short-lived wave per block of items would make the
  // lane-per-record kernel wait on wave launches rather than decode.
  int cus = 256;
  (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, g->device);
  const int grid = std::max(1, std::min(8 * cus, method == 0 ? g->grid(I.count) : g->grid((int64_t)I.count * 64)));
  for (int32_t p = 0; p <= passes; ++p) {  // pass 0 warms up
    if (p == 1) HIP_TRY(hipEventRecord(a, g->stream));
    k_fill<unsigned long long><<<1, 64, 0, g->stream>>>(g->u64s, 1, 0ull);
    if (method == 0)
      k_ts_bench<0><<<grid, kBlock, 0, g->stream>>>(g->S, I, g->vdeg, g->u64s);
    else
      k_ts_bench<1><<<grid, kBlock, 0, g->stream>>>(g->S, I, g->vdeg, g->u64s);
    HIP_TRY(hipGetLastError());
  }
  HIP_TRY(hipEventRecord(b, g->stream));
  HIP_TRY(hipEventSynchronize(b));
  float t = 0;
  HIP_TRY(hipEventElapsedTime(&t, a, b));
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
  *ms = (double)t / passes;
  HIP_TRY(hipMemcpy(sum, g->u64s, sizeof *sum, hipMemcpyDeviceToHost));
  return g->check_err();
}

// Scan tiles for m counts (k_scan_tiles).
static int64_t scan_tiles(int64_t m) { return (m + 1 + kScanTile - 1) / kScanTile; }

// List state (defined with the graph coloring lists below).
static int list_budget(uint64_t need, bool held, bool* ok);
static uint64_t item_scan_bytes(const tga_graph* g);
static int item_scan_alloc(tga_graph* g);
static int scan_counts(tga_graph* g, const uint32_t* c, int64_t m, uint64_t* off, uint64_t* tsum = nullptr);
static int item_scan(tga_graph* g);

// Builds the contact-time cache (g->ct, g->coff) on first use; *ok is false,
// with nothing built, when it does not fit list_budget.
static int ensure_ctimes(tga_graph* g, bool* ok) {
  const uint64_t contacts = (uint64_t)std::max<int64_t>(g->info.contacts, 0);
  int rc = list_budget(contacts * 8 + (uint64_t)(g->n + 1) * 8 + item_scan_bytes(g), g->ct != nullptr, ok);
  if (rc || !*ok || g->ct) return rc;
  AllocTxn txn(g);
  if ((rc = item_scan_alloc(g)) || (rc = g->alloc(g->coff, (size_t)g->n + 1)) ||
      (rc = scan_counts(g, (const uint32_t*)g->vdeg, g->n, g->coff)) || (rc = g->alloc(g->ct, (size_t)contacts)))
    return rc;
  const Window A{TGA_FULL_WINDOW_START, TGA_FULL_WINDOW_END, g->ab, g->ae};
  LAUNCH_STATIC(k_ctimes, A, g->ct, (const uint64_t*)g->coff);
  txn.commit();
  return TGA_OK;
}

extern "C" int tga_degree_centrality(tga_graph* g, tga_window w, int32_t* count, int32_t* centrality, uint32_t flags) {
  if (!g || !count || bad_window(w)) return fail(TGA_EINVAL, "tga_degree_centrality: bad arguments");
  HIP_TRY(hipSetDevice(g->device));
  const Window W = g->window(w);
  k_fill<int32_t><<<g->grid(g->n), kBlock, 0, g->stream>>>(g->i0, g->n, 0);
  int32_t* centr = centrality ? g->i1 : nullptr;
  if (centr) {
    // Vertices with no decoded record keep the formula's value for a zero count.
    const int32_t c0 = (w.start == w.end) ? 0 : 1;
    k_fill<int32_t><<<g->grid(g->n), kBlock, 0, g->stream>>>(centr, g->n, c0);
  }
  // Long records: one wave per record over the contact-time cache (k_dc_rec)
  // when the cache fits list_budget, else one thread decoding each record.
  bool rec = false;
  if (centr) {
    int rc = ensure_ctimes(g, &rec);
    if (rc) return rc;
  }
  LAUNCH_STATIC(k_dc, W, g->i0, centr);
  if (centr) {
    for (int c = 0; c < kClasses; ++c) {
      if (!g->n_big[c]) continue;
      if (rec) {
        k_dc_rec<<<(g->n_big[c] + kBlock / 64 - 1) / (kBlock / 64), kBlock, 0, g->stream>>>(
            g->big[c], g->n_big[c], g->vdeg, g->coff, g->ct, W, centr);
        continue;
      }
      const int blocks = (g->n_big[c] + 63) / 64;
      if (c == 0) k_dc_long<0><<<blocks, 64, 0, g->stream>>>(g->S, g->big[c], g->n_big[c], W, centr);
      if (c == 1) k_dc_long<1><<<blocks, 64, 0, g->stream>>>(g->S, g->big[c], g->n_big[c], W, centr);
      if (c == 2) k_dc_long<2><<<blocks, 64, 0, g->stream>>>(g->S, g->big[c], g->n_big[c], W, centr);
      if (c == 3) k_dc_long<3><<<blocks, 64, 0, g->stream>>>(g->S, g->big[c], g->n_big[c], W, centr);
    }
    HIP_TRY(hipGetLastError());
  }
  int rc = g->out_copy(count, g->i0, (size_t)g->n * 4, flags);
  if (!rc && centrality) rc = g->out_copy(centrality, centr, (size_t)g->n * 4, flags);
  if (rc) return rc;
  return g->check_err();
}

extern "C" int tga_degree_centrality_batch(tga_graph* g, const tga_window* windows, int32_t nq, int32_t* count,
                                           int32_t* centrality, uint32_t flags) {
  if (!g || !windows || !count || nq < 1 || nq > 64)
    return fail(TGA_EINVAL, "tga_degree_centrality_batch: bad arguments (1 <= nq <= 64)");
  for (int32_t q = 0; q < nq; ++q)
    if (bad_window(windows[q])) return fail(TGA_EINVAL, "tga_degree_centrality_batch: window start > end");
  HIP_TRY(hipSetDevice(g->device));
  const int64_t n = g->n;
  bool ok = false;
  int rc = ensure_ctimes(g, &ok);
  if (rc) return rc;
  if (!ok) {
    // No room for the cache: one call per window.
    for (int32_t q = 0; q < nq; ++q)
      if ((rc = tga_degree_centrality(g, windows[q], count + q * n, centrality ? centrality + q * n : nullptr, flags)))
        return rc;
    return TGA_OK;
  }
  if (!g->dcb_win && (rc = g->alloc(g->dcb_win, 2 * 64))) return rc;
  int64_t hw[2 * 64];
  for (int32_t q = 0; q < nq; ++q) {
    hw[2 * q] = windows[q].start;
    hw[2 * q + 1] = windows[q].end;
  }
  HIP_TRY(hipMemcpyAsync(g->dcb_win, hw, sizeof(int64_t) * 2 * (size_t)nq, hipMemcpyHostToDevice, g->stream));
  int32_t *cnt = count, *cen = centrality;
  if (!(flags & TGA_OUT_DEVICE)) {
    const int32_t rows = centrality ? 2 * nq : nq;
    if ((rc = g->rows_buffer(g->dcb_out, g->dcb_rows, rows))) return rc;
    cnt = g->dcb_out;
    cen = centrality ? g->dcb_out + (int64_t)nq * n : nullptr;
  }
  const Window A{TGA_FULL_WINDOW_START, TGA_FULL_WINDOW_END, g->ab, g->ae};
  g->kbegin();
  k_dcb<<<g->grid(n * 64), kBlock, 0, g->stream>>>(A, (int32_t)n, g->vinfo, g->vdeg, g->coff, g->ct, g->dcb_win, nq,
                                                 cnt, cen);
  for (int c = 0; c < kClasses; ++c)
    if (g->n_big[c])
      k_dcb_long<<<g->grid((int64_t)g->n_big[c] * nq * 64), kBlock, 0, g->stream>>>(
          g->big[c], g->n_big[c], A, (int32_t)n, g->vdeg, g->coff, g->ct, g->dcb_win, nq, cnt, cen);
  g->kend();
  HIP_TRY(hipGetLastError());
  if (!(flags & TGA_OUT_DEVICE)) {
    HIP_TRY(hipMemcpyAsync(count, cnt, sizeof(int32_t) * (size_t)nq * (size_t)n, hipMemcpyDeviceToHost, g->stream));
    if (centrality)
      HIP_TRY(hipMemcpyAsync(centrality, cen, sizeof(int32_t) * (size_t)nq * (size_t)n, hipMemcpyDeviceToHost,
                             g->stream));
  }
  return g->check_err();
}

// Device memory freed at scope exit (build-time temporaries).
struct DevTmp {
  void* p = nullptr;
  ~DevTmp() {
    if (p) (void)hipFree(p);
  }
  template <class T>
  hipError_t get(T*& out, size_t count) {
    hipError_t e = hipMalloc(&p, std::max<size_t>(count, 1) * sizeof(T));
    out = (T*)p;
    return e;
  }
};

static int read_u64(tga_graph* g, const uint64_t* dev, uint64_t* host) {
  HIP_TRY(hipMemcpyAsync(host, dev, sizeof(uint64_t), hipMemcpyDeviceToHost, g->stream));
  HIP_TRY(hipStreamSynchronize(g->stream));
  return TGA_OK;
}

// Builds the batched PageRank run structure (k_runs) once per handle; *ok is
// false, with nothing built, when it and the batch state do not fit
// list_budget (the contact-time cache is needed too).
static int ensure_prb(tga_graph* g, bool* ok) {
  if (g->prb_runs) {
    *ok = true;
    return TGA_OK;
  }
  int rc = ensure_ctimes(g, ok);
  if (rc || !*ok) return rc;
  const int64_t n = g->n, m = n + g->nck;
  const uint64_t cap = (uint64_t)std::max<int64_t>(g->info.contacts, 1);
  const int end_bit = 32 - __builtin_clz((uint32_t)n | 1u);
  size_t tmp_bytes = 0;
  HIP_TRY(sort_pairs_u32_u64(nullptr, tmp_bytes, nullptr, nullptr, nullptr, nullptr, (size_t)cap, end_bit,
                             g->stream));
  // temporaries 40 B per run + sort storage; kept: 16 B per run, 8 + 8 + 4 B
  // per vertex, batch state 1.8 KB per vertex (64 columns), segments
  const uint64_t need = cap * (56 + 8 + 16) + cap / 2 * 256 + tmp_bytes +
                        (uint64_t)n * (20 + 64 * (4 + 4 + 4 + 8 + 8) + 24 + 40) + 64 * 1024;
  if ((rc = list_budget(need, false, ok)) || !*ok) return rc;
  const Window A{TGA_FULL_WINDOW_START, TGA_FULL_WINDOW_END, g->ab, g->ae};
  hipStream_t st = g->stream;
  AllocTxn txn(g);
  if ((rc = item_scan_alloc(g))) return rc;
  k_fill<uint32_t><<<g->grid(m), kBlock, 0, st>>>(g->lcnt, m, 0u);
  LAUNCH_STATIC(k_runs, A, g->lcnt, (const uint64_t*)nullptr, (int32_t*)nullptr, (int32_t*)nullptr,
                (int32_t*)nullptr);
  if ((rc = item_scan(g))) return rc;
  uint64_t nr = 0;
  if ((rc = read_u64(g, g->loff + m, &nr))) return rc;
  if (nr > cap) return fail(TGA_EFORMAT, "neighbor runs exceed the contact count");
  DevTmp t_rx, t_rr, t_rk, t_k1, t_keys, t_vals, t_sk, t_sv, t_sort, t_nseg;
  int32_t *rx, *rr, *rk, *k1;
  uint32_t *keys, *sk, *nseg;
  uint64_t *vals, *sv;
  uint8_t* sort_tmp;
  HIP_TRY(t_rx.get(rx, nr));
  HIP_TRY(t_rr.get(rr, nr));
  HIP_TRY(t_rk.get(rk, nr));
  HIP_TRY(t_k1.get(k1, nr));
  HIP_TRY(t_keys.get(keys, nr));
  HIP_TRY(t_vals.get(vals, nr));
  HIP_TRY(t_sk.get(sk, nr));
  HIP_TRY(t_sv.get(sv, nr));
  HIP_TRY(t_sort.get(sort_tmp, tmp_bytes));
  HIP_TRY(t_nseg.get(nseg, (size_t)n));
  LAUNCH_STATIC(k_runs, A, (uint32_t*)nullptr, g->loff, rx, rr, rk);
  k_run_keys<<<g->grid((int64_t)nr), kBlock, 0, st>>>((int64_t)nr, rx, rr, rk, g->vdeg, A, (int32_t)n, k1, keys,
                                                      vals);
  HIP_TRY(hipGetLastError());
  size_t tb = tmp_bytes;
  HIP_TRY(sort_pairs_u32_u64(sort_tmp, tb, keys, sk, vals, sv, (size_t)nr, end_bit, st));
  if ((rc = g->alloc(g->prb_in_off, (size_t)n + 1)) || (rc = g->alloc(g->prb_seg_off, (size_t)n + 1))) return rc;
  k_key_bounds<<<g->grid(n + 1), kBlock, 0, st>>>(sk, nr, (int32_t)n, g->prb_in_off);
  HIP_TRY(hipGetLastError());
  uint64_t nv = 0;
  if ((rc = read_u64(g, g->prb_in_off + n, &nv))) return rc;
  int4* runs = nullptr;
  if ((rc = g->alloc(runs, (size_t)nv)) || (rc = g->alloc(g->prb_f0, (size_t)nv))) return rc;
  k_run_gather<<<g->grid((int64_t)nv), kBlock, 0, st>>>(nv, sv, rx, rk, k1, g->coff, g->ct, runs, g->prb_f0);
  {
    DevTmp t_flag, t_moff, t_mts;
    uint32_t* flag;
    uint64_t *moff, *mts;
    HIP_TRY(t_flag.get(flag, nv));
    HIP_TRY(t_moff.get(moff, nv + 1));
    HIP_TRY(t_mts.get(mts, (size_t)scan_tiles((int64_t)nv)));
    k_multi_flag<<<g->grid((int64_t)nv), kBlock, 0, st>>>(nv, runs, flag);
    HIP_TRY(hipGetLastError());
    if ((rc = scan_counts(g, flag, (int64_t)nv, moff, mts))) return rc;
    uint64_t nmulti = 0;
    if ((rc = read_u64(g, moff + nv, &nmulti))) return rc;
    if ((rc = g->alloc(g->prb_midx, (size_t)nv)) || (rc = g->alloc(g->prb_mlist, (size_t)nmulti))) return rc;
    k_multi_list<<<g->grid((int64_t)nv), kBlock, 0, st>>>(nv, runs, moff, g->prb_midx, g->prb_mlist);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(st));  // before the temporaries go
    g->prb_nmulti = (int32_t)nmulti;
  }
  {
    // out-runs for batched HITS: fill order, with the sorted index of each
    DevTmp t_inv;
    int32_t* inv;
    HIP_TRY(t_inv.get(inv, nr));
    k_fill<int32_t><<<g->grid((int64_t)nr), kBlock, 0, st>>>(inv, (int64_t)nr, -1);
    k_run_inv<<<g->grid((int64_t)nv), kBlock, 0, st>>>(nv, sv, inv);
    if ((rc = g->alloc(g->prb_out_runs, (size_t)nr)) || (rc = g->alloc(g->prb_obeg, (size_t)n)) ||
        (rc = g->alloc(g->prb_olen, (size_t)n)) || (rc = g->alloc(g->prb_oseg_off, (size_t)n + 1)) ||
        (rc = g->alloc(g->prb_omrows, (size_t)n)) || (rc = g->alloc(g->prb_onm, 1)))
      return rc;
    k_out_runs<<<g->grid((int64_t)nr), kBlock, 0, st>>>((int64_t)nr, rr, inv, runs, g->prb_out_runs);
    k_out_rows<<<g->grid(n), kBlock, 0, st>>>(A, (int32_t)n, g->vinfo, g->ck, g->loff, g->prb_obeg, g->prb_olen, nseg);
    HIP_TRY(hipGetLastError());
    if ((rc = scan_counts(g, nseg, n, g->prb_oseg_off))) return rc;
    uint64_t nos = 0;
    if ((rc = read_u64(g, g->prb_oseg_off + n, &nos))) return rc;
    if ((rc = g->alloc(g->prb_oseg_row, (size_t)nos)) || (rc = g->alloc(g->prb_oseg_beg, (size_t)nos)) ||
        (rc = g->alloc(g->prb_oseg_len, (size_t)nos)))
      return rc;
    k_fill<uint32_t><<<1, kBlock, 0, st>>>(g->prb_onm, 1, 0u);
    k_oseg_fill<<<g->grid(n), kBlock, 0, st>>>((int32_t)n, g->prb_obeg, g->prb_olen, g->prb_oseg_off,
                                               g->prb_oseg_row, g->prb_oseg_beg, g->prb_oseg_len, g->prb_omrows,
                                               g->prb_onm);
    HIP_TRY(hipGetLastError());
    uint32_t onm = 0;
    if ((rc = g->read_counts(g->prb_onm, &onm, 1))) return rc;  // synchronizes before the temporaries go
    g->prb_noseg = (int64_t)nos;
    g->prb_nomrows = (int32_t)onm;
  }
  k_seg_count<<<g->grid(n), kBlock, 0, st>>>((int32_t)n, g->prb_in_off, nseg);
  HIP_TRY(hipGetLastError());
  if ((rc = scan_counts(g, nseg, n, g->prb_seg_off))) return rc;
  uint64_t ns = 0;
  if ((rc = read_u64(g, g->prb_seg_off + n, &ns))) return rc;
  if ((rc = g->alloc(g->prb_seg_row, (size_t)ns)) || (rc = g->alloc(g->prb_seg_beg, (size_t)ns)) ||
      (rc = g->alloc(g->prb_seg_len, (size_t)ns)) || (rc = g->alloc(g->prb_mrows, (size_t)n)) ||
      (rc = g->alloc(g->prb_nm, 2)) || (rc = g->alloc(g->prb_long, (size_t)n)))
    return rc;
  k_fill<uint32_t><<<1, kBlock, 0, st>>>(g->prb_nm, 2, 0u);
  k_seg_fill<<<g->grid(n), kBlock, 0, st>>>((int32_t)n, g->prb_in_off, g->prb_seg_off, g->prb_seg_row,
                                            g->prb_seg_beg, g->prb_seg_len, g->prb_mrows, g->prb_nm);
  k_long_list<<<g->grid(n), kBlock, 0, st>>>(A, (int32_t)n, g->vdeg, g->prb_long, g->prb_nm + 1);
  HIP_TRY(hipGetLastError());
  uint32_t hn[2] = {0, 0};
  if ((rc = g->read_counts(g->prb_nm, hn, 2))) return rc;
  g->prb_nseg = (int64_t)ns;
  g->prb_nmrows = (int32_t)hn[0];
  g->prb_nlong = (int32_t)hn[1];
  if (g->prb_nlong) {
    // items: at most contacts / kPrbDuChunk + one per long record
    const size_t cap_items = (size_t)(cap / kPrbDuChunk) + (size_t)g->prb_nlong;
    if ((rc = g->alloc(g->prb_du_items, cap_items))) return rc;
    k_fill<uint32_t><<<1, kBlock, 0, st>>>(g->prb_nm + 1, 1, 0u);  // slot 1 is free once read
    k_prb_du_items<<<g->grid(g->prb_nlong), kBlock, 0, st>>>(g->prb_long, g->prb_nlong, g->vdeg, g->prb_du_items,
                                                             g->prb_nm + 1);
    HIP_TRY(hipGetLastError());
    uint32_t ni = 0;
    if ((rc = g->read_counts(g->prb_nm + 1, &ni, 1))) return rc;
    g->prb_ndu_items = (int32_t)ni;
  }
  g->prb_runs = runs;  // built
  txn.commit();
  return TGA_OK;
}

extern "C" int tga_pagerank_batch(tga_graph* g, const tga_window* windows, int32_t nq, int32_t iters, float damping,
                                  float* rank, uint32_t flags) {
  if (!g || !windows || !rank || nq < 1 || nq > 64 || iters < 0)
    return fail(TGA_EINVAL, "tga_pagerank_batch: bad arguments (1 <= nq <= 64)");
  for (int32_t q = 0; q < nq; ++q)
    if (bad_window(windows[q])) return fail(TGA_EINVAL, "tga_pagerank_batch: window start > end");
  HIP_TRY(hipSetDevice(g->device));
  const int64_t n = g->n;
  const int64_t N = g->ae - g->ab;
  if (N <= 0) return fail(TGA_EINVAL, "tga_pagerank_batch: no active vertices");
  bool ok = false;
  int rc = ensure_prb(g, &ok);
  if (rc) return rc;
  if (!ok) {
    // No room for the run structure: one call per window.
    for (int32_t q = 0; q < nq; ++q)
      if ((rc = tga_pagerank(g, windows[q], iters, damping, 0.0f, rank + q * n, nullptr, flags))) return rc;
    return TGA_OK;
  }
  if (!g->prb_du) {
    AllocTxn txn(g);
    if ((rc = g->alloc(g->prb_du, (size_t)n * 64)) || (rc = g->alloc(g->prb_rank, (size_t)n * 64)) ||
        (rc = g->alloc(g->prb_coef, (size_t)n * 64)) ||
        (rc = g->alloc(g->prb_partial, (size_t)std::max<int64_t>(g->prb_nseg, 1) * 64)) ||
        (rc = g->alloc(g->prb_sums, 64)) || (rc = g->alloc(g->prb_wm, (size_t)g->prb_nmulti * 64)))
      return rc;
    // shared with batched HITS, which may have allocated them first
    if ((!g->prb_acc && (rc = g->alloc(g->prb_acc, (size_t)n * 64))) ||
        (!g->prb_part && (rc = g->alloc(g->prb_part, (size_t)kPrbRedBlocks * 64))) ||
        (!g->prb_win && (rc = g->alloc(g->prb_win, 2 * 64))))
      return rc;
    txn.commit();
  }
  int64_t hw[2 * 64];
  for (int32_t q = 0; q < nq; ++q) {
    hw[2 * q] = windows[q].start;
    hw[2 * q + 1] = windows[q].end;
  }
  hipStream_t st = g->stream;
  HIP_TRY(hipMemcpyAsync(g->prb_win, hw, sizeof(int64_t) * 2 * (size_t)nq, hipMemcpyHostToDevice, st));
  float* out = rank;
  if (!(flags & TGA_OUT_DEVICE)) {
    if ((rc = g->rows_buffer(g->prb_out, g->prb_out_rows, nq))) return rc;
    out = g->prb_out;
  }
  const Window A{TGA_FULL_WINDOW_START, TGA_FULL_WINDOW_END, g->ab, g->ae};
  const float inv = 1.0f / (float)N;
  const float base = (1.0f - damping) / (float)N;
  g->kbegin();
  k_prb_du<<<g->grid(n * 64), kBlock, 0, st>>>(A, (int32_t)n, g->vdeg, g->coff, g->ct, g->prb_win, nq, g->prb_du,
                                                ~0ull);
  if (g->prb_ndu_items)
    k_prb_du_long<<<g->grid((int64_t)g->prb_ndu_items * 64), kBlock, 0, st>>>(
        g->prb_du_items, g->prb_ndu_items, g->coff, g->vdeg, g->ct, g->prb_win, nq, g->prb_du, ~0ull);
  if (g->prb_nmulti)
    k_prb_wmulti<<<g->grid((int64_t)g->prb_nmulti * 64), kBlock, 0, st>>>(
        g->prb_mlist, g->prb_nmulti, g->prb_runs, g->coff, g->ct, g->prb_win, nq, g->prb_wm, ~0ull);
  k_prb_init<<<g->grid(n * 64), kBlock, 0, st>>>(A, (int32_t)n, inv, g->prb_rank, ~0ull);
  HIP_TRY(hipGetLastError());
  // The iterations replay one captured graph (cached per batch shape).
  auto iterations = [&]() -> int {
  for (int32_t it = 0; it < iters; ++it) {
    if (it == 0)
      k_prb_coef<<<g->grid(n * 64), kBlock, 0, st>>>(A, (int32_t)n, g->prb_rank, g->prb_du, damping, g->prb_coef,
                                                     (double)base, g->prb_acc);
    else
      k_prb_step<<<g->grid(n * 64), kBlock, 0, st>>>(A, (int32_t)n, g->prb_sums, g->prb_du, damping, g->prb_coef,
                                                     (double)base, g->prb_acc);
    if (g->prb_nseg)
      k_prb_pull<<<g->grid(g->prb_nseg * 64), kBlock, 0, st>>>(
          g->prb_nseg, g->prb_seg_row, g->prb_seg_beg, g->prb_seg_len, g->prb_seg_off, g->prb_runs, g->prb_f0,
          g->prb_midx, g->prb_wm, g->prb_coef, g->prb_win, nq, (double)base, g->prb_acc, g->prb_partial);
    if (g->prb_nmrows)
      k_prb_fix<<<g->grid((int64_t)g->prb_nmrows * 64), kBlock, 0, st>>>(g->prb_mrows, g->prb_nm, g->prb_seg_off,
                                                                       g->prb_partial, (double)base, g->prb_acc);
    k_prb_colsum<<<kPrbRedBlocks, kBlock, 0, st>>>((int32_t)n, g->prb_acc, g->prb_part);
    k_prb_colsum2<<<1, 64, 0, st>>>(g->prb_part, kPrbRedBlocks, g->prb_sums);
  }
  HIP_TRY(hipGetLastError());
  return TGA_OK;
  };
  uint32_t dbits;
  std::memcpy(&dbits, &damping, sizeof dbits);
  if ((rc = g->replay(g->gr_prb, (uint64_t)nq | (uint64_t)(uint32_t)iters << 8 | (uint64_t)dbits << 32, iterations)))
    return rc;
  k_prb_out<<<g->grid(n * 64), kBlock, 0, st>>>((int32_t)n, nq, g->prb_acc, iters > 0 ? g->prb_sums : nullptr,
                                                g->prb_rank, out);
  g->kend();
  HIP_TRY(hipGetLastError());
  if (!(flags & TGA_OUT_DEVICE))
    HIP_TRY(hipMemcpyAsync(rank, out, sizeof(float) * (size_t)nq * (size_t)n, hipMemcpyDeviceToHost, st));
  return g->check_err();
}

extern "C" int tga_connected_components_batch(tga_graph* g, const tga_window* windows, int32_t nq, int32_t* label,
                                              uint32_t flags) {
  if (!g || !windows || !label || nq < 1 || nq > 64)
    return fail(TGA_EINVAL, "tga_connected_components_batch: bad arguments (1 <= nq <= 64)");
  for (int32_t q = 0; q < nq; ++q)
    if (bad_window(windows[q])) return fail(TGA_EINVAL, "tga_connected_components_batch: window start > end");
  HIP_TRY(hipSetDevice(g->device));
  const int64_t n = g->n;
  bool ok = false;
  int rc = ensure_prb(g, &ok);
  if (rc) return rc;
  if (!ok) {
    // No room for the run structure: one call per window.
    for (int32_t q = 0; q < nq; ++q)
      if ((rc = tga_connected_components(g, windows[q], label + q * n, flags))) return rc;
    return TGA_OK;
  }
  if (!g->ccb_label) {
    AllocTxn txn(g);
    if ((rc = g->alloc(g->ccb_label, (size_t)n * 64)) ||
        (rc = g->alloc(g->ccb_partial, (size_t)std::max<int64_t>(g->prb_nseg, 1) * 64)) ||
        (rc = g->alloc(g->ccb_mask, (size_t)std::max<int32_t>(g->prb_nmulti, 1))) ||
        (rc = g->alloc(g->ccb_changed, kCcbGroup + 1)))
      return rc;
    if (!g->prb_win && (rc = g->alloc(g->prb_win, 2 * 64))) return rc;
    txn.commit();
  }
  int64_t hw[2 * 64];
  for (int32_t q = 0; q < nq; ++q) {
    hw[2 * q] = windows[q].start;
    hw[2 * q + 1] = windows[q].end;
  }
  hipStream_t st = g->stream;
  HIP_TRY(hipMemcpyAsync(g->prb_win, hw, sizeof(int64_t) * 2 * (size_t)nq, hipMemcpyHostToDevice, st));
  int32_t* out = label;
  if (!(flags & TGA_OUT_DEVICE)) {
    if ((rc = g->rows_buffer(g->ccb_out, g->ccb_out_rows, nq))) return rc;
    out = g->ccb_out;
  }
  const Window A{TGA_FULL_WINDOW_START, TGA_FULL_WINDOW_END, g->ab, g->ae};
  g->kbegin();
  if (g->prb_nmulti)
    k_ccb_mask<<<g->grid((int64_t)g->prb_nmulti * 64), kBlock, 0, st>>>(
        g->prb_mlist, g->prb_nmulti, g->prb_runs, g->coff, g->ct, g->prb_win, nq, g->ccb_mask, ~0ull);
  k_ccb_init<<<g->grid(n * 64), kBlock, 0, st>>>(A, (int32_t)n, g->ccb_label, ~0ull);
  HIP_TRY(hipGetLastError());
  // Rounds run in groups of kCcbGroup, replayed from one captured graph:
  // flag r of a group records whether round r lowered a label, and the
  // kernels of round r + 1 return at once when it did not (the labels are
  // then the fixpoint), so the host reads one flag per group instead of
  // synchronising every round.
  auto rounds = [&]() -> int {
    for (int r = 0; r < kCcbGroup; ++r) {
      k_fill<int32_t><<<1, kBlock, 0, st>>>(g->ccb_changed + r + 1, 1, 0);
      const int32_t* run_if = r ? g->ccb_changed + r : nullptr;
      k_ccb_pull<<<g->grid(g->prb_nseg * 64), kBlock, 0, st>>>(
          g->prb_nseg, g->prb_seg_row, g->prb_seg_beg, g->prb_seg_len, g->prb_seg_off, g->prb_runs, g->prb_midx,
          g->ccb_mask, g->prb_win, nq, g->ccb_label, g->ccb_partial, g->ccb_changed + r + 1, run_if, nullptr);
      if (g->prb_nmrows)
        k_ccb_fix<<<g->grid((int64_t)g->prb_nmrows * 64), kBlock, 0, st>>>(
            g->prb_mrows, g->prb_nm, g->prb_seg_off, g->ccb_partial, nq, g->ccb_label, g->ccb_changed + r + 1,
            run_if, nullptr);
      HIP_TRY(hipGetLastError());
    }
    return TGA_OK;
  };
  for (int64_t round = 0; round < n && g->prb_nseg; round += kCcbGroup) {
    if ((rc = g->replay(g->gr_ccb, (uint64_t)nq, rounds))) return rc;
    int32_t changed = 0;
    HIP_TRY(hipMemcpyAsync(&changed, g->ccb_changed + kCcbGroup, sizeof changed, hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    if (!changed) break;
  }
  k_ccb_out<<<g->grid(n * 64), kBlock, 0, st>>>((int32_t)n, nq, g->ccb_label, out);
  g->kend();
  HIP_TRY(hipGetLastError());
  if (!(flags & TGA_OUT_DEVICE))
    HIP_TRY(hipMemcpyAsync(label, out, sizeof(int32_t) * (size_t)nq * (size_t)n, hipMemcpyDeviceToHost, st));
  return g->check_err();
}

extern "C" int tga_hits_batch(tga_graph* g, const tga_window* windows, int32_t nq, int32_t iters, float* hub,
                              float* auth, uint32_t flags) {
  if (!g || !windows || !hub || !auth || nq < 1 || nq > 64 || iters < 0)
    return fail(TGA_EINVAL, "tga_hits_batch: bad arguments (1 <= nq <= 64)");
  for (int32_t q = 0; q < nq; ++q)
    if (bad_window(windows[q])) return fail(TGA_EINVAL, "tga_hits_batch: window start > end");
  HIP_TRY(hipSetDevice(g->device));
  const int64_t n = g->n;
  bool ok = false;
  int rc = ensure_prb(g, &ok);
  if (rc) return rc;
  if (!ok) {
    // No room for the run structure: one call per window.
    for (int32_t q = 0; q < nq; ++q)
      if ((rc = tga_hits(g, windows[q], iters, hub + q * n, auth + q * n, flags))) return rc;
    return TGA_OK;
  }
  if (!g->hb_hub) {
    AllocTxn txn(g);
    if ((rc = g->alloc(g->hb_hub, (size_t)n * 64)) || (rc = g->alloc(g->hb_auth, (size_t)n * 64)) ||
        (rc = g->alloc(g->hb_partial,
                       (size_t)std::max<int64_t>(std::max(g->prb_nseg, g->prb_noseg), 1) * 64)) ||
        (rc = g->alloc(g->hb_norm, 64)) || (rc = g->alloc(g->hb_mask, (size_t)std::max<int32_t>(g->prb_nmulti, 1))))
      return rc;
    if (!g->prb_win && (rc = g->alloc(g->prb_win, 2 * 64))) return rc;
    if (!g->prb_acc && (rc = g->alloc(g->prb_acc, (size_t)n * 64))) return rc;
    if (!g->prb_part && (rc = g->alloc(g->prb_part, (size_t)kPrbRedBlocks * 64))) return rc;
    txn.commit();
  }
  int64_t hw[2 * 64];
  for (int32_t q = 0; q < nq; ++q) {
    hw[2 * q] = windows[q].start;
    hw[2 * q + 1] = windows[q].end;
  }
  hipStream_t st = g->stream;
  HIP_TRY(hipMemcpyAsync(g->prb_win, hw, sizeof(int64_t) * 2 * (size_t)nq, hipMemcpyHostToDevice, st));
  float *oh = hub, *oa = auth;
  if (!(flags & TGA_OUT_DEVICE)) {
    if ((rc = g->rows_buffer(g->hb_out, g->hb_out_rows, 2 * nq))) return rc;
    oh = g->hb_out;
    oa = g->hb_out + (int64_t)nq * n;
  }
  const Window A{TGA_FULL_WINDOW_START, TGA_FULL_WINDOW_END, g->ab, g->ae};
  g->kbegin();
  if (g->prb_nmulti)
    k_ccb_mask<<<g->grid((int64_t)g->prb_nmulti * 64), kBlock, 0, st>>>(
        g->prb_mlist, g->prb_nmulti, g->prb_runs, g->coff, g->ct, g->prb_win, nq, g->hb_mask, ~0ull);
  k_cols_init<<<g->grid(n * 64), kBlock, 0, st>>>(A, (int32_t)n, 1.0f, g->hb_hub, g->hb_auth);
  HIP_TRY(hipGetLastError());
  auto iterations = [&]() -> int {
  for (int32_t it = 0; it < iters; ++it) {
    for (int step = 0; step < 2; ++step) {  // 0: hub from authority over out-runs; 1: authority from hub, in-runs
      k_fill<double><<<g->grid(n * 64), kBlock, 0, st>>>(g->prb_acc, n * 64, 0.0);
      if (step == 0) {
        if (g->prb_noseg)
          k_hitsb_pull<1><<<g->grid(g->prb_noseg * 64), kBlock, 0, st>>>(
              g->prb_noseg, g->prb_oseg_row, g->prb_oseg_beg, g->prb_oseg_len, g->prb_oseg_off, g->prb_out_runs,
              g->prb_midx, g->hb_mask, g->prb_win, nq, g->hb_auth, g->prb_acc, g->hb_partial);
        if (g->prb_nomrows)
          k_prb_fix<<<g->grid((int64_t)g->prb_nomrows * 64), kBlock, 0, st>>>(
              g->prb_omrows, g->prb_onm, g->prb_oseg_off, g->hb_partial, 0.0, g->prb_acc);
      } else {
        if (g->prb_nseg)
          k_hitsb_pull<0><<<g->grid(g->prb_nseg * 64), kBlock, 0, st>>>(
              g->prb_nseg, g->prb_seg_row, g->prb_seg_beg, g->prb_seg_len, g->prb_seg_off, g->prb_runs, g->prb_midx,
              g->hb_mask, g->prb_win, nq, g->hb_hub, g->prb_acc, g->hb_partial);
        if (g->prb_nmrows)
          k_prb_fix<<<g->grid((int64_t)g->prb_nmrows * 64), kBlock, 0, st>>>(
              g->prb_mrows, g->prb_nm, g->prb_seg_off, g->hb_partial, 0.0, g->prb_acc);
      }
      k_colsq<<<kPrbRedBlocks, kBlock, 0, st>>>((int32_t)n, g->prb_acc, g->prb_part);
      k_colsq2<<<1, 64, 0, st>>>(g->prb_part, kPrbRedBlocks, g->hb_norm);
      k_colscale<<<g->grid(n * 64), kBlock, 0, st>>>((int32_t)n, g->prb_acc, g->hb_norm,
                                                      step == 0 ? g->hb_hub : g->hb_auth);
      HIP_TRY(hipGetLastError());
    }
  }
  return TGA_OK;
  };
  if ((rc = g->replay(g->gr_hb, (uint64_t)nq | (uint64_t)(uint32_t)iters << 8, iterations))) return rc;
  k_cols_out<<<g->grid(n * 64), kBlock, 0, st>>>((int32_t)n, nq, g->hb_hub, oh);
  k_cols_out<<<g->grid(n * 64), kBlock, 0, st>>>((int32_t)n, nq, g->hb_auth, oa);
  g->kend();
  HIP_TRY(hipGetLastError());
  if (!(flags & TGA_OUT_DEVICE)) {
    HIP_TRY(hipMemcpyAsync(hub, oh, sizeof(float) * (size_t)nq * (size_t)n, hipMemcpyDeviceToHost, st));
    HIP_TRY(hipMemcpyAsync(auth, oa, sizeof(float) * (size_t)nq * (size_t)n, hipMemcpyDeviceToHost, st));
  }
  return g->check_err();
}

static uint32_t total_items(const uint32_t (&hc)[kGroups + 1]) {
  uint32_t t = 0;
  for (int c = 0; c < kGroups; ++c) t += hc[c];
  return t;
}

extern "C" int tga_bfs(tga_graph* g, tga_window w, int32_t src, int32_t* level, uint32_t flags) {
  if (!g || !level || bad_window(w) || src < 0 || src >= g->n) return fail(TGA_EINVAL, "tga_bfs: bad arguments");
  HIP_TRY(hipSetDevice(g->device));
  const Window W = g->window(w);
  k_fill<int32_t><<<g->grid(g->n), kBlock, 0, g->stream>>>(g->i0, g->n, kInf);
  const int32_t zero = 0;
  HIP_TRY(hipMemcpyAsync(g->i0 + src, &zero, 4, hipMemcpyHostToDevice, g->stream));
  // Small levels travel as vertex lists (two scratch arrays of n entries;
  // level 0 is the source), large ones as level values found by a scan of
  // the active range, which also queues their records in id order so that
  // the lanes of a wave read neighbouring stream bytes.  A level is expanded
  // from its list when it has one and holds under 1/32 of the active
  // vertices; a level's kernel builds the next list when the level's work
  // items (a long record counts once per chunk, so a hub never does) are
  // under 1/4096 of them, so that the next level stays small enough for the
  // swaps' waits.  (netflow-like, 2^30 contacts: BFS 12.3 ms with scans
  // only, 70 ms with lists only, 11.1 ms this way; 1% window 0.67 / 0.19.)
  int32_t* cur = g->q0;
  int32_t* nxt = g->i2;
  HIP_TRY(hipMemcpyAsync(cur, &src, 4, hipMemcpyHostToDevice, g->stream));
  const uint64_t active = (uint64_t)(g->ae - g->ab);
  uint32_t nf = 1;
  bool listed = true;  // cur holds this level's vertices (nf of them)
  for (int32_t lvl = 0;; ++lvl) {
    uint32_t hc[kGroups + 1];
    const bool from_list = listed && (uint64_t)nf * 32 < active;
    int rc = from_list ? g->expand(cur, 0, nf, Sel{kSelAll, nullptr, 0}, false, false, hc)
                       : g->expand(nullptr, g->ab, (uint32_t)active, Sel{kSelEq, g->i0, lvl}, false, false, hc);
    if (rc) return rc;
    const uint64_t items = total_items(hc);
    if (items == 0) break;
    if (items * 4096 < active) {
      uint32_t* nc = g->next_count();
      LAUNCH_QUEUED(k_bfsl, hc, W, g->i0, lvl, nxt, nc);
      if ((rc = g->read_next_count(&nf))) return rc;
      std::swap(cur, nxt);
      listed = true;
      if (nf == 0) break;
    } else {
      LAUNCH_QUEUED(k_bfs, hc, W, g->i0, lvl);
      listed = false;
    }
  }
  int rc = g->out_copy(level, g->i0, (size_t)g->n * 4, flags);
  if (rc) return rc;
  return g->check_err();
}

extern "C" int tga_bfs_batch(tga_graph* g, const tga_window* windows, const int32_t* src, int32_t nq, int32_t* levels,
                             uint32_t flags) {
  if (!g || !windows || !src || !levels || nq < 1 || nq > kMaxQueries)
    return fail(TGA_EINVAL, "tga_bfs_batch: bad arguments (1 <= nq <= 64)");
  for (int32_t q = 0; q < nq; ++q)
    if (bad_window(windows[q]) || src[q] < 0 || src[q] >= g->n)
      return fail(TGA_EINVAL, "tga_bfs_batch: bad window or source of query " + std::to_string(q));
  HIP_TRY(hipSetDevice(g->device));
  const int64_t n = g->n;
  int rc;
  if (!g->bq_front) {
    AllocTxn txn(g);
    if ((rc = g->alloc(g->bq_front, (size_t)n)) || (rc = g->alloc(g->bq_vis, (size_t)n)) ||
        (rc = g->alloc(g->bq_next, (size_t)n)) || (rc = g->alloc(g->bq_bound, kMaxBounds)) ||
        (rc = g->alloc(g->bq_mask, kMaxBounds + 1)) || (rc = g->alloc(g->bq_src, kMaxQueries)))
      return rc;
    txn.commit();
  }
  int32_t* lv = levels;
  if (!(flags & TGA_OUT_DEVICE)) {
    if ((rc = g->rows_buffer(g->bq_levels, g->bq_levels_rows, nq))) return rc;
    lv = g->bq_levels;
  }
  // Window table: sorted boundaries (each window [s, e] contributes s and
  // e + 1) and the query mask of every elementary interval between them.
  std::vector<int64_t> b;
  for (int32_t q = 0; q < nq; ++q) {
    b.push_back(windows[q].start);
    b.push_back(windows[q].end + 1);
  }
  std::sort(b.begin(), b.end());
  b.erase(std::unique(b.begin(), b.end()), b.end());
  std::vector<uint64_t> m(b.size() + 1, 0);
  for (size_t i = 1; i <= b.size(); ++i)  // interval [b[i-1], b[i]) (b[nb] = +inf)
    for (int32_t q = 0; q < nq; ++q)
      if (windows[q].start <= b[i - 1] && b[i - 1] <= windows[q].end) m[i] |= 1ull << q;
  HIP_TRY(hipMemcpyAsync(g->bq_bound, b.data(), b.size() * 8, hipMemcpyHostToDevice, g->stream));
  HIP_TRY(hipMemcpyAsync(g->bq_mask, m.data(), m.size() * 8, hipMemcpyHostToDevice, g->stream));
  HIP_TRY(hipMemcpyAsync(g->bq_src, src, (size_t)nq * 4, hipMemcpyHostToDevice, g->stream));
  const WinTable tab{g->bq_bound, g->bq_mask, (int32_t)b.size()};
  k_fill<int32_t><<<g->grid((int64_t)nq * n), kBlock, 0, g->stream>>>(lv, (int64_t)nq * n, kInf);
  k_fill<uint64_t><<<g->grid(n), kBlock, 0, g->stream>>>(g->bq_front, n, 0);
  k_fill<uint64_t><<<g->grid(n), kBlock, 0, g->stream>>>(g->bq_vis, n, 0);
  k_fill<uint64_t><<<g->grid(n), kBlock, 0, g->stream>>>(g->bq_next, n, 0);
  k_bfsb_start<<<1, kMaxQueries, 0, g->stream>>>((int32_t)n, g->bq_src, nq, g->bq_front, g->bq_vis, lv);
  HIP_TRY(hipGetLastError());
  const Window W = g->window(tga_window{TGA_FULL_WINDOW_START, TGA_FULL_WINDOW_END});
  for (int32_t lvl = 0;; ++lvl) {
    uint32_t hc[kGroups + 1];
    Sel sel{kSelMask, nullptr, 0};
    sel.m = g->bq_front;
    if ((rc = g->expand(nullptr, g->ab, (uint32_t)(g->ae - g->ab), sel, false, false, hc))) return rc;
    if (total_items(hc) == 0) break;
    LAUNCH_QUEUED(k_bfsb, hc, W, g->bq_front, g->bq_vis, g->bq_next, tab);
    k_bfsb_finish<<<g->grid(n), kBlock, 0, g->stream>>>((int32_t)n, g->bq_front, g->bq_vis, g->bq_next, lv, lvl,
                                                        nullptr, nullptr);
    HIP_TRY(hipGetLastError());
  }
  if (!(flags & TGA_OUT_DEVICE))
    HIP_TRY(hipMemcpyAsync(levels, lv, (size_t)nq * (size_t)n * 4, hipMemcpyDeviceToHost, g->stream));
  return g->check_err();
}

// ------------------------------------------------------------ BFS server --
//
// Continuous batching of BFS queries over the batched kernels: up to 64 lanes;
// queries are admitted into free lanes between steps, every step advances all
// running lanes by one level in one decode pass, and a lane whose frontier
// empties is finished (its level row is final) until its result is taken.
struct tga_bfs_server : LaneServer {
  uint64_t *front = nullptr, *vis = nullptr, *next = nullptr, *mask = nullptr;
  int64_t* bound = nullptr;
  int32_t *src = nullptr, *start = nullptr, *levels = nullptr;
  unsigned long long* alive = nullptr;
  int32_t nb = 0;
  int32_t step = 0;
  int32_t h_src[kMaxQueries] = {};
  int32_t h_start[kMaxQueries] = {};
  using LaneServer::LaneServer;
};

extern "C" int tga_bfs_server_open(tga_graph* g, tga_bfs_server** out) {
  if (!g || !out) return fail(TGA_EINVAL, "tga_bfs_server_open: NULL argument");
  *out = nullptr;
  HIP_TRY(hipSetDevice(g->device));
  auto* s = new tga_bfs_server(g);
  const size_t n = (size_t)g->n;
  int rc;
  if ((rc = s->alloc(s->front, n)) || (rc = s->alloc(s->vis, n)) || (rc = s->alloc(s->next, n)) ||
      (rc = s->alloc(s->mask, kMaxBounds + 1)) || (rc = s->alloc(s->bound, kMaxBounds)) ||
      (rc = s->alloc(s->src, kMaxQueries)) || (rc = s->alloc(s->start, kMaxQueries)) ||
      (rc = s->alloc(s->levels, (size_t)kMaxQueries * n)) || (rc = s->alloc(s->alive, 1))) {
    delete s;
    return rc;
  }
  k_fill<uint64_t><<<g->grid(g->n), kBlock, 0, g->stream>>>(s->front, g->n, 0);
  k_fill<uint64_t><<<g->grid(g->n), kBlock, 0, g->stream>>>(s->vis, g->n, 0);
  k_fill<uint64_t><<<g->grid(g->n), kBlock, 0, g->stream>>>(s->next, g->n, 0);
  if (hipGetLastError() != hipSuccess || hipStreamSynchronize(g->stream) != hipSuccess) {
    delete s;
    return fail(TGA_EDEVICE, "tga_bfs_server_open: initialisation failed");
  }
  *out = s;
  return TGA_OK;
}

extern "C" void tga_bfs_server_close(tga_bfs_server* s) { delete s; }

extern "C" int tga_bfs_server_admit(tga_bfs_server* s, int32_t src, tga_window w, int32_t* lane) {
  if (!s || !s->g || !lane || bad_window(w) || src < 0 || src >= s->g->n)
    return fail(TGA_EINVAL, "tga_bfs_server_admit: bad arguments (or the handle is closed)");
  const int q = s->free_lane();
  if (q < 0) return fail(TGA_EBUSY, "tga_bfs_server_admit: all 64 lanes are busy");
  s->h_src[q] = src;
  s->h_win[q] = w;
  s->pending |= 1ull << q;
  *lane = q;
  return TGA_OK;
}

extern "C" int tga_bfs_server_step(tga_bfs_server* s, uint64_t* finished) {
  if (!s || !s->g || !finished) return fail(TGA_EINVAL, "tga_bfs_server_step: NULL argument (or the handle is closed)");
  tga_graph* g = s->g;
  HIP_TRY(hipSetDevice(g->device));
  const int64_t n = g->n;
  *finished = 0;
  if (s->pending) {
    for (uint64_t b = s->pending; b; b &= b - 1) s->h_start[__builtin_ctzll(b)] = s->step;
    s->running |= s->pending;
    // window table of the running lanes (see tga_bfs_batch)
    std::vector<int64_t> bd;
    for (uint64_t b = s->running; b; b &= b - 1) {
      const int q = __builtin_ctzll(b);
      bd.push_back(s->h_win[q].start);
      bd.push_back(s->h_win[q].end + 1);
    }
    std::sort(bd.begin(), bd.end());
    bd.erase(std::unique(bd.begin(), bd.end()), bd.end());
    std::vector<uint64_t> m(bd.size() + 1, 0);
    for (size_t i = 1; i <= bd.size(); ++i)
      for (uint64_t b = s->running; b; b &= b - 1) {
        const int q = __builtin_ctzll(b);
        if (s->h_win[q].start <= bd[i - 1] && bd[i - 1] <= s->h_win[q].end) m[i] |= 1ull << q;
      }
    s->nb = (int32_t)bd.size();
    HIP_TRY(hipMemcpyAsync(s->bound, bd.data(), bd.size() * 8, hipMemcpyHostToDevice, g->stream));
    HIP_TRY(hipMemcpyAsync(s->mask, m.data(), m.size() * 8, hipMemcpyHostToDevice, g->stream));
    HIP_TRY(hipMemcpyAsync(s->src, s->h_src, sizeof s->h_src, hipMemcpyHostToDevice, g->stream));
    HIP_TRY(hipMemcpyAsync(s->start, s->h_start, sizeof s->h_start, hipMemcpyHostToDevice, g->stream));
    k_bfsb_admit<<<g->grid(n), kBlock, 0, g->stream>>>((int32_t)n, s->pending, s->src, s->vis, s->front, s->levels);
    k_bfsb_admit_src<<<1, 64, 0, g->stream>>>((int32_t)n, s->pending, s->src, s->vis, s->front, s->levels);
    HIP_TRY(hipGetLastError());
    s->pending = 0;
  }
  if (!s->running) return TGA_OK;
  uint32_t hc[kGroups + 1];
  Sel sel{kSelMask, nullptr, 0};
  sel.m = s->front;
  int rc = g->expand(nullptr, g->ab, (uint32_t)(g->ae - g->ab), sel, false, false, hc);
  if (rc) return rc;
  uint64_t alive = 0;
  if (total_items(hc) > 0) {
    const WinTable tab{s->bound, s->mask, s->nb};
    const Window W = g->window(tga_window{TGA_FULL_WINDOW_START, TGA_FULL_WINDOW_END});
    LAUNCH_QUEUED(k_bfsb, hc, W, s->front, s->vis, s->next, tab);
    HIP_TRY(hipMemsetAsync(s->alive, 0, 8, g->stream));
    k_bfsb_finish<<<g->grid(n), kBlock, 0, g->stream>>>((int32_t)n, s->front, s->vis, s->next, s->levels, s->step,
                                                        s->start, s->alive);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipMemcpyAsync(&alive, s->alive, 8, hipMemcpyDeviceToHost, g->stream));
  }
  if ((rc = g->check_err())) return rc;
  ++s->step;
  *finished = s->running & ~alive;
  s->done |= *finished;
  s->running &= alive;
  return TGA_OK;
}

extern "C" int tga_bfs_server_result(tga_bfs_server* s, int32_t lane, int32_t* level, uint32_t flags) {
  if (!s || !s->g || !level || lane < 0 || lane >= kMaxQueries || !(s->done >> lane & 1))
    return fail(TGA_EINVAL, "tga_bfs_server_result: lane has no finished query (or the handle is closed)");
  tga_graph* g = s->g;
  HIP_TRY(hipSetDevice(g->device));
  int rc = g->out_copy(level, s->levels + (int64_t)lane * g->n, (size_t)g->n * 4, flags);
  if (rc) return rc;
  HIP_TRY(hipStreamSynchronize(g->stream));
  s->done &= ~(1ull << lane);
  return TGA_OK;
}

// ------------------------------------------------ CC and PageRank servers ---
//
// Continuous admission into the batched column state of
// tga_connected_components_batch / tga_pagerank_batch: all 64 lanes are
// processed every step (free lanes hold empty windows), and a query admitted
// between steps starts in its own lane at the next one while the others go
// on.  Lanes are independent columns, so every query's result is the one its
// batched (and plain) call returns.

// Window table of a server's lanes: queries of running or pending lanes,
// empty windows [1, 0] elsewhere.
// The table goes through a pinned buffer, rewritten only once its last
// upload has been done (the step does not wait for the device).
static int upload_windows(LaneServer* s, int64_t* win) {
  if (!s->h_wtab) {
    HIP_TRY(hipHostMalloc((void**)&s->h_wtab, sizeof(int64_t) * 2 * 64, hipHostMallocDefault));
    HIP_TRY(hipEventCreateWithFlags(&s->ev_wtab, hipEventDisableTiming));
  } else {
    HIP_TRY(hipEventSynchronize(s->ev_wtab));
  }
  const uint64_t live = s->running | s->pending | s->done;
  for (int q = 0; q < 64; ++q) {
    const bool on = live >> q & 1;
    s->h_wtab[2 * q] = on ? s->h_win[q].start : 1;
    s->h_wtab[2 * q + 1] = on ? s->h_win[q].end : 0;
  }
  HIP_TRY(hipMemcpyAsync(win, s->h_wtab, sizeof(int64_t) * 2 * 64, hipMemcpyHostToDevice, s->g->stream));
  HIP_TRY(hipEventRecord(s->ev_wtab, s->g->stream));
  return TGA_OK;
}

template <class T>
static int lane_result(LaneServer* s, int32_t lane, const T* cols, T*& stage, T* out, uint32_t flags) {
  tga_graph* g = s->g;
  HIP_TRY(hipSetDevice(g->device));
  T* dst = out;
  if (!(flags & TGA_OUT_DEVICE)) {
    if (!stage) {
      const int rc = s->alloc(stage, (size_t)g->n);
      if (rc) return rc;
    }
    dst = stage;
  }
  k_col_out<T><<<g->grid(g->n), kBlock, 0, g->stream>>>((int32_t)g->n, lane, cols, dst);
  HIP_TRY(hipGetLastError());
  if (!(flags & TGA_OUT_DEVICE))
    HIP_TRY(hipMemcpyAsync(out, dst, sizeof(T) * (size_t)g->n, hipMemcpyDeviceToHost, g->stream));
  HIP_TRY(hipStreamSynchronize(g->stream));
  s->done &= ~(1ull << lane);
  return TGA_OK;
}

struct tga_cc_server : LaneServer {
  int32_t *label = nullptr, *partial = nullptr, *changed = nullptr, *stage = nullptr;
  uint64_t* mask = nullptr;
  unsigned long long* lane_changed = nullptr;  // per round of a group: lanes that lowered a label
  int64_t* win = nullptr;
  tga_graph::Graph gr;
  using LaneServer::LaneServer;
  ~tga_cc_server() override {
    if (g) (void)hipStreamSynchronize(g->stream);
    if (gr.exec) (void)hipGraphExecDestroy(gr.exec);
  }
};

extern "C" int tga_cc_server_open(tga_graph* g, tga_cc_server** out) {
  if (!g || !out) return fail(TGA_EINVAL, "tga_cc_server_open: NULL argument");
  *out = nullptr;
  HIP_TRY(hipSetDevice(g->device));
  bool ok = false;
  int rc = ensure_prb(g, &ok);
  if (rc) return rc;
  if (!ok) return fail(TGA_ENOMEM, "tga_cc_server_open: no room for the run structure");
  auto* s = new tga_cc_server(g);
  const int64_t n = g->n;
  if ((rc = s->alloc(s->label, (size_t)n * 64)) ||
      (rc = s->alloc(s->partial, (size_t)std::max<int64_t>(g->prb_nseg, 1) * 64)) ||
      (rc = s->alloc(s->mask, (size_t)std::max<int32_t>(g->prb_nmulti, 1))) ||
      (rc = s->alloc(s->changed, kCcbGroup + 1)) || (rc = s->alloc(s->lane_changed, kCcbGroup)) ||
      (rc = s->alloc(s->win, 2 * 64))) {
    delete s;
    return rc;
  }
  const Window A{TGA_FULL_WINDOW_START, TGA_FULL_WINDOW_END, g->ab, g->ae};
  k_fill<uint64_t><<<g->grid(g->prb_nmulti), kBlock, 0, g->stream>>>(s->mask, std::max<int32_t>(g->prb_nmulti, 1), 0);
  k_ccb_init<<<g->grid(n * 64), kBlock, 0, g->stream>>>(A, (int32_t)n, s->label, ~0ull);
  if (hipGetLastError() != hipSuccess || (rc = upload_windows(s, s->win))) {
    delete s;
    return rc ? rc : fail(TGA_EDEVICE, "tga_cc_server_open: initialisation failed");
  }
  *out = s;
  return TGA_OK;
}

extern "C" void tga_cc_server_close(tga_cc_server* s) { delete s; }

extern "C" int tga_cc_server_admit(tga_cc_server* s, tga_window w, int32_t* lane) {
  if (!s || !s->g || !lane || bad_window(w))
    return fail(TGA_EINVAL, "tga_cc_server_admit: bad arguments (or the handle is closed)");
  const int q = s->free_lane();
  if (q < 0) return fail(TGA_EBUSY, "tga_cc_server_admit: all 64 lanes are busy");
  s->h_win[q] = w;
  s->pending |= 1ull << q;
  *lane = q;
  return TGA_OK;
}

// One step: admits the queued queries (their link masks and initial labels),
// then runs kCcbGroup rounds from the server's captured graph.  A lane whose
// labels no round of the group lowered is at its fixpoint: it finishes.
extern "C" int tga_cc_server_step(tga_cc_server* s, uint64_t* finished) {
  if (!s || !s->g || !finished) return fail(TGA_EINVAL, "tga_cc_server_step: NULL argument (or the handle is closed)");
  tga_graph* g = s->g;
  HIP_TRY(hipSetDevice(g->device));
  hipStream_t st = g->stream;
  const int64_t n = g->n;
  *finished = 0;
  int rc;
  if (s->pending) {
    if ((rc = upload_windows(s, s->win))) return rc;
    const Window A{TGA_FULL_WINDOW_START, TGA_FULL_WINDOW_END, g->ab, g->ae};
    g->kbegin();
    if (g->prb_nmulti)
      k_ccb_mask<<<g->grid((int64_t)g->prb_nmulti * 64), kBlock, 0, st>>>(
          g->prb_mlist, g->prb_nmulti, g->prb_runs, g->coff, g->ct, s->win, 64, s->mask, s->pending);
    k_ccb_init<<<g->grid(n * 64), kBlock, 0, st>>>(A, (int32_t)n, s->label, s->pending);
    g->kend();
    HIP_TRY(hipGetLastError());
    s->running |= s->pending;
    s->pending = 0;
  }
  if (!s->running) return TGA_OK;
  uint64_t conv = ~0ull;  // no segments: every window's labels are the initial ones
  if (g->prb_nseg) {
    auto rounds = [&]() -> int {
      for (int r = 0; r < kCcbGroup; ++r) {
        k_fill<int32_t><<<1, kBlock, 0, st>>>(s->changed + r + 1, 1, 0);
        k_fill<unsigned long long><<<1, kBlock, 0, st>>>(s->lane_changed + r, 1, 0ull);
        const int32_t* run_if = r ? s->changed + r : nullptr;
        k_ccb_pull<<<g->grid(g->prb_nseg * 64), kBlock, 0, st>>>(
            g->prb_nseg, g->prb_seg_row, g->prb_seg_beg, g->prb_seg_len, g->prb_seg_off, g->prb_runs, g->prb_midx,
            s->mask, s->win, 64, s->label, s->partial, s->changed + r + 1, run_if, s->lane_changed + r);
        if (g->prb_nmrows)
          k_ccb_fix<<<g->grid((int64_t)g->prb_nmrows * 64), kBlock, 0, st>>>(
              g->prb_mrows, g->prb_nm, g->prb_seg_off, s->partial, 64, s->label, s->changed + r + 1, run_if,
              s->lane_changed + r);
        HIP_TRY(hipGetLastError());
      }
      return TGA_OK;
    };
    g->kbegin();
    if ((rc = g->replay(s->gr, 1, rounds))) return rc;
    g->kend();
    unsigned long long lc[kCcbGroup];
    HIP_TRY(hipMemcpyAsync(lc, s->lane_changed, sizeof lc, hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    conv = 0;
    for (int r = 0; r < kCcbGroup; ++r) conv |= ~(uint64_t)lc[r];
  }
  *finished = s->running & conv;
  s->done |= *finished;
  s->running &= ~*finished;
  return TGA_OK;
}

extern "C" int tga_cc_server_result(tga_cc_server* s, int32_t lane, int32_t* label, uint32_t flags) {
  if (!s || !s->g || !label || lane < 0 || lane >= 64 || !(s->done >> lane & 1))
    return fail(TGA_EINVAL, "tga_cc_server_result: lane has no finished query (or the handle is closed)");
  return lane_result<int32_t>(s, lane, s->label, s->stage, label, flags);
}

struct tga_pr_server : LaneServer {
  float damping = 0.85f;
  int32_t* du = nullptr;
  float *rank = nullptr, *coef = nullptr, *wm = nullptr, *stage = nullptr;
  double *acc = nullptr, *partial = nullptr, *sums = nullptr, *part = nullptr;
  int64_t* win = nullptr;
  int32_t h_iters[64] = {};
  int32_t h_it[64] = {};
  int32_t h_pend_iters[64] = {};
  using LaneServer::LaneServer;
};

extern "C" int tga_pr_server_open(tga_graph* g, float damping, tga_pr_server** out) {
  if (!g || !out) return fail(TGA_EINVAL, "tga_pr_server_open: NULL argument");
  *out = nullptr;
  if (g->ae - g->ab <= 0) return fail(TGA_EINVAL, "tga_pr_server_open: no active vertices");
  HIP_TRY(hipSetDevice(g->device));
  bool ok = false;
  int rc = ensure_prb(g, &ok);
  if (rc) return rc;
  if (!ok) return fail(TGA_ENOMEM, "tga_pr_server_open: no room for the run structure");
  auto* s = new tga_pr_server(g);
  s->damping = damping;
  const int64_t n = g->n;
  if ((rc = s->alloc(s->du, (size_t)n * 64)) || (rc = s->alloc(s->rank, (size_t)n * 64)) ||
      (rc = s->alloc(s->coef, (size_t)n * 64)) || (rc = s->alloc(s->acc, (size_t)n * 64)) ||
      (rc = s->alloc(s->wm, (size_t)std::max<int32_t>(g->prb_nmulti, 1) * 64)) ||
      (rc = s->alloc(s->partial, (size_t)std::max<int64_t>(g->prb_nseg, 1) * 64)) ||
      (rc = s->alloc(s->sums, 64)) || (rc = s->alloc(s->part, (size_t)kPrbRedBlocks * 64)) ||
      (rc = s->alloc(s->win, 2 * 64))) {
    delete s;
    return rc;
  }
  hipStream_t st = g->stream;
  k_fill<int32_t><<<g->grid(n * 64), kBlock, 0, st>>>(s->du, n * 64, 0);
  k_fill<float><<<g->grid(n * 64), kBlock, 0, st>>>(s->rank, n * 64, 0.0f);
  k_fill<double><<<g->grid(n * 64), kBlock, 0, st>>>(s->acc, n * 64, 0.0);
  k_fill<float><<<g->grid((int64_t)std::max<int32_t>(g->prb_nmulti, 1) * 64), kBlock, 0, st>>>(
      s->wm, (int64_t)std::max<int32_t>(g->prb_nmulti, 1) * 64, 0.0f);
  k_fill<double><<<1, kBlock, 0, st>>>(s->sums, 64, 0.0);
  if (hipGetLastError() != hipSuccess || (rc = upload_windows(s, s->win))) {
    delete s;
    return rc ? rc : fail(TGA_EDEVICE, "tga_pr_server_open: initialisation failed");
  }
  *out = s;
  return TGA_OK;
}

extern "C" void tga_pr_server_close(tga_pr_server* s) { delete s; }

extern "C" int tga_pr_server_admit(tga_pr_server* s, tga_window w, int32_t iters, int32_t* lane) {
  if (!s || !s->g || !lane || bad_window(w) || iters < 0)
    return fail(TGA_EINVAL, "tga_pr_server_admit: bad arguments (or the handle is closed)");
  const int q = s->free_lane();
  if (q < 0) return fail(TGA_EBUSY, "tga_pr_server_admit: all 64 lanes are busy");
  s->h_win[q] = w;
  s->h_pend_iters[q] = iters;
  s->pending |= 1ull << q;
  *lane = q;
  return TGA_OK;
}

// One step: admits the queued queries (window degrees, multi-contact
// weights, initial ranks), then runs one iteration for every running lane;
// lanes that reach their iteration count keep their ranks and finish (the
// counts are known on the host: the device is waited for only so that a
// step's time is its own, as for the other servers).
extern "C" int tga_pr_server_step(tga_pr_server* s, uint64_t* finished) {
  if (!s || !s->g || !finished) return fail(TGA_EINVAL, "tga_pr_server_step: NULL argument (or the handle is closed)");
  tga_graph* g = s->g;
  HIP_TRY(hipSetDevice(g->device));
  hipStream_t st = g->stream;
  const int64_t n = g->n;
  const int64_t N = g->ae - g->ab;
  const Window A{TGA_FULL_WINDOW_START, TGA_FULL_WINDOW_END, g->ab, g->ae};
  const float damping = s->damping;
  const float inv = 1.0f / (float)N;
  const float base = (1.0f - damping) / (float)N;
  *finished = 0;
  int rc;
  g->kbegin();
  if (s->pending) {
    if ((rc = upload_windows(s, s->win))) return rc;
    const uint64_t P = s->pending;
    k_prb_du<<<g->grid(n * 64), kBlock, 0, st>>>(A, (int32_t)n, g->vdeg, g->coff, g->ct, s->win, 64, s->du, P);
    if (g->prb_ndu_items)
      k_prb_du_long<<<g->grid((int64_t)g->prb_ndu_items * 64), kBlock, 0, st>>>(
          g->prb_du_items, g->prb_ndu_items, g->coff, g->vdeg, g->ct, s->win, 64, s->du, P);
    if (g->prb_nmulti)
      k_prb_wmulti<<<g->grid((int64_t)g->prb_nmulti * 64), kBlock, 0, st>>>(
          g->prb_mlist, g->prb_nmulti, g->prb_runs, g->coff, g->ct, s->win, 64, s->wm, P);
    k_prb_init<<<g->grid(n * 64), kBlock, 0, st>>>(A, (int32_t)n, inv, s->rank, P);
    HIP_TRY(hipGetLastError());
    for (uint64_t b = P; b; b &= b - 1) {
      const int q = __builtin_ctzll(b);
      s->h_iters[q] = s->h_pend_iters[q];
      s->h_it[q] = 0;
      if (s->h_iters[q] == 0) *finished |= 1ull << q;  // the initial ranks are the result
    }
    s->running |= P & ~*finished;
    s->pending = 0;
  }
  if (s->running) {
    uint64_t fresh = 0, last = 0;
    for (uint64_t b = s->running; b; b &= b - 1) {
      const int q = __builtin_ctzll(b);
      if (s->h_it[q] == 0) fresh |= 1ull << q;
      if (++s->h_it[q] == s->h_iters[q]) last |= 1ull << q;
    }
    k_prb_coef_step<<<g->grid(n * 64), kBlock, 0, st>>>(A, (int32_t)n, s->rank, s->sums, s->du, damping, s->coef,
                                                        (double)base, s->acc, fresh);
    if (g->prb_nseg)
      k_prb_pull<<<g->grid(g->prb_nseg * 64), kBlock, 0, st>>>(
          g->prb_nseg, g->prb_seg_row, g->prb_seg_beg, g->prb_seg_len, g->prb_seg_off, g->prb_runs, g->prb_f0,
          g->prb_midx, s->wm, s->coef, s->win, 64, (double)base, s->acc, s->partial);
    if (g->prb_nmrows)
      k_prb_fix<<<g->grid((int64_t)g->prb_nmrows * 64), kBlock, 0, st>>>(g->prb_mrows, g->prb_nm, g->prb_seg_off,
                                                                       s->partial, (double)base, s->acc);
    k_prb_colsum<<<kPrbRedBlocks, kBlock, 0, st>>>((int32_t)n, s->acc, s->part);
    k_prb_colsum2<<<1, 64, 0, st>>>(s->part, kPrbRedBlocks, s->sums);
    if (last) k_prb_fin<<<g->grid(n), kBlock, 0, st>>>((int32_t)n, last, s->acc, s->sums, s->rank);
    HIP_TRY(hipGetLastError());
    *finished |= last;
    s->running &= ~last;
  }
  g->kend();
  HIP_TRY(hipStreamSynchronize(st));
  s->done |= *finished;
  return TGA_OK;
}

extern "C" int tga_pr_server_result(tga_pr_server* s, int32_t lane, float* rank, uint32_t flags) {
  if (!s || !s->g || !rank || lane < 0 || lane >= 64 || !(s->done >> lane & 1))
    return fail(TGA_EINVAL, "tga_pr_server_result: lane has no finished query (or the handle is closed)");
  return lane_result<float>(s, lane, s->rank, s->stage, rank, flags);
}

extern "C" int tga_earliest_arrival(tga_graph* g, tga_window w, int32_t src, int32_t* arrival, uint32_t flags) {
  if (!g || !arrival || bad_window(w) || src < 0 || src >= g->n)
    return fail(TGA_EINVAL, "tga_earliest_arrival: bad arguments");
  HIP_TRY(hipSetDevice(g->device));
  const Window W = g->window(w);
  k_fill<int32_t><<<g->grid(g->n), kBlock, 0, g->stream>>>(g->i0, g->n, kInf);
  k_fill<int32_t><<<g->grid(g->n), kBlock, 0, g->stream>>>(g->i1, g->n, kInf);
  const int32_t a0 = (int32_t)std::max<int64_t>(0, std::min<int64_t>(w.start, INT32_MAX));
  const int32_t start_stamp = -1;  // round r expands the vertices stamped r - 1
  HIP_TRY(hipMemcpyAsync(g->i0 + src, &a0, 4, hipMemcpyHostToDevice, g->stream));
  HIP_TRY(hipMemcpyAsync(g->i1 + src, &start_stamp, 4, hipMemcpyHostToDevice, g->stream));
  for (int32_t round = 0;; ++round) {
    uint32_t hc[kGroups + 1];
    int rc = g->expand(nullptr, g->ab, (uint32_t)(g->ae - g->ab), Sel{kSelEq, g->i1, round - 1}, false, false, hc);
    if (rc) return rc;
    if (total_items(hc) == 0) break;
    LAUNCH_QUEUED(k_ea, hc, W, g->i0, g->i1, round, g->runkey);
    if (g->nck > 0) {
      k_ea_fix<<<g->grid(g->nck), kBlock, 0, g->stream>>>(g->ck, (int32_t)g->nck, g->runkey, W, g->i0, g->i1, round);
      HIP_TRY(hipGetLastError());
    }
  }
  int rc = g->out_copy(arrival, g->i0, (size_t)g->n * 4, flags);
  if (rc) return rc;
  return g->check_err();
}

extern "C" int tga_latest_departure(tga_graph* g, tga_window w, int32_t target, int32_t* departure,
                                    int32_t* rounds, uint32_t flags) {
  if (!g || !departure || bad_window(w) || target < 0 || target >= g->n)
    return fail(TGA_EINVAL, "tga_latest_departure: bad arguments");
  HIP_TRY(hipSetDevice(g->device));
  const Window W = g->window(w);
  const int32_t d0 = (int32_t)std::max<int64_t>(INT32_MIN, std::min<int64_t>(w.end, INT32_MAX));
  k_fill<int32_t><<<g->grid(g->n), kBlock, 0, g->stream>>>(g->i0, g->n, kLdUnreached);
  HIP_TRY(hipMemcpyAsync(g->i0 + target, &d0, 4, hipMemcpyHostToDevice, g->stream));
  uint32_t* changed = (uint32_t*)g->u64s;
  int32_t round = 0;
  for (uint32_t more = 1; more; ++round) {
    HIP_TRY(hipMemsetAsync(changed, 0, sizeof(uint32_t), g->stream));
    LAUNCH_STATIC(k_ld, W, g->i0, changed);
    HIP_TRY(hipMemcpyAsync(&more, changed, sizeof more, hipMemcpyDeviceToHost, g->stream));
    HIP_TRY(hipStreamSynchronize(g->stream));
  }
  if (rounds) *rounds = round;
  int rc = g->out_copy(departure, g->i0, (size_t)g->n * 4, flags);
  if (rc) return rc;
  return g->check_err();
}

extern "C" int tga_connected_components(tga_graph* g, tga_window w, int32_t* label, uint32_t flags) {
  if (!g || !label || bad_window(w)) return fail(TGA_EINVAL, "tga_connected_components: bad arguments");
  HIP_TRY(hipSetDevice(g->device));
  const Window W = g->window(w);
  k_init_labels<<<g->grid(g->n), kBlock, 0, g->stream>>>(W, (int32_t)g->n, g->i0, g->i1);
  // Rounds whose frontier holds under 1/1024 of the active vertices list the
  // vertices they stamp (k_ccl, into two scratch arrays of n entries), and
  // the next round expands that list instead of scanning the range while it
  // stays under that size.  Larger rounds count only (k_cc): the list
  // kernel's per-contact ballot slows a large round, and a long list in
  // discovery order queues records whose stream bytes lie far apart
  // (netflow-like 2^30 contacts, lists below 1/32: 37.5 -> 54.9 ms).
  const uint64_t active = (uint64_t)(g->ae - g->ab);
  uint32_t nf = (uint32_t)active;  // every active vertex starts stamped -1
  int32_t* cur = g->q0;
  int32_t* nxt = g->i2;
  bool listed = false;
  for (int32_t round = 0; nf > 0; ++round) {
    uint32_t* nc = g->next_count();
    if (g->dense(nf)) {
      LAUNCH_STATIC(k_cc, W, g->i0, g->i1, round, g->i1, nc);
      listed = false;
    } else {
      const bool small = (uint64_t)nf * 1024 < active;
      uint32_t hc[kGroups + 1];
      int rc = listed && small
                   ? g->expand(cur, 0, nf, Sel{kSelEq, g->i1, round - 1}, false, false, hc)
                   : g->expand(nullptr, g->ab, (uint32_t)active, Sel{kSelEq, g->i1, round - 1}, false, false, hc);
      if (rc) return rc;
      if (total_items(hc) == 0) break;
      if (small) {
        LAUNCH_QUEUED(k_ccl, hc, W, g->i0, g->i1, round, nxt, nc, (uint32_t)g->n);
        std::swap(cur, nxt);
      } else {
        LAUNCH_QUEUED(k_cc, hc, W, g->i0, g->i1, round, nullptr, nc);
      }
      listed = small;
    }
    int rc = g->read_next_count(&nf);
    if (rc) return rc;
    if (listed && nf > (uint32_t)g->n) listed = false;  // the list overflowed
  }
  int rc = g->out_copy(label, g->i0, (size_t)g->n * 4, flags);
  if (rc) return rc;
  return g->check_err();
}

// java.util.Random(seed).nextInt(bound) (Java SE API, class Random), the GC
// priority scheme of BaselineRunner.java:392-396.
static void java_random_ints(int64_t seed, int32_t bound, int64_t count, int32_t* out) {
  const uint64_t mult = 0x5DEECE66DULL, addend = 0xBULL, mask = (1ULL << 48) - 1;
  uint64_t s = ((uint64_t)seed ^ mult) & mask;
  auto next31 = [&]() -> int32_t {
    s = (s * mult + addend) & mask;
    return (int32_t)(s >> 17);
  };
  for (int64_t i = 0; i < count; ++i) {
    int32_t r = next31();
    const int32_t m = bound - 1;
    if ((bound & m) == 0) {
      r = (int32_t)(((int64_t)bound * (int64_t)r) >> 31);
    } else {
      for (int32_t u = r; u - (r = u % bound) + m < 0; u = next31()) {
      }
    }
    out[i] = r;
  }
}

// Whether list state of `need` device bytes may be built: at most
// TGA_LIST_MAX_BYTES from the environment, else a quarter of the free device
// memory (always, when `held` says it is allocated already).
static int list_budget(uint64_t need, bool held, bool* ok) {
  const char* env = std::getenv("TGA_LIST_MAX_BYTES");
  uint64_t budget = UINT64_MAX;
  if (env && *env) {
    budget = std::strtoull(env, nullptr, 10);
  } else if (!held) {
    size_t fr = 0, tot = 0;
    HIP_TRY(hipMemGetInfo(&fr, &tot));
    budget = fr / 4;
  }
  *ok = need <= budget;
  return TGA_OK;
}


// Exclusive scan of c[0, m) into off[0, m]; the tile sums go to tsum
// (scan_tiles(m) entries), or to g->ltsum when m <= n + nck (item_scan_alloc).
static int scan_counts(tga_graph* g, const uint32_t* c, int64_t m, uint64_t* off, uint64_t* tsum) {
  const int64_t tiles = scan_tiles(m);
  if (!tsum) tsum = g->ltsum;
  hipStream_t st = g->stream;
  k_scan_tiles<<<(int)tiles, kBlock, 0, st>>>(c, m, tsum);
  k_scan_top<<<1, kBlock, 0, st>>>(tsum, tiles);
  k_scan_apply<<<(int)tiles, kBlock, 0, st>>>(c, m, tsum, off);
  HIP_TRY(hipGetLastError());
  return TGA_OK;
}

// Exclusive scan of the per-item counts g->lcnt into g->loff.
static int item_scan(tga_graph* g) { return scan_counts(g, g->lcnt, g->n + g->nck, g->loff); }

static int item_scan_alloc(tga_graph* g) {
  if (g->lcnt) return TGA_OK;
  const int64_t m = g->n + g->nck;
  int rc;
  AllocTxn txn(g);
  if ((rc = g->alloc(g->lcnt, (size_t)m)) || (rc = g->alloc(g->loff, (size_t)m + 1)) ||
      (rc = g->alloc(g->ltsum, (size_t)scan_tiles(m))))
    return rc;
  txn.commit();
  return TGA_OK;
}

static uint64_t item_scan_bytes(const tga_graph* g) {
  const int64_t m = g->n + g->nck;
  return g->lcnt ? 0 : (uint64_t)m * 12 + 8 + (uint64_t)scan_tiles(m) * 8;
}

// Graph coloring over higher-priority lists (k_gc_list, k_gc_round) into
// g->i0, which holds kInf everywhere on entry.  Returns 1, having done
// nothing, when the lists do not fit list_budget.
static int gc_lists(tga_graph* g, const Window& W, const int32_t* prio, int32_t rounds) {
  const int64_t n = g->n;
  const uint64_t entries = (uint64_t)std::max<int64_t>(g->info.contacts, 0);
  const uint64_t need = entries * 4 + (uint64_t)n * (12 + 2 * 16) + 8 + item_scan_bytes(g);
  bool ok = false;
  int rc = list_budget(need, g->gc_hp != nullptr, &ok);
  if (rc) return rc;
  if (!ok) return 1;
  if ((rc = item_scan_alloc(g))) return rc;
  if (!g->gc_hp) {
    AllocTxn txn(g);
    if ((rc = g->alloc(g->gc_hp, (size_t)entries)) || (rc = g->alloc(g->gc_lbeg, (size_t)n)) ||
        (rc = g->alloc(g->gc_llen, (size_t)n)) || (rc = g->alloc(g->gc_ctr, 2)))
      return rc;
    for (auto& l : g->gc_list)
      if ((rc = g->alloc(l, (size_t)n))) return rc;
    txn.commit();
  }
  hipStream_t st = g->stream;
  k_fill<uint32_t><<<g->grid(n + g->nck), kBlock, 0, st>>>(g->lcnt, n + g->nck, 0u);
  LAUNCH_STATIC(k_gc_list, W, prio, g->lcnt, (const uint64_t*)nullptr, (int32_t*)nullptr);
  if ((rc = item_scan(g))) return rc;
  LAUNCH_STATIC(k_gc_list, W, prio, (uint32_t*)nullptr, g->loff, g->gc_hp);
  k_gc_ranges<<<g->grid(g->ae - g->ab), kBlock, 0, st>>>(g->ab, g->ae, (int32_t)n, g->vinfo, g->ck, g->loff,
                                                        g->gc_hp, g->gc_lbeg, g->gc_llen, g->gc_list[0], g->gc_ctr);
  HIP_TRY(hipGetLastError());
  // Every 16 rounds the colored candidates are dropped and the count read
  // back, to stop early and to size the grid.
  int a = 0;
  uint32_t blocks = (uint32_t)std::min(g->grid(g->ae - g->ab), 2048);
  for (int32_t r = 0; r < rounds; ++r) {
    k_gc_round<<<blocks, kBlock, 0, st>>>(g->gc_list[a], g->gc_ctr + a, g->gc_hp, g->gc_lbeg, g->gc_llen, g->i0, r);
    HIP_TRY(hipGetLastError());
    if ((r & 15) == 15 && r + 1 < rounds) {
      k_fill<uint32_t><<<1, kBlock, 0, st>>>(g->gc_ctr + (a ^ 1), 1, 0u);
      k_gc_compact<<<blocks, kBlock, 0, st>>>(g->gc_list[a], g->gc_ctr + a, g->gc_list[a ^ 1], g->gc_ctr + (a ^ 1));
      HIP_TRY(hipGetLastError());
      a ^= 1;
      uint32_t h = 0;
      if ((rc = g->read_counts(g->gc_ctr + a, &h, 1))) return rc;
      if (h == 0) break;
      blocks = std::min<uint32_t>(2048, (h + kBlock - 1) / kBlock);
    }
  }
  return TGA_OK;
}

extern "C" int tga_graph_coloring(tga_graph* g, tga_window w, int32_t rounds, const int32_t* priority, int32_t* color,
                                  uint32_t flags) {
  if (!g || !color || bad_window(w) || rounds < 0) return fail(TGA_EINVAL, "tga_graph_coloring: bad arguments");
  HIP_TRY(hipSetDevice(g->device));
  const Window W = g->window(w);
  const int32_t* prio = nullptr;
  if (priority) {
    HIP_TRY(hipMemcpyAsync(g->i1, priority, (size_t)g->n * 4,
                           (flags & TGA_OUT_DEVICE) ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, g->stream));
    prio = g->i1;
  } else {
    if (!g->prio) {
      std::vector<int32_t> h((size_t)g->n);
      java_random_ints(123, 256, g->n, h.data());
      AllocTxn txn(g);
      int rc = g->upload(g->prio, h);
      if (rc) return rc;
      txn.commit();
    }
    prio = g->prio;
  }
  k_fill<int32_t><<<g->grid(g->n), kBlock, 0, g->stream>>>(g->i0, g->n, kInf);
  if (rounds > 0) {
    const int rc = gc_lists(g, W, prio, rounds);
    if (rc == TGA_OK) {
      const int rc2 = g->out_copy(color, g->i0, (size_t)g->n * 4, flags);
      return rc2 ? rc2 : g->check_err();
    }
    if (rc != 1) return rc;
  }
  // Decoding rounds (the lists do not fit).
  k_fill<int32_t><<<g->grid(g->n), kBlock, 0, g->stream>>>(g->i2, g->n, 0);
  int32_t* blocker = g->q0;  // cached blocker per vertex (-1: none)
  k_fill<int32_t><<<g->grid(g->n), kBlock, 0, g->stream>>>(blocker, g->n, -1);
  for (int32_t round = 0; round < rounds; ++round) {
    // Worklist: the active vertices still uncolored (vertices without
    // contacts included: they are colored in the first round) whose cached
    // blocker, if any, got colored.  Empty only when every vertex is colored:
    // the highest-priority uncolored vertex never has an uncolored blocker.
    uint32_t hc[kGroups + 1];
    Sel sel{kSelGc, g->i0, 0};
    sel.b = blocker;
    int rc = g->expand(nullptr, g->ab, (uint32_t)(g->ae - g->ab), sel, true, true, hc);
    if (rc) return rc;
    if (total_items(hc) == 0) break;
    LAUNCH_QUEUED(k_gc, hc, W, prio, g->i0, g->i2, blocker, round);
    if (hc[kGroups] > 0) {
      k_gc_finish<<<g->grid(hc[kGroups]), kBlock, 0, g->stream>>>(g->xq_big, hc[kGroups], W, g->i0, g->i2, round);
      HIP_TRY(hipGetLastError());
    }
  }
  int rc = g->out_copy(color, g->i0, (size_t)g->n * 4, flags);
  if (rc) return rc;
  return g->check_err();
}

// Batched graph coloring (see k_gcb_level): up to 64 windows, one sweep over
// priority levels.  Levels are the distinct priority values, highest first
// (at most kGcbLevels; with more distinct values, runs of consecutive values
// share a level and the passes inside it settle their order too).
constexpr int kGcbLevels = 512;
constexpr int kGcbPasses = 4;  // passes per level in the replayed sweep (later ones over inner segments)

// Levels of a priority array: gcb_segs holds, level by level, all
// out-segments of the level's rows, then its inner segments (k_gcb_inner);
// gcb_level_off[2l] .. [2l + 1] .. [2l + 2] bound them.
static int gcb_levels(tga_graph* g, const std::vector<int32_t>& prio, const int32_t* dprio) {
  const int64_t nos = g->prb_noseg;
  const int64_t n = g->n;
  if (g->h_oseg_row.size() != (size_t)nos) {
    g->h_oseg_row.resize((size_t)nos);
    if (nos)
      HIP_TRY(hipMemcpy(g->h_oseg_row.data(), g->prb_oseg_row, (size_t)nos * 4, hipMemcpyDeviceToHost));
  }
  // level of every priority value: distinct values, descending, grouped so
  // that there are at most kGcbLevels levels
  std::vector<int32_t> vals(prio);
  std::sort(vals.begin(), vals.end(), std::greater<int32_t>());
  vals.erase(std::unique(vals.begin(), vals.end()), vals.end());
  const size_t per = std::max<size_t>(1, (vals.size() + kGcbLevels - 1) / kGcbLevels);
  auto level_of = [&](int32_t p) {
    return (int32_t)((size_t)(std::lower_bound(vals.begin(), vals.end(), p, std::greater<int32_t>()) - vals.begin()) /
                     per);
  };
  const int64_t nl = vals.empty() ? 0 : level_of(vals.back()) + 1;
  std::vector<int32_t> vlevel((size_t)n);
  for (int64_t v = 0; v < n; ++v) vlevel[(size_t)v] = level_of(prio[(size_t)v]);
  // segment lengths give each segment's pieces (list entries s << 3 | piece)
  std::vector<int32_t> slen((size_t)nos);
  if (nos) HIP_TRY(hipMemcpy(slen.data(), g->prb_oseg_len, (size_t)nos * 4, hipMemcpyDeviceToHost));
  auto pieces = [&](int64_t s) { return (slen[(size_t)s] + kGcbPiece - 1) / kGcbPiece; };
  int64_t total = 0;
  for (int64_t s = 0; s < nos; ++s) total += pieces(s);
  if (!g->gcb_segs) {  // every piece at most twice (all, inner); the inner flags use nos of them first
    AllocTxn txn(g);
    int rc;
    if ((rc = g->alloc(g->gcb_segs, (size_t)std::max<int64_t>(2 * std::max(total, nos), 1))) ||
        (rc = g->alloc(g->gcb_flags, (size_t)kGcbLevels * (kGcbPasses + 1))))
      return rc;
    txn.commit();
  }
  // inner flags into the second half of gcb_segs, levels into i2
  HIP_TRY(hipMemcpyAsync(g->i2, vlevel.data(), (size_t)n * 4, hipMemcpyHostToDevice, g->stream));
  std::vector<int32_t> inner((size_t)nos, 0);
  if (nos) {
    k_gcb_inner<<<g->grid(nos * 64), kBlock, 0, g->stream>>>(nos, g->prb_oseg_row, g->prb_oseg_beg, g->prb_oseg_len,
                                                           g->prb_out_runs, dprio, g->i2, g->gcb_segs + std::max(total, nos));
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipMemcpyAsync(inner.data(), g->gcb_segs + std::max(total, nos), (size_t)nos * 4, hipMemcpyDeviceToHost,
                           g->stream));
  }
  HIP_TRY(hipStreamSynchronize(g->stream));
  std::vector<int64_t> cnt_all((size_t)nl, 0), cnt_in((size_t)nl, 0);
  for (int64_t s = 0; s < nos; ++s) {
    const int32_t l = vlevel[(size_t)g->h_oseg_row[(size_t)s]];
    cnt_all[(size_t)l] += pieces(s);
    cnt_in[(size_t)l] += inner[(size_t)s] != 0 ? pieces(s) : 0;
  }
  std::vector<int64_t> off(2 * (size_t)nl + 1, 0);
  for (int64_t l = 0; l < nl; ++l) {
    off[2 * (size_t)l + 1] = off[2 * (size_t)l] + cnt_all[(size_t)l];
    off[2 * (size_t)l + 2] = off[2 * (size_t)l + 1] + cnt_in[(size_t)l];
  }
  std::vector<int32_t> segs((size_t)off.back());
  std::vector<int64_t> at_all((size_t)nl), at_in((size_t)nl);
  for (int64_t l = 0; l < nl; ++l) {
    at_all[(size_t)l] = off[2 * (size_t)l];
    at_in[(size_t)l] = off[2 * (size_t)l + 1];
  }
  for (int64_t s = 0; s < nos; ++s) {
    const int32_t l = vlevel[(size_t)g->h_oseg_row[(size_t)s]];
    for (int32_t p = 0; p < pieces(s); ++p) {
      const int32_t e = (int32_t)(s << kGcbPieceBits) | p;
      segs[(size_t)at_all[(size_t)l]++] = e;
      if (inner[(size_t)s]) segs[(size_t)at_in[(size_t)l]++] = e;
    }
  }
  if (!segs.empty())
    HIP_TRY(hipMemcpyAsync(g->gcb_segs, segs.data(), segs.size() * 4, hipMemcpyHostToDevice, g->stream));
  HIP_TRY(hipStreamSynchronize(g->stream));  // before the host vectors go
  g->gcb_level_off = std::move(off);
  return TGA_OK;
}

extern "C" int tga_graph_coloring_batch(tga_graph* g, const tga_window* windows, int32_t nq, int32_t rounds,
                                        const int32_t* priority, int32_t* color, uint32_t flags) {
  if (!g || !windows || !color || nq < 1 || nq > 64 || rounds < 0)
    return fail(TGA_EINVAL, "tga_graph_coloring_batch: bad arguments (1 <= nq <= 64)");
  for (int32_t q = 0; q < nq; ++q)
    if (bad_window(windows[q])) return fail(TGA_EINVAL, "tga_graph_coloring_batch: window start > end");
  HIP_TRY(hipSetDevice(g->device));
  const int64_t n = g->n;
  bool ok = false;
  int rc = ensure_prb(g, &ok);
  if (rc) return rc;
  if (!ok) {
    // No room for the run structure: one call per window.
    for (int32_t q = 0; q < nq; ++q)
      if ((rc = tga_graph_coloring(g, windows[q], rounds, priority, color + q * n, flags))) return rc;
    return TGA_OK;
  }
  if (!g->ccb_label) {  // the [vertex][64] int32 columns and window masks of batched CC
    AllocTxn txn(g);
    if ((rc = g->alloc(g->ccb_label, (size_t)n * 64)) ||
        (rc = g->alloc(g->ccb_partial, (size_t)std::max<int64_t>(g->prb_nseg, 1) * 64)) ||
        (rc = g->alloc(g->ccb_mask, (size_t)std::max<int32_t>(g->prb_nmulti, 1))) ||
        (rc = g->alloc(g->ccb_changed, kCcbGroup + 1)))
      return rc;
    if (!g->prb_win && (rc = g->alloc(g->prb_win, 2 * 64))) return rc;
    txn.commit();
  }
  hipStream_t st = g->stream;
  // priorities: the caller's, or java.util.Random(123).nextInt(256) per vertex
  std::vector<int32_t> hp((size_t)n);
  const int32_t* dprio = nullptr;
  if (priority) {
    if (flags & TGA_OUT_DEVICE) {
      HIP_TRY(hipMemcpyAsync(hp.data(), priority, (size_t)n * 4, hipMemcpyDeviceToHost, st));
      HIP_TRY(hipStreamSynchronize(st));
    } else {
      std::memcpy(hp.data(), priority, (size_t)n * 4);
    }
    HIP_TRY(hipMemcpyAsync(g->i1, hp.data(), (size_t)n * 4, hipMemcpyHostToDevice, st));
    dprio = g->i1;
    g->gcb_default_levels = false;
    if ((rc = gcb_levels(g, hp, dprio))) return rc;
  } else {
    if (!g->prio) {
      java_random_ints(123, 256, n, hp.data());
      AllocTxn txn(g);
      if ((rc = g->upload(g->prio, hp))) return rc;
      txn.commit();
    }
    dprio = g->prio;
    if (!g->gcb_default_levels) {
      HIP_TRY(hipMemcpy(hp.data(), g->prio, (size_t)n * 4, hipMemcpyDeviceToHost));
      if ((rc = gcb_levels(g, hp, dprio))) return rc;
      g->gcb_default_levels = true;
    }
  }
  int64_t hw[2 * 64];
  for (int32_t q = 0; q < nq; ++q) {
    hw[2 * q] = windows[q].start;
    hw[2 * q + 1] = windows[q].end;
  }
  HIP_TRY(hipMemcpyAsync(g->prb_win, hw, sizeof(int64_t) * 2 * (size_t)nq, hipMemcpyHostToDevice, st));
  int32_t* out = color;
  if (!(flags & TGA_OUT_DEVICE)) {
    if ((rc = g->rows_buffer(g->ccb_out, g->ccb_out_rows, nq))) return rc;
    out = g->ccb_out;
  }
  const Window A{TGA_FULL_WINDOW_START, TGA_FULL_WINDOW_END, g->ab, g->ae};
  const std::vector<int64_t>& off = g->gcb_level_off;
  const int64_t nl = ((int64_t)off.size() - 1) / 2;
  g->kbegin();
  if (g->prb_nmulti)
    k_ccb_mask<<<g->grid((int64_t)g->prb_nmulti * 64), kBlock, 0, st>>>(
        g->prb_mlist, g->prb_nmulti, g->prb_runs, g->coff, g->ct, g->prb_win, nq, g->ccb_mask, ~0ull);
  k_fill<int32_t><<<g->grid(n * 64), kBlock, 0, st>>>(g->ccb_label, n * 64, 0);
  HIP_TRY(hipGetLastError());
  int32_t* const fl = g->gcb_flags;
  // Pass k of level l: k == 1 over all its segments, k > 1 over its inner
  // segments, raising flag k; with `chain`, pass k > 1 runs only if pass
  // k - 1 raised a color.
  auto pass = [&](int64_t l, int k, bool chain) -> int {
    const int64_t a = off[2 * (size_t)l], b = off[2 * (size_t)l + 1], c = off[2 * (size_t)l + 2];
    const int64_t s0 = k == 1 ? a : b, ns = k == 1 ? b - a : c - b;
    if (ns == 0) return TGA_OK;
    int32_t* f = fl + (size_t)l * (kGcbPasses + 1);
    k_gcb_level<<<g->grid(ns * 64), kBlock, 0, st>>>(g->gcb_segs + s0, (int32_t)ns, g->prb_oseg_row, g->prb_oseg_beg,
                                                      g->prb_oseg_len, g->prb_out_runs, g->prb_midx, g->ccb_mask,
                                                      g->prb_win, nq, dprio, g->ccb_label, f + k,
                                                      chain && k > 1 ? f + k - 1 : nullptr);
    HIP_TRY(hipGetLastError());
    return TGA_OK;
  };
  auto has_segs = [&](int64_t l) { return off[2 * (size_t)l + 1] > off[2 * (size_t)l]; };
  auto sweep = [&]() -> int {
    // every flag starts at 0, so levels without segments read as settled
    k_fill<int32_t><<<g->grid((int64_t)kGcbLevels * (kGcbPasses + 1)), kBlock, 0, st>>>(
        fl, (int64_t)kGcbLevels * (kGcbPasses + 1), 0);
    for (int64_t l = 0; l < nl; ++l)
      for (int k = 1; k <= kGcbPasses && has_segs(l); ++k)
        if ((rc = pass(l, k, true))) return rc;
    return TGA_OK;
  };
  if (priority) {
    if ((rc = sweep())) return rc;
  } else if ((rc = g->replay(g->gr_gcb, (uint64_t)nq, sweep))) {
    return rc;
  }
  // A level still raising colors in its last pass (an outranking chain of
  // more than kGcbPasses - 1 vertices inside one level) is finished here,
  // one pass at a time, and the levels after it are swept again.
  std::vector<int32_t> hf((size_t)kGcbLevels * (kGcbPasses + 1));
  HIP_TRY(hipMemcpyAsync(hf.data(), fl, hf.size() * 4, hipMemcpyDeviceToHost, st));
  HIP_TRY(hipStreamSynchronize(st));
  int64_t first = -1;
  for (int64_t l = 0; l < nl && first < 0; ++l)
    if (has_segs(l) && hf[(size_t)l * (kGcbPasses + 1) + kGcbPasses]) first = l;
  if (first >= 0) {
    for (int64_t l = first; l < nl; ++l) {
      if (!has_segs(l)) continue;
      // A later level saw too small colors of this one: a pass over all its
      // segments, then inner passes until one raises nothing.  Colors only
      // rise and are bounded by the longest outranking path, so passes end;
      // the bound turns a broken invariant into an error.
      int32_t* f = fl + (size_t)l * (kGcbPasses + 1);
      if (l > first) {
        k_fill<int32_t><<<1, 64, 0, st>>>(f + 1, 1, 0);
        if ((rc = pass(l, 1, false))) return rc;
      }
      for (int64_t it = 0;; ++it) {
        if (it > n) return fail(TGA_EFORMAT, "tga_graph_coloring_batch: a level did not settle");
        k_fill<int32_t><<<1, 64, 0, st>>>(f + 2, 1, 0);
        if ((rc = pass(l, 2, false))) return rc;
        int32_t f1 = 0;
        HIP_TRY(hipMemcpyAsync(&f1, f + 2, 4, hipMemcpyDeviceToHost, st));
        HIP_TRY(hipStreamSynchronize(st));
        if (!f1) break;
      }
    }
  }
  k_gcb_out<<<g->grid(n * nq), kBlock, 0, st>>>(A, (int32_t)n, nq, rounds, g->ccb_label, out);
  g->kend();
  HIP_TRY(hipGetLastError());
  if (!(flags & TGA_OUT_DEVICE))
    HIP_TRY(hipMemcpyAsync(color, out, sizeof(int32_t) * (size_t)nq * (size_t)n, hipMemcpyDeviceToHost, st));
  return g->check_err();
}

static int reduce(tga_graph* g, int mode, const double* x, const float* old, const double* scale, double* out,
                  bool sqrt_out) {
  const int64_t n = g->n;
  if (mode == 0) k_reduce1<0><<<kRedBlocks, kBlock, 0, g->stream>>>(x, old, scale, n, g->partial);
  if (mode == 1) k_reduce1<1><<<kRedBlocks, kBlock, 0, g->stream>>>(x, old, scale, n, g->partial);
  if (mode == 2) k_reduce1<2><<<kRedBlocks, kBlock, 0, g->stream>>>(x, old, scale, n, g->partial);
  k_reduce2<<<1, kBlock, 0, g->stream>>>(g->partial, kRedBlocks, out, sqrt_out ? 1 : 0);
  HIP_TRY(hipGetLastError());
  return TGA_OK;
}

// Decodes the neighbor pairs of window W (k_pairs) and sorts them by r.
// Returns 1, having done nothing, when they do not fit list_budget.  *extra
// receives the position of the split-run slots, after the items' pairs, and
// *count the number of pairs and slots.
static int build_pairs(tga_graph* g, const Window& W, uint64_t* extra, uint64_t* count_out) {
  const int64_t n = g->n, nck = g->nck, m = n + nck;
  const uint64_t cap = (uint64_t)std::max<int64_t>(g->info.contacts, 0) + (uint64_t)nck;
  const int end_bit = 32 - __builtin_clz((uint32_t)n | 1u);  // keys <= n
  size_t tmp_bytes = 0;
  HIP_TRY(sort_pairs_u32_u64(nullptr, tmp_bytes, nullptr, nullptr, nullptr, nullptr, (size_t)cap, end_bit,
                             g->stream));
  for (int mode = 0; mode < 3; ++mode) {
    size_t b = 0;
    HIP_TRY(reduce_rows(mode, nullptr, b, nullptr, nullptr, nullptr, nullptr, (int32_t)n, 0, (size_t)cap, nullptr,
                        nullptr, nullptr, g->stream));
    tmp_bytes = std::max(tmp_bytes, b);
  }
  const uint64_t need = cap * 24 + tmp_bytes + (uint64_t)(n + 2) * 12 + 4 + item_scan_bytes(g);
  bool ok = false;
  int rc = list_budget(need, g->pk != nullptr, &ok);
  if (rc) return rc;
  if (!ok) return 1;
  if ((rc = item_scan_alloc(g))) return rc;
  if (!g->pk) {
    AllocTxn txn(g);
    uint8_t* tmp = nullptr;
    if ((rc = g->alloc(g->pk, (size_t)cap)) || (rc = g->alloc(g->pv, (size_t)cap)) ||
        (rc = g->alloc(g->sk, (size_t)cap)) || (rc = g->alloc(g->sv, (size_t)cap)) ||
        (rc = g->alloc(g->ukeys, (size_t)n + 2)) || (rc = g->alloc(g->usums, (size_t)n + 2)) ||
        (rc = g->alloc(g->nruns, 1)) || (rc = g->alloc(tmp, tmp_bytes)))
      return rc;
    g->sort_tmp = tmp;
    g->sort_tmp_bytes = tmp_bytes;
    txn.commit();
  }
  hipStream_t st = g->stream;
  k_fill<uint32_t><<<g->grid(m), kBlock, 0, st>>>(g->lcnt, m, 0u);
  LAUNCH_STATIC(k_pairs, W, g->lcnt, (const uint64_t*)nullptr, (uint32_t*)nullptr, (uint64_t*)nullptr, g->runkey);
  if ((rc = item_scan(g))) return rc;
  uint64_t T = 0;
  HIP_TRY(hipMemcpyAsync(&T, g->loff + m, sizeof T, hipMemcpyDeviceToHost, st));
  HIP_TRY(hipStreamSynchronize(st));
  if (T + (uint64_t)nck > cap) return fail(TGA_EFORMAT, "neighbor pairs exceed the contact count");
  LAUNCH_STATIC(k_pairs, W, (uint32_t*)nullptr, g->loff, g->pk, g->pv, g->runkey);
  if (nck > 0) {
    k_pair_fix<<<g->grid(nck), kBlock, 0, st>>>(g->ck, (int32_t)nck, g->runkey, W, (int32_t)n, g->pk + T, g->pv + T);
    HIP_TRY(hipGetLastError());
  }
  const uint64_t count = T + (uint64_t)nck;
  size_t tb = g->sort_tmp_bytes;
  HIP_TRY(sort_pairs_u32_u64(g->sort_tmp, tb, g->pk, g->sk, g->pv, g->sv, (size_t)count, end_bit, st));
  *extra = T;
  *count_out = count;
  return TGA_OK;
}

// acc[k] = base + (row sum of k) for the rows of pairs [begin, begin + count)
// (add: acc[k] += row sum); see reduce_rows for the modes.
static int pull_rows(tga_graph* g, int mode, const float* vec, uint64_t begin, uint64_t count, double base, bool add) {
  if (count == 0) return TGA_OK;
  size_t tb = g->sort_tmp_bytes;
  HIP_TRY(reduce_rows(mode, g->sort_tmp, tb, g->sk, mode == 2 ? g->pv : g->sv, g->pk, vec, (int32_t)g->n,
                      (size_t)begin, (size_t)count, g->ukeys, g->usums, g->nruns, g->stream));
  k_scatter_rows<<<g->grid(g->n + 2), kBlock, 0, g->stream>>>(g->ukeys, g->usums, g->nruns, (int32_t)g->n, base,
                                                              add ? 1 : 0, g->acc);
  HIP_TRY(hipGetLastError());
  return TGA_OK;
}

extern "C" int tga_pagerank(tga_graph* g, tga_window w, int32_t iters, float damping, float tol, float* rank,
                            int32_t* iters_run, uint32_t flags) {
  if (!g || !rank || bad_window(w) || iters < 0) return fail(TGA_EINVAL, "tga_pagerank: bad arguments");
  HIP_TRY(hipSetDevice(g->device));
  const Window W = g->window(w);
  const int64_t N = g->ae - g->ab;
  if (N <= 0) return fail(TGA_EINVAL, "tga_pagerank: no active vertices");
  const bool full = w.start <= TGA_FULL_WINDOW_START && w.end >= TGA_FULL_WINDOW_END;
  int32_t* du = g->i0;
  if (full) {
    k_copy_deg<<<g->grid(g->n), kBlock, 0, g->stream>>>(W, (int32_t)g->n, g->vdeg, du);
  } else {
    k_fill<int32_t><<<g->grid(g->n), kBlock, 0, g->stream>>>(du, g->n, 0);
    LAUNCH_STATIC(k_wdeg, W, du);
  }
  const float inv = 1.0f / (float)N;
  const float base = (1.0f - damping) / (float)N;
  k_fill_active<float><<<g->grid(g->n), kBlock, 0, g->stream>>>(W, (int32_t)g->n, g->f0, inv, 0.0f);
  uint64_t extra = 0, npairs = 0;
  const int lp = iters > 0 ? build_pairs(g, W, &extra, &npairs) : 1;
  if (lp < 0) return lp;
  int32_t it = 0;
  if (lp == 0 && iters > 0) {
    // the first iteration's coefficients and accumulators; later ones come
    // from k_pr_next
    k_pr_coef<<<g->grid(g->n), kBlock, 0, g->stream>>>(W, (int32_t)g->n, g->f0, du, damping, g->f1);
    k_fill_active<double><<<g->grid(g->n), kBlock, 0, g->stream>>>(W, (int32_t)g->n, g->acc, (double)base, 0.0);
  }
  for (; it < iters; ++it) {
    if (lp == 0) {
      // row sums over the sorted pairs (f1 holds the coefficients)
      int rc = pull_rows(g, 0, g->f1, 0, npairs, (double)base, false);
      if (rc) return rc;
    } else {
      k_fill_active<double><<<g->grid(g->n), kBlock, 0, g->stream>>>(W, (int32_t)g->n, g->acc, (double)base, 0.0);
      LAUNCH_STATIC(k_pr_push, W, g->f0, du, damping, g->acc, g->runkey);
      if (g->nck > 0)
        k_pr_fix<<<g->grid(g->nck), kBlock, 0, g->stream>>>(g->ck, (int32_t)g->nck, g->runkey, W, g->f0, du,
                                                            damping, g->acc);
    }
    int rc = reduce(g, 0, g->acc, nullptr, nullptr, g->scal, false);
    if (!rc && tol > 0) rc = reduce(g, 2, g->acc, g->f0, g->scal, g->scal + 1, false);
    if (rc) return rc;
    if (lp == 0)
      k_pr_next<<<g->grid(g->n), kBlock, 0, g->stream>>>(W, (int32_t)g->n, g->acc, g->scal, g->f0, du, damping, g->f1,
                                                         (double)base);
    else
      k_scale_store<<<g->grid(g->n), kBlock, 0, g->stream>>>(g->acc, g->scal, g->n, g->f0);
    HIP_TRY(hipGetLastError());
    if (tol > 0) {
      double diff = 0;
      HIP_TRY(hipMemcpyAsync(&diff, g->scal + 1, sizeof diff, hipMemcpyDeviceToHost, g->stream));
      HIP_TRY(hipStreamSynchronize(g->stream));
      if (diff < (double)tol) {
        ++it;
        break;
      }
    }
  }
  if (iters_run) *iters_run = it;
  int rc = g->out_copy(rank, g->f0, (size_t)g->n * 4, flags);
  if (rc) return rc;
  return g->check_err();
}

extern "C" int tga_hits(tga_graph* g, tga_window w, int32_t iters, float* hub, float* auth, uint32_t flags) {
  if (!g || !hub || !auth || bad_window(w) || iters < 0) return fail(TGA_EINVAL, "tga_hits: bad arguments");
  HIP_TRY(hipSetDevice(g->device));
  const Window W = g->window(w);
  k_fill_active<float><<<g->grid(g->n), kBlock, 0, g->stream>>>(W, (int32_t)g->n, g->f0, 1.0f, 0.0f);  // hub
  k_fill_active<float><<<g->grid(g->n), kBlock, 0, g->stream>>>(W, (int32_t)g->n, g->f1, 1.0f, 0.0f);  // auth
  uint64_t extra = 0, npairs = 0;
  const int lp = iters > 0 ? build_pairs(g, W, &extra, &npairs) : 1;
  if (lp < 0) return lp;
  const int32_t n = (int32_t)g->n;
  for (int32_t it = 0; it < iters; ++it) {
    k_fill<double><<<g->grid(n), kBlock, 0, g->stream>>>(g->acc, n, 0.0);
    if (lp == 0) {
      // hub: out-rows in fill order, then the split-run slots
      int rc = pull_rows(g, 2, g->f1, 0, extra, 0.0, false);
      if (!rc) rc = pull_rows(g, 2, g->f1, extra, npairs - extra, 0.0, true);
      if (rc) return rc;
    } else {
      LAUNCH_STATIC(k_hits_pull, W, g->f1, g->acc, g->runkey);
      if (g->nck > 0)
        k_hits_fix<<<g->grid(g->nck), kBlock, 0, g->stream>>>(g->ck, (int32_t)g->nck, g->runkey, W, g->f1, g->acc, 0);
    }
    int rc = reduce(g, 1, g->acc, nullptr, nullptr, g->scal, true);
    if (rc) return rc;
    k_scale_store<<<g->grid(n), kBlock, 0, g->stream>>>(g->acc, g->scal, n, g->f0);
    k_fill<double><<<g->grid(n), kBlock, 0, g->stream>>>(g->acc, n, 0.0);
    if (lp == 0) {
      rc = pull_rows(g, 1, g->f0, 0, npairs, 0.0, false);  // authority: sorted in-rows
      if (rc) return rc;
    } else {
      LAUNCH_STATIC(k_hits_push, W, g->f0, g->acc, g->runkey);
      if (g->nck > 0)
        k_hits_fix<<<g->grid(g->nck), kBlock, 0, g->stream>>>(g->ck, (int32_t)g->nck, g->runkey, W, g->f0, g->acc, 1);
    }
    rc = reduce(g, 1, g->acc, nullptr, nullptr, g->scal, true);
    if (rc) return rc;
    k_scale_store<<<g->grid(g->n), kBlock, 0, g->stream>>>(g->acc, g->scal, g->n, g->f1);
    HIP_TRY(hipGetLastError());
  }
  int rc = g->out_copy(hub, g->f0, (size_t)g->n * 4, flags);
  if (!rc) rc = g->out_copy(auth, g->f1, (size_t)g->n * 4, flags);
  if (rc) return rc;
  return g->check_err();
}

extern "C" int tga_debug_fail_alloc(tga_graph* g, int64_t nth) {
  if (!g || nth < 0) return fail(TGA_EINVAL, "tga_debug_fail_alloc: bad arguments");
  g->fail_alloc_at = nth;
  return TGA_OK;
}

extern "C" int tga_profile(tga_graph* g, int enable) {
  if (!g) return fail(TGA_EINVAL, "tga_profile: NULL graph");
  g->harvest();
  g->prof = enable != 0;
  return TGA_OK;
}

extern "C" int tga_profile_read(tga_graph* g, double* kernel_ms, int64_t* launches, int reset) {
  if (!g) return fail(TGA_EINVAL, "tga_profile_read: NULL graph");
  g->harvest();
  if (kernel_ms) *kernel_ms = g->prof_ms;
  if (launches) *launches = g->prof_launches;
  if (reset) {
    g->prof_ms = 0;
    g->prof_launches = 0;
  }
  return TGA_OK;
}
