// Device decode core for gfx950: bit peeks, instantaneous codes, Elias-Fano
// offsets, the successor-list state machine and decode checkpoints.
//
// HBM layout (built once by the loader, files on disk stay unchanged):
//  * .graph and .timestamps are stored as 64-bit words with their bytes
//    swapped, so one aligned load yields 64 stream bits MSB-first (the
//    stream's own bit order, dsiutils / bfs.cu:166-175).  A 64-bit window at
//    any bit position is a funnel shift of two adjacent words.
//  * Elias-Fano lower bits are stored the same way (they are MSB-first,
//    bfs.cu:543-552); upper bits keep the file's little-endian words
//    (LSB-first, bfs.cu:305-307, :446-448).
//  * The select directory samples the position of every 64th one of the
//    upper bits; a 64-one block spanning more than kSpanBits bits stores all
//    64 positions instead (bit 63 of the sample flags it).  It plays the role
//    of SimpleSelect (bfs.cu:281-541) but is built on the host at load time.
//
// Codes are decoded arithmetically from a 64-bit window with a count of
// leading zeros instead of the reference's 16-bit lookup tables
// (bfs.cu:202-256, :590-639); every codeword this format produces is at most
// 51 bits long (zeta_2 of an int2nat'ed int32 difference).  The sequential
// sub-streams (timestamps, record headers, and the residuals and multiples of
// every level) are read through register bit buffers that load one 64-bit
// word per 64 bits consumed, ahead of use; the rare ones (intervals, copy
// blocks) use two-word peeks.
//
// Everything here is __host__ __device__: the kernels use it on the GPU, and
// a host build of the same code walks whole graphs on the CPU under address
// and undefined-behaviour sanitizers (tests/host_decode).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define TGA_HD __host__ __device__ __forceinline__

namespace tga {

TGA_HD uint32_t clz64(uint64_t x) { return x ? (uint32_t)__builtin_clzll(x) : 64u; }
TGA_HD uint32_t popc64(uint64_t x) { return (uint32_t)__builtin_popcountll(x); }
TGA_HD int32_t min32(int32_t a, int32_t b) { return a < b ? a : b; }
TGA_HD uint32_t minu32(uint32_t a, uint32_t b) { return a < b ? a : b; }

constexpr int32_t kInf = 0x7fffffff;
constexpr int kMaxDepth = 3;          // maxrefcount (BVMultiGraph.java:163)
constexpr int kWindow = 7;            // windowsize (BVMultiGraph.java:169)
constexpr uint32_t kSpanBits = 1024;  // select blocks wider than this are stored exactly
// Contacts between consecutive decode checkpoints of a long record, by the
// record's reference-chain depth: deeper records cost more per contact (their
// copies walk the referenced lists), so their chunks are shorter.  The stride
// of records without references (`flat`) is chosen per graph at open: short
// chunks balance a small graph's pass, long ones cut the per-chunk reads of a
// large graph's.
constexpr int32_t kFlatStrideMin = 128, kFlatStrideMax = 1024;
TGA_HD constexpr int32_t chunk_stride(int depth, int32_t flat) { return depth == 0 ? flat : (depth == 1 ? 16 : 8); }
// About 2^20 flat items per pass: contacts / 2^20, a power of two in
// [kFlatStrideMin, kFlatStrideMax] (128 up to 2^27 contacts, 1024 from 2^30).
inline int32_t flat_stride_for(int64_t contacts) {
  int32_t s = kFlatStrideMin;
  while (s < kFlatStrideMax && ((int64_t)s << 21) <= contacts) s <<= 1;
  return s;
}

// Neighbor-access checksum term of one contact (x, r, t): the checksum of a
// pass is the sum of the terms of its contacts modulo 2^64, so it does not
// depend on the order contacts are visited in.  The term is one multiply of
// the time and a key of the pair (r * K1 once per run, x * K2 once per
// record: the decode, not the checksum, is what the pass measures):
//   h = (t ^ r * 0x9E3779B1 ^ x * 0x85EBCA77) * 0x7FEB352D   (all uint32)
TGA_HD uint32_t contact_key_x(int32_t x) { return (uint32_t)x * 0x85EBCA77u; }
TGA_HD uint32_t contact_key_r(uint32_t kx, int32_t r) { return ((uint32_t)r * 0x9E3779B1u) ^ kx; }
TGA_HD uint64_t contact_mix(uint32_t kr, int32_t t) {
  return ((uint32_t)t ^ kr) * 0x7FEB352Du;
}
TGA_HD uint64_t contact_term(uint32_t kx, int32_t r, int32_t t) { return contact_mix(contact_key_r(kx, r), t); }
TGA_HD uint64_t contact_hash(int32_t x, int32_t r, int32_t t) { return contact_term(contact_key_x(x), r, t); }

// Error bits reported through the per-handle device error word.
enum : int32_t {
  kErrOverrun = 1,      // a cursor ran past the end of a stream
  kErrNeighbor = 2,     // decoded neighbor outside [0, n)
  kErrDepth = 4,        // reference chain deeper than the decoder or reference outside the window
  kErrExhausted = 8,    // record yields fewer successors than its outdegree
  kErrCode = 16,        // codeword longer than the format allows
  kErrBlocks = 32,      // copy blocks inconsistent with the referenced list
};

struct EFDir {
  const uint64_t* lower;  // byte-swapped words, MSB-first fields of width l
  const uint64_t* upper;  // little-endian words
  const uint64_t* samp;   // position of one #64k, or (1<<63) | spill base
  const uint64_t* spill;
  uint64_t upper_words;
  int32_t l;
};

struct Ckpt;

struct Streams {
  const uint64_t* g;  // graph stream, byte-swapped words (+5 zero words)
  const uint64_t* t;  // timestamp stream, byte-swapped words (+5 zero words)
  uint64_t g_bits, t_bits;
  uint32_t g_last, t_last;  // index of the last padding word of each stream
  // Word i of the graph stream is g[i << g_sh]: 0 for the stream itself; a
  // walk over a record's column copy (run_item) sets g to the record's first
  // word, g_sh to 6 and g_bits / g_last to the copy's extent.
  uint32_t g_sh;
  EFDir gi, ti;
  int32_t n;       // vertices; ids are [0, n)
  int32_t min_ts;  // .timestamps header value
  int32_t flat_stride;  // checkpoint stride of records without references (chunk_stride)
  int32_t burst;        // contacts per consume phase of a walk (kBurstDefault unless tuned)
  int32_t* err;    // sticky error word
  // Decode checkpoints of long records: vinfo[x] >= 0 is the first
  // checkpoint of x, checkpoint j covering contacts [j * stride, (j + 1) *
  // stride) with the stride stored in the checkpoint.
  const Ckpt* ck;
  const int32_t* vinfo;
  // Record start table: vpos[2x] and vpos[2x + 1] are the bit offsets of
  // vertex x in .graph and .timestamps (the Elias-Fano getOffset of both
  // indexes, bfs.cu:543-565, evaluated once at load), so that starting a
  // record costs one 16-byte load instead of two select walks.
  const uint64_t* vpos;
};

TGA_HD void raise_err(const Streams& S, int32_t bit) {
#if defined(__HIP_DEVICE_COMPILE__)
  atomicOr(S.err, bit);
#else
  __atomic_fetch_or(S.err, bit, __ATOMIC_RELAXED);
#endif
}

// 64 stream bits starting at bit `pos` (MSB-first), word i at w[i << sh].
TGA_HD uint64_t peek64(const uint64_t* w, uint64_t pos, uint32_t sh = 0) {
  const uint64_t i = pos >> 6;
  const uint32_t s = (uint32_t)pos & 63u;
  const uint64_t a = w[i << sh];
  const uint64_t b = w[(i + 1) << sh];
  return s ? (a << s) | (b >> (64u - s)) : a;
}

TGA_HD int64_t nat2int(uint64_t v) { return (v & 1) ? -(int64_t)(v >> 1) - 1 : (int64_t)(v >> 1); }

// Codeword decoders on a 64-bit window `p`; `len` receives the codeword
// length.  A window too short for the code (corrupt stream) flags kErrCode
// and yields 0 with length 1.
TGA_HD uint64_t dec_gamma(const Streams& S, uint64_t p, uint32_t& len) {
  const uint32_t z = clz64(p);
  if (z > 31) {
    raise_err(S, kErrCode);
    len = 1;
    return 0;
  }
  len = 2 * z + 1;
  return ((p << z) >> (63 - z)) - 1;
}

TGA_HD uint32_t dec_unary(const Streams& S, uint64_t p, uint32_t& len) {
  const uint32_t z = clz64(p);
  if (z > 62) {
    raise_err(S, kErrCode);
    len = 1;
    return 0;
  }
  len = z + 1;
  return z;
}

template <int K>
TGA_HD uint64_t dec_zeta(const Streams& S, uint64_t p, uint32_t& len) {
  const uint32_t h = clz64(p);
  const uint32_t nb = h * K + K - 1;  // bits of the short form
  if (h + nb + 2 > 64) {
    raise_err(S, kErrCode);
    len = 1;
    return 0;
  }
  const uint64_t left = 1ull << (h * K);
  const uint64_t m = (p << (h + 1)) >> (64 - nb);
  if (m < left) {
    len = h + 1 + nb;
    return m + left - 1;
  }
  len = h + 2 + nb;
  return (m << 1) + ((p << (h + 1 + nb)) >> 63) - 1;
}

// Two-word-peek cursor reads (bound checked against the stream length).
struct Rd {
  const uint64_t* w;
  uint64_t bits;
  uint32_t sh;  // word i at w[i << sh]
};

TGA_HD uint64_t safe_peek(const Streams& S, Rd r, uint64_t pos) {
  if (pos >= r.bits) {
    raise_err(S, kErrOverrun);
    return ~0ull;  // all ones: every code decodes to 0 with length 1
  }
  return peek64(r.w, pos, r.sh);
}

TGA_HD uint32_t rd_unary(const Streams& S, Rd r, uint64_t& pos) {
  uint32_t len;
  const uint32_t v = dec_unary(S, safe_peek(S, r, pos), len);
  pos += len;
  return v;
}

TGA_HD uint64_t rd_gamma(const Streams& S, Rd r, uint64_t& pos) {
  uint32_t len;
  const uint64_t v = dec_gamma(S, safe_peek(S, r, pos), len);
  pos += len;
  return v;
}

template <int K>
TGA_HD uint64_t rd_zeta(const Streams& S, Rd r, uint64_t& pos) {
  uint32_t len;
  const uint64_t v = dec_zeta<K>(S, safe_peek(S, r, pos), len);
  pos += len;
  return v;
}

// Register bit buffer over one sequentially read sub-stream: words
// a = w[idx], b = w[idx + 1] and c = w[idx + 2] are held, `sh` is the bit
// offset into a.  Advancing past a loads the word after c, which is needed
// only 64 or more bits later, so the load's latency overlaps the decoding of
// the codewords in between.  The stream pointer and the index of its last
// padding word are passed in (they are the same for every buffer over one
// stream); stream word arrays carry five zero words of padding, and reads
// past `last` return 0 (every code then fails its length check).  Word i of
// the stream is w[i << ws] (ws: Streams::g_sh).
struct BitBuf {
  uint64_t a, b, c;
  uint32_t idx;  // word index of a
  uint32_t sh;   // bit offset into a

  TGA_HD void init(const uint64_t* w, uint32_t last, uint64_t pos, uint32_t ws = 0) {
    uint64_t i = pos >> 6;
    if (i + 2 > last) i = last - 2;
    idx = (uint32_t)i;
    sh = (uint32_t)pos & 63u;
    a = w[(uint64_t)idx << ws];
    b = w[(uint64_t)(idx + 1) << ws];
    c = w[(uint64_t)(idx + 2) << ws];
  }
  // A buffer that will not be read (its position is recorded as 0).
  TGA_HD void clear() {
    a = b = c = 0;
    idx = sh = 0;
  }
  TGA_HD uint64_t peek() const { return sh ? (a << sh) | (b >> (64u - sh)) : a; }
  TGA_HD void skip(uint32_t n, const uint64_t* w, uint32_t last, uint32_t ws = 0) {  // n <= 64
    sh += n;
    if (sh >= 64) {
      sh -= 64;
      a = b;
      b = c;
      ++idx;
      c = w[(uint64_t)minu32(idx + 2, last) << ws];  // past the end: the last padding word
    }
  }
  TGA_HD uint64_t pos() const { return (uint64_t)idx * 64 + sh; }
};

// Pipelined bit buffer for the sub-streams read about one codeword per
// contact (timestamps, and the top-level residuals and multiples).  Words
// a, b, c = w[idx..idx+2] are always held, and up to two look-ahead words
// d, e = w[idx+3], w[idx+4] (`nv` of them).  Once per contact the walk calls
// step() before decoding: it merges the word requested by the previous step
// and requests the next missing look-ahead word.  Every wait for loaded data
// then falls at the top of a contact, on loads issued a contact earlier: the
// wave's memory counter is shared by its lanes and drains in order, so a load
// consumed in the contact that issues it would stall the whole wave on every
// lane's refills.  A shift with no look-ahead word left (more than one word
// consumed in one contact) loads the missing word itself.
struct PBuf {
  uint64_t a, b, c, d, e;
  uint64_t pf;    // word pfi, requested by the last step
  uint32_t idx;   // word index of a
  uint32_t sh;    // bit offset into a
  uint32_t pfi;   // index of the word in pf (~0: none)
  uint32_t nv;    // look-ahead words held (0..2)

  TGA_HD void init(const uint64_t* w, uint32_t last, uint64_t pos) {
    uint64_t i = pos >> 6;
    if (i + 4 > last) i = last - 4;
    idx = (uint32_t)i;
    sh = (uint32_t)pos & 63u;
    a = w[idx];
    b = w[idx + 1];
    c = w[idx + 2];
    d = w[idx + 3];
    e = w[idx + 4];
    nv = 2;
    pf = 0;
    pfi = ~0u;
  }
  // A buffer that will not be read: its look-ahead counts as full, so steps
  // request nothing.
  TGA_HD void clear() {
    a = b = c = d = e = pf = 0;
    idx = sh = 0;
    nv = 2;
    pfi = ~0u;
  }
  TGA_HD uint64_t peek() const { return sh ? (a << sh) | (b >> (64u - sh)) : a; }
  TGA_HD void skip(uint32_t n, const uint64_t* w, uint32_t last) {  // n <= 64
    sh += n;
    if (sh >= 64) {
      sh -= 64;
      a = b;
      b = c;
      if (nv) {
        c = d;
        d = e;
        --nv;
      } else {
        c = w[minu32(idx + 3, last)];
      }
      ++idx;
    }
  }
  TGA_HD void step(const uint64_t* w, uint32_t last) {
    // Merge, branch-free: pf fills the first missing look-ahead slot.
    const uint32_t m = (pfi == idx + 3 + nv);
    d = (m & (nv == 0)) ? pf : d;
    e = (m & (nv == 1)) ? pf : e;
    nv += m;
    pfi = nv < 2 ? idx + 3 + nv : ~0u;
    if (nv < 2) pf = w[minu32(pfi, last)];
  }
  TGA_HD uint64_t pos() const { return (uint64_t)idx * 64 + sh; }
};

// Buffered readers of the graph stream (g_*) and the timestamp stream (t_*).
template <class B>
TGA_HD void g_init(const Streams& S, B& b, uint64_t pos) {
  if (pos >= S.g_bits) raise_err(S, kErrOverrun);
  b.init(S.g, S.g_last, pos, S.g_sh);
}
template <class B>
TGA_HD uint64_t g_gamma(const Streams& S, B& b) {
  uint32_t len;
  const uint64_t v = dec_gamma(S, b.peek(), len);
  b.skip(len, S.g, S.g_last, S.g_sh);
  return v;
}
template <class B>
TGA_HD uint32_t g_unary(const Streams& S, B& b) {
  uint32_t len;
  const uint32_t v = dec_unary(S, b.peek(), len);
  b.skip(len, S.g, S.g_last, S.g_sh);
  return v;
}
template <int K, class B>
TGA_HD uint64_t g_zeta(const Streams& S, B& b) {
  uint32_t len;
  const uint64_t v = dec_zeta<K>(S, b.peek(), len);
  b.skip(len, S.g, S.g_last, S.g_sh);
  return v;
}
template <class B>
TGA_HD void t_init(const Streams& S, B& b, uint64_t pos) {
  if (pos >= S.t_bits) raise_err(S, kErrOverrun);
  b.init(S.t, S.t_last, pos);
}
template <class B>
TGA_HD uint64_t t_zeta2(const Streams& S, B& b) {
  uint32_t len;
  const uint64_t v = dec_zeta<2>(S, b.peek(), len);
  b.skip(len, S.t, S.t_last);
  return v;
}

// Position of the r-th (0-based) set bit of w (r < popcount(w)).
TGA_HD uint32_t select_in_word(uint64_t w, uint32_t r) {
  uint32_t pos = 0;
  uint32_t c = popc64(w & 0xffffffffull);
  if (r >= c) { r -= c; w >>= 32; pos += 32; }
  c = popc64(w & 0xffffull);
  if (r >= c) { r -= c; w >>= 16; pos += 16; }
  c = popc64(w & 0xffull);
  if (r >= c) { r -= c; w >>= 8; pos += 8; }
  c = popc64(w & 0xfull);
  if (r >= c) { r -= c; w >>= 4; pos += 4; }
  c = popc64(w & 0x3ull);
  if (r >= c) { r -= c; w >>= 2; pos += 2; }
  c = (uint32_t)(w & 1ull);
  if (r >= c) { pos += 1; }
  return pos;
}

// select1(i): position of the i-th one of the upper bits.
TGA_HD uint64_t ef_select(const Streams& S, const EFDir& d, uint64_t i) {
  const uint64_t s = d.samp[i >> 6];
  if (s >> 63) return d.spill[(s & ~(1ull << 63)) + (i & 63)];
  uint32_t r = (uint32_t)(i & 63);
  uint64_t w = s >> 6;
  uint64_t word = d.upper[w] & (~0ull << (s & 63));
  for (uint32_t guard = 0;; ++guard) {
    const uint32_t c = popc64(word);
    if (r < c) break;
    r -= c;
    ++w;
    if (w >= d.upper_words || guard > (kSpanBits / 64 + 2)) {
      raise_err(S, kErrOverrun);
      return 0;
    }
    word = d.upper[w];
  }
  return w * 64 + select_in_word(word, r);
}

// getOffset (bfs.cu:543-565): ((select1(i) - i) << l) | low_i.
TGA_HD uint64_t ef_get(const Streams& S, const EFDir& d, uint64_t i) {
  const uint64_t hi = ef_select(S, d, i) - i;
  uint64_t lo = 0;
  if (d.l > 0) lo = peek64(d.lower, i * (uint64_t)d.l) >> (64 - d.l);
  return (hi << d.l) | lo;
}

// End of the timestamps of x: where those of x + 1 start.
TGA_HD uint64_t rec_tend(const Streams& S, int32_t x) {
  return x + 1 < S.n ? S.vpos[2 * (size_t)x + 3] : S.t_bits;
}

// Start offsets of the record of x in both streams.
TGA_HD void rec_pos(const Streams& S, int32_t x, uint64_t& gpos, uint64_t& tpos) {
  const ulonglong2 p = reinterpret_cast<const ulonglong2*>(S.vpos)[x];
  gpos = p.x;
  tpos = p.y;
}

// ---------------------------------------------------------------------------
// Successor-list state machine.
//
// A record (BVMultiGraph.java:738-880 / bfs.cu:1218-1633) is the sorted merge
// of four sources: values copied from the reference list (copy/skip blocks,
// implicit last block), multiples runs (value repeated len times), intervals
// (consecutive values) and zeta_3 residuals.  Each source keeps its head; the
// merge returns the smallest head.  The copy source of level L pulls values
// from level L+1, which decodes the referenced vertex the same way, so a
// decoder for chains of depth MAXD holds MAXD + 1 levels.  MAXD = 0 is the
// flat decoder for records without a reference.  A valid encoding never
// yields the same value from two sources (multiples are pulled out before
// copying, BVMultiGraph.java:1644-1647), so heads are distinct unless kInf.
struct Level {
  uint64_t mpos, ipos, rpos, bpos;       // cursors of the four sub-streams
  int32_t x;                              // vertex whose list this level yields
  int32_t hm, mrem, mleft;                // multiples: head, copies left of head, runs not read
  int32_t hi, irem, ileft;                // intervals: head, values left incl. head, intervals not read
  int32_t hr, rleft;                      // residuals: head, residuals not read
  int32_t hc, cleft, bleft, bidx, brem;   // copies: head, copies left, blocks not read, block index, left in block
  int32_t kpos;                           // values this level has yielded
  int32_t pad_;
};

// Residuals and multiples of every level are read through bit buffers (their
// Level cursors rpos / mpos are stale while decoding and are written back
// only into checkpoints); intervals and copy blocks, which are rare, are read
// directly.  The top level yields one value per contact and uses pipelined
// buffers; the levels of a reference chain yield values in bursts and refill
// on demand.
template <int MAXD>
struct Dec {
  Level lv[MAXD + 1];
  BitBuf rb0, mb0;                       // level 0: residuals, multiples (and the header while parsed)
  BitBuf rb[MAXD > 0 ? MAXD : 1];        // level L >= 1 at [L - 1]
  BitBuf mb[MAXD > 0 ? MAXD : 1];
};

template <int L, int MAXD>
TGA_HD auto& rbuf(Dec<MAXD>& D) {
  if constexpr (L == 0) return D.rb0; else return D.rb[L - 1];
}
template <int L, int MAXD>
TGA_HD auto& mbuf(Dec<MAXD>& D) {
  if constexpr (L == 0) return D.mb0; else return D.mb[L - 1];
}
template <int L, int MAXD>
TGA_HD const auto& rbuf(const Dec<MAXD>& D) {
  if constexpr (L == 0) return D.rb0; else return D.rb[L - 1];
}
template <int L, int MAXD>
TGA_HD const auto& mbuf(const Dec<MAXD>& D) {
  if constexpr (L == 0) return D.mb0; else return D.mb[L - 1];
}

// Saved decoder state at a chunk boundary of a long list (see engine.hip).
//
// Chunks hold a fixed number of contacts, so a run of contacts to one
// neighbor may cross a chunk boundary.  Such a "split run" gets a slot (the
// index of the first chunk it enters); `cont_*` names the split run a chunk
// starts inside, `tail_*` the one it ends inside (slot -1: none).  Rules that
// act once per distinct neighbor (first window contact of a pair) defer the
// contacts of split runs to their slot and resolve them after the pass.
struct Ckpt {
  int32_t x;          // vertex
  int32_t k0;         // first contact of the chunk
  int32_t k1;         // one past its last contact
  int32_t depth;      // levels in use - 1
  uint64_t tpos;      // timestamp cursor
  int32_t tprev;      // timestamp before contact k0
  int32_t nchunk;     // chunks of x (stored in every chunk of x)
  int32_t cont_slot;  // split run continuing into this chunk, or -1
  int32_t cont_r;     // its neighbor
  int32_t tail_slot;  // split run continuing out of this chunk, or -1
  int32_t tail_r;     // its neighbor
  int32_t stride;     // contacts per chunk of x
  int32_t pad_;
  Level lv[kMaxDepth + 1];
};

// Split-run context of the item being walked.
struct Split {
  int32_t cont_slot, cont_r, tail_slot, tail_r;
  // Slot that contact (r) of this item is deferred to, or -1.
  TGA_HD int32_t slot_of(int32_t r) const {
    if (r == cont_r && cont_slot >= 0) return cont_slot;
    if (r == tail_r && tail_slot >= 0) return tail_slot;
    return -1;
  }
};

template <int L, int MAXD>
TGA_HD int32_t dec_next(const Streams& S, Dec<MAXD>& D);

// Restores one level from a saved copy: the buffers of the sub-streams that
// still have codes to read are refilled at their saved positions.
template <class RB, class MB>
TGA_HD void lv_restore(const Streams& S, Level& v, RB& rb, MB& mb, const Level& src) {
  v = src;
  if (v.rleft > 0) g_init(S, rb, v.rpos); else rb.clear();
  if (v.mleft > 0) g_init(S, mb, v.mpos); else mb.clear();
}

// Copies one level, with its sub-stream positions, into a checkpoint.
template <class RB, class MB>
TGA_HD void lv_save(Level& dst, const Level& v, const RB& rb, const MB& mb) {
  dst = v;
  dst.rpos = rb.pos();
  dst.mpos = mb.pos();
}

// Restores levels L..MAXD of D from saved levels src[L - LL ..] (those within
// the saved chain depth).
template <int LL, int L, int MAXD>
TGA_HD void restore_levels(const Streams& S, Dec<MAXD>& D, const Level* src, int32_t depth) {
  if (L - LL <= depth) lv_restore(S, D.lv[L], rbuf<L>(D), mbuf<L>(D), src[L - LL]);
  if constexpr (L < MAXD) restore_levels<LL, L + 1, MAXD>(S, D, src, depth);
}

template <int L, int MAXD>
TGA_HD void save_levels(Level* dst, const Dec<MAXD>& D) {
  lv_save(dst[L], D.lv[L], rbuf<L>(D), mbuf<L>(D));
  if constexpr (L < MAXD) save_levels<L + 1, MAXD>(dst, D);
}

// Skips toward value `target` of level LL by restoring the checkpoint of the
// referenced record at or before it, when that record is long enough to have
// checkpoints.  Returns how many values the jump passed over.
template <int LL, int MAXD>
TGA_HD int32_t dec_jump(const Streams& S, Dec<MAXD>& D, int32_t target) {
  if (!S.vinfo) return 0;
  const int32_t first = S.vinfo[D.lv[LL].x];
  if (first < 0) return 0;
  const int32_t stride = S.ck[first].stride;
  int32_t j = target / stride;
  const int32_t nch = S.ck[first].nchunk;
  if (j >= nch) j = nch - 1;
  const int32_t from = D.lv[LL].kpos;
  if (j * stride <= from) return 0;
  const Ckpt* c = S.ck + first + j;
  restore_levels<LL, LL, MAXD>(S, D, c->lv, c->depth);
  return D.lv[LL].kpos - from;
}

// Pulls the next value of level L's copy source from level L+1.
template <int L, int MAXD>
TGA_HD int32_t dec_masked(const Streams& S, Dec<MAXD>& D) {
  if constexpr (L >= MAXD) {
    return kInf;
  } else {
    Level& v = D.lv[L];
    const Rd g{S.g, S.g_bits, S.g_sh};
    for (;;) {
      if (v.cleft <= 0) return kInf;
      if (v.brem == 0) {
        ++v.bidx;
        if (v.bleft > 0) {
          v.brem = (int32_t)rd_gamma(S, g, v.bpos) + (v.bidx ? 1 : 0);
          --v.bleft;
          // A long skip block tries, once, to jump through the referenced
          // record's checkpoints.
          if ((v.bidx & 1) && v.brem > 32) v.brem -= dec_jump<L + 1, MAXD>(S, D, D.lv[L + 1].kpos + v.brem);
        } else if ((v.bidx & 1) == 0) {
          v.brem = kInf;  // implicit last block: copy the rest
        } else {
          raise_err(S, kErrBlocks);
          v.cleft = 0;
          return kInf;
        }
        continue;
      }
      const int32_t val = dec_next<L + 1, MAXD>(S, D);
      if (val == kInf) {
        raise_err(S, kErrBlocks);
        v.cleft = 0;
        return kInf;
      }
      --v.brem;
      if ((v.bidx & 1) == 0) {
        --v.cleft;
        return val;
      }
    }
  }
}

template <int L, int MAXD>
TGA_HD int32_t dec_next(const Streams& S, Dec<MAXD>& D) {
  Level& v = D.lv[L];
  const Rd g{S.g, S.g_bits, S.g_sh};
  // The deepest level has no copy source.
  const int32_t m = (L < MAXD) ? min32(min32(v.hm, v.hi), min32(v.hr, v.hc)) : min32(min32(v.hm, v.hi), v.hr);
  if (m == kInf) return kInf;
  ++v.kpos;
  if (m == v.hr) {
    if (v.rleft > 0) {
      v.hr += (int32_t)g_zeta<3>(S, rbuf<L>(D)) + 1;
      --v.rleft;
    } else {
      v.hr = kInf;
    }
  } else if (m == v.hm) {
    if (--v.mrem == 0) {
      if (v.mleft > 0) {
        v.hm += (int32_t)g_gamma(S, mbuf<L>(D)) + 1;
        v.mrem = (int32_t)g_gamma(S, mbuf<L>(D)) + 2;
        --v.mleft;
      } else {
        v.hm = kInf;
      }
    }
  } else if (m == v.hi) {
    if (--v.irem == 0) {
      if (v.ileft > 0) {
        v.hi += (int32_t)rd_gamma(S, g, v.ipos) + 2;
        v.irem = (int32_t)rd_gamma(S, g, v.ipos) + 4;
        --v.ileft;
      } else {
        v.hi = kInf;
      }
    } else {
      ++v.hi;
    }
  } else if constexpr (L < MAXD) {
    v.hc = dec_masked<L, MAXD>(S, D);
  }
  return m;
}

// Next successor without consuming it.
template <int MAXD>
TGA_HD int32_t dec_peek(const Dec<MAXD>& D) {
  const Level& v = D.lv[0];
  return (MAXD > 0) ? min32(min32(v.hm, v.hi), min32(v.hr, v.hc)) : min32(min32(v.hm, v.hi), v.hr);
}

// Parses the record of x at bit offset `gpos` into level L (and the levels of
// its reference chain).  Returns the outdegree d; `depth` receives the
// deepest level used.
template <int L, int MAXD>
TGA_HD int32_t dec_init(const Streams& S, Dec<MAXD>& D, int32_t x, uint64_t gpos,
                                            int32_t& depth) {
  Level& v = D.lv[L];
  // The header is parsed through the residual buffer, which ends up at the
  // first residual; when the record has multiples, the multiples buffer is
  // forked from it at the first run.
  auto& h = rbuf<L>(D);
  v.x = x;
  v.hm = v.hi = v.hr = v.hc = kInf;
  v.mrem = v.mleft = v.irem = v.ileft = v.rleft = 0;
  v.cleft = v.bleft = v.brem = 0;
  v.kpos = 0;
  v.bidx = -1;
  v.mpos = v.ipos = v.rpos = v.bpos = gpos;
  mbuf<L>(D).clear();
  g_init(S, h, gpos);
  const int32_t d = (int32_t)g_gamma(S, h);
  if (d <= 0) return 0;
  if (L > depth) depth = L;
  const uint32_t r = g_unary(S, h);
  int32_t copied = 0;
  if (r > 0) {
    if constexpr (L >= MAXD) {
      raise_err(S, kErrDepth);
    } else {
      if (r > (uint32_t)kWindow || (int32_t)r > x) {
        raise_err(S, kErrDepth);
      } else {
        const int32_t bc = (int32_t)g_gamma(S, h);
        v.bpos = h.pos();
        int32_t total = 0, even = 0;
        for (int32_t i = 0; i < bc; ++i) {
          if (h.pos() >= S.g_bits) {  // malformed: stop at the end of the stream
            raise_err(S, kErrOverrun);
            break;
          }
          const int32_t b = (int32_t)g_gamma(S, h) + (i ? 1 : 0);
          total += b;
          if ((i & 1) == 0) even += b;
        }
        const int32_t y = x - (int32_t)r;
        const int32_t dref = dec_init<L + 1, MAXD>(S, D, y, S.vpos[2 * (size_t)y], depth);
        copied = even + ((bc & 1) ? 0 : dref - total);
        if (copied < 0 || copied > d) {
          raise_err(S, kErrBlocks);
          copied = 0;
        }
        v.cleft = copied;
        v.bleft = bc;
        v.hc = dec_masked<L, MAXD>(S, D);
      }
    }
  }
  int32_t extra = d - copied;
  if (extra > 0) {  // multiples
    const int32_t mc = (int32_t)g_gamma(S, h);
    if (mc > 0) {
      mbuf<L>(D) = h;
      for (int32_t i = 0; i < mc; ++i) {
        if (extra <= 0) {  // more runs than contacts: malformed
          raise_err(S, kErrBlocks);
          break;
        }
        g_gamma(S, h);
        extra -= (int32_t)g_gamma(S, h) + 2;
      }
      v.hm = (int32_t)(nat2int(g_gamma(S, mbuf<L>(D))) + x);
      v.mrem = (int32_t)g_gamma(S, mbuf<L>(D)) + 2;
      v.mleft = mc - 1;
    }
  }
  if (extra > 0) {  // intervals
    const int32_t ic = (int32_t)g_gamma(S, h);
    if (ic > 0) {
      const Rd g{S.g, S.g_bits, S.g_sh};
      uint64_t first = h.pos();
      for (int32_t i = 0; i < ic; ++i) {
        if (extra <= 0) {  // more intervals than contacts: malformed
          raise_err(S, kErrBlocks);
          break;
        }
        g_gamma(S, h);
        extra -= (int32_t)g_gamma(S, h) + 4;
      }
      v.hi = (int32_t)(nat2int(rd_gamma(S, g, first)) + x);
      v.irem = (int32_t)rd_gamma(S, g, first) + 4;
      v.ileft = ic - 1;
      v.ipos = first;
    }
  }
  if (extra > 0) {  // residuals
    v.hr = (int32_t)(nat2int(g_zeta<3>(S, h)) + x);
    v.rleft = extra - 1;
  }
  return d;
}

// Start word of a record without references: what dec_init learns by
// parsing the header and skipping the multiples and intervals sections
// (their codes are read again, one run at a time, as the walk reaches them),
// packed into 64 bits at open so that a walk can start the three sub-streams
// directly: d (12 bits) | residuals (12) | bit offset of the intervals
// section + 1, 0 when the record has none (20) | bit offset of the first
// residual (20), offsets from the record's first bit.  kNoStart when the
// record has a reference, does not fit the fields or is malformed (the walk
// then parses it with dec_init, which reports the error).
constexpr uint64_t kNoStart = ~0ull;
TGA_HD uint32_t gamma_len(uint32_t v) { return 2u * (31u - (uint32_t)__builtin_clz(v + 1u)) + 1u; }

TGA_HD uint64_t flat_start_word(const Streams& S, uint64_t gpos) {
  const Rd g{S.g, S.g_bits, S.g_sh};
  uint64_t p = gpos;
  const uint64_t d = rd_gamma(S, g, p);
  if (d == 0 || d >= 4096) return kNoStart;
  if (rd_unary(S, g, p) != 0) return kNoStart;
  int64_t extra = (int64_t)d;
  const uint64_t mc = rd_gamma(S, g, p);
  for (uint64_t i = 0; i < mc; ++i) {
    if (extra <= 0 || p >= S.g_bits) return kNoStart;
    rd_gamma(S, g, p);
    extra -= (int64_t)rd_gamma(S, g, p) + 2;
  }
  uint64_t ip = 0, rp = 0;
  if (extra > 0) {
    ip = p - gpos + 1;
    const uint64_t ic = rd_gamma(S, g, p);
    for (uint64_t i = 0; i < ic; ++i) {
      if (extra <= 0 || p >= S.g_bits) return kNoStart;
      rd_gamma(S, g, p);
      extra -= (int64_t)rd_gamma(S, g, p) + 4;
    }
  }
  if (extra > 0) rp = p - gpos;
  if (extra < 0 || ip >= (1u << 20) || rp >= (1u << 20)) return kNoStart;
  return d | (uint64_t)extra << 12 | ip << 24 | rp << 44;
}

// dec_init of a record without references from its start word (Streams
// positions relative to the record's first bit): the same level state, the
// header and the skipped sections left unread.
template <int MAXD>
TGA_HD int32_t dec_init_flat(const Streams& S, Dec<MAXD>& D, int32_t x, uint64_t w) {
  Level& v = D.lv[0];
  auto& h = rbuf<0>(D);
  v.x = x;
  v.hm = v.hi = v.hr = v.hc = kInf;
  v.mrem = v.mleft = v.irem = v.ileft = v.rleft = 0;
  v.cleft = v.bleft = v.brem = 0;
  v.kpos = 0;
  v.bidx = -1;
  v.mpos = v.ipos = v.rpos = v.bpos = 0;
  mbuf<0>(D).clear();
  const uint32_t d = (uint32_t)w & 4095u;
  const int32_t nres = (int32_t)(w >> 12) & 4095;
  const uint64_t ip = (w >> 24) & 0xFFFFFu, rp = w >> 44;
  g_init(S, h, gamma_len(d) + 1);  // past d and the reference 0: the multiples count
  const int32_t mc = (int32_t)g_gamma(S, h);
  if (mc > 0) {
    mbuf<0>(D) = h;
    v.hm = (int32_t)(nat2int(g_gamma(S, mbuf<0>(D))) + x);
    v.mrem = (int32_t)g_gamma(S, mbuf<0>(D)) + 2;
    v.mleft = mc - 1;
  }
  if (ip) {
    const Rd g{S.g, S.g_bits, S.g_sh};
    uint64_t p = ip - 1;
    const int32_t ic = (int32_t)rd_gamma(S, g, p);
    if (ic > 0) {
      v.hi = (int32_t)(nat2int(rd_gamma(S, g, p)) + x);
      v.irem = (int32_t)rd_gamma(S, g, p) + 4;
      v.ileft = ic - 1;
      v.ipos = p;
    }
  }
  if (nres > 0) {
    g_init(S, h, rp);
    v.hr = (int32_t)(nat2int(g_zeta<3>(S, h)) + x);
    v.rleft = nres - 1;
  }
  return (int32_t)d;
}

// Timestamp cursor: zeta_2 of int2nat deltas, reset to the header value per
// vertex (EvolvingMultiGraph.java:131-144; bfs.cu:1444, :1597).
struct TsCursor {
  PBuf b;
  int32_t prev;
};

TGA_HD int32_t ts_next(const Streams& S, TsCursor& c) {
  c.prev = (int32_t)((uint32_t)c.prev + (uint32_t)nat2int(t_zeta2(S, c.b)));
  return c.prev;
}

// Saves the decoder at a chunk boundary.
template <int MAXD>
TGA_HD void ckpt_save(Ckpt* c, const Dec<MAXD>& D, int32_t depth, const TsCursor& tc, int32_t x,
                                          int32_t k0) {
  c->x = x;
  c->k0 = k0;
  c->depth = depth;
  c->tpos = tc.b.pos();
  c->tprev = tc.prev;
  save_levels<0, MAXD>(c->lv, D);
}

// Next run of the top level's successor sequence: a value and how many
// consecutive contacts carry it (cnt).  Only multiples repeat a value (they
// are pulled out before copying, so no other source yields a multiple's
// value), so a run is the rest of a multiple, or one value of the other
// sources.  kInf when the list is exhausted.
template <int MAXD>
TGA_HD int32_t dec_next_run(const Streams& S, Dec<MAXD>& D, int32_t& cnt) {
  Level& v = D.lv[0];
  const int32_t o = (MAXD > 0) ? min32(min32(v.hi, v.hr), v.hc) : min32(v.hi, v.hr);
  if (v.hm < o) {
    const int32_t m = v.hm;
    cnt = v.mrem;
    v.kpos += cnt;
    if (v.mleft > 0) {
      v.hm += (int32_t)g_gamma(S, mbuf<0>(D)) + 1;
      v.mrem = (int32_t)g_gamma(S, mbuf<0>(D)) + 2;
      --v.mleft;
    } else {
      v.hm = kInf;
      v.mrem = 0;
    }
    return m;
  }
  cnt = 1;
  return dec_next<0, MAXD>(S, D);
}

// Per-lane LDS window over the timestamp stream, read by the walk.
//
// A lane's timestamps are one sequential stream of zeta_2 codes (about 14 to
// 28 bits a contact).  Read eight bytes at a time from HBM, each lane would
// touch each 128-byte line of its stream a dozen times over hundreds of
// contacts, and with ~10^5 lanes in flight the L2 evicts the line between
// touches (a pass fetched ~9x its stream bytes).  Instead the window is
// topped up once per burst of contacts in whole 16-byte units (a
// global_load_dwordx4 each, the units of one top-up back to back, so a line is
// fetched once), and the codes are decoded from LDS.
//
// Layout: kStageDw stream dwords per lane in stream order at slot
// (dword index mod kStageDw), slots kStageDw and kStageDw + 1 mirroring slots
// 0 and 1 so a 96-bit read never wraps; slot i of a lane at s[i * stride]
// (lanes interleaved: every lane of a wave hits its own bank; stride a power
// of two, at least 8).
//
// The units come either from the linear stream (uw = the stream words,
// ustride 1, positions absolute) or from a column copy of the item's codes
// made at open (uw = the item's first unit, ustride = the units between two
// of its consecutive units, positions relative to the item's first code,
// nbits its code bits): there the 64 items a wave walks together have their
// units interleaved, so the wave's top-up loads fall into a few adjacent
// lines instead of 64 separate streams.
constexpr int32_t kStageDw = 16;  // staged dwords per lane (power of two, multiple of 4)
constexpr uint32_t kNoEnd = ~0u;
struct TsStage {
  uint32_t* s;
  int32_t stride;
  uint32_t pos;    // bit offset of the next code from relative dword 0 (the first unit's first dword)
  uint32_t wend;   // relative dword index one past the last staged dword
  int32_t wtop;    // 32 * wend - 64: a code at pos >= wtop may reach past the staged dwords
  int32_t prev;    // the timestamp before the next code
  uint32_t uleft;  // units holding codes of the item not yet staged
  uint32_t pend;   // pos one past the item's last code (kNoEnd: not known), checked when the walk ends
  const uint64_t* uw = nullptr;  // next unit to stage (word pairs); nullptr: the linear stream (stage_init)
  uint32_t ustride = 1;          // unit pitch of the source, in units
  uint32_t nbits = 0;            // column source: the item's code bits
  // The item's graph column copy (run_item), or nullptr.
  const uint64_t* gw = nullptr;
  uint32_t gbits = 0;            // the record's graph bits
};

// One unit's four dwords, each word's high dword first.
struct Unit {
  uint32_t d0, d1, d2, d3;
};
TGA_HD Unit load_unit(const uint64_t* w) {
#if defined(__HIP_DEVICE_COMPILE__)
  const uint4 v = *reinterpret_cast<const uint4*>(w);
  return Unit{v.y, v.x, v.w, v.z};
#else
  return Unit{(uint32_t)(w[0] >> 32), (uint32_t)w[0], (uint32_t)(w[1] >> 32), (uint32_t)w[1]};
#endif
}
// Writes a loaded unit at relative dword wend.
TGA_HD void put_unit(TsStage& st, Unit u) {
  const uint32_t d0 = u.d0, d1 = u.d1, d2 = u.d2, d3 = u.d3;
  const uint32_t slot = st.wend & (kStageDw - 1);
  uint32_t* p = st.s + (size_t)slot * st.stride;
  p[0] = d0;
  p[st.stride] = d1;
  p[2 * st.stride] = d2;
  p[3 * st.stride] = d3;
  if (slot == 0) {
    st.s[(size_t)kStageDw * st.stride] = d0;
    st.s[(size_t)(kStageDw + 1) * st.stride] = d1;
  }
  st.wend += 4;
  st.wtop += 128;
}
// Stages the next unit (source words uw[0], uw[1]) at relative dword wend
// with one 16-byte load.  Only units whose first dword lies before the item's
// end are staged (uleft); in the linear stream the last word of such a unit
// lies within the stream's five words of padding, and a column holds every
// unit its item's codes touch, so no bound check is needed.
TGA_HD void stage_unit(TsStage& st) {
  const Unit u = load_unit(st.uw);
  st.uw += 2 * (size_t)st.ustride;
  --st.uleft;
  put_unit(st, u);
}

// Starts the window at bit `pos` for codes ending before bit `end` (the
// first unit holds the code at pos).
// `exact`: the item's codes end at `end` (a whole record; a chunk's codes end
// at the next chunk's cursor, before `end`).
TGA_HD void stage_init(const Streams& S, TsStage& st, uint64_t pos, uint64_t end, int32_t prev, bool exact) {
  if (pos >= S.t_bits) raise_err(S, kErrOverrun);
  const uint64_t base = (pos >> 5) & ~(uint64_t)3;  // stream dword of the first unit
  const uint64_t lim = (end + 31) >> 5;             // one past the item's last dword
  st.uw = S.t + (base >> 1);
  st.ustride = 1;
  st.uleft = lim > base ? (uint32_t)((lim - base + 3) >> 2) : 0u;
  st.pend = exact && end - base * 32 < kNoEnd ? (uint32_t)(end - base * 32) : kNoEnd;
  st.pos = (uint32_t)(pos - base * 32);
  st.wend = 0;
  st.wtop = -64;
  st.prev = prev;
}

// Starts the window at the first code of a column source (uw, ustride and
// nbits set by the caller).
TGA_HD void stage_init_col(TsStage& st, int32_t prev) {
  st.uleft = (st.nbits + 127) >> 7;
  st.pend = st.nbits;
  st.pos = 0;
  st.wend = 0;
  st.wtop = -64;
  st.prev = prev;
}

// Tops the window up: whole units while they fit behind the current dword
// and hold codes of the item.  A 64-bit peek near the item's end may then see
// stale dwords after its code, which no valid code depends on.  Positions
// count from the item's first unit: an item holds at most a checkpoint
// stride of contacts (at most 4096, of at most 96 bits even in a corrupt
// stream), far below the 2^31 bits the 32-bit positions could hold.
//
// Units are loaded two at a time, both loads issued before either is
// written, so a top-up waits for memory once per pair.
TGA_HD void stage_fill(const Streams&, TsStage& st) {
  const uint32_t d = st.pos >> 5;
  uint32_t k = (d + kStageDw - st.wend) >> 2;  // units that fit behind the current dword
  k = minu32(k, st.uleft);
  while (k >= 2) {
    const Unit a = load_unit(st.uw);
    const Unit b = load_unit(st.uw + 2 * (size_t)st.ustride);
    st.uw += 4 * (size_t)st.ustride;
    st.uleft -= 2;
    put_unit(st, a);
    put_unit(st, b);
    k -= 2;
  }
  if (k) stage_unit(st);
}

// The 32-bit funnel shift ({hi, lo} >> (n mod 32)) mod 2^32: one
// v_alignbit_b32 on the device.
TGA_HD uint32_t funnel32(uint32_t hi, uint32_t lo, uint32_t n) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __builtin_amdgcn_alignbit(hi, lo, n);
#else
  return (uint32_t)((((uint64_t)hi << 32) | lo) >> (n & 31u));
#endif
}

// A valid zeta_2 code of the timestamp stream codes an int2nat'ed int32
// difference, below 2^32, so its unary prefix has at most 16 zeros.
constexpr uint32_t kZeta2MaxPrefix = 16;

// Decodes the zeta_2 codeword at the start of the 64-bit window (hi, lo) and
// returns the int2nat inverse of its value modulo 2^32; `len` receives the
// codeword length.  A valid code has a prefix of at most kZeta2MaxPrefix; a
// longer one (a corrupt stream) still decodes to some value and length, and
// the walk detects it by the item's codes not ending where the item ends.
//
// With h the prefix zeros, the code is h zeros, a one, the 2h + 1 bits m of
// the short form and, when m >= 4^h (the bit s after the one is set), one
// more bit b (BVMultiGraph's minimal binary coding, bfs.cu:640-665).  Let B
// be the len - h = 2h + 2 + s bits after the zeros, read as a number: the
// short form has B = 2^(2h+1) + m and the value m + 4^h - 1, the long form
// B = 2^(2h+2) + 2m + b and the value 2m + b - 1.  Both are B - 4^(h+s) - 1.
// The value needs 33 bits (two int32 timestamps can differ by up to
// 2^32 - 1), and its int2nat inverse is taken modulo 2^32, as the
// timestamps are.
TGA_HD int32_t zeta2_delta(uint32_t hi, uint32_t lo, uint32_t& len) {
  const uint32_t h = (uint32_t)__builtin_clz(hi | 1u);  // hi == 0: h = 31, an invalid prefix
  const uint32_t s = ((hi << h) >> 30) & 1u;
  const uint32_t tl = 3 * h + s;
  len = tl + 2;
  const uint64_t p = ((uint64_t)hi << 32) | lo;
  const uint64_t b = p >> ((62u - tl) & 63u);
  const uint64_t v = b - (1ull << ((2 * (h + s)) & 63u)) - 1u;
  return (int32_t)((uint32_t)(v >> 1) ^ (0u - ((uint32_t)v & 1u)));
}

// Next timestamp: zeta_2 of the int2nat delta from the previous one
// (EvolvingMultiGraph.java:131-144; bfs.cu:640-665), decoded from the window.
//
// The window's 64 bits at bit offset pos are two funnel shifts of three
// staged dwords: with t = pos - 1, dwords t/32 .. t/32 + 2 shifted by
// (-pos) mod 32.  When pos is a multiple of 32 the shift is 0 and the first
// dword (one before the current one, possibly already replaced) drops out;
// otherwise t/32 is pos's own dword.  At pos = 0, t/32 maps to slot 15 and
// the next two to the mirror slots of 0 and 1.
TGA_HD int32_t stage_ts_next(const Streams& S, TsStage& st) {
  // (pos >> 5) + 3 > wend: the 96 bits read may pass the staged dwords
  if ((int32_t)st.pos >= st.wtop) stage_fill(S, st);  // codes longer than the burst's share
  const uint32_t t = st.pos - 1u;
  // Slot (t >> 5) mod kStageDw at byte offset slot << lg (lg = log2 of the
  // slot pitch in bytes, >= 5): computed as (t << (lg - 5)) masked to the
  // slot bits, whose low bits (t mod 32) fall below them.
  const uint32_t lg = (uint32_t)__builtin_ctz((uint32_t)st.stride) + 2u;
  const uint32_t off = (t << (lg - 5u)) & ((uint32_t)(kStageDw - 1) << lg);
  const uint32_t* p = reinterpret_cast<const uint32_t*>(reinterpret_cast<const char*>(st.s) + off);
  const uint32_t w0 = p[0], w1 = p[st.stride], w2 = p[2 * st.stride];
  const uint32_t hi = funnel32(w0, w1, ~t), lo = funnel32(w1, w2, ~t);
  uint32_t len;
  const int32_t d = zeta2_delta(hi, lo, len);
  st.pos += len;
  st.prev = (int32_t)((uint32_t)st.prev + (uint32_t)d);
  return st.prev;
}

// Per-lane ring of decoded runs (value, contacts) between the two phases of
// a walk; entry i of a lane is at e[i * stride], stride a power of two (in the kernels: LDS, the
// lanes of a block interleaved so that lanes at any ring positions touch
// distinct banks).
constexpr int32_t kRing = 4;  // runs held (power of two)
// Contacts consumed per refill at most (Streams::burst, set per graph by the
// autotuner): long bursts amortise the refill over long runs of one neighbor;
// a burst ends early when the ring's runs are used up.
constexpr int32_t kBurstDefault = 4, kBurstMax = 64;
struct RunRing {
  int2* e;
  int32_t stride;
};

// Actions may define ring_val(v): the value the walk hands to contact() for
// neighbor v (a per-neighbor key computed once per run of contacts when the
// run is decoded); without it contact() gets the neighbor itself.
template <class Act>
TGA_HD auto act_ring_val(Act& a, int32_t v, int) -> decltype(a.ring_val(v)) {
  return a.ring_val(v);
}
template <class Act>
TGA_HD int32_t act_ring_val(Act&, int32_t v, long) {
  return v;
}

// Walks contacts [k, kend) of one work item and hands (x, r, t, k) to
// act.contact(); a false return stops the walk.
//
// The walk alternates two phases so that every lane of a wave runs the same
// code at the same time: a refill that decodes neighbor runs into the ring
// until it covers the next burst of contacts (the sub-stream merge and its
// branches run once per run, not once per contact), and a consume phase that
// decodes one timestamp per contact and takes the neighbor from the current
// run.  On graphs whose pairs repeat (runs of tens of contacts) the refill
// takes one or two passes per burst.  The consume phase takes the contacts
// the ring holds, at most a burst, so it never waits on the ring; the
// timestamp codes' validity is checked once per walk, by where they end.
template <int MAXD, class Act>
TGA_HD void walk(const Streams& S, Dec<MAXD>& D, TsStage& ts, int32_t x, int32_t k, int32_t kend, Act& act,
                 RunRing ring) {
  // Ring positions in units of ring.stride (a power of two), so an entry's
  // offset is the position masked: tail - head = stride * runs held.
  const uint32_t rstep = (uint32_t)ring.stride, rmask = (uint32_t)(kRing - 1) * rstep;
  uint32_t head = 0, tail = 0;
  int32_t avail = 0;            // contacts held: the current run's and the ring's
  int32_t r = 0, c = 0;         // current run: neighbor (or its ring_val), contacts left in it
  bool more = true;             // the record has runs left
  bool full = true;             // every contact of [k, kend) walked
  while (k < kend) {
    while (more && avail < S.burst && tail - head < (uint32_t)kRing * rstep) {
      int32_t cnt;
      const int32_t v = dec_next_run<MAXD>(S, D, cnt);
      if ((uint32_t)v >= (uint32_t)S.n) {
        if (v != kInf) raise_err(S, kErrNeighbor);
        more = false;
        break;
      }
      ring.e[tail & rmask] = make_int2(act_ring_val(act, v, 0), cnt);
      tail += rstep;
      avail += cnt;
    }
    const int32_t n = min32(min32(kend - k, S.burst), avail);
    if (n <= 0) {
      raise_err(S, kErrExhausted);  // fewer successors than the outdegree
      break;
    }
    stage_fill(S, ts);
    for (int32_t j = 0; j < n; ++j) {
      if (c == 0) {
        const int2 e = ring.e[head & rmask];
        head += rstep;
        r = e.x;
        c = e.y;
      }
      --c;
      const int32_t t = stage_ts_next(S, ts);
      if (!act.contact(x, r, t, k + j)) {
        kend = k;  // stop
        full = false;
        break;
      }
    }
    avail -= n;
    k += n;
  }
  // Valid codes of a whole item end exactly at its end (a stream's last
  // record may carry up to 7 bits of byte padding after its codes).
  if (full && ts.pend != kNoEnd && (ts.pos > ts.pend || ts.pend - ts.pos > 7)) raise_err(S, kErrCode);
}

// Actions that keep per-item state define item(code); it runs first.
template <class Act>
TGA_HD auto act_item(Act& a, int32_t code, int) -> decltype(a.item(code), void()) {
  a.item(code);
}
template <class Act>
TGA_HD void act_item(Act&, int32_t, long) {}

// Runs one work item.  code >= 0: the whole record of vertex `code`;
// code < 0: chunk -code-1 of the checkpoint table.  act.begin(x, chunked,
// split) may veto the item; act.end(x, chunked) runs after the walk.  The
// timestamps come from the linear stream unless ts.uw names the item's
// column copy; a whole record without references is read from its graph
// column copy when ts.gw names one (ts.gbits: the record's bits; the copy
// holds the record's start word, then ceil(gbits / 64) + 5 words of the
// stream's bits from the record's first on, so every read the decoder makes
// stays within it).
template <int MAXD, class Act>
TGA_HD void run_item(const Streams& S0, const Ckpt* ck, int32_t code, Act& act, RunRing ring, TsStage ts) {
  Streams S = S0;
  Dec<MAXD> D;
  int32_t x, k = 0, kend;
  const bool chunked = code < 0;
  act_item(act, code, 0);
  if (!chunked) {
    x = code;
    if (!act.begin(x, false, Split{-1, -1, -1, -1})) return;
    int32_t depth = 0;
    uint64_t gpos = 0, tpos = 0;
    uint64_t sw = kNoStart;
    if (MAXD == 0 && ts.gw) {
      const uint32_t words = (ts.gbits + 63) / 64;
      sw = ts.gw[0];
      S.g = ts.gw + 64;
      S.g_sh = 6;
      S.g_last = words + 4;
      S.g_bits = 64ull * (words + 4);
    } else {
      rec_pos(S, x, gpos, tpos);
    }
    kend = sw != kNoStart ? dec_init_flat<MAXD>(S, D, x, sw) : dec_init<0, MAXD>(S, D, x, gpos, depth);
    if (kend == 0) {
      act.end(x, false);
      return;
    }
    if (ts.uw)
      stage_init_col(ts, S.min_ts);
    else
      stage_init(S, ts, tpos, rec_tend(S, x), S.min_ts, true);
  } else {
    const Ckpt* c = ck + (-code - 1);
    x = c->x;
    if (!act.begin(x, true, Split{c->cont_slot, c->cont_r, c->tail_slot, c->tail_r})) return;
    k = c->k0;
    kend = c->k1;
    restore_levels<0, 0, MAXD>(S, D, c->lv, c->depth);
    if (ts.uw)
      stage_init_col(ts, c->tprev);
    else
      stage_init(S, ts, c->tpos, rec_tend(S, x), c->tprev, false);
  }
  walk<MAXD>(S, D, ts, x, k, kend, act, ring);
  act.end(x, chunked);
}

// Decodes the long record of x once and saves a checkpoint every
// chunk_stride(MAXD, S.flat_stride) contacts into out[0 .. nchunk).  A run of one neighbor
// crossing a boundary gets a split-run slot: `slot0` + the index of the first
// chunk it continues into (see Ckpt).  Returns the number of chunks.
template <int MAXD>
TGA_HD int32_t ckpt_build_record(const Streams& S, int32_t x, int64_t slot0, Ckpt* out) {
  const int32_t kStride = chunk_stride(MAXD, S.flat_stride);
  Dec<MAXD> D;
  TsCursor tc;
  int32_t depth = 0;
  uint64_t gpos, tpos;
  rec_pos(S, x, gpos, tpos);
  const int32_t d = dec_init<0, MAXD>(S, D, x, gpos, depth);
  t_init(S, tc.b, tpos);
  tc.prev = S.min_ts;
  ckpt_save<MAXD>(out, D, depth, tc, x, 0);
  out[0].cont_slot = -1;
  out[0].cont_r = -1;
  int32_t nch = 1, prev = -1, run_slot = -1;
  for (int32_t k = 0; k < d; ++k) {
    if (k > 0 && k % kStride == 0) {
      const int32_t nr = dec_peek<MAXD>(D);
      const bool cont = nr == prev;
      if (cont && run_slot < 0) run_slot = (int32_t)slot0 + nch;
      out[nch - 1].k1 = k;
      out[nch - 1].tail_slot = cont ? run_slot : -1;
      out[nch - 1].tail_r = prev;
      ckpt_save<MAXD>(out + nch, D, depth, tc, x, k);
      out[nch].cont_slot = cont ? run_slot : -1;
      out[nch].cont_r = nr;
      ++nch;
    }
    const int32_t r = dec_next<0, MAXD>(S, D);
    ts_next(S, tc);
    if (r != prev) run_slot = -1;
    prev = r;
  }
  out[nch - 1].k1 = d;
  out[nch - 1].tail_slot = -1;
  out[nch - 1].tail_r = -1;
  for (int32_t j = 0; j < nch; ++j) {
    out[j].nchunk = nch;
    out[j].stride = kStride;
  }
  return nch;
}

}  // namespace tga
