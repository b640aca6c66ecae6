// MI355X engine: loader, decode directories, algorithm kernels and the C ABI
// of include/tga.h.
//
// Every algorithm is one or more passes of `visit_vertex` (decode.hpp) over a
// set of source vertices, with the per-contact rule of the algorithm plugged
// in as a small functor.  Frontier algorithms (BFS, CC, GC, earliest arrival)
// keep their frontier as a device queue; the host reads one 4-byte count per
// level to size the next launch (the reference copies the queue size back per
// level too, bfs.cu:2028).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <string>
#include <vector>

#include "../../../include/tga.h"
#include "decode.hpp"

using namespace tga;

namespace {

thread_local std::string g_err;

int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}

#define HIP_TRY(call)                                                                 \
  do {                                                                                \
    hipError_t e_ = (call);                                                           \
    if (e_ != hipSuccess)                                                             \
      return fail(TGA_EDEVICE, std::string(#call " failed: ") + hipGetErrorString(e_)); \
  } while (0)

constexpr int kBlock = 256;

struct Window {
  int64_t ws, we;
  int32_t ab, ae;  // active range [ab, ae)
  __device__ __forceinline__ bool in(int32_t t) const { return (int64_t)t >= ws && (int64_t)t <= we; }
  __device__ __forceinline__ bool active(int32_t v) const { return v >= ab && v < ae; }
};

__device__ __forceinline__ uint64_t fmix64(uint64_t k) {
  k ^= k >> 33;
  k *= 0xff51afd7ed558ccdULL;
  k ^= k >> 33;
  k *= 0xc4ceb9fe1a85ec53ULL;
  k ^= k >> 33;
  return k;
}

__device__ __forceinline__ uint64_t contact_hash(int32_t x, int32_t r, int32_t t) {
  return fmix64((uint64_t)(uint32_t)x * 0x9E3779B97F4A7C15ULL ^ (uint64_t)(uint32_t)r * 0xC2B2AE3D27D4EB4FULL ^
                (uint64_t)(uint32_t)t * 0x165667B19E3779F9ULL);
}

template <class T>
__device__ __forceinline__ T wave_sum(T v) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// Index i of a vertex set that is either an explicit list or a range.
__device__ __forceinline__ int32_t item(const int32_t* list, int32_t base, int64_t i) {
  return list ? list[i] : base + (int32_t)i;
}

// Appends v to a device queue.
__device__ __forceinline__ void push(int32_t* q, uint32_t* qn, int32_t v) { q[atomicAdd(qn, 1u)] = v; }

// ---------------------------------------------------------------- kernels ---

// Outdegree of every vertex from its record header; validates the index.
__global__ void __launch_bounds__(kBlock) k_outdeg(Streams S, unsigned long long* total, unsigned long long* maxd) {
  const Rd g{S.g, S.g_bits};
  unsigned long long sum = 0, mx = 0;
  for (int64_t x = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; x < S.n; x += (int64_t)gridDim.x * blockDim.x) {
    uint64_t pos = ef_get(S, S.gi, (uint64_t)x);
    const uint64_t d = rd_gamma(S, g, pos);
    ef_get(S, S.ti, (uint64_t)x);
    sum += d;
    mx = d > mx ? d : mx;
  }
  sum = wave_sum(sum);
  for (int off = 32; off > 0; off >>= 1) {
    const unsigned long long o = __shfl_xor(mx, off, 64);
    mx = o > mx ? o : mx;
  }
  if ((threadIdx.x & 63) == 0) {
    atomicAdd(total, sum);
    atomicMax(maxd, mx);
  }
}

struct RaAct {
  uint64_t cs = 0;
  uint64_t cnt = 0;
  __device__ bool begin(int32_t, int32_t d) {
    cnt += (uint64_t)d;
    return true;
  }
  __device__ bool contact(int32_t x, int32_t r, int32_t t) {
    cs += contact_hash(x, r, t);
    return true;
  }
};

__global__ void __launch_bounds__(kBlock) k_ra(Streams S, const int32_t* list, int32_t base, int64_t count,
                                               unsigned long long* out) {
  RaAct a;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < count; i += (int64_t)gridDim.x * blockDim.x) {
    const int32_t x = item(list, base, i);
    if ((uint32_t)x >= (uint32_t)S.n) {
      raise_err(S, kErrNeighbor);
      continue;
    }
    visit_vertex(S, x, a);
  }
  const unsigned long long cs = wave_sum((unsigned long long)a.cs);
  const unsigned long long cnt = wave_sum((unsigned long long)a.cnt);
  if ((threadIdx.x & 63) == 0) {
    atomicAdd(&out[0], cs);
    atomicAdd(&out[1], cnt);
  }
}

struct DcAct {
  Window W;
  int32_t deg = 0;
  int64_t avg = 0;
  __device__ bool begin(int32_t, int32_t) { return true; }
  __device__ bool contact(int32_t, int32_t r, int32_t t) {
    if (W.in(t) && W.active(r)) {
      ++deg;
      avg += ((int64_t)t - avg) / deg;  // truncating, as Java's long division
    }
    return true;
  }
};

#pragma clang fp contract(off)
__device__ int32_t dc_formula(int32_t deg, int64_t avg, int64_t ws, int64_t we) {
  if (ws == we) return deg;
  const double c = 0.85 * (double)deg + 0.15 * (double)deg * (double)(avg - ws) / (double)(we - ws) + 1.0;
  return (int32_t)c;
}
#pragma clang fp contract(on)

__global__ void __launch_bounds__(kBlock) k_dc(Streams S, Window W, int32_t* count, int32_t* centr) {
  for (int64_t x = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; x < S.n; x += (int64_t)gridDim.x * blockDim.x) {
    DcAct a{W};
    if (W.active((int32_t)x)) visit_vertex(S, (int32_t)x, a);
    count[x] = a.deg;
    if (centr) centr[x] = dc_formula(a.deg, a.avg, W.ws, W.we);
  }
}

struct BfsAct {
  Window W;
  int32_t* level;
  int32_t* next;
  uint32_t* nsize;
  int32_t lvl;
  __device__ bool begin(int32_t, int32_t) { return true; }
  __device__ bool contact(int32_t, int32_t r, int32_t t) {
    if (W.in(t) && W.active(r) && level[r] == kInf && atomicCAS(&level[r], kInf, lvl + 1) == kInf)
      push(next, nsize, r);
    return true;
  }
};

__global__ void __launch_bounds__(kBlock) k_bfs(Streams S, Window W, const int32_t* front, int64_t fsize,
                                                int32_t* level, int32_t* next, uint32_t* nsize, int32_t lvl) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < fsize; i += (int64_t)gridDim.x * blockDim.x) {
    const int32_t x = front[i];
    if (!W.active(x)) continue;
    BfsAct a{W, level, next, nsize, lvl};
    visit_vertex(S, x, a);
  }
}

// Min-label propagation: a vertex whose label drops is queued once per round.
struct CcAct {
  Window W;
  int32_t lx;
  int32_t* label;
  int32_t* stamp;
  int32_t round;
  int32_t* next;
  uint32_t* nsize;
  __device__ bool begin(int32_t, int32_t) { return true; }
  __device__ bool contact(int32_t, int32_t r, int32_t t) {
    if (W.in(t) && W.active(r) && lx < label[r]) {
      const int32_t old = atomicMin(&label[r], lx);
      if (lx < old && atomicMax(&stamp[r], round) < round) push(next, nsize, r);
    }
    return true;
  }
};

__global__ void __launch_bounds__(kBlock) k_cc(Streams S, Window W, const int32_t* front, int32_t base, int64_t fsize,
                                               int32_t* label, int32_t* stamp, int32_t round, int32_t* next,
                                               uint32_t* nsize) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < fsize; i += (int64_t)gridDim.x * blockDim.x) {
    const int32_t x = item(front, base, i);
    if (!W.active(x)) continue;
    CcAct a{W, label[x], label, stamp, round, next, nsize};
    visit_vertex(S, x, a);
  }
}

// Graph coloring round: a neighbor blocks when it has higher priority and is
// uncolored at the start of the round (its color is kInf, or `round` if it was
// colored earlier in this same round).
struct GcAct {
  Window W;
  const int32_t* prio;
  const int32_t* color;
  int32_t v, pv, round;
  bool blocked = false;
  __device__ bool begin(int32_t, int32_t) { return true; }
  __device__ bool contact(int32_t, int32_t u, int32_t t) {
    if (W.in(t) && W.active(u)) {
      const int32_t pu = prio[u];
      if ((pu > pv || (pu == pv && u < v)) && color[u] >= round) {
        blocked = true;
        return false;
      }
    }
    return true;
  }
};

__global__ void __launch_bounds__(kBlock) k_gc(Streams S, Window W, const int32_t* work, int32_t base, int64_t wsize,
                                               const int32_t* prio, int32_t* color, int32_t round, int32_t* next,
                                               uint32_t* nsize) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < wsize; i += (int64_t)gridDim.x * blockDim.x) {
    const int32_t v = item(work, base, i);
    if (!W.active(v) || color[v] != kInf) continue;
    GcAct a{W, prio, color, v, prio[v], round};
    visit_vertex(S, v, a);
    if (!a.blocked)
      color[v] = round;
    else
      push(next, nsize, v);
  }
}

// Earliest arrival (bfs.cu:1452-1471 rule): first window contact per pair.
struct EaAct {
  Window W;
  int32_t ax;
  int32_t* arr;
  int32_t* stamp;
  int32_t round;
  int32_t* next;
  uint32_t* nsize;
  int32_t last = -1;
  __device__ bool begin(int32_t, int32_t) { return true; }
  __device__ bool contact(int32_t, int32_t r, int32_t t) {
    if (!W.in(t) || r == last) return true;
    last = r;
    if (W.active(r) && ax <= t && t < arr[r]) {
      const int32_t old = atomicMin(&arr[r], t);
      if (t < old && atomicMax(&stamp[r], round) < round) push(next, nsize, r);
    }
    return true;
  }
};

__global__ void __launch_bounds__(kBlock) k_ea(Streams S, Window W, const int32_t* front, int64_t fsize, int32_t* arr,
                                               int32_t* stamp, int32_t round, int32_t* next, uint32_t* nsize) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < fsize; i += (int64_t)gridDim.x * blockDim.x) {
    const int32_t x = front[i];
    if (!W.active(x)) continue;
    EaAct a{W, arr[x], arr, stamp, round, next, nsize};
    visit_vertex(S, x, a);
  }
}

// Window contact count of every active vertex (PageRank divisor).
struct WdegAct {
  Window W;
  int32_t c = 0;
  __device__ bool begin(int32_t, int32_t) { return true; }
  __device__ bool contact(int32_t, int32_t, int32_t t) {
    c += W.in(t) ? 1 : 0;
    return true;
  }
};

__global__ void __launch_bounds__(kBlock) k_wdeg(Streams S, Window W, int32_t* du, int full_window) {
  const Rd g{S.g, S.g_bits};
  for (int64_t x = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; x < S.n; x += (int64_t)gridDim.x * blockDim.x) {
    int32_t c = 0;
    if (W.active((int32_t)x)) {
      if (full_window) {
        uint64_t pos = ef_get(S, S.gi, (uint64_t)x);
        c = (int32_t)rd_gamma(S, g, pos);
      } else {
        WdegAct a{W};
        visit_vertex(S, (int32_t)x, a);
        c = a.c;
      }
    }
    du[x] = c;
  }
}

__global__ void __launch_bounds__(kBlock) k_fill_active(Window W, int32_t n, double* acc, double inside) {
  for (int64_t v = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; v < n; v += (int64_t)gridDim.x * blockDim.x)
    acc[v] = W.active((int32_t)v) ? inside : 0.0;
}

template <class T>
__global__ void __launch_bounds__(kBlock) k_fill(T* a, int64_t n, T val) {
  for (int64_t v = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; v < n; v += (int64_t)gridDim.x * blockDim.x)
    a[v] = val;
}

__global__ void __launch_bounds__(kBlock) k_fill_active_f32(Window W, int32_t n, float* a, float inside) {
  for (int64_t v = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; v < n; v += (int64_t)gridDim.x * blockDim.x)
    a[v] = W.active((int32_t)v) ? inside : 0.0f;
}

struct PrAct {
  Window W;
  float coef;
  double* acc;
  int32_t last = -1;
  __device__ bool begin(int32_t, int32_t) { return true; }
  __device__ bool contact(int32_t, int32_t r, int32_t t) {
    if (!W.in(t) || r == last) return true;
    last = r;
    if (W.active(r)) {
      const float w = 1.0f - expf(-0.1f * (float)t);
      atomicAdd(&acc[r], (double)(coef * w));
    }
    return true;
  }
};

__global__ void __launch_bounds__(kBlock) k_pr_push(Streams S, Window W, const float* rank, const int32_t* du,
                                                    float damping, double* acc) {
  for (int64_t u = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; u < S.n; u += (int64_t)gridDim.x * blockDim.x) {
    if (!W.active((int32_t)u) || du[u] == 0) continue;
    PrAct a{W, damping * rank[u] / (float)du[u], acc};
    visit_vertex(S, (int32_t)u, a);
  }
}

struct HubAct {
  Window W;
  const float* auth;
  double acc = 0;
  int32_t last = -1;
  __device__ bool begin(int32_t, int32_t) { return true; }
  __device__ bool contact(int32_t, int32_t r, int32_t t) {
    if (!W.in(t) || r == last) return true;
    last = r;
    if (W.active(r)) acc += (double)auth[r];
    return true;
  }
};

__global__ void __launch_bounds__(kBlock) k_hits_pull(Streams S, Window W, const float* auth, double* acc) {
  for (int64_t u = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; u < S.n; u += (int64_t)gridDim.x * blockDim.x) {
    HubAct a{W, auth};
    if (W.active((int32_t)u)) visit_vertex(S, (int32_t)u, a);
    acc[u] = a.acc;
  }
}

struct AuthAct {
  Window W;
  double h;
  double* acc;
  int32_t last = -1;
  __device__ bool begin(int32_t, int32_t) { return true; }
  __device__ bool contact(int32_t, int32_t r, int32_t t) {
    if (!W.in(t) || r == last) return true;
    last = r;
    if (W.active(r)) atomicAdd(&acc[r], h);
    return true;
  }
};

__global__ void __launch_bounds__(kBlock) k_hits_push(Streams S, Window W, const float* hub, double* acc) {
  for (int64_t u = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; u < S.n; u += (int64_t)gridDim.x * blockDim.x) {
    if (!W.active((int32_t)u)) continue;
    AuthAct a{W, (double)hub[u], acc};
    visit_vertex(S, (int32_t)u, a);
  }
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

// ----------------------------------------------------------------- loader ---

bool read_file(const std::string& path, std::vector<uint8_t>& out) {
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
// plus two zero words of padding.
std::vector<uint64_t> be_words(const std::vector<uint8_t>& b) {
  std::vector<uint64_t> w((b.size() + 7) / 8 + 2, 0);
  for (size_t i = 0; i < b.size(); ++i) w[i >> 3] |= (uint64_t)b[i] << (56 - 8 * (i & 7));
  return w;
}

std::vector<uint64_t> le_words(const std::vector<uint8_t>& b) {
  std::vector<uint64_t> w((b.size() + 7) / 8 + 2, 0);
  std::memcpy(w.data(), b.data(), b.size());
  return w;
}

int msb64_host(uint64_t v) { return 63 - __builtin_clzll(v); }

// Sampled select directory over the upper bits (see decode.hpp).
void build_select(const std::vector<uint64_t>& up, size_t words, uint64_t& ones, std::vector<uint64_t>& samp,
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

}  // namespace

struct tga_graph {
  int device = 0;
  hipStream_t stream = nullptr;
  int64_t n = 0;
  int32_t ab = 0, ae = 0;
  int32_t min_ts = 0;
  tga_graph_info info{};
  Streams S{};
  std::vector<void*> allocs;
  int32_t* d_err = nullptr;
  // scratch (n-sized unless stated)
  int32_t *q0 = nullptr, *q1 = nullptr, *i0 = nullptr, *i1 = nullptr;
  uint32_t* cnt = nullptr;  // 4 counters
  unsigned long long* u64s = nullptr;  // 4 values
  double* acc = nullptr;
  double* partial = nullptr;  // kRedBlocks
  double* scal = nullptr;     // 4 scalars
  float *f0 = nullptr, *f1 = nullptr;
  int32_t* prio = nullptr;  // default GC priorities, built on first use
  int64_t device_bytes = 0;
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

  template <class T>
  int alloc(T*& p, size_t count) {
    void* q = nullptr;
    hipError_t e = hipMalloc(&q, std::max<size_t>(count, 1) * sizeof(T));
    if (e != hipSuccess) return fail(TGA_ENOMEM, std::string("hipMalloc: ") + hipGetErrorString(e));
    allocs.push_back(q);
    device_bytes += (int64_t)(std::max<size_t>(count, 1) * sizeof(T));
    p = (T*)q;
    return TGA_OK;
  }
  template <class T>
  int upload(const T*& dst, const std::vector<T>& src) {
    T* p = nullptr;
    int rc = alloc(p, src.size());
    if (rc) return rc;
    HIP_TRY(hipMemcpy(p, src.data(), src.size() * sizeof(T), hipMemcpyHostToDevice));
    dst = p;
    return TGA_OK;
  }
  int grid(int64_t items) const {
    const int64_t b = (items + kBlock - 1) / kBlock;
    return (int)std::max<int64_t>(1, std::min<int64_t>(b, 1 << 16));
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
  int read_count(uint32_t* dev, uint32_t& host) {
    HIP_TRY(hipMemcpyAsync(&host, dev, sizeof(uint32_t), hipMemcpyDeviceToHost, stream));
    HIP_TRY(hipStreamSynchronize(stream));
    return TGA_OK;
  }
  ~tga_graph() {
    for (hipEvent_t e : ev) (void)hipEventDestroy(e);
    for (void* p : allocs) (void)hipFree(p);
    if (stream) (void)hipStreamDestroy(stream);
  }
};

extern "C" const char* tga_last_error(void) { return g_err.c_str(); }

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

  const std::vector<uint64_t> gw = be_words(graph), tw = be_words(ts), glw = be_words(glo), tlw = be_words(tlo);
  const std::vector<uint64_t> guw = le_words(gup), tuw = le_words(tup);
  uint64_t gones = 0, tones = 0;
  std::vector<uint64_t> gsamp, gspill, tsamp, tspill;
  build_select(guw, (gup.size() + 7) / 8, gones, gsamp, gspill);
  build_select(tuw, (tup.size() + 7) / 8, tones, tsamp, tspill);
  if (gones < (uint64_t)n || tones < (uint64_t)n)
    return fail(TGA_EFORMAT, "GPU index holds fewer entries than MAX_LABEL + 1");
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
  if (hipStreamCreate(&g->stream) != hipSuccess) return bail(fail(TGA_EDEVICE, "hipStreamCreate failed"));
  g->n = n;
  g->ab = (int32_t)std::max<int64_t>(0, desc ? desc->active_begin : 0);
  g->ae = (int32_t)((desc && desc->active_end > 0) ? std::min<int64_t>(desc->active_end, n) : n);
  g->min_ts = (int32_t)min_ts;

  Streams& S = g->S;
  if ((rc = g->upload(S.g, gw)) || (rc = g->upload(S.t, tw)) || (rc = g->upload(S.gi.lower, glw)) ||
      (rc = g->upload(S.gi.upper, guw)) || (rc = g->upload(S.gi.samp, gsamp)) || (rc = g->upload(S.gi.spill, gspill)) ||
      (rc = g->upload(S.ti.lower, tlw)) || (rc = g->upload(S.ti.upper, tuw)) || (rc = g->upload(S.ti.samp, tsamp)) ||
      (rc = g->upload(S.ti.spill, tspill)))
    return bail(rc);
  S.g_bits = (uint64_t)graph.size() * 8;
  S.t_bits = (uint64_t)ts.size() * 8;
  S.gi.upper_words = (gup.size() + 7) / 8;
  S.ti.upper_words = (tup.size() + 7) / 8;
  S.gi.l = lg;
  S.ti.l = lt;
  S.n = (int32_t)n;
  S.min_ts = (int32_t)min_ts;
  if ((rc = g->alloc(g->d_err, 1))) return bail(rc);
  S.err = g->d_err;
  if (hipMemset(g->d_err, 0, sizeof(int32_t)) != hipSuccess) return bail(fail(TGA_EDEVICE, "hipMemset failed"));

  if ((rc = g->alloc(g->q0, (size_t)n)) || (rc = g->alloc(g->q1, (size_t)n)) || (rc = g->alloc(g->i0, (size_t)n)) ||
      (rc = g->alloc(g->i1, (size_t)n)) || (rc = g->alloc(g->cnt, 4)) || (rc = g->alloc(g->u64s, 4)) ||
      (rc = g->alloc(g->acc, (size_t)n)) || (rc = g->alloc(g->partial, kRedBlocks)) || (rc = g->alloc(g->scal, 4)) ||
      (rc = g->alloc(g->f0, (size_t)n)) || (rc = g->alloc(g->f1, (size_t)n)))
    return bail(rc);

  // Header pass: contacts and maximum outdegree; validates the index.
  if (hipMemsetAsync(g->u64s, 0, 4 * sizeof(unsigned long long), g->stream) != hipSuccess)
    return bail(fail(TGA_EDEVICE, "hipMemsetAsync failed"));
  k_outdeg<<<g->grid(n), kBlock, 0, g->stream>>>(S, g->u64s, g->u64s + 1);
  unsigned long long hv[2] = {0, 0};
  if (hipMemcpyAsync(hv, g->u64s, sizeof hv, hipMemcpyDeviceToHost, g->stream) != hipSuccess ||
      hipStreamSynchronize(g->stream) != hipSuccess)
    return bail(fail(TGA_EDEVICE, "header pass failed"));
  if ((rc = g->check_err())) return bail(rc);

  tga_graph_info& I = g->info;
  I.n = n;
  I.contacts = (int64_t)hv[0];
  I.max_degree = (int64_t)hv[1];
  I.min_timestamp = min_ts;
  I.graph_bytes = graph.size();
  I.timestamps_bytes = ts.size();
  I.index_bytes = glo.size() + gup.size() + tlo.size() + tup.size();
  I.l_graph = lg;
  I.l_timestamps = lt;
  I.active_begin = g->ab;
  I.active_end = g->ae;
  I.chunks = n;
  I.device_bytes = g->device_bytes;
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
  return TGA_OK;
}

static bool bad_window(tga_window w) { return w.start > w.end; }

extern "C" int tga_neighbor_access(tga_graph* g, const int32_t* verts, int64_t nverts, uint64_t* checksum,
                                   int64_t* contacts) {
  if (!g || (nverts > 0 && !verts) || nverts < 0) return fail(TGA_EINVAL, "tga_neighbor_access: bad arguments");
  HIP_TRY(hipSetDevice(g->device));
  const int32_t* list = nullptr;
  int64_t count = g->ae - g->ab;
  if (nverts > 0) {
    if (nverts > g->n) return fail(TGA_EINVAL, "tga_neighbor_access: more vertices than the graph holds");
    HIP_TRY(hipMemcpyAsync(g->q0, verts, (size_t)nverts * sizeof(int32_t), hipMemcpyHostToDevice, g->stream));
    list = g->q0;
    count = nverts;
  }
  HIP_TRY(hipMemsetAsync(g->u64s, 0, 2 * sizeof(unsigned long long), g->stream));
  g->kbegin();
  k_ra<<<g->grid(count), kBlock, 0, g->stream>>>(g->S, list, g->ab, count, g->u64s);
  g->kend();
  HIP_TRY(hipGetLastError());
  unsigned long long hv[2];
  HIP_TRY(hipMemcpyAsync(hv, g->u64s, sizeof hv, hipMemcpyDeviceToHost, g->stream));
  int rc = g->check_err();
  if (rc) return rc;
  if (checksum) *checksum = hv[0];
  if (contacts) *contacts = (int64_t)hv[1];
  return TGA_OK;
}

extern "C" int tga_degree_centrality(tga_graph* g, tga_window w, int32_t* count, int32_t* centrality, uint32_t flags) {
  if (!g || !count || bad_window(w)) return fail(TGA_EINVAL, "tga_degree_centrality: bad arguments");
  HIP_TRY(hipSetDevice(g->device));
  g->kbegin();
  k_dc<<<g->grid(g->n), kBlock, 0, g->stream>>>(g->S, g->window(w), g->i0, centrality ? g->i1 : nullptr);
  g->kend();
  HIP_TRY(hipGetLastError());
  int rc = g->out_copy(count, g->i0, (size_t)g->n * 4, flags);
  if (!rc && centrality) rc = g->out_copy(centrality, g->i1, (size_t)g->n * 4, flags);
  if (rc) return rc;
  return g->check_err();
}

extern "C" int tga_bfs(tga_graph* g, tga_window w, int32_t src, int32_t* level, uint32_t flags) {
  if (!g || !level || bad_window(w) || src < 0 || src >= g->n) return fail(TGA_EINVAL, "tga_bfs: bad arguments");
  HIP_TRY(hipSetDevice(g->device));
  const Window W = g->window(w);
  k_fill<int32_t><<<g->grid(g->n), kBlock, 0, g->stream>>>(g->i0, g->n, kInf);
  const int32_t zero = 0;
  HIP_TRY(hipMemcpyAsync(g->i0 + src, &zero, 4, hipMemcpyHostToDevice, g->stream));
  HIP_TRY(hipMemcpyAsync(g->q0, &src, 4, hipMemcpyHostToDevice, g->stream));
  int32_t *cur = g->q0, *nxt = g->q1;
  uint32_t fsize = 1;
  for (int32_t lvl = 0; fsize > 0; ++lvl) {
    HIP_TRY(hipMemsetAsync(g->cnt, 0, sizeof(uint32_t), g->stream));
    g->kbegin();
    k_bfs<<<g->grid(fsize), kBlock, 0, g->stream>>>(g->S, W, cur, fsize, g->i0, nxt, g->cnt, lvl);
    g->kend();
    HIP_TRY(hipGetLastError());
    int rc = g->read_count(g->cnt, fsize);
    if (rc) return rc;
    std::swap(cur, nxt);
  }
  int rc = g->out_copy(level, g->i0, (size_t)g->n * 4, flags);
  if (rc) return rc;
  return g->check_err();
}

extern "C" int tga_earliest_arrival(tga_graph* g, tga_window w, int32_t src, int32_t* arrival, uint32_t flags) {
  if (!g || !arrival || bad_window(w) || src < 0 || src >= g->n)
    return fail(TGA_EINVAL, "tga_earliest_arrival: bad arguments");
  HIP_TRY(hipSetDevice(g->device));
  const Window W = g->window(w);
  k_fill<int32_t><<<g->grid(g->n), kBlock, 0, g->stream>>>(g->i0, g->n, kInf);
  k_fill<int32_t><<<g->grid(g->n), kBlock, 0, g->stream>>>(g->i1, g->n, -1);
  const int32_t a0 = (int32_t)std::max<int64_t>(0, std::min<int64_t>(w.start, INT32_MAX));
  HIP_TRY(hipMemcpyAsync(g->i0 + src, &a0, 4, hipMemcpyHostToDevice, g->stream));
  HIP_TRY(hipMemcpyAsync(g->q0, &src, 4, hipMemcpyHostToDevice, g->stream));
  int32_t *cur = g->q0, *nxt = g->q1;
  uint32_t fsize = 1;
  for (int32_t round = 0; fsize > 0; ++round) {
    HIP_TRY(hipMemsetAsync(g->cnt, 0, sizeof(uint32_t), g->stream));
    g->kbegin();
    k_ea<<<g->grid(fsize), kBlock, 0, g->stream>>>(g->S, W, cur, fsize, g->i0, g->i1, round, nxt, g->cnt);
    g->kend();
    HIP_TRY(hipGetLastError());
    int rc = g->read_count(g->cnt, fsize);
    if (rc) return rc;
    std::swap(cur, nxt);
  }
  int rc = g->out_copy(arrival, g->i0, (size_t)g->n * 4, flags);
  if (rc) return rc;
  return g->check_err();
}

__global__ void k_init_labels(Window W, int32_t n, int32_t* label, int32_t* stamp) {
  for (int64_t v = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; v < n; v += (int64_t)gridDim.x * blockDim.x) {
    label[v] = W.active((int32_t)v) ? (int32_t)v : -1;
    stamp[v] = -1;
  }
}

extern "C" int tga_connected_components(tga_graph* g, tga_window w, int32_t* label, uint32_t flags) {
  if (!g || !label || bad_window(w)) return fail(TGA_EINVAL, "tga_connected_components: bad arguments");
  HIP_TRY(hipSetDevice(g->device));
  const Window W = g->window(w);
  k_init_labels<<<g->grid(g->n), kBlock, 0, g->stream>>>(W, (int32_t)g->n, g->i0, g->i1);
  int32_t *cur = nullptr, *nxt = g->q0;
  uint32_t fsize = (uint32_t)(g->ae - g->ab);
  for (int32_t round = 0; fsize > 0; ++round) {
    HIP_TRY(hipMemsetAsync(g->cnt, 0, sizeof(uint32_t), g->stream));
    g->kbegin();
    k_cc<<<g->grid(fsize), kBlock, 0, g->stream>>>(g->S, W, cur, g->ab, fsize, g->i0, g->i1, round, nxt, g->cnt);
    g->kend();
    HIP_TRY(hipGetLastError());
    int rc = g->read_count(g->cnt, fsize);
    if (rc) return rc;
    cur = nxt;
    nxt = (nxt == g->q0) ? g->q1 : g->q0;
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
      int rc = g->alloc(g->prio, (size_t)g->n);
      if (rc) return rc;
      HIP_TRY(hipMemcpy(g->prio, h.data(), h.size() * 4, hipMemcpyHostToDevice));
    }
    prio = g->prio;
  }
  k_fill<int32_t><<<g->grid(g->n), kBlock, 0, g->stream>>>(g->i0, g->n, kInf);
  int32_t *cur = nullptr, *nxt = g->q0;
  uint32_t wsize = (uint32_t)(g->ae - g->ab);
  for (int32_t round = 0; round < rounds && wsize > 0; ++round) {
    HIP_TRY(hipMemsetAsync(g->cnt, 0, sizeof(uint32_t), g->stream));
    g->kbegin();
    k_gc<<<g->grid(wsize), kBlock, 0, g->stream>>>(g->S, W, cur, g->ab, wsize, prio, g->i0, round, nxt, g->cnt);
    g->kend();
    HIP_TRY(hipGetLastError());
    int rc = g->read_count(g->cnt, wsize);
    if (rc) return rc;
    cur = nxt;
    nxt = (nxt == g->q0) ? g->q1 : g->q0;
  }
  int rc = g->out_copy(color, g->i0, (size_t)g->n * 4, flags);
  if (rc) return rc;
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

extern "C" int tga_pagerank(tga_graph* g, tga_window w, int32_t iters, float damping, float tol, float* rank,
                            int32_t* iters_run, uint32_t flags) {
  if (!g || !rank || bad_window(w) || iters < 0) return fail(TGA_EINVAL, "tga_pagerank: bad arguments");
  HIP_TRY(hipSetDevice(g->device));
  const Window W = g->window(w);
  const int64_t N = g->ae - g->ab;
  if (N <= 0) return fail(TGA_EINVAL, "tga_pagerank: no active vertices");
  const bool full = w.start <= TGA_FULL_WINDOW_START && w.end >= TGA_FULL_WINDOW_END;
  g->kbegin();
  k_wdeg<<<g->grid(g->n), kBlock, 0, g->stream>>>(g->S, W, g->i0, full ? 1 : 0);
  g->kend();
  const float inv = 1.0f / (float)N;
  const float base = (1.0f - damping) / (float)N;
  k_fill_active_f32<<<g->grid(g->n), kBlock, 0, g->stream>>>(W, (int32_t)g->n, g->f0, inv);
  int32_t it = 0;
  for (; it < iters; ++it) {
    k_fill_active<<<g->grid(g->n), kBlock, 0, g->stream>>>(W, (int32_t)g->n, g->acc, (double)base);
    g->kbegin();
    k_pr_push<<<g->grid(g->n), kBlock, 0, g->stream>>>(g->S, W, g->f0, g->i0, damping, g->acc);
    g->kend();
    int rc = reduce(g, 0, g->acc, nullptr, nullptr, g->scal, false);
    if (!rc && tol > 0) rc = reduce(g, 2, g->acc, g->f0, g->scal, g->scal + 1, false);
    if (rc) return rc;
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
  k_fill_active_f32<<<g->grid(g->n), kBlock, 0, g->stream>>>(W, (int32_t)g->n, g->f0, 1.0f);  // hub
  k_fill_active_f32<<<g->grid(g->n), kBlock, 0, g->stream>>>(W, (int32_t)g->n, g->f1, 1.0f);  // auth
  for (int32_t it = 0; it < iters; ++it) {
    g->kbegin();
    k_hits_pull<<<g->grid(g->n), kBlock, 0, g->stream>>>(g->S, W, g->f1, g->acc);
    g->kend();
    int rc = reduce(g, 1, g->acc, nullptr, nullptr, g->scal, true);
    if (rc) return rc;
    k_scale_store<<<g->grid(g->n), kBlock, 0, g->stream>>>(g->acc, g->scal, g->n, g->f0);
    k_fill<double><<<g->grid(g->n), kBlock, 0, g->stream>>>(g->acc, g->n, 0.0);
    g->kbegin();
    k_hits_push<<<g->grid(g->n), kBlock, 0, g->stream>>>(g->S, W, g->f0, g->acc);
    g->kend();
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
