// GPU writers of the per-contact and per-vertex ChronoGraph streams
// (include/tga_encode.h): the timestamp stream, the record-offset stream and
// the GPU Elias-Fano index files.
//
// Every code's length depends on its own value only, so a stream is written
// in three steps: the code lengths, an exclusive scan of them into bit
// positions (rocPRIM), and one thread per code OR-ing its bits into the
// 64-bit words its position falls in (a code of at most 64 bits touches one
// or two words; neighbours share the boundary words, hence the atomics).
// Words hold the stream MSB-first, as the decoders read it; a final pass
// swaps each word's bytes into file order.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cstdint>
#include <string>
#include <rocprim/device/device_scan.hpp>
#include <rocprim/iterator/counting_iterator.hpp>
#include <rocprim/iterator/transform_iterator.hpp>

#include "../../../include/tga.h"
#include "../../../include/tga_encode.h"

namespace tga {
int set_error(int code, const std::string& msg);  // engine.hip: tga_last_error's text
}  // namespace tga

struct tga_stream {
  int device = 0;
  uint64_t bytes = 0;
  uint8_t* data = nullptr;  // device bytes
};

namespace {

constexpr int kBlock = 256;

int err(int code, const std::string& msg) { return tga::set_error(code, msg); }

#define ENC_TRY(call)                                                                  \
  do {                                                                                 \
    hipError_t e_ = (call);                                                            \
    if (e_ != hipSuccess) return err(TGA_EDEVICE, std::string(#call " failed: ") + hipGetErrorString(e_)); \
  } while (0)

int grid_for(int64_t n) { return (int)std::max<int64_t>(1, std::min<int64_t>((n + kBlock - 1) / kBlock, 1 << 16)); }

__host__ __device__ inline uint64_t int2nat(int64_t x) {
  return x >= 0 ? (uint64_t)x << 1 : ((uint64_t)(-(x + 1)) << 1) | 1u;
}
__host__ __device__ inline uint32_t msb64(uint64_t v) { return 63u - (uint32_t)__builtin_clzll(v); }  // v > 0

// zeta_2 of value v as (code, len): h = floor(msb(v + 1) / 2) zeros, a one,
// then the minimal binary code of v + 1 - 4^h over [0, 3 * 4^h): the low
// 2h + 1 bits of v + 1 - 4^h when that is below 4^h, else v + 1 in 2h + 2
// bits (dsiutils' writeZeta, as the host BitWriter::zeta).
__host__ __device__ inline void zeta2(uint64_t v, uint64_t& code, uint32_t& len) {
  const uint64_t x = v + 1;
  const uint32_t h = msb64(x) / 2;
  const uint64_t left = 1ull << (2 * h);
  const uint32_t plen = (x - left < left) ? 2 * h + 1 : 2 * h + 2;
  const uint64_t payload = (x - left < left) ? x - left : x;
  len = h + 1 + plen;
  code = (1ull << plen) | payload;
}

// gamma of value v: m = msb(v + 1) zeros, then v + 1 in m + 1 bits.
__host__ __device__ inline void gamma(uint64_t v, uint64_t& code, uint32_t& len) {
  code = v + 1;
  len = 2 * msb64(code) + 1;
}

// ORs the low `len` bits of code (len <= 64) into the MSB-first word array at
// bit position pos.
__device__ inline void put_bits(unsigned long long* w, uint64_t pos, uint64_t code, uint32_t len) {
  if (len == 0) return;
  const uint64_t i = pos >> 6;
  const uint32_t off = (uint32_t)(pos & 63);
  if (off + len <= 64) {
    atomicOr(&w[i], (unsigned long long)(code << (64 - off - len)));
  } else {
    const uint32_t spill = off + len - 64;  // bits going to word i + 1 (1..63)
    atomicOr(&w[i], (unsigned long long)(code >> spill));
    atomicOr(&w[i + 1], (unsigned long long)(code << (64 - spill)));
  }
}

// The time delta coded for contact i: from the previous contact of the same
// source, or from min_ts at a source's first contact.
struct TsDelta {
  const int32_t* src;
  const int64_t* t;
  int64_t m, tmin;
  __host__ __device__ int64_t operator()(int64_t i) const {
    const int64_t prev = (i == 0 || src[i] != src[i - 1]) ? tmin : t[i - 1];
    return t[i] - prev;
  }
};
struct TsLen {
  TsDelta d;
  __host__ __device__ uint64_t operator()(int64_t i) const {
    if (i >= d.m) return 0;
    uint64_t c;
    uint32_t len;
    zeta2(int2nat(d(i)), c, len);
    return len;
  }
};
struct GapLen {
  const uint64_t* v;
  int64_t count;
  __host__ __device__ uint64_t operator()(int64_t i) const {
    if (i >= count) return 0;
    uint64_t c;
    uint32_t len;
    gamma(v[i] - (i ? v[i - 1] : 0), c, len);
    return len;
  }
};

// The first contact that breaks the input contract (sources sorted in
// [0, n), times in int32 and not below min_ts), into *bad (atomicMin).
__global__ void __launch_bounds__(kBlock) k_ts_check(const int32_t* src, const int64_t* t, int64_t m, int64_t n,
                                                     int64_t tmin, unsigned long long* bad) {
  for (int64_t i = blockIdx.x * (int64_t)kBlock + threadIdx.x; i < m; i += (int64_t)gridDim.x * kBlock) {
    const bool ok = src[i] >= 0 && src[i] < n && (i == 0 || src[i] >= src[i - 1]) && t[i] >= INT32_MIN &&
                    t[i] <= INT32_MAX && t[i] >= tmin;
    if (!ok) atomicMin(bad, (unsigned long long)i);
  }
}

__global__ void __launch_bounds__(kBlock) k_ts_write(TsDelta d, const uint64_t* pos, unsigned long long* w) {
  for (int64_t i = blockIdx.x * (int64_t)kBlock + threadIdx.x; i < d.m; i += (int64_t)gridDim.x * kBlock) {
    uint64_t c;
    uint32_t len;
    zeta2(int2nat(d(i)), c, len);
    put_bits(w, pos[i], c, len);
  }
}

// off[x] = pos[first contact of x] for x in [0, n] (pos[m] for vertices past
// the last contact): contact i starts the vertices src[i - 1] + 1 .. src[i].
__global__ void __launch_bounds__(kBlock) k_vertex_off(const int32_t* src, int64_t m, int64_t n, const uint64_t* pos,
                                                       uint64_t* off) {
  for (int64_t i = blockIdx.x * (int64_t)kBlock + threadIdx.x; i <= m; i += (int64_t)gridDim.x * kBlock) {
    const int64_t lo = i == 0 ? 0 : (int64_t)src[i - 1] + 1;
    const int64_t hi = i == m ? n : (int64_t)src[i];
    for (int64_t x = lo; x <= hi; ++x) off[x] = pos[i];
  }
}

__global__ void __launch_bounds__(kBlock) k_gap_write(const uint64_t* v, int64_t count, const uint64_t* pos,
                                                      unsigned long long* w) {
  for (int64_t i = blockIdx.x * (int64_t)kBlock + threadIdx.x; i < count; i += (int64_t)gridDim.x * kBlock) {
    uint64_t c;
    uint32_t len;
    gamma(v[i] - (i ? v[i - 1] : 0), c, len);
    put_bits(w, pos[i], c, len);
  }
}

__global__ void __launch_bounds__(kBlock) k_ef_write(const uint64_t* v, int64_t count, int32_t l,
                                                     unsigned long long* lower, uint32_t* upper) {
  for (int64_t i = blockIdx.x * (int64_t)kBlock + threadIdx.x; i < count; i += (int64_t)gridDim.x * kBlock) {
    const uint64_t x = v[i];
    if (l > 0) put_bits(lower, (uint64_t)i * (uint64_t)l, x & ((1ull << l) - 1), (uint32_t)l);
    const uint64_t p = (x >> l) + (uint64_t)i;
    atomicOr(&upper[p >> 5], 1u << (p & 31));
  }
}

// Word i's bytes in stream order (most significant first).
__global__ void __launch_bounds__(kBlock) k_words_to_bytes(const uint64_t* w, int64_t nwords, uint64_t* out) {
  for (int64_t i = blockIdx.x * (int64_t)kBlock + threadIdx.x; i < nwords; i += (int64_t)gridDim.x * kBlock)
    out[i] = __builtin_bswap64(w[i]);
}

// Device buffers freed on scope exit.
struct Buf {
  void* p = nullptr;
  ~Buf() {
    if (p) (void)hipFree(p);
  }
};

int dev_alloc(Buf& b, size_t bytes) {
  if (hipMalloc(&b.p, bytes ? bytes : 1) != hipSuccess) {
    b.p = nullptr;
    (void)hipGetLastError();
    return err(TGA_ENOMEM, "device allocation of " + std::to_string(bytes) + " bytes failed");
  }
  return TGA_OK;
}

// Exclusive scan of lens(i), i in [0, count], from `init`: pos[0..count].
template <class F>
int scan_positions(F lens, int64_t count, uint64_t init, uint64_t* pos, hipStream_t st) {
  const auto in = rocprim::make_transform_iterator(rocprim::counting_iterator<int64_t>(0), lens);
  size_t tmp_bytes = 0;
  ENC_TRY(rocprim::exclusive_scan(nullptr, tmp_bytes, in, pos, init, (size_t)count + 1, rocprim::plus<uint64_t>(), st));
  Buf tmp;
  if (int rc = dev_alloc(tmp, tmp_bytes)) return rc;
  ENC_TRY(rocprim::exclusive_scan(tmp.p, tmp_bytes, in, pos, init, (size_t)count + 1, rocprim::plus<uint64_t>(), st));
  ENC_TRY(hipStreamSynchronize(st));
  return TGA_OK;
}

// A stream of `bytes` bytes from `nwords` MSB-first words (nwords * 8 >= bytes;
// the bytes past the written bits are zero).
int make_stream(int device, const unsigned long long* w, int64_t nwords, uint64_t bytes, hipStream_t st,
                tga_stream** out) {
  auto* s = new tga_stream();
  s->device = device;
  s->bytes = bytes;
  if (hipMalloc((void**)&s->data, (size_t)nwords * 8 + 8) != hipSuccess) {
    (void)hipGetLastError();
    delete s;
    return err(TGA_ENOMEM, "device allocation of the encoded stream failed");
  }
  if (nwords > 0)
    k_words_to_bytes<<<grid_for(nwords), kBlock, 0, st>>>((const uint64_t*)w, nwords, (uint64_t*)s->data);
  if (hipGetLastError() != hipSuccess || hipStreamSynchronize(st) != hipSuccess) {
    (void)hipFree(s->data);
    delete s;
    return err(TGA_EDEVICE, "encoding kernels failed");
  }
  *out = s;
  return TGA_OK;
}

struct Stream {  // a stream for the encoders' kernels, destroyed on scope exit
  hipStream_t s = nullptr;
  ~Stream() {
    if (s) (void)hipStreamDestroy(s);
  }
};

template <class T>
int upload(Buf& b, const T* host, size_t count, hipStream_t st) {
  if (int rc = dev_alloc(b, count * sizeof(T))) return rc;
  if (count) ENC_TRY(hipMemcpyAsync(b.p, host, count * sizeof(T), hipMemcpyHostToDevice, st));
  return TGA_OK;
}

}  // namespace

extern "C" int tga_encode_timestamps(const int32_t* src, const int64_t* t, int64_t m, int64_t min_ts, int64_t n,
                                     int device, tga_stream** out, uint64_t* offsets) {
  if (!src || !t || m < 0 || n < 1 || !out) return err(TGA_EINVAL, "tga_encode_timestamps: bad arguments");
  *out = nullptr;
  ENC_TRY(hipSetDevice(device));
  Stream st;
  ENC_TRY(hipStreamCreate(&st.s));
  Buf d_src, d_t, d_pos, d_w, d_off, d_bad;
  int rc = upload(d_src, src, (size_t)m, st.s);
  if (!rc) rc = upload(d_t, t, (size_t)m, st.s);
  if (!rc) rc = dev_alloc(d_pos, ((size_t)m + 1) * 8);
  if (!rc) rc = dev_alloc(d_bad, 8);
  if (rc) return rc;
  {
    // the order the decoders assume, and the int32 times
    unsigned long long bad = ~0ull;
    ENC_TRY(hipMemcpyAsync(d_bad.p, &bad, 8, hipMemcpyHostToDevice, st.s));
    if (m > 0)
      k_ts_check<<<grid_for(m), kBlock, 0, st.s>>>((const int32_t*)d_src.p, (const int64_t*)d_t.p, m, n, min_ts,
                                                    (unsigned long long*)d_bad.p);
    ENC_TRY(hipGetLastError());
    ENC_TRY(hipMemcpyAsync(&bad, d_bad.p, 8, hipMemcpyDeviceToHost, st.s));
    ENC_TRY(hipStreamSynchronize(st.s));
    if (bad != ~0ull)
      return err(TGA_EINVAL, "tga_encode_timestamps: contact " + std::to_string(bad) +
                                 " has its source out of [0, n) or out of order, or its time outside int32 or below "
                                 "min_ts");
  }
  const TsDelta d{(const int32_t*)d_src.p, (const int64_t*)d_t.p, m, min_ts};
  if ((rc = scan_positions(TsLen{d}, m, 64, (uint64_t*)d_pos.p, st.s))) return rc;
  uint64_t bits = 0;
  ENC_TRY(hipMemcpy(&bits, (const uint64_t*)d_pos.p + m, 8, hipMemcpyDeviceToHost));
  const int64_t nwords = (int64_t)((bits + 63) / 64);
  if ((rc = dev_alloc(d_w, (size_t)nwords * 8 + 8))) return rc;
  ENC_TRY(hipMemsetAsync(d_w.p, 0, (size_t)nwords * 8 + 8, st.s));
  auto* w = (unsigned long long*)d_w.p;
  const uint64_t head = (uint64_t)min_ts;  // the 64-bit header, MSB-first
  ENC_TRY(hipMemcpyAsync(w, &head, 8, hipMemcpyHostToDevice, st.s));
  if (m > 0) k_ts_write<<<grid_for(m), kBlock, 0, st.s>>>(d, (const uint64_t*)d_pos.p, w);
  ENC_TRY(hipGetLastError());
  if (offsets) {
    if ((rc = dev_alloc(d_off, ((size_t)n + 1) * 8))) return rc;
    k_vertex_off<<<grid_for(m + 1), kBlock, 0, st.s>>>((const int32_t*)d_src.p, m, n, (const uint64_t*)d_pos.p,
                                                        (uint64_t*)d_off.p);
    ENC_TRY(hipGetLastError());
    ENC_TRY(hipMemcpyAsync(offsets, d_off.p, ((size_t)n + 1) * 8, hipMemcpyDeviceToHost, st.s));
  }
  return make_stream(device, w, nwords, (bits + 7) / 8, st.s, out);
}

extern "C" int tga_encode_gamma_gaps(const uint64_t* values, int64_t count, int device, tga_stream** out) {
  if (!values || count < 0 || !out) return err(TGA_EINVAL, "tga_encode_gamma_gaps: bad arguments");
  *out = nullptr;
  for (int64_t i = 0; i < count; ++i) {
    const uint64_t prev = i ? values[i - 1] : 0;
    if (values[i] < prev) return err(TGA_EINVAL, "tga_encode_gamma_gaps: values decrease at " + std::to_string(i));
    if (values[i] - prev >= (1ull << 31))  // codes of at most 63 bits
      return err(TGA_ERANGE, "tga_encode_gamma_gaps: a gap of 2^31 or more at " + std::to_string(i));
  }
  ENC_TRY(hipSetDevice(device));
  Stream st;
  ENC_TRY(hipStreamCreate(&st.s));
  Buf d_v, d_pos, d_w;
  int rc = upload(d_v, values, (size_t)count, st.s);
  if (!rc) rc = dev_alloc(d_pos, ((size_t)count + 1) * 8);
  if (rc) return rc;
  if ((rc = scan_positions(GapLen{(const uint64_t*)d_v.p, count}, count, 0, (uint64_t*)d_pos.p, st.s))) return rc;
  uint64_t bits = 0;
  ENC_TRY(hipMemcpy(&bits, (const uint64_t*)d_pos.p + count, 8, hipMemcpyDeviceToHost));
  const int64_t nwords = (int64_t)((bits + 63) / 64);
  if ((rc = dev_alloc(d_w, (size_t)nwords * 8 + 8))) return rc;
  ENC_TRY(hipMemsetAsync(d_w.p, 0, (size_t)nwords * 8 + 8, st.s));
  auto* w = (unsigned long long*)d_w.p;
  if (count > 0)
    k_gap_write<<<grid_for(count), kBlock, 0, st.s>>>((const uint64_t*)d_v.p, count, (const uint64_t*)d_pos.p, w);
  ENC_TRY(hipGetLastError());
  return make_stream(device, w, nwords, (bits + 7) / 8, st.s, out);
}

extern "C" int tga_encode_ef_index(const uint64_t* values, int64_t count, int32_t l, int device, tga_stream** lower,
                                   tga_stream** upper) {
  if (!values || count < 0 || l < 0 || l > 63 || !lower || !upper)
    return err(TGA_EINVAL, "tga_encode_ef_index: bad arguments");
  *lower = *upper = nullptr;
  for (int64_t i = 1; i < count; ++i)
    if (values[i] < values[i - 1]) return err(TGA_EINVAL, "tga_encode_ef_index: values decrease at " + std::to_string(i));
  ENC_TRY(hipSetDevice(device));
  Stream st;
  ENC_TRY(hipStreamCreate(&st.s));
  // File sizes of the host index writer: the bytes holding the bits, rounded
  // up to a multiple of 8, plus 8 zero bytes.
  const uint64_t lower_bits = (uint64_t)count * (uint64_t)l;
  const uint64_t lower_bytes = ((lower_bits + 7) / 8 + 7) / 8 * 8 + 8;
  const uint64_t upper_bits = (count ? (values[count - 1] >> l) : 0) + (uint64_t)count;
  const uint64_t upper_bytes = ((upper_bits + 7) / 8 + 7) / 8 * 8 + 8;
  Buf d_v, d_lo, d_up;
  int rc = upload(d_v, values, (size_t)count, st.s);
  if (!rc) rc = dev_alloc(d_lo, lower_bytes);
  if (!rc) rc = dev_alloc(d_up, upper_bytes);
  if (rc) return rc;
  ENC_TRY(hipMemsetAsync(d_lo.p, 0, lower_bytes, st.s));
  ENC_TRY(hipMemsetAsync(d_up.p, 0, upper_bytes, st.s));
  if (count > 0)
    k_ef_write<<<grid_for(count), kBlock, 0, st.s>>>((const uint64_t*)d_v.p, count, l, (unsigned long long*)d_lo.p,
                                                       (uint32_t*)d_up.p);
  ENC_TRY(hipGetLastError());
  // lower: MSB-first words to bytes; upper: already in file order (LSB-first
  // bits in little-endian words)
  if ((rc = make_stream(device, (const unsigned long long*)d_lo.p, (int64_t)(lower_bytes / 8), lower_bytes, st.s,
                        lower)))
    return rc;
  auto* s = new tga_stream();
  s->device = device;
  s->bytes = upper_bytes;
  if (hipMalloc((void**)&s->data, upper_bytes) != hipSuccess ||
      hipMemcpyAsync(s->data, d_up.p, upper_bytes, hipMemcpyDeviceToDevice, st.s) != hipSuccess ||
      hipStreamSynchronize(st.s) != hipSuccess) {
    (void)hipGetLastError();
    if (s->data) (void)hipFree(s->data);
    delete s;
    tga_stream_free(*lower);
    *lower = nullptr;
    return err(TGA_EDEVICE, "tga_encode_ef_index: upper-bits stream failed");
  }
  *upper = s;
  return TGA_OK;
}

extern "C" uint64_t tga_stream_bytes(const tga_stream* s) { return s ? s->bytes : 0; }

extern "C" int tga_stream_read(const tga_stream* s, uint8_t* out, uint64_t cap) {
  if (!s || (!out && s->bytes)) return err(TGA_EINVAL, "tga_stream_read: bad arguments");
  if (cap < s->bytes) return err(TGA_ERANGE, "tga_stream_read: buffer smaller than the stream");
  ENC_TRY(hipSetDevice(s->device));
  if (s->bytes) ENC_TRY(hipMemcpy(out, s->data, s->bytes, hipMemcpyDeviceToHost));
  return TGA_OK;
}

extern "C" void tga_stream_free(tga_stream* s) {
  if (!s) return;
  if (s->data && hipSetDevice(s->device) == hipSuccess) (void)hipFree(s->data);
  delete s;
}
