// rocPRIM primitives over the neighbor pairs of PageRank and HITS: the radix
// sort that groups them by destination, and the per-row sums of every
// iteration (a deterministic reduce-by-key, load-balanced over rows of any
// length).  Kept in their own translation unit: rocPRIM's templates are slow
// to compile.
//
// A pair is key r (uint32) and value (x << 32 | float bits of its weight w).
#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_reduce_by_key.hpp>
#include <rocprim/iterator/counting_iterator.hpp>
#include <rocprim/iterator/transform_iterator.hpp>

namespace tga {

// Sorts keys_in/vals_in[0, n) by the key bits [0, end_bit) into keys_out/
// vals_out (stable).  With tmp == nullptr only sets tmp_bytes.
hipError_t sort_pairs_u32_u64(void* tmp, size_t& tmp_bytes, const uint32_t* keys_in, uint32_t* keys_out,
                              const uint64_t* vals_in, uint64_t* vals_out, size_t n, int end_bit,
                              hipStream_t stream) {
  return rocprim::radix_sort_pairs(tmp, tmp_bytes, keys_in, keys_out, vals_in, vals_out, n, 0, end_bit, stream);
}

namespace {

__host__ __device__ inline float low_float(uint64_t v) { return __builtin_bit_cast(float, (uint32_t)v); }

// PageRank in-row term of pair i: (double)(coef[x] * w).
struct PrTerm {
  const uint64_t* val;
  const float* coef;
  __host__ __device__ double operator()(uint64_t i) const {
    const uint64_t e = val[i];
    return (double)(coef[(uint32_t)(e >> 32)] * low_float(e));
  }
};

// HITS in-row term of pair i: (double)vec[x].
struct XTerm {
  const uint64_t* val;
  const float* vec;
  __host__ __device__ double operator()(uint64_t i) const { return (double)vec[(uint32_t)(val[i] >> 32)]; }
};

// HITS out-row term of pair i: (double)vec[r], 0 for an empty slot (r >= n).
struct RTerm {
  const uint32_t* key;
  const float* vec;
  uint32_t n;
  __host__ __device__ double operator()(uint64_t i) const {
    const uint32_t r = key[i];
    return r < n ? (double)vec[r] : 0.0;
  }
};

// Key x of pair i (out rows).
struct XKey {
  const uint64_t* val;
  __host__ __device__ uint32_t operator()(uint64_t i) const { return (uint32_t)(val[i] >> 32); }
};

template <class K, class V>
hipError_t rbk(void* tmp, size_t& tmp_bytes, K keys, V vals, size_t count, uint32_t* ukeys, double* sums,
               uint32_t* runs, hipStream_t stream) {
  return rocprim::deterministic_reduce_by_key(tmp, tmp_bytes, keys, vals, count, ukeys, sums, runs,
                                              rocprim::plus<double>(), rocprim::equal_to<uint32_t>(), stream);
}

}  // namespace

// Row sums over the pairs [begin, begin + count): every run of equal keys
// yields its key in ukeys and the sum of its terms in sums (in pair order),
// and *runs receives the number of runs.  mode 0: keys sk (sorted r), terms
// (double)(vec[x] * w) (PageRank); mode 1: keys sk, terms (double)vec[x]
// (HITS authority); mode 2: keys x of val in fill order, terms (double)vec[r]
// over key pk (HITS hub).  With tmp == nullptr only sets tmp_bytes.
hipError_t reduce_rows(int mode, void* tmp, size_t& tmp_bytes, const uint32_t* sk, const uint64_t* val,
                       const uint32_t* pk, const float* vec, int32_t n, size_t begin, size_t count, uint32_t* ukeys,
                       double* sums, uint32_t* runs, hipStream_t stream) {
  const rocprim::counting_iterator<uint64_t> idx(begin);
  if (mode == 0)
    return rbk(tmp, tmp_bytes, sk + begin, rocprim::make_transform_iterator(idx, PrTerm{val, vec}), count, ukeys,
               sums, runs, stream);
  if (mode == 1)
    return rbk(tmp, tmp_bytes, sk + begin, rocprim::make_transform_iterator(idx, XTerm{val, vec}), count, ukeys,
               sums, runs, stream);
  return rbk(tmp, tmp_bytes, rocprim::make_transform_iterator(idx, XKey{val}),
             rocprim::make_transform_iterator(idx, RTerm{pk, vec, (uint32_t)n}), count, ukeys, sums, runs, stream);
}

}  // namespace tga
