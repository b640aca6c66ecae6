/* GPU writers of the ChronoGraph/WebGraph streams that are independent per
 * contact or per vertex once their bit lengths are prefix-summed: the
 * timestamp stream, the record-offset stream and the GPU Elias-Fano index
 * files.  The graph stream itself (the record planner: reference choice,
 * copy blocks, multiples, intervals, residuals) stays on the host, in the
 * format kit (tga_format.h), because its reference choice is sequential in
 * the chain-depth limit.
 *
 * Every writer produces exactly the bytes the format kit's host encoder
 * writes for the same input (tgaf_encode, tgaf_build_gpu_index).  Inputs are
 * host arrays; the encoded stream is held on the device until read.
 * Functions return the status codes of tga.h (TGA_OK, TGA_EINVAL,
 * TGA_ERANGE, TGA_ENOMEM, TGA_EDEVICE); tga_last_error() describes the last
 * failure.
 */
#ifndef TGA_ENCODE_H
#define TGA_ENCODE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* An encoded byte stream held in device memory. */
typedef struct tga_stream tga_stream;

/* The .timestamps stream: a 64-bit header holding min_ts, then for every
 * contact in order the zeta_2 code of int2nat(t - previous), the previous
 * time being min_ts at the first contact of each source vertex.  Replaces the
 * timestamp writer of EvolvingMultiGraph.storeTimestampsAndIndex and its
 * TimeStampsWriter (third_party/ChronoGraph/.../EvolvingMultiGraph.java:113-175,
 * :346-376), the stream the reference's readZeta_k2 decodes
 * (gpu_cuda/shared/bfs.cu:640-665).
 *   src, t:   m contacts sorted by (source, destination, time), t in int32
 *             range, src in [0, n)
 *   offsets:  if not NULL, n + 1 entries: the bit offset of every vertex's
 *             first code (of the next code for a vertex with none), then the
 *             stream's length in bits */
int tga_encode_timestamps(const int32_t* src, const int64_t* t, int64_t m, int64_t min_ts, int64_t n, int device,
                          tga_stream** out, uint64_t* offsets);

/* The .offsets stream: gamma codes of values[0] and of every difference
 * values[i] - values[i - 1] (count codes; values non-decreasing).  Replaces
 * the offsets writer of BVMultiGraph's compression (writeOffset of the gaps
 * between record bit offsets, third_party/ChronoGraph/.../BVMultiGraph.java:359-365,
 * :1912, :1983). */
int tga_encode_gamma_gaps(const uint64_t* values, int64_t count, int device, tga_stream** out);

/* The GPU index files of one stream: lower = the low l bits of every value,
 * MSB-first, back to back; upper = bit (values[i] >> l) + i set for every i,
 * LSB-first within bytes; each zero padded to a multiple of 8 bytes plus 8
 * bytes (values non-decreasing, 0 <= l < 64).  Replaces the index writer the
 * reference's getOffset reads (gpu_cuda/shared/bfs.cu:543-565). */
int tga_encode_ef_index(const uint64_t* values, int64_t count, int32_t l, int device, tga_stream** lower,
                        tga_stream** upper);

/* Length of an encoded stream in bytes. */
uint64_t tga_stream_bytes(const tga_stream* s);

/* Copies the stream into out (cap bytes, at least tga_stream_bytes). */
int tga_stream_read(const tga_stream* s, uint8_t* out, uint64_t cap);

void tga_stream_free(tga_stream* s);

#ifdef __cplusplus
}
#endif

#endif /* TGA_ENCODE_H */
