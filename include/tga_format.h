/*
 * tga_format.h — C ABI of the host-side format kit (libtga_format.so).
 *
 * The format kit writes and checks the on-disk files that the engine
 * (libtga.so, tga.h) reads unchanged:
 *
 *   <dir>/<name>.graph        BV multigraph successor records
 *   <dir>/<name>.offsets      gamma-coded record bit-offset gaps
 *   <dir>/<name>.properties   Java-properties text with the coding parameters
 *   <dir>/<name>.timestamps   64-bit minTimestamp header + zeta_2 deltas
 *   <dir>/graph_lower_bits.bin, graph_upper_bits.bin,
 *   <dir>/timestamp_lower_bits.bin, timestamp_upper_bits.bin
 *                              the GPU Elias-Fano offset indexes
 *
 * Each entry point returns 0 on success and a negative TGAF_E* code on
 * failure; tgaf_last_error() describes the last failure of the calling
 * thread.  No entry point exits the process.
 */
#ifndef TGA_FORMAT_H
#define TGA_FORMAT_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
  TGAF_OK = 0,
  TGAF_EINVAL = -1,   /* bad argument */
  TGAF_EIO = -2,      /* file could not be opened, read or written */
  TGAF_EFORMAT = -3,  /* malformed input (unsorted sources, short line, ...) */
  TGAF_ERANGE = -4,   /* value outside what the format or the GPU reader supports */
  TGAF_ENOMEM = -5
};

typedef struct tgaf_stats {
  int64_t nodes;              /* n = largest id + 1 (MAX_LABEL + 1) */
  int64_t arcs;               /* contacts */
  int64_t min_timestamp;      /* value of the .timestamps header */
  uint64_t graph_bits;        /* bits of .graph payload */
  uint64_t timestamps_bits;   /* bits of .timestamps payload, header included */
  int64_t copied_arcs;        /* arcs reproduced from a reference list */
  int64_t multiple_arcs;      /* arcs inside multiples runs */
  int64_t intervalised_arcs;  /* arcs inside intervals */
  int64_t residual_arcs;      /* arcs coded as zeta_3 residuals */
  int64_t records_with_ref;   /* records with r > 0 */
  int32_t max_ref_depth;      /* deepest realised reference chain (<= 3) */
  int32_t l_graph;            /* Elias-Fano lower-bit width of the graph index */
  int32_t l_timestamps;       /* Elias-Fano lower-bit width of the timestamp index */
  int32_t pad_;
} tgaf_stats;

const char* tgaf_last_error(void);

/* Encode m contacts (src[i], dst[i], t[i]) into <dir>/<name>.* plus the four
 * GPU index files.  Contacts must be sorted by (src, dst); within a pair the
 * given order is the timestamp order (ArcListASCIIEvolvingGraph.java:160-215,
 * EvolvingMultiGraph.java:202-285).  n_min, when larger than max id + 1, adds
 * trailing vertices without contacts.  Window 7, max reference count 3,
 * minimum interval 4, zeta_3 residuals, zeta_2 timestamps, aggregation 1
 * (BVMultiGraph.java:159-184; BaselineRunner.java:17-18). */
int tgaf_encode(const int32_t* src, const int32_t* dst, const int64_t* t, int64_t m, int64_t n_min,
                const char* dir, const char* name, tgaf_stats* stats);

/* The graph half of tgaf_encode: <dir>/<name>.graph and .properties only, the
 * record bit offsets of the graph stream into goff (n + 1 entries, n =
 * max(n_min, largest id + 1)).  The timestamp stream, .offsets and the four
 * GPU index files are then written on the GPU from goff and the contacts
 * (tga_encode.h), byte for byte as tgaf_encode writes them. */
int tgaf_encode_graph(const int32_t* src, const int32_t* dst, int64_t m, int64_t n_min, const char* dir,
                      const char* name, uint64_t* goff, tgaf_stats* stats);

/* Rebuild the four GPU index files of an existing dataset from its .graph,
 * .offsets and .timestamps (the reference ships no generator for them). */
int tgaf_build_gpu_index(const char* dir, const char* name, tgaf_stats* stats);

/* Synthetic workloads, written into caller-allocated arrays of m contacts and
 * sorted by (src, dst, t).  Deterministic for a given seed.
 *  R-MAT (a, b, c, 1-a-b-c) over 2^scale vertices; each sampled pair carries a
 *  Geometric(p_geo) number of contacts (support 1, 2, ...), the last pair
 *  truncated so the total is exactly m; t uniform in [0, t_span). */
int tgaf_gen_rmat(int32_t scale, double a, double b, double c, double p_geo, int64_t t_span, uint64_t seed,
                  int64_t m, int32_t* src, int32_t* dst, int64_t* t);
/*  Chung-Lu with weights w_i ~ (i+1)^(-1/(exponent-1)) on both endpoints, m
 *  independent contacts; timestamps i.i.d. from the diurnal intensity
 *  1 + 0.5 sin(2 pi t / period) on [0, t_span). */
int tgaf_gen_chung_lu(int64_t n, double exponent, int64_t t_span, int64_t period, uint64_t seed,
                      int64_t m, int32_t* src, int32_t* dst, int64_t* t);
/*  Netflow-like: pairs with both endpoints from the same power-law weights,
 *  a Pareto (alpha 2) number of contacts per pair, round(count_min *
 *  U^(-1/2)) with mean about 2 count_min, repeated pairs merged, the last
 *  pair cut so the total is exactly m; each pair's contacts a burst from a
 *  uniform start with Pareto (alpha 1.5) gaps of at least gap_min ticks,
 *  folded into [0, t_span).  With TGAF_PROGRESS set in the environment,
 *  progress lines go to stderr (also for tgaf_encode). */
int tgaf_gen_netflow(int64_t n, double exponent, double count_min, double gap_min, int64_t t_span, uint64_t seed,
                     int64_t m, int32_t* src, int32_t* dst, int64_t* t);

/* The wrapping 64-bit sum over contacts of h = (t ^ r*0x9E3779B1 ^
 * x*0x85EBCA77) * 0x7FEB352D (uint32 arithmetic, x, r, t as unsigned 32-bit
 * values): the checksum tga_neighbor_access returns for the encoded list. */
int tgaf_contact_checksum(const int32_t* src, const int32_t* dst, const int64_t* t, int64_t m, uint64_t* out);

/* Raw 4-column text ("src dst 1 t" per line), gzip-compressed. */
int tgaf_write_raw_gz(const char* path, const int32_t* src, const int32_t* dst, const int64_t* t, int64_t m);
/* Reads a raw .gz with the reference parser's rules (>= 4 numeric columns,
 * t in column 4, sources non-decreasing).  Two calls: with src == NULL the
 * number of contacts is stored in *m and nothing else is written. */
int tgaf_read_raw_gz(const char* path, int headers, int64_t* m, int32_t* src, int32_t* dst, int64_t* t);

/* The 65536-entry 16-bit-window decode tables of the reference
 * (gpu_cuda/decode_tables/gammaInA.txt, zetaInA.txt): entry w is
 * (codeword length << 16) | value when a whole codeword starting at the most
 * significant bit of w fits in 16 bits, else 0.  kind 0 = gamma, 1 = zeta_3. */
int tgaf_decode_table(int kind, int32_t* out65536);

/* Java's java.util.Random(seed).nextInt(bound) sequence (the GC priority
 * scheme of BaselineRunner.java:392-396), count values into out. */
int tgaf_java_random_ints(int64_t seed, int32_t bound, int64_t count, int32_t* out);

/* Writes a sequence of codes MSB-first into out (cap bytes): kinds[i] is 0 for
 * unary, 1 for gamma, k in [2, 63] for zeta_k, 64 for a 64-bit integer
 * (the .timestamps header).  *nbits receives the number of bits written. */
int tgaf_encode_codes(const int32_t* kinds, const uint64_t* values, int64_t n, uint8_t* out, int64_t cap,
                      uint64_t* nbits);

#ifdef __cplusplus
}
#endif
#endif
