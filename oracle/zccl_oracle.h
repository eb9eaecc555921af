/*
 * zccl_oracle.h -- CPU restatement of the SZp "fastmode1" codec and the hZCCL
 * homomorphic sum.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP path:
 * tests/, __graft_entry__.smoke() and the cpu_baseline leg of bench.py may load
 * it; the product (zccl_amd/) never links, calls or falls back to it.
 *
 * Every function restates the reference algorithm from its source and cites
 * the reference file:line it follows (paths relative to the reference root).
 * Arithmetic is integer-exact except quantize (FP32 x FP32 product evaluated
 * exactly in FP64, then truncated) and dequantize (FP32 convert + FP32 mul);
 * the file is built with -ffp-contract=off and without -ffast-math.
 */
#ifndef ZCCL_ORACLE_H
#define ZCCL_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Stream geometry for (n, L, T).  Segments are the seeded runs of the
 * Lorenzo chain: T regular chunks of tb = n/T elements, plus (when
 * R = n % T > 0) a remainder segment of R elements whose bytes are appended to
 * chunk T-1's payload.  ZCCL/src/ZCCL_float.c:533-544, 672-679. */
typedef struct zo_geom {
    uint64_t n;
    int32_t L, T;
    uint64_t tb, R;
    int32_t S;          /* number of segments: T or T+1 */
    uint64_t nbT;       /* blocks in a regular chunk: ceil((tb-1)/L) */
    uint64_t nbR;       /* blocks in the remainder segment: ceil((R-1)/L), 0 if R==0 */
    uint64_t nblocks;   /* T*nbT + nbR */
} zo_geom;

/* Status codes (negative). */
#define ZO_OK 0
#define ZO_EINVAL (-1)   /* n < T, L < 1, T < 1 ... */
#define ZO_ENOSPC (-2)   /* output capacity too small */
#define ZO_ECORRUPT (-3) /* width byte > 32 or stream overruns its size */

int zo_geometry(uint64_t n, int L, int T, zo_geom *g);

/* Bit width stored for a block whose max |delta| is m: (int)log2f((float)m)+1
 * as computed by glibc 2.35, restated with clz and a threshold table. */
int zo_width(uint32_t m);
void zo_width_array(const uint32_t *m, uint64_t n, uint8_t *b);
/* Bytes of one encoded block of n deltas at width b (width byte included). */
uint64_t zo_block_bytes(int b, uint32_t n);
/* Worst-case stream size for (n, L, T). */
uint64_t zo_worst_case_bytes(uint64_t n, int L, int T);

/* Quantize / dequantize closed forms. */
void zo_quantize(const float *x, uint64_t n, float eb, int32_t *q);
void zo_dequantize(const int32_t *q, uint64_t n, float eb, float *x);

/* Encode / decode single blocks (exposed for packer tests). */
uint64_t zo_encode_block(const int32_t *d, uint32_t n, uint8_t *out);
int64_t zo_decode_block(const uint8_t *in, uint64_t avail, uint32_t n, int32_t *d);

/* Whole streams.  Return stream size (>= 0) or a negative status. */
int64_t zo_encode_q(const int32_t *q, uint64_t n, int L, int T, uint8_t *out, uint64_t cap);
int64_t zo_compress(const float *x, uint64_t n, float eb, int L, int T, uint8_t *out, uint64_t cap);
int zo_decode_q(const uint8_t *in, uint64_t in_size, uint64_t n, int L, int T, int32_t *q);
int zo_decompress(const uint8_t *in, uint64_t in_size, uint64_t n, float eb, int L, int T, float *x);

/* k-operand homomorphic sum of streams with equal (n, L, T): seeds and block
 * deltas add in int32 (wrap), every block re-encoded canonically. */
int64_t zo_homsum(const uint8_t *const *ins, const uint64_t *in_sizes, int k, uint64_t n, int L,
    int T, uint8_t *out, uint64_t cap);

/* Decompress-operate-compress reduction of one segment over P ranks
 * (xs in ring order): acc = the last rank's FP32 partial sum; final (may be
 * null) = what every rank holds after the compressed allgather. */
int zo_doc_chain(const float *const *xs, int P, uint64_t n, float eb, int L, int T, float *acc, float *final);

/* Parse a stream into its block index: widths[nblocks], byte offset of every
 * block's width byte (block_off[nblocks]) and of every segment's seed
 * (seed_off[S]); *total receives the stream size implied by the parse. */
int zo_parse(const uint8_t *in, uint64_t in_size, uint64_t n, int L, int T, uint8_t *widths,
    uint64_t *block_off, uint64_t *seed_off, uint64_t *total);

/* Naive reference encoders used only to pin the packers: one bit at a time. */
uint64_t zo_encode_block_bitwise(const int32_t *d, uint32_t n, uint8_t *out);

/* 3-D SZp codec (SZp_compress3D_fast_openmp / SZp_decompress3D_fast_openmp):
 * dims d1 (slowest) x d2 x d3, cubic blocks of side Bs, absolute bound eb,
 * a T-way split of the volume (sub-volume t compressed by thread t). */
typedef struct zo_part3 {
    uint64_t ox, oy, oz;  /* first plane of the sub-volume along each axis */
    uint64_t sx, sy, sz;  /* planes along each axis */
} zo_part3;

void zo_dims_create(int T, int dims[3]);
int zo_partition3d(uint64_t d1, uint64_t d2, uint64_t d3, int T, zo_part3 *parts);
void zo_quantize3d(const float *x, uint64_t n, double eb, int32_t *q);
void zo_dequantize3d(const int32_t *q, uint64_t n, double eb, float *x);
uint64_t zo_worst_case_bytes3d(uint64_t d1, uint64_t d2, uint64_t d3, int Bs, int T);
int64_t zo_compress3d(const float *x, uint64_t d1, uint64_t d2, uint64_t d3, int Bs, double eb, int T,
    uint8_t *out, uint64_t cap);
int zo_decompress3d(const uint8_t *in, uint64_t in_size, uint64_t d1, uint64_t d2, uint64_t d3, int Bs,
    double eb, int T, float *x);

#ifdef __cplusplus
}
#endif
#endif
