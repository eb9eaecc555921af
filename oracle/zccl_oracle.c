/*
 * zccl_oracle.c -- CPU restatement of SZp fastmode1 compress / decompress and
 * the hZCCL homomorphic sum.  TEST INFRASTRUCTURE ONLY (the checker): the
 * product path never links or calls this file.
 *
 * Reference algorithm followed (paths relative to the reference root):
 *   compressor      ZCCL/src/ZCCL_float.c:501-807  (ZCCL_float_openmp_threadblock_arg)
 *   decompressor    ZCCL/src/ZCCLd_float.c:538-754 (ZCCL_float_decompress_openmp_threadblock_arg)
 *   homomorphic add ZCCL/src/ZCCLd_float.c:988-1679 (ZCCL_float_homomophic_add_openmp_threadblock)
 *   bit packing     ZCCL/src/ZCCL_TypeManager.c:12-91, 118-421 (save), 423-822 (extract),
 *                   824-848 / 906-950 (sign bitmap)
 *
 * Pinning of each function (where the check lives):
 *   zo_width            exhaustive glibc log2f probe (tests/golden/width_probe.c) and
 *                       test_oracle.py::test_width_*;
 *   zo_quantize,
 *   zo_dequantize       closed forms checked in test_oracle.py against numpy float64 and
 *                       float32 arithmetic, and through decompress(compress(x)) == fixtures;
 *   zo_encode_block,
 *   zo_decode_block     compiled-reference fixtures + naive bit-writer + round trip;
 *   zo_encode_q,
 *   zo_compress         byte equality with the compiled reference and committed fixtures
 *                       over T in {1,2,3,5,8,16}, L in {1,7,32,36,128}, ragged n;
 *   zo_decode_q,
 *   zo_decompress       equality with the compiled reference and the closed form
 *                       fl32(fl32(q)*eb), which is independent of T and L;
 *   zo_homsum           equality with chains of the compiled reference's pairwise add and
 *                       with encode_q(sum of q) (canonical-encoding invariant);
 *   zo_parse            offsets cross-checked against the sizes the encoder reports.
 */
#include "zccl_oracle.h"

#include <limits.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ---------------------------------------------------------------- geometry */

/* tb, R, full/partial block counts: ZCCL/src/ZCCL_float.c:541-544 and the
 * remainder section :672-679 (segment of R elements appended to chunk T-1). */
int zo_geometry(uint64_t n, int L, int T, zo_geom *g)
{
    if (L < 1 || T < 1 || n < (uint64_t) T)
        return ZO_EINVAL;
    g->n = n;
    g->L = L;
    g->T = T;
    g->tb = n / (uint64_t) T;
    g->R = n % (uint64_t) T;
    g->S = T + (g->R > 0 ? 1 : 0);
    g->nbT = (g->tb - 1 + (uint64_t) L - 1) / (uint64_t) L;
    g->nbR = g->R > 0 ? (g->R - 1 + (uint64_t) L - 1) / (uint64_t) L : 0;
    g->nblocks = (uint64_t) T * g->nbT + g->nbR;
    return ZO_OK;
}

/* First element and length of segment s. */
static void seg_range(const zo_geom *g, int s, uint64_t *a, uint64_t *len)
{
    if (s < g->T) {
        *a = (uint64_t) s * g->tb;
        *len = g->tb;
    } else {
        *a = (uint64_t) g->T * g->tb;
        *len = g->R;
    }
}

static uint64_t seg_nblocks(const zo_geom *g, int s) { return s < g->T ? g->nbT : g->nbR; }

/* Element count of block j of a segment of length len. */
static uint32_t block_len(const zo_geom *g, uint64_t len, uint64_t j)
{
    uint64_t left = len - 1 - j * (uint64_t) g->L;
    return (uint32_t) (left < (uint64_t) g->L ? left : (uint64_t) g->L);
}

/* ------------------------------------------------------------------- width */

/* (int)log2f((float)m) + 1 for m > 0 under glibc 2.35 (ZCCL/src/ZCCL_float.c:606,
 * ZCCL/src/ZCCLd_float.c:1205).  For widths w = 32 - clz(m) >= 21 the float
 * conversion plus correctly rounded log2f reach w exactly once m >= 2^w - c_w;
 * c_w below was measured exhaustively over all m in [1, 2^32). */
static const uint32_t zo_bump_c[12] = {1, 2, 5, 11, 22, 45, 91, 183, 367, 735, 1471, 2943};

int zo_width(uint32_t m)
{
    if (m == 0)
        return 0;
    int w = 32 - __builtin_clz(m);
    if (w >= 21 && (uint64_t) m >= ((uint64_t) 1 << w) - zo_bump_c[w - 21])
        w++;
    return w;
}

void zo_width_array(const uint32_t *m, uint64_t n, uint8_t *b)
{
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < (int64_t) n; i++)
        b[i] = (uint8_t) zo_width(m[i]);
}

/* Size of one block: width byte, sign bitmap, B high bytes per element, then
 * the r-bit low fields.  ZCCL/src/ZCCL_TypeManager.c:20-29 and
 * ZCCL/src/ZCCL_float.c:606-622. */
uint64_t zo_block_bytes(int b, uint32_t n)
{
    if (b == 0)
        return 1;
    uint64_t B = (uint64_t) (b >> 3), r = (uint64_t) (b & 7);
    return 1 + ((uint64_t) n + 7) / 8 + B * n + (r * n + 7) / 8;
}

uint64_t zo_worst_case_bytes(uint64_t n, int L, int T)
{
    zo_geom g;
    if (zo_geometry(n, L, T, &g) != ZO_OK)
        return 0;
    /* every block at width 32 */
    uint64_t tot = 8 * (uint64_t) T + 4 * (uint64_t) g.S;
    for (int s = 0; s < g.S; s++) {
        uint64_t a, len;
        seg_range(&g, s, &a, &len);
        uint64_t nb = seg_nblocks(&g, s);
        if (nb == 0)
            continue;
        uint64_t nfull = (len - 1) / (uint64_t) L;
        uint32_t tail = (uint32_t) ((len - 1) % (uint64_t) L);
        tot += nfull * zo_block_bytes(32, (uint32_t) L);
        if (tail)
            tot += zo_block_bytes(32, tail);
    }
    return tot;
}

/* ---------------------------------------------------------- quantization */

/* q = (int)((double)x * (double)(1.0f/eb)): the FP32 division happens first
 * (int / float in C), the widened product is exact in FP64, and the conversion
 * truncates toward zero.  x86 cvttsd2si yields INT_MIN for NaN and for
 * products outside (-2^31-1, 2^31).  ZCCL/src/ZCCL_float.c:540, 563, 580. */
static inline int32_t quant1(float x, double inv)
{
    double p = (double) x * inv;
    if (p > -2147483649.0 && p < 2147483648.0)
        return (int32_t) p;
    return INT32_MIN;
}

static inline double inv_bound(float eb)
{
    volatile float one = 1.0f; /* keep the division in FP32 */
    float inv = one / eb;
    return (double) inv;
}

void zo_quantize(const float *x, uint64_t n, float eb, int32_t *q)
{
    double inv = inv_bound(eb);
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < (int64_t) n; i++)
        q[i] = quant1(x[i], inv);
}

/* x' = (float)q * eb in FP32.  ZCCL/src/ZCCLd_float.c:588, 603, 624. */
void zo_dequantize(const int32_t *q, uint64_t n, float eb, float *x)
{
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < (int64_t) n; i++) {
        float f = (float) q[i];
        x[i] = f * eb;
    }
}

/* --------------------------------------------------------- block packing */

static inline uint32_t mag32(int32_t d) { return d < 0 ? (uint32_t) 0 - (uint32_t) d : (uint32_t) d; }

/* Block = width byte; if width b > 0: sign bitmap (bit 0x80>>(j&7) of byte
 * j>>3 set iff d_j < 0, ZCCL/src/ZCCL_TypeManager.c:824-848); B = b>>3 high
 * bytes per element, element-major little-endian bytes of |d_j|>>r
 * (:34-47); then the low r = b&7 bits of every |d_j| as one MSB-first bit
 * stream, zero-padded to a byte (:48-88 and the r-bit packers :118-421). */
/* Sign bitmap, high bytes and low bits of n deltas at width b > 0, written
 * at p (the block without its width byte); returns the bytes written. */
static uint64_t pack_fields(const int32_t *d, uint32_t n, int b, uint8_t *p0)
{
    uint8_t *p = p0;
    uint32_t sb = (n + 7) / 8;
    memset(p, 0, sb);
    for (uint32_t j = 0; j < n; j++)
        if (d[j] < 0)
            p[j >> 3] |= (uint8_t) (0x80u >> (j & 7));
    p += sb;
    int B = b >> 3, r = b & 7;
    if (B > 0) {
        for (uint32_t j = 0; j < n; j++) {
            uint32_t v = mag32(d[j]) >> r;
            for (int k = 0; k < B; k++)
                p[(uint64_t) j * B + k] = (uint8_t) (v >> (8 * k));
        }
        p += (uint64_t) B * n;
    }
    if (r > 0) {
        uint32_t mask = (1u << r) - 1;
        uint64_t acc = 0;
        int nbits = 0;
        for (uint32_t j = 0; j < n; j++) {
            acc = (acc << r) | (mag32(d[j]) & mask);
            nbits += r;
            while (nbits >= 8) {
                *p++ = (uint8_t) (acc >> (nbits - 8));
                nbits -= 8;
            }
        }
        if (nbits > 0)
            *p++ = (uint8_t) (acc << (8 - nbits));
    }
    return (uint64_t) (p - p0);
}

uint64_t zo_encode_block(const int32_t *d, uint32_t n, uint8_t *out)
{
    uint32_t m = 0;
    for (uint32_t j = 0; j < n; j++) {
        uint32_t v = mag32(d[j]);
        if (v > m)
            m = v;
    }
    int b = zo_width(m);
    out[0] = (uint8_t) b;
    if (b == 0)
        return 1;
    return 1 + pack_fields(d, n, b, out + 1);
}

/* Inverse of zo_encode_block (ZCCL/src/ZCCL_TypeManager.c:423-517, 906-950;
 * sign applied as 0 - |d| in unsigned arithmetic, ZCCL/src/ZCCLd_float.c:617-623).
 * Returns bytes consumed or ZO_ECORRUPT. */
/* Inverse of pack_fields: n deltas at width b > 0 from p. */
static void unpack_fields(const uint8_t *p, uint32_t n, int b, int32_t *d)
{
    const uint8_t *sg = p;
    const uint8_t *hi = sg + (n + 7) / 8;
    int B = b >> 3, r = b & 7;
    const uint8_t *lo = hi + (uint64_t) B * n;
    for (uint32_t j = 0; j < n; j++) {
        uint32_t v = 0;
        for (int k = 0; k < B; k++)
            v |= (uint32_t) hi[(uint64_t) j * B + k] << (8 * k);
        v <<= r;
        if (r > 0) {
            uint64_t bit = (uint64_t) j * r;
            uint32_t f = 0;
            for (int t = 0; t < r; t++, bit++)
                f = (f << 1) | ((lo[bit >> 3] >> (7 - (bit & 7))) & 1u);
            v |= f;
        }
        int neg = (sg[j >> 3] >> (7 - (j & 7))) & 1;
        d[j] = neg ? (int32_t) ((uint32_t) 0 - v) : (int32_t) v;
    }
}

int64_t zo_decode_block(const uint8_t *in, uint64_t avail, uint32_t n, int32_t *d)
{
    if (avail < 1)
        return ZO_ECORRUPT;
    int b = in[0];
    if (b == 0) {
        memset(d, 0, (size_t) n * sizeof(int32_t));
        return 1;
    }
    if (b > 32)
        return ZO_ECORRUPT;
    uint64_t sz = zo_block_bytes(b, n);
    if (sz > avail)
        return ZO_ECORRUPT;
    unpack_fields(in + 1, n, b, d);
    return (int64_t) sz;
}

/* Independent bit-at-a-time writer for the same layout, used to pin the
 * packers: each field is emitted MSB first into a zeroed bit array. */
static void put_bits(uint8_t *buf, uint64_t *pos, uint32_t v, int nbits)
{
    for (int t = nbits - 1; t >= 0; t--, (*pos)++)
        if ((v >> t) & 1u)
            buf[*pos >> 3] |= (uint8_t) (0x80u >> (*pos & 7));
}

uint64_t zo_encode_block_bitwise(const int32_t *d, uint32_t n, uint8_t *out)
{
    uint32_t m = 0;
    for (uint32_t j = 0; j < n; j++)
        if (mag32(d[j]) > m)
            m = mag32(d[j]);
    int b = zo_width(m);
    uint64_t sz = zo_block_bytes(b, n);
    memset(out, 0, sz);
    out[0] = (uint8_t) b;
    if (b == 0)
        return 1;
    uint64_t pos = 8;
    for (uint32_t j = 0; j < n; j++)
        put_bits(out, &pos, d[j] < 0, 1);
    pos = 8 * (1 + ((uint64_t) n + 7) / 8);
    int B = b >> 3, r = b & 7;
    for (uint32_t j = 0; j < n; j++)
        for (int k = 0; k < B; k++)
            put_bits(out, &pos, (mag32(d[j]) >> (r + 8 * k)) & 0xFFu, 8);
    for (uint32_t j = 0; j < n; j++)
        put_bits(out, &pos, mag32(d[j]) & ((1u << r) - 1), r);
    return sz;
}

/* --------------------------------------------------------------- streams */

/* Encode one segment (seed + blocks of Lorenzo deltas) of q into out.
 * Deltas chain across block boundaries and restart only at the seed
 * (ZCCL/src/ZCCL_float.c:563-568, 576-604).  Returns bytes written. */
static uint64_t encode_segment(const zo_geom *g, int s, const int32_t *q, uint8_t *out)
{
    uint64_t a, len;
    seg_range(g, s, &a, &len);
    int32_t seed = q[a];
    memcpy(out, &seed, 4);
    uint64_t pos = 4;
    uint64_t nb = seg_nblocks(g, s);
    int32_t *d = (int32_t *) malloc((size_t) g->L * sizeof(int32_t));
    for (uint64_t j = 0; j < nb; j++) {
        uint32_t n = block_len(g, len, j);
        uint64_t e0 = a + 1 + j * (uint64_t) g->L;
        for (uint32_t t = 0; t < n; t++)
            d[t] = (int32_t) ((uint32_t) q[e0 + t] - (uint32_t) q[e0 + t - 1]);
        pos += zo_encode_block(d, n, out + pos);
    }
    free(d);
    return pos;
}

/* Assemble per-segment payloads into the wire format: T u64 chunk offsets
 * (relative to the end of the header), chunk payloads in order, the remainder
 * segment appended to chunk T-1.  ZCCL/src/ZCCL_float.c:777-791. */
static int64_t assemble(const zo_geom *g, uint8_t **segbuf, const uint64_t *segsz, uint8_t *out,
    uint64_t cap)
{
    uint64_t total = 8 * (uint64_t) g->T;
    for (int s = 0; s < g->S; s++)
        total += segsz[s];
    if (total > cap)
        return ZO_ENOSPC;
    uint64_t off = 0;
    for (int t = 0; t < g->T; t++) {
        memcpy(out + 8 * (uint64_t) t, &off, 8);
        off += segsz[t];
    }
    uint64_t pos = 8 * (uint64_t) g->T;
    for (int s = 0; s < g->S; s++) {
        memcpy(out + pos, segbuf[s], segsz[s]);
        pos += segsz[s];
    }
    return (int64_t) total;
}

static uint64_t seg_worst(const zo_geom *g, int s)
{
    uint64_t a, len;
    seg_range(g, s, &a, &len);
    return 4 + (len + seg_nblocks(g, s)) * 5 + 16;
}

int64_t zo_encode_q(const int32_t *q, uint64_t n, int L, int T, uint8_t *out, uint64_t cap)
{
    zo_geom g;
    if (zo_geometry(n, L, T, &g) != ZO_OK)
        return ZO_EINVAL;
    uint8_t **segbuf = (uint8_t **) calloc((size_t) g.S, sizeof(uint8_t *));
    uint64_t *segsz = (uint64_t *) calloc((size_t) g.S, sizeof(uint64_t));
#pragma omp parallel for schedule(dynamic, 1)
    for (int s = 0; s < g.S; s++) {
        segbuf[s] = (uint8_t *) malloc((size_t) seg_worst(&g, s));
        segsz[s] = encode_segment(&g, s, q, segbuf[s]);
    }
    int64_t rc = assemble(&g, segbuf, segsz, out, cap);
    for (int s = 0; s < g.S; s++)
        free(segbuf[s]);
    free(segbuf);
    free(segsz);
    return rc;
}

int64_t zo_compress(const float *x, uint64_t n, float eb, int L, int T, uint8_t *out, uint64_t cap)
{
    zo_geom g;
    if (zo_geometry(n, L, T, &g) != ZO_OK)
        return ZO_EINVAL;
    int32_t *q = (int32_t *) malloc((size_t) n * sizeof(int32_t));
    zo_quantize(x, n, eb, q);
    int64_t rc = zo_encode_q(q, n, L, T, out, cap);
    free(q);
    return rc;
}

/* Walk one segment starting at byte pos: seed then blocks.  Optionally
 * records widths / block offsets (global block index g0 onward) and the
 * integer values (running prefix of deltas from the seed,
 * ZCCL/src/ZCCLd_float.c:581-628).  Returns the byte position after the
 * segment or a negative status. */
static int64_t walk_segment(const zo_geom *g, int s, const uint8_t *in, uint64_t in_size,
    uint64_t pos, uint64_t g0, uint8_t *widths, uint64_t *block_off, int32_t *q)
{
    uint64_t a, len;
    seg_range(g, s, &a, &len);
    if (pos + 4 > in_size)
        return ZO_ECORRUPT;
    int32_t prior;
    memcpy(&prior, in + pos, 4);
    pos += 4;
    if (q)
        q[a] = prior;
    uint64_t nb = seg_nblocks(g, s);
    int32_t *d = (int32_t *) malloc((size_t) g->L * sizeof(int32_t));
    for (uint64_t j = 0; j < nb; j++) {
        uint32_t n = block_len(g, len, j);
        if (widths)
            widths[g0 + j] = pos < in_size ? in[pos] : 0;
        if (block_off)
            block_off[g0 + j] = pos;
        int64_t used = zo_decode_block(in + pos, pos < in_size ? in_size - pos : 0, n, d);
        if (used < 0) {
            free(d);
            return used;
        }
        if (q) {
            uint64_t e0 = a + 1 + j * (uint64_t) g->L;
            for (uint32_t t = 0; t < n; t++) {
                prior = (int32_t) ((uint32_t) prior + (uint32_t) d[t]);
                q[e0 + t] = prior;
            }
        }
        pos += (uint64_t) used;
    }
    free(d);
    return (int64_t) pos;
}

static uint64_t seg_first_block(const zo_geom *g, int s) { return (uint64_t) s * g->nbT; }

int zo_parse(const uint8_t *in, uint64_t in_size, uint64_t n, int L, int T, uint8_t *widths,
    uint64_t *block_off, uint64_t *seed_off, uint64_t *total)
{
    zo_geom g;
    if (zo_geometry(n, L, T, &g) != ZO_OK)
        return ZO_EINVAL;
    if (in_size < 8 * (uint64_t) T)
        return ZO_ECORRUPT;
    int rc = ZO_OK;
    uint64_t end_last = 0;
#pragma omp parallel for schedule(dynamic, 1) reduction(min : rc)
    for (int t = 0; t < g.T; t++) {
        uint64_t h;
        memcpy(&h, in + 8 * (uint64_t) t, 8);
        uint64_t pos = 8 * (uint64_t) g.T + h;
        if (seed_off)
            seed_off[t] = pos;
        int64_t e = walk_segment(&g, t, in, in_size, pos, seg_first_block(&g, t), widths, block_off,
            NULL);
        if (e >= 0 && t == g.T - 1 && g.R > 0) {
            if (seed_off)
                seed_off[g.T] = (uint64_t) e;
            e = walk_segment(&g, g.T, in, in_size, (uint64_t) e, seg_first_block(&g, g.T), widths,
                block_off, NULL);
        }
        if (e < 0)
            rc = (int) e;
        else if (t == g.T - 1)
            end_last = (uint64_t) e;
    }
    if (total)
        *total = end_last;
    return rc;
}

int zo_decode_q(const uint8_t *in, uint64_t in_size, uint64_t n, int L, int T, int32_t *q)
{
    zo_geom g;
    if (zo_geometry(n, L, T, &g) != ZO_OK)
        return ZO_EINVAL;
    if (in_size < 8 * (uint64_t) T)
        return ZO_ECORRUPT;
    int rc = ZO_OK;
#pragma omp parallel for schedule(dynamic, 1) reduction(min : rc)
    for (int t = 0; t < g.T; t++) {
        uint64_t h;
        memcpy(&h, in + 8 * (uint64_t) t, 8);
        int64_t e = walk_segment(&g, t, in, in_size, 8 * (uint64_t) g.T + h, 0, NULL, NULL, q);
        if (e >= 0 && t == g.T - 1 && g.R > 0)
            e = walk_segment(&g, g.T, in, in_size, (uint64_t) e, 0, NULL, NULL, q);
        if (e < 0)
            rc = (int) e;
    }
    return rc;
}

int zo_decompress(const uint8_t *in, uint64_t in_size, uint64_t n, float eb, int L, int T, float *x)
{
    int32_t *q = (int32_t *) malloc((size_t) n * sizeof(int32_t));
    int rc = zo_decode_q(in, in_size, n, L, T, q);
    if (rc == ZO_OK)
        zo_dequantize(q, n, eb, x);
    free(q);
    return rc;
}

/* The decompress-operate-compress (DOC) reduction of one segment, the
 * reference's non-homomorphic ring MPI_Allreduce_ZCCL_RI2_mt_oa_record
 * (ZCCL/src/ZCCL_ring.c:979-1108): the first rank's values travel
 * compressed (:1049-1055), each receiver decompresses them (:1066-1067) and
 * adds its own values in FP32 (MPI_Reduce_local, MPI_SUM, :1068-1069) before
 * compressing the partial sum for the next rank.  xs[0..P-1] are the P ranks'
 * copies of the segment in ring order (xs[0] starts the chain); acc receives
 * the last rank's FP32 partial sum.  final, if not null, receives what every
 * rank holds after the compressed allgather (:1094-1105, :814-978): the
 * decompression of acc's compression. */
int zo_doc_chain(const float *const *xs, int P, uint64_t n, float eb, int L, int T, float *acc, float *final)
{
    zo_geom g;
    if (P < 1 || zo_geometry(n, L, T, &g) != ZO_OK)
        return ZO_EINVAL;
    const uint64_t cap = zo_worst_case_bytes(n, L, T);
    uint8_t *buf = (uint8_t *) malloc((size_t) cap);
    float *dec = (float *) malloc((size_t) n * sizeof(float));
    int rc = ZO_OK;
    memcpy(acc, xs[0], (size_t) n * sizeof(float));
    for (int m = 1; m < P && rc == ZO_OK; m++) {
        int64_t sz = zo_compress(acc, n, eb, L, T, buf, cap);
        if (sz < 0) {
            rc = (int) sz;
            break;
        }
        rc = zo_decompress(buf, (uint64_t) sz, n, eb, L, T, dec);
        for (uint64_t i = 0; rc == ZO_OK && i < n; i++)
            acc[i] = dec[i] + xs[m][i];
    }
    if (rc == ZO_OK && final) {
        int64_t sz = zo_compress(acc, n, eb, L, T, buf, cap);
        rc = sz < 0 ? (int) sz : zo_decompress(buf, (uint64_t) sz, n, eb, L, T, final);
    }
    free(dec);
    free(buf);
    return rc;
}

/* k-operand homomorphic sum.  The reference adds two streams block by block
 * (ZCCL/src/ZCCLd_float.c:1072 seeds, :1099-1219 blocks: both zero -> zero
 * block; one zero -> copy the other's bytes; else decode, add in int32,
 * recompute width, re-encode).  Every case yields the canonical encoding of
 * the summed deltas, so the k-ary form below decodes each operand's block,
 * accumulates in uint32 (int32 wrap) and re-encodes. */
int64_t zo_homsum(const uint8_t *const *ins, const uint64_t *in_sizes, int k, uint64_t n, int L,
    int T, uint8_t *out, uint64_t cap)
{
    zo_geom g;
    if (k < 1 || zo_geometry(n, L, T, &g) != ZO_OK)
        return ZO_EINVAL;
    for (int r = 0; r < k; r++)
        if (in_sizes[r] < 8 * (uint64_t) T)
            return ZO_ECORRUPT;
    uint8_t **segbuf = (uint8_t **) calloc((size_t) g.S, sizeof(uint8_t *));
    uint64_t *segsz = (uint64_t *) calloc((size_t) g.S, sizeof(uint64_t));
    int rc = ZO_OK;
#pragma omp parallel for schedule(dynamic, 1) reduction(min : rc)
    for (int t = 0; t < g.T; t++) {
        int lrc = ZO_OK;
        uint64_t *pos = (uint64_t *) malloc((size_t) k * sizeof(uint64_t));
        int32_t *d = (int32_t *) malloc((size_t) L * sizeof(int32_t));
        uint32_t *acc = (uint32_t *) malloc((size_t) L * sizeof(uint32_t));
        for (int r = 0; r < k; r++) {
            uint64_t h;
            memcpy(&h, ins[r] + 8 * (uint64_t) t, 8);
            pos[r] = 8 * (uint64_t) T + h;
        }
        int last = (t == g.T - 1);
        for (int s = t; s <= (last ? g.S - 1 : t) && lrc == ZO_OK; s++) {
            uint64_t a, len;
            seg_range(&g, s, &a, &len);
            uint8_t *ob = (uint8_t *) malloc((size_t) seg_worst(&g, s));
            segbuf[s] = ob;
            uint32_t seed = 0;
            for (int r = 0; r < k; r++) {
                int32_t v;
                if (pos[r] + 4 > in_sizes[r]) {
                    lrc = ZO_ECORRUPT;
                    break;
                }
                memcpy(&v, ins[r] + pos[r], 4);
                pos[r] += 4;
                seed += (uint32_t) v;
            }
            if (lrc != ZO_OK)
                break;
            memcpy(ob, &seed, 4);
            uint64_t op = 4;
            uint64_t nb = seg_nblocks(&g, s);
            for (uint64_t j = 0; j < nb && lrc == ZO_OK; j++) {
                uint32_t bn = block_len(&g, len, j);
                memset(acc, 0, (size_t) bn * sizeof(uint32_t));
                for (int r = 0; r < k; r++) {
                    int64_t used = zo_decode_block(ins[r] + pos[r],
                        pos[r] < in_sizes[r] ? in_sizes[r] - pos[r] : 0, bn, d);
                    if (used < 0) {
                        lrc = (int) used;
                        break;
                    }
                    pos[r] += (uint64_t) used;
                    for (uint32_t e = 0; e < bn; e++)
                        acc[e] += (uint32_t) d[e];
                }
                if (lrc != ZO_OK)
                    break;
                op += zo_encode_block((const int32_t *) acc, bn, ob + op);
            }
            segsz[s] = op;
        }
        free(pos);
        free(d);
        free(acc);
        if (lrc < rc)
            rc = lrc;
    }
    int64_t res = rc;
    if (rc == ZO_OK)
        res = assemble(&g, segbuf, segsz, out, cap);
    for (int s = 0; s < g.S; s++)
        free(segbuf[s]);
    free(segbuf);
    free(segsz);
    return res;
}

/* ------------------------------------------------------------ 3-D SZp codec */

/* Factors T into 3 dims the way My_Dims_create does (ZCCL/src/ZCCL_multidim.c:
 * 100-140): the prime factors of T, largest first, each multiplied into the
 * currently smallest dim (the first one on ties); then the dims are sorted in
 * descending order. */
void zo_dims_create(int T, int dims[3])
{
    int f[64], nf = 0, rem = T;
    dims[0] = dims[1] = dims[2] = 1;
    for (int i = 2; i <= rem; i++)
        while (rem % i == 0 && nf < 64) {
            f[nf++] = i;
            rem /= i;
        }
    for (int a = 0; a < nf; a++)
        for (int b = a + 1; b < nf; b++)
            if (f[b] > f[a]) {
                int t = f[a];
                f[a] = f[b];
                f[b] = t;
            }
    for (int i = 0; i < nf; i++) {
        int m = 0;
        for (int j = 1; j < 3; j++)
            if (dims[j] < dims[m])
                m = j;
        dims[m] *= f[i];
    }
    for (int a = 0; a < 3; a++)
        for (int b = a + 1; b < 3; b++)
            if (dims[b] > dims[a]) {
                int t = dims[a];
                dims[a] = dims[b];
                dims[b] = t;
            }
}

/* Sub-volume t of the T-way split: the first (d mod td) parts along an axis get
 * one more plane; parts are numbered x-major (ZCCL/src/ZCCL_float.c:1901-1935).
 * ZO_EINVAL when a part would be empty. */
int zo_partition3d(uint64_t d1, uint64_t d2, uint64_t d3, int T, zo_part3 *parts)
{
    int td[3];
    if (T < 1 || d1 < 1 || d2 < 1 || d3 < 1)
        return ZO_EINVAL;
    zo_dims_create(T, td);
    if (d1 < (uint64_t) td[0] || d2 < (uint64_t) td[1] || d3 < (uint64_t) td[2])
        return ZO_EINVAL;
    uint64_t bx = d1 / td[0], rx = d1 % td[0], by = d2 / td[1], ry = d2 % td[1], bz = d3 / td[2], rz = d3 % td[2];
    int r = 0;
    uint64_t ox = 0;
    for (int px = 0; px < td[0]; px++) {
        uint64_t sx = bx + ((uint64_t) px < rx ? 1 : 0), oy = 0;
        for (int py = 0; py < td[1]; py++) {
            uint64_t sy = by + ((uint64_t) py < ry ? 1 : 0), oz = 0;
            for (int pz = 0; pz < td[2]; pz++) {
                uint64_t sz = bz + ((uint64_t) pz < rz ? 1 : 0);
                zo_part3 p = {ox, oy, oz, sx, sy, sz};
                parts[r++] = p;
                oz += sz;
            }
            oy += sy;
        }
        ox += sx;
    }
    return ZO_OK;
}

/* q = (int)floor((double)x * (0.5 / eb) + 0.5), product and sum rounded
 * separately (SZ_quantize, ZCCL/include/ZCCL_float.h:57-61; the reference is
 * built without FMA).  Out-of-range and NaN give INT_MIN as x86 cvttsd2si. */
static inline int32_t quant3(float x, double inv)
{
    double p = (double) x * inv;
    p = floor(p + 0.5);
    if (p >= -2147483648.0 && p < 2147483648.0)
        return (int32_t) p;
    return INT32_MIN;
}

void zo_quantize3d(const float *x, uint64_t n, double eb, int32_t *q)
{
    double inv = 0.5 / eb;
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < (int64_t) n; i++)
        q[i] = quant3(x[i], inv);
}

/* x' = (float)((double)q * (2 * eb)) (ZCCL/src/ZCCLd_float.c:2885, 3030). */
void zo_dequantize3d(const int32_t *q, uint64_t n, double eb, float *x)
{
    double tw = 2 * eb;
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < (int64_t) n; i++)
        x[i] = (float) ((double) q[i] * tw);
}

static uint64_t nblk(uint64_t s, int Bs) { return (s - 1) / (uint64_t) Bs + 1; }

uint64_t zo_worst_case_bytes3d(uint64_t d1, uint64_t d2, uint64_t d3, int Bs, int T)
{
    zo_part3 *parts = (zo_part3 *) malloc(sizeof(zo_part3) * (size_t) (T > 0 ? T : 1));
    if (Bs < 1 || zo_partition3d(d1, d2, d3, T, parts) != ZO_OK) {
        free(parts);
        return 0;
    }
    uint64_t tot = 8ull * T;
    for (int t = 0; t < T; t++) {
        const zo_part3 *p = &parts[t];
        uint64_t n = p->sx * p->sy * p->sz;
        tot += nblk(p->sx, Bs) * nblk(p->sy, Bs) * nblk(p->sz, Bs);  /* width bytes */
        tot += 4 * n;                                                /* B = 4 high bytes */
        /* sign bytes: at most one partial byte per block beyond n/8 */
        tot += n / 8 + nblk(p->sx, Bs) * nblk(p->sy, Bs) * nblk(p->sz, Bs);
    }
    free(parts);
    return tot;
}

/* Residual of sub-volume p at local (lx, ly, lz): the 3-D difference
 * dx dy dz q with q = 0 outside the sub-volume.  The reference forms it with
 * running buffers (prevLeft, prevRow, prevSlice; ZCCL/src/ZCCL_float.c:
 * 1967-1997): err_1 = q - q(z-1), err_2 = err_1 - err_1(y-1),
 * err_3 = err_2 - err_2(x-1), in 32-bit wrap-around arithmetic. */
static int32_t resid3(const int32_t *Q, uint64_t d2, uint64_t d3, const zo_part3 *p, uint64_t lx, uint64_t ly,
    uint64_t lz)
{
    uint32_t acc = 0;
    for (int a = 0; a < 2; a++)
        for (int b = 0; b < 2; b++)
            for (int c = 0; c < 2; c++) {
                if ((a && lx == 0) || (b && ly == 0) || (c && lz == 0))
                    continue;
                uint64_t X = p->ox + lx - a, Y = p->oy + ly - b, Z = p->oz + lz - c;
                uint32_t v = (uint32_t) Q[(X * d2 + Y) * d3 + Z];
                acc += ((a + b + c) & 1) ? 0u - v : v;
            }
    return (int32_t) acc;
}

/* SZp_compress3D_fast_openmp (ZCCL/src/ZCCL_float.c:1862-2048) with a team of
 * T: header of T u64 payload offsets (relative to the end of the header),
 * then per sub-volume the width byte of every block (32 - clz(max |r|), 0
 * when all residuals are 0) followed by the sign bitmap and fixed-length bits
 * of every block of nonzero width, blocks in x, y, z order and values in
 * (i, j, k) order inside a block. */
int64_t zo_compress3d(const float *x, uint64_t d1, uint64_t d2, uint64_t d3, int Bs, double eb, int T,
    uint8_t *out, uint64_t cap)
{
    if (Bs < 1)
        return ZO_EINVAL;
    zo_part3 *parts = (zo_part3 *) malloc(sizeof(zo_part3) * (size_t) (T > 0 ? T : 1));
    if (zo_partition3d(d1, d2, d3, T, parts) != ZO_OK) {
        free(parts);
        return ZO_EINVAL;
    }
    uint64_t need = zo_worst_case_bytes3d(d1, d2, d3, Bs, T);
    uint64_t n = d1 * d2 * d3;
    int32_t *Q = (int32_t *) malloc(sizeof(int32_t) * n);
    int32_t *res = (int32_t *) malloc(sizeof(int32_t) * (size_t) Bs * Bs * Bs);
    uint8_t *tmp = (uint8_t *) malloc(need);
    zo_quantize3d(x, n, eb, Q);
    uint64_t pos = 8ull * T;
    for (int t = 0; t < T; t++) {
        const zo_part3 *p = &parts[t];
        uint64_t b1 = nblk(p->sx, Bs), b2 = nblk(p->sy, Bs), b3 = nblk(p->sz, Bs);
        uint8_t *wid = tmp + pos;
        uint64_t q = pos + b1 * b2 * b3;
        uint64_t bi = 0;
        for (uint64_t bx = 0; bx < b1; bx++)
            for (uint64_t by = 0; by < b2; by++)
                for (uint64_t bz = 0; bz < b3; bz++, bi++) {
                    uint64_t ex = p->sx - bx * Bs < (uint64_t) Bs ? p->sx - bx * Bs : (uint64_t) Bs;
                    uint64_t ey = p->sy - by * Bs < (uint64_t) Bs ? p->sy - by * Bs : (uint64_t) Bs;
                    uint64_t ez = p->sz - bz * Bs < (uint64_t) Bs ? p->sz - bz * Bs : (uint64_t) Bs;
                    uint32_t m = 0, cnt = 0;
                    for (uint64_t i = 0; i < ex; i++)
                        for (uint64_t j = 0; j < ey; j++)
                            for (uint64_t k = 0; k < ez; k++) {
                                int32_t r = resid3(Q, d2, d3, p, bx * Bs + i, by * Bs + j, bz * Bs + k);
                                res[cnt++] = r;
                                if (mag32(r) > m)
                                    m = mag32(r);
                            }
                    int w = m ? 32 - __builtin_clz(m) : 0;
                    wid[bi] = (uint8_t) w;
                    if (w)
                        q += pack_fields(res, cnt, w, tmp + q);
                }
        uint64_t off = pos - 8ull * T;
        memcpy(tmp + 8ull * t, &off, 8);
        pos = q;
    }
    int64_t ret = (int64_t) pos;
    if (pos > cap)
        ret = ZO_ENOSPC;
    else
        memcpy(out, tmp, pos);
    free(tmp);
    free(res);
    free(Q);
    free(parts);
    return ret;
}

/* SZp_decompress3D_fast_openmp (ZCCL/src/ZCCLd_float.c:2873-3040): residuals
 * of every block, then per sub-volume the inclusive prefix sum over x, y and z
 * (the reference's colSum / row_acc / prefix buffers, :2996-3027), then
 * x' = (float)((double)q * 2eb). */
int zo_decompress3d(const uint8_t *in, uint64_t in_size, uint64_t d1, uint64_t d2, uint64_t d3, int Bs,
    double eb, int T, float *x)
{
    if (Bs < 1)
        return ZO_EINVAL;
    zo_part3 *parts = (zo_part3 *) malloc(sizeof(zo_part3) * (size_t) (T > 0 ? T : 1));
    if (zo_partition3d(d1, d2, d3, T, parts) != ZO_OK || in_size < 8ull * T) {
        free(parts);
        return ZO_EINVAL;
    }
    int rc = ZO_OK;
    int32_t *res = (int32_t *) malloc(sizeof(int32_t) * (size_t) Bs * Bs * Bs);
    double tw = 2 * eb;
    for (int t = 0; t < T && rc == ZO_OK; t++) {
        const zo_part3 *p = &parts[t];
        uint64_t off;
        memcpy(&off, in + 8ull * t, 8);
        uint64_t b1 = nblk(p->sx, Bs), b2 = nblk(p->sy, Bs), b3 = nblk(p->sz, Bs), nb = b1 * b2 * b3;
        uint64_t pos = 8ull * T + off;
        if (pos + nb > in_size) {
            rc = ZO_ECORRUPT;
            break;
        }
        const uint8_t *wid = in + pos;
        pos += nb;
        uint64_t nsub = p->sx * p->sy * p->sz;
        uint32_t *R = (uint32_t *) calloc(nsub, sizeof(uint32_t));
        uint64_t bi = 0;
        for (uint64_t bx = 0; bx < b1 && rc == ZO_OK; bx++)
            for (uint64_t by = 0; by < b2 && rc == ZO_OK; by++)
                for (uint64_t bz = 0; bz < b3; bz++, bi++) {
                    int w = wid[bi];
                    if (!w)
                        continue;
                    uint64_t ex = p->sx - bx * Bs < (uint64_t) Bs ? p->sx - bx * Bs : (uint64_t) Bs;
                    uint64_t ey = p->sy - by * Bs < (uint64_t) Bs ? p->sy - by * Bs : (uint64_t) Bs;
                    uint64_t ez = p->sz - bz * Bs < (uint64_t) Bs ? p->sz - bz * Bs : (uint64_t) Bs;
                    uint32_t cnt = (uint32_t) (ex * ey * ez);
                    uint64_t sz = zo_block_bytes(w, cnt) - 1;
                    if (w > 32 || pos + sz > in_size) {
                        rc = ZO_ECORRUPT;
                        break;
                    }
                    unpack_fields(in + pos, cnt, w, res);
                    pos += sz;
                    uint32_t c = 0;
                    for (uint64_t i = 0; i < ex; i++)
                        for (uint64_t j = 0; j < ey; j++)
                            for (uint64_t k = 0; k < ez; k++)
                                R[((bx * Bs + i) * p->sy + by * Bs + j) * p->sz + bz * Bs + k] = (uint32_t) res[c++];
                }
        if (rc == ZO_OK) {
            /* inclusive prefix sums along z, then y, then x (wrap-around) */
            for (uint64_t i = 0; i < p->sx; i++)
                for (uint64_t j = 0; j < p->sy; j++)
                    for (uint64_t k = 1; k < p->sz; k++)
                        R[(i * p->sy + j) * p->sz + k] += R[(i * p->sy + j) * p->sz + k - 1];
            for (uint64_t i = 0; i < p->sx; i++)
                for (uint64_t j = 1; j < p->sy; j++)
                    for (uint64_t k = 0; k < p->sz; k++)
                        R[(i * p->sy + j) * p->sz + k] += R[(i * p->sy + j - 1) * p->sz + k];
            for (uint64_t i = 1; i < p->sx; i++)
                for (uint64_t j = 0; j < p->sy; j++)
                    for (uint64_t k = 0; k < p->sz; k++)
                        R[(i * p->sy + j) * p->sz + k] += R[((i - 1) * p->sy + j) * p->sz + k];
            for (uint64_t i = 0; i < p->sx; i++)
                for (uint64_t j = 0; j < p->sy; j++)
                    for (uint64_t k = 0; k < p->sz; k++)
                        x[((p->ox + i) * d2 + p->oy + j) * d3 + p->oz + k] =
                            (float) ((double) (int32_t) R[(i * p->sy + j) * p->sz + k] * tw);
        }
        free(R);
    }
    free(res);
    free(parts);
    return rc;
}
