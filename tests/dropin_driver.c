/* A C caller of the drop-in entry points, shaped like the reference's
 * examples/testfloat_homocompress_fastmode1_arg.c:134-282: compress two
 * fields into 4N-byte caller buffers, add them homomorphically (the size is
 * incremented, so the caller zeroes it first), decompress.  Built with gcc
 * against include/zccl_amd.h and linked to libzccl_amd.so; the GPU test
 * tests/test_gpu_dropin_c.py checks what it writes against the oracle.
 *
 *   dropin_driver N blockSize nchunks out_prefix
 *
 * writes out_prefix.{a,b}.f32 (inputs), .ca, .cb, .cs (streams), .ds.f32
 * (decompressed sum) and .meta (eb and sizes, text). */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "zccl_amd.h"

static int write_file(const char *prefix, const char *ext, const void *p, size_t n)
{
    char path[4096];
    snprintf(path, sizeof(path), "%s.%s", prefix, ext);
    FILE *f = fopen(path, "wb");
    if (!f)
        return -1;
    size_t w = fwrite(p, 1, n, f);
    fclose(f);
    return w == n ? 0 : -1;
}

static int check(const char *what)
{
    int e = zccl_last_error();
    if (e) {
        fprintf(stderr, "%s: zccl_last_error %d\n", what, e);
        return 1;
    }
    return 0;
}

int main(int argc, char **argv)
{
    if (argc != 5) {
        fprintf(stderr, "usage: %s N blockSize nchunks out_prefix\n", argv[0]);
        return 2;
    }
    const size_t n = (size_t) strtoull(argv[1], NULL, 10);
    const int L = atoi(argv[2]);
    const int T = atoi(argv[3]);
    const char *pre = argv[4];
    zccl_set_num_chunks(T);
    float *a = malloc(n * sizeof(float)), *b = malloc(n * sizeof(float));
    float *ds = calloc(n, sizeof(float));
    unsigned char *ca = calloc(4 * n, 1), *cb = calloc(4 * n, 1), *cs = calloc(4 * n, 1);
    if (!a || !b || !ds || !ca || !cb || !cs)
        return 3;
    /* two smooth fields with a little deterministic noise */
    unsigned int s = 12345u;
    float lo = 0.0f, hi = 0.0f;
    for (size_t i = 0; i < n; i++) {
        s = s * 1664525u + 1013904223u;
        const float noise = ((float) (s >> 8) / 16777216.0f - 0.5f) * 1e-3f;
        a[i] = sinf((float) i * 1e-3f) + 0.5f * sinf((float) i * 3.7e-5f) + noise;
        b[i] = 0.8f * cosf((float) i * 2e-3f) + noise * 0.5f;
        const float m = a[i] > b[i] ? a[i] : b[i], k = a[i] < b[i] ? a[i] : b[i];
        if (i == 0 || m > hi)
            hi = m;
        if (i == 0 || k < lo)
            lo = k;
    }
    const float eb = (hi - lo) * 1e-3f;
    size_t sa = 0, sb = 0, ss = 0;
    ZCCL_float_openmp_threadblock_arg(ca, a, &sa, eb, n, L);
    if (check("compress a"))
        return 4;
    ZCCL_float_openmp_threadblock_arg(cb, b, &sb, eb, n, L);
    if (check("compress b"))
        return 4;
    ZCCL_float_homomophic_add_openmp_threadblock(cs, &ss, n, eb, L, ca, cb);
    if (check("homomorphic add"))
        return 4;
    ZCCL_float_decompress_openmp_threadblock_arg(ds, n, eb, L, cs);
    if (check("decompress"))
        return 4;
    double worst = 0.0;
    for (size_t i = 0; i < n; i++) {
        const double e = fabs((double) ds[i] - ((double) a[i] + (double) b[i]));
        if (e > worst)
            worst = e;
    }
    char meta[256];
    snprintf(meta, sizeof(meta), "%.9g %zu %zu %zu %.9g\n", (double) eb, sa, sb, ss, worst);
    if (write_file(pre, "a.f32", a, n * 4) || write_file(pre, "b.f32", b, n * 4) || write_file(pre, "ca", ca, sa) ||
        write_file(pre, "cb", cb, sb) || write_file(pre, "cs", cs, ss) || write_file(pre, "ds.f32", ds, n * 4) ||
        write_file(pre, "meta", meta, strlen(meta)))
        return 5;
    printf("dropin_driver n=%zu L=%d T=%d eb=%g sizes %zu %zu %zu max err %g\n", n, L, T, (double) eb, sa, sb, ss,
        worst);
    free(a);
    free(b);
    free(ds);
    free(ca);
    free(cb);
    free(cs);
    return 0;
}
