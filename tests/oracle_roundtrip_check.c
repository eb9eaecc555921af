/* oracle_roundtrip_check.c -- the oracle's whole-stream functions on random
 * inputs in exactly-sized heap buffers, built with AddressSanitizer and
 * UndefinedBehaviorSanitizer by tests/test_sanitizers.py: compress, parse,
 * decompress (|x - x'| <= eb + 2^-21 |x|), a k-operand homomorphic sum
 * (decompresses to within k eb + 2^-21 sum |x|) and the DOC chain.  Prints
 * "ok <cases>". */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "zccl_oracle.h"

static unsigned g_seed = 7;
static unsigned rnd(void)
{
    g_seed = g_seed * 1664525u + 1013904223u;
    return g_seed >> 8;
}

static uint8_t *exact_copy(const uint8_t *p, int64_t n)
{
    uint8_t *q = malloc((size_t) n);
    memcpy(q, p, (size_t) n);
    return q;
}

int main(void)
{
    long cases = 0;
    const uint64_t sizes[] = {1, 5, 36, 37, 999, 4096, 70001};
    const int Ls[] = {1, 8, 32, 36, 128};
    const int Ts[] = {1, 3, 16};
    for (size_t si = 0; si < sizeof(sizes) / sizeof(sizes[0]); si++)
        for (size_t li = 0; li < sizeof(Ls) / sizeof(Ls[0]); li++)
            for (size_t ti = 0; ti < sizeof(Ts) / sizeof(Ts[0]); ti++) {
                const uint64_t n = sizes[si];
                const int L = Ls[li], T = Ts[ti], k = 1 + (int) (rnd() % 5);
                if (n < (uint64_t) T)
                    continue;
                const float eb = 1e-3f * (float) (1 + rnd() % 50);
                const uint64_t cap = zo_worst_case_bytes(n, L, T);
                float *xs[5];
                uint8_t *cs[5];
                uint64_t szs[5];
                for (int r = 0; r < k; r++) {
                    xs[r] = malloc(n * sizeof(float));
                    const float amp = (float) (1u << (rnd() % 20));
                    for (uint64_t i = 0; i < n; i++)
                        xs[r][i] = amp * sinf(0.003f * (float) i + (float) r) + (float) (rnd() % 100) * 0.01f;
                    uint8_t *tmp = malloc(cap);
                    const int64_t sz = zo_compress(xs[r], n, eb, L, T, tmp, cap);
                    if (sz <= 0) {
                        printf("FAIL compress\n");
                        return 1;
                    }
                    cs[r] = exact_copy(tmp, sz);
                    szs[r] = (uint64_t) sz;
                    free(tmp);
                    float *y = malloc(n * sizeof(float));
                    if (zo_decompress(cs[r], szs[r], n, eb, L, T, y)) {
                        printf("FAIL decompress\n");
                        return 1;
                    }
                    for (uint64_t i = 0; i < n; i++)
                        if (fabs((double) y[i] - (double) xs[r][i]) > eb + ldexp(fabs((double) xs[r][i]), -21)) {
                            printf("FAIL bound n=%llu L=%d T=%d i=%llu\n", (unsigned long long) n, L, T,
                                (unsigned long long) i);
                            return 1;
                        }
                    free(y);
                    {
                        zo_geom g;
                        if (zo_geometry(n, L, T, &g)) {
                            printf("FAIL geometry\n");
                            return 1;
                        }
                        uint8_t *w = malloc(g.nblocks ? g.nblocks : 1);
                        uint64_t *bo = malloc((g.nblocks ? g.nblocks : 1) * sizeof(uint64_t));
                        uint64_t *so = malloc((size_t) g.S * sizeof(uint64_t));
                        uint64_t tot = 0;
                        if (zo_parse(cs[r], szs[r], n, L, T, w, bo, so, &tot) || tot != szs[r]) {
                            printf("FAIL parse\n");
                            return 1;
                        }
                        free(w);
                        free(bo);
                        free(so);
                    }
                }
                uint8_t *tmp = malloc(cap);
                const int64_t hs = zo_homsum((const uint8_t *const *) cs, szs, k, n, L, T, tmp, cap);
                if (hs <= 0) {
                    printf("FAIL homsum\n");
                    return 1;
                }
                uint8_t *h = exact_copy(tmp, hs);
                free(tmp);
                float *y = malloc(n * sizeof(float));
                if (zo_decompress(h, (uint64_t) hs, n, eb, L, T, y)) {
                    printf("FAIL decompress sum\n");
                    return 1;
                }
                for (uint64_t i = 0; i < n; i++) {
                    double s = 0, a = 0;
                    for (int r = 0; r < k; r++) {
                        s += xs[r][i];
                        a += fabs((double) xs[r][i]);
                    }
                    if (fabs((double) y[i] - s) > k * eb + ldexp(a, -21) + ldexp(fabs(s), -22)) {
                        printf("FAIL sum bound\n");
                        return 1;
                    }
                }
                float *acc = malloc(n * sizeof(float)), *fin = malloc(n * sizeof(float));
                if (zo_doc_chain((const float *const *) xs, k, n, eb, L, T, acc, fin)) {
                    printf("FAIL doc chain\n");
                    return 1;
                }
                free(acc);
                free(fin);
                free(y);
                free(h);
                for (int r = 0; r < k; r++) {
                    free(xs[r]);
                    free(cs[r]);
                }
                cases++;
            }
    printf("ok %ld\n", cases);
    return 0;
}
