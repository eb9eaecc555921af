/*
 * width_probe.c -- exhaustive probe of the block bit width the reference
 * stores, (int)log2f((float)m) + 1 (ZCCL/src/ZCCL_float.c:606), over every
 * m in [1, 2^32), against the system libm.
 *
 * For every w = 32 - clz(m) it reports the smallest m whose stored width is
 * w+1 (or none), and checks that the stored width is w or w+1 and monotone in
 * m.  Output (JSON on stdout) is committed as tests/golden/width_thresholds.json
 * and pins zo_width() in the oracle and the device width function.
 *
 * Build: gcc -O2 -fopenmp width_probe.c -o width_probe -lm
 */
#include <math.h>
#include <stdint.h>
#include <stdio.h>

static int ref_width(uint32_t m) { return (int) (log2f((float) m)) + 1; }

int main(void)
{
    uint64_t first_bump[33];
    int bad = 0;
    for (int w = 0; w <= 32; w++)
        first_bump[w] = 0;
    for (int w = 1; w <= 32; w++) {
        uint64_t lo = (uint64_t) 1 << (w - 1), hi = (uint64_t) 1 << w;
        uint64_t fb = hi; /* smallest m in [lo, hi) with width w+1 */
        int local_bad = 0;
#pragma omp parallel for reduction(min : fb) reduction(+ : local_bad) schedule(static)
        for (int64_t m = (int64_t) lo; m < (int64_t) hi; m++) {
            int b = ref_width((uint32_t) m);
            if (b == w + 1) {
                if ((uint64_t) m < fb)
                    fb = (uint64_t) m;
            } else if (b != w) {
                local_bad++;
            }
        }
        /* monotone: every m >= fb in the octave must be w+1 */
        if (fb < hi) {
#pragma omp parallel for reduction(+ : local_bad) schedule(static)
            for (int64_t m = (int64_t) fb; m < (int64_t) hi; m++)
                if (ref_width((uint32_t) m) != w + 1)
                    local_bad++;
        }
        first_bump[w] = fb;
        bad += local_bad;
    }
    printf("{\n  \"libm\": \"glibc log2f\",\n  \"monotone_and_w_or_w_plus_1\": %s,\n",
        bad == 0 ? "true" : "false");
    printf("  \"bump_c\": {");
    int first = 1;
    for (int w = 1; w <= 32; w++) {
        uint64_t hi = (uint64_t) 1 << w;
        if (first_bump[w] < hi) {
            printf("%s\"%d\": %llu", first ? "" : ", ", w, (unsigned long long) (hi - first_bump[w]));
            first = 0;
        }
    }
    printf("}\n}\n");
    return bad != 0;
}
