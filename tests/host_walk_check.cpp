// host_walk_check.cpp -- the drop-in layer's host stream walk
// (zccl_internal.h host_stream_bytes) on oracle streams held in heap buffers
// of exactly the stream's size, built with AddressSanitizer and
// UndefinedBehaviorSanitizer by tests/test_sanitizers.py: the walk returns
// the stream size and reads no byte past it; a width byte above 32 in the
// last chunk makes it return 0.  Prints "ok <cases>".
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "zccl_internal.h"
extern "C" {
#include "zccl_oracle.h"
}

using namespace zk;

int main()
{
    long cases = 0;
    unsigned seed = 1;
    auto rnd = [&]() {
        seed = seed * 1664525u + 1013904223u;
        return seed >> 8;
    };
    const uint64_t sizes[] = {1, 2, 7, 36, 37, 100, 1000, 4097, 65537, 300001};
    const int Ls[] = {1, 7, 8, 32, 36, 128};
    const int Ts[] = {1, 2, 3, 8, 16};
    for (uint64_t n : sizes)
        for (int L : Ls)
            for (int T : Ts) {
                if (n < (uint64_t) T)
                    continue;
                Geom g;
                if (!make_geom(n, (uint32_t) L, (uint32_t) T, &g))
                    continue;
                std::vector<float> x(n);
                const float amp = (float) (1u << (rnd() % 24));
                for (uint64_t i = 0; i < n; i++)
                    x[i] = amp * std::sin(0.01f * (float) i) + (float) (rnd() % 1000) * 1e-3f;
                const float eb = 1e-3f * (float) (1 + rnd() % 100);
                const uint64_t cap = zo_worst_case_bytes(n, L, T);
                std::vector<uint8_t> tmp(cap);
                const int64_t sz = zo_compress(x.data(), n, eb, L, T, tmp.data(), cap);
                if (sz <= 0) {
                    std::printf("FAIL oracle compress n=%llu L=%d T=%d\n", (unsigned long long) n, L, T);
                    return 1;
                }
                // exactly the stream's bytes on the heap: any read past them is an ASan error
                uint8_t* exact = (uint8_t*) std::malloc((size_t) sz);
                std::memcpy(exact, tmp.data(), (size_t) sz);
                const uint64_t got = host_stream_bytes(exact, g);
                if (got != (uint64_t) sz) {
                    std::printf("FAIL walk n=%llu L=%d T=%d: %llu != %lld\n", (unsigned long long) n, L, T,
                        (unsigned long long) got, (long long) sz);
                    return 1;
                }
                // a width byte above 32 in the last chunk's first block stops the walk
                uint64_t h = 0;
                std::memcpy(&h, exact + 8ull * (T - 1), 8);
                const uint64_t first = 8ull * T + h + 4;
                if (first < (uint64_t) sz && g.nbT > 0) {
                    const uint8_t keep = exact[first];
                    exact[first] = (uint8_t) (33 + rnd() % 200);
                    if (host_stream_bytes(exact, g) != 0) {
                        std::printf("FAIL corrupt width accepted n=%llu L=%d T=%d\n", (unsigned long long) n, L, T);
                        return 1;
                    }
                    exact[first] = keep;
                }
                std::free(exact);
                cases++;
            }
    std::printf("ok %ld\n", cases);
    return 0;
}
