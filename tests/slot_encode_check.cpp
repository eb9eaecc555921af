// slot_encode_check.cpp -- host check of the slot-layout block encoder.
//
// The sum kernel for the reference's default block length encodes each block
// with emit_unit (zccl_amd/csrc/zccl_internal.h): one writer per byte, fields
// of the next unit handed over by a lane shuffle.  This program runs the same
// functions on the CPU, unit by unit with the neighbour's fields as the
// shuffle would deliver them, over every width 0..32 (the widths the kernel
// assigns: those of the oracle's zo_width), many block lengths and random
// deltas, and compares the block bytes with the oracle's zo_encode_block.
// Every block is encoded twice into buffers pre-filled with different
// sentinels: equal results show every byte of the block was written, and the
// bytes around the block must keep their sentinel.  Built and run by
// tests/test_bounds.py; prints "ok <blocks>" or the first difference.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "zccl_internal.h"
#include "zccl_oracle.h"

using namespace zk;

static uint64_t g_state = 0x9E3779B97F4A7C15ull;
static uint32_t rnd32()
{
    g_state = g_state * 6364136223846793005ull + 1442695040888963407ull;
    return (uint32_t) (g_state >> 32);
}

// Deltas of one block whose largest magnitude has exactly `bits` bits
// (bits = 32: magnitudes up to 2^31, INT_MIN included).
static void make_block(uint32_t nj, uint32_t bits, std::vector<int32_t>& d)
{
    d.assign(nj, 0);
    if (bits == 0)
        return;
    const uint64_t top = 1ull << (bits - 1);
    for (uint32_t i = 0; i < nj; i++) {
        uint64_t m = (uint64_t) rnd32() & ((top << 1) - 1);
        if (rnd32() % 4 == 0)
            m = 0;
        if (m > (1ull << 31))
            m = 1ull << 31;
        const bool neg = rnd32() & 1;
        d[i] = (int32_t) (neg ? 0u - (uint32_t) m : (uint32_t) m);
    }
    // one value at the top of the range so the width is exactly `bits`
    const uint32_t k = rnd32() % nj;
    const uint64_t m = bits == 32 ? (1ull << 31) : (top | ((uint64_t) rnd32() & (top - 1)));
    d[k] = (int32_t) ((rnd32() & 1) ? 0u - (uint32_t) m : (uint32_t) m);
}

static void encode_slots(const std::vector<int32_t>& d, uint32_t nj, uint32_t U, uint32_t w, uint8_t* buf,
    uint32_t bo)
{
    std::vector<UnitFields> f(U);
    std::vector<uint32_t> cnt(U);
    for (uint32_t u = 0; u < U; u++) {
        uint32_t v[4] = {0, 0, 0, 0};
        cnt[u] = 4 * u < nj ? (nj - 4 * u < 4 ? nj - 4 * u : 4u) : 0u;
        for (uint32_t c = 0; c < cnt[u]; c++)
            v[c] = (uint32_t) d[4 * u + c];
        f[u] = unit_fields(w, v);
    }
    for (uint32_t u = 0; u < U; u++) {
        // __shfl_down(x, 1) hands each lane the next lane's value
        const uint32_t pnib = u + 1 < U ? f[u + 1].nib : 0xFFFFFFFFu;
        const uint32_t plv = u + 1 < U ? f[u + 1].lv : 0xFFFFFFFFu;
        if (u == 0 && cnt[u])
            buf[bo] = (uint8_t) w;  // the kernel's unit-0 lane writes the width byte
        emit_unit(buf, bo + 1, w, nj, u, U, cnt[u], f[u], pnib, plv);
    }
}

int main()
{
    const uint32_t Ls[] = {36, 32, 13, 16, 17, 35, 37, 64, 100, 143, 144};
    long blocks = 0;
    std::vector<int32_t> d;
    for (uint32_t L : Ls) {
        const uint32_t U = (L + 3) / 4;
        for (uint32_t nj = 1; nj <= L; nj++) {
            for (uint32_t bits = 0; bits <= 32; bits++) {
                for (int rep = 0; rep < 6; rep++) {
                    make_block(nj, bits, d);
                    uint32_t mx = 0;
                    for (int32_t v : d)
                        mx = mx > mag32((uint32_t) v) ? mx : mag32((uint32_t) v);
                    const uint32_t w = (uint32_t) zo_width(mx);
                    std::vector<uint8_t> want(block_bytes(32, nj) + 64);
                    const uint64_t nw = zo_encode_block(d.data(), nj, want.data());
                    if (nw != block_bytes(w, nj)) {
                        std::printf("FAIL size L=%u nj=%u w=%u oracle %llu formula %u\n", L, nj, w,
                            (unsigned long long) nw, block_bytes(w, nj));
                        return 1;
                    }
                    const uint32_t bo = 8 + rnd32() % 8, cap = bo + block_bytes(32, nj) + 32;
                    std::vector<uint8_t> a(cap, 0x00), b(cap, 0xFF);
                    encode_slots(d, nj, U, w, a.data(), bo);
                    encode_slots(d, nj, U, w, b.data(), bo);
                    for (uint32_t i = 0; i < cap; i++) {
                        const bool in = i >= bo && i < bo + nw;
                        const bool ok = in ? (a[i] == want[i - bo] && b[i] == want[i - bo]) : (a[i] == 0x00 && b[i] == 0xFF);
                        if (!ok) {
                            std::printf("FAIL L=%u nj=%u w=%u byte %d (%s): got %02x/%02x want %02x\n", L, nj, w,
                                (int) i - (int) bo, in ? "inside" : "outside", a[i], b[i], in ? want[i - bo] : 0);
                            return 1;
                        }
                    }
                    blocks++;
                }
            }
        }
    }
    std::printf("ok %ld\n", blocks);
    return 0;
}
