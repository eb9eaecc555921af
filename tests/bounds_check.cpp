// bounds_check.cpp -- host-side replay of the tile kernels' index arithmetic.
//
// For many stream geometries (n, L, T) and every tiling the kernels use, walk
// every tile, thread and unit exactly as zccl_kernels.hip computes them and
// check that each global, LDS and workspace index stays inside the buffer the
// host allocates for it.  Built and run by tests/test_bounds.py with g++
// against zccl_amd/csrc/zccl_internal.h; prints "ok <cases>" or the first
// violation.
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "zccl_internal.h"

using namespace zk;

static long g_cases = 0;

#define REQUIRE(cond, ...)                                                                         \
    do {                                                                                           \
        if (!(cond)) {                                                                             \
            std::printf("FAIL %s:%d: %s | ", __FILE__, __LINE__, #cond);                           \
            std::printf(__VA_ARGS__);                                                              \
            std::printf("\n");                                                                     \
            std::exit(1);                                                                          \
        }                                                                                          \
    } while (0)

// LDS bytes the launchers request (mirrors zccl_kernels.hip).
static uint64_t compress_lds(const Tiling& tl) { return align_up(tl.tile_bytes_max + 64, 16); }
static uint64_t stage_bytes(const Tiling& tl, int op)
{
    if (op != 1)
        return align_up(tl.tile_bytes_max + 64, 16);
    uint64_t need = align_up(tl.tile_bytes_max + 48, 32);
    return need < 17 * 1024 + 512 ? 17 * 1024 + 512 : need;
}

// Staging plans of the sum: for k operands of any tile size up to the worst
// case and any alignment, every operand's chunks plus slack lie inside its
// round's region (a half, or the whole stage), operands of a round do not
// overlap, and rounds cover the operands in order.
static void check_stage_plans(const Tiling& tl, int k, uint32_t seed)
{
    const uint32_t stage = (uint32_t) stage_bytes(tl, 1), half = stage / 2;
    REQUIRE(stage % 32 == 0, "stage %u", stage);
    uint32_t tot[kMaxOperands], shift[kMaxOperands];
    uint64_t st = seed * 2654435761ull + 1;
    auto rnd = [&](uint32_t m) { st = st * 6364136223846793005ull + 1442695040888963407ull; return (uint32_t) ((st >> 33) % m); };
    for (int r = 0; r < k; r++) {
        const uint32_t mode = rnd(4);
        tot[r] = mode == 0 ? tl.tile_bytes_max : mode == 1 ? 0u : rnd(tl.tile_bytes_max + 1);
        shift[r] = rnd(16);
    }
    StagePlan sp;
    plan_stage(tot, shift, k, half, sp);
    REQUIRE(sp.nrounds >= 1 && sp.nrounds <= k, "rounds %d", sp.nrounds);
    int r = 0;
    for (int g = 0; g < sp.nrounds; g++) {
        REQUIRE(sp.gend[g] > r && sp.gend[g] <= k, "round end");
        const uint32_t lo = sp.whole[g] ? 0u : (uint32_t) (g & 1) * half;
        const uint32_t hi = sp.whole[g] ? stage : lo + half;
        uint32_t prev_end = lo;
        for (; r < sp.gend[g]; r++) {
            const uint32_t need = stage_need(shift[r], tot[r]);
            REQUIRE(sp.nch[r] == stage_chunks(shift[r], tot[r]), "chunks");
            REQUIRE(tot[r] == 0 || 16 * sp.nch[r] >= shift[r] + tot[r], "chunks cover the bytes");
            REQUIRE(sp.lds[r] >= prev_end && sp.lds[r] + need <= hi, "operand %d in round %d: [%u,%u) of [%u,%u)", r,
                g, sp.lds[r], sp.lds[r] + need, lo, hi);
            REQUIRE(sp.lds[r] % 16 == 0, "aligned");
            // the decoder reads up to 16 bytes past the operand's bytes
            REQUIRE(tot[r] == 0 || sp.lds[r] + shift[r] + tot[r] + 16 <= sp.lds[r] + need, "slack");
            prev_end = sp.lds[r] + need;
        }
        if (sp.whole[g])
            REQUIRE(sp.gend[g] == (g ? sp.gend[g - 1] : 0) + 1, "a whole-stage round holds one operand");
    }
    REQUIRE(r == k, "every operand staged");
    g_cases++;
}

static void check_tiling(const Geom& g, uint32_t threads, int op)
{
    const Tiling tl = make_tiling(g, threads);
    REQUIRE(tl.threads >= 64 && tl.threads <= threads && tl.threads % 64 == 0,
        "threads %u", tl.threads);
    REQUIRE(tl.G <= kMaxTileBlocks && tl.G == tl.TG * g.IG, "G %u", tl.G);
    REQUIRE((uint64_t) tl.G * g.U <= (uint64_t) kUnitsPerThread * tl.threads, "units %u", tl.G * g.U);
    // workspace sized by the host for exactly ntiles
    const uint64_t ws_bytes = workspace_bytes(tl.ntiles);
    const uint64_t cap = workspace_capacity(ws_bytes);
    REQUIRE(cap >= tl.ntiles, "cap %lu ntiles %lu", (unsigned long) cap, (unsigned long) tl.ntiles);
    const uint64_t status_words = (ws_bytes - sizeof(WsHeader)) / 8;
    for (uint32_t par = 0; par < 2; par++) {
        // status of tile t and the cleared entries of the other array
        uint64_t hi_status = par * cap * kStatusStride + (tl.ntiles - 1) * kStatusStride;
        uint64_t hi_other = (par ^ 1u) * cap * kStatusStride + (cap - 1) * kStatusStride;
        REQUIRE(hi_status < status_words && hi_other < status_words, "status words");
    }
    if (op == 0) {
        // compress stages the tile's values and the one before them in its
        // encode buffer (up to 15 bytes of alignment shift in front); a
        // partial unit reads up to 12 bytes past them
        const uint64_t reach = 15 + 4ull * ((uint64_t) tl.G * g.L + 1) + 12;
        REQUIRE(reach <= compress_lds(tl), "compress staging %lu of %lu", (unsigned long) reach,
            (unsigned long) compress_lds(tl));
    }
    uint64_t blocks_seen = 0, groups_seen = 0;
    std::vector<uint8_t> seg_seen(g.S, 0);
    for (uint64_t t = 0; t < tl.ntiles; t++) {
        const Tile x = tile_of(g, tl, t);
        REQUIRE(x.s < g.S, "segment %u", x.s);
        REQUIRE(x.ngrp >= 1 && x.ngrp <= tl.TG, "ngrp %u", x.ngrp);
        REQUIRE(x.grp0 + x.ngrp <= g.ngroups, "groups");
        REQUIRE(x.g0 + x.nblk <= g.nblocks, "blocks");
        // widths are read as dwords of 4 blocks: stays inside the index
        REQUIRE(x.g0 + 4ull * ((x.nblk + 3) / 4) <= g.ix_bytes, "width dwords");
        REQUIRE(x.lt < x.seg_tiles && x.seg_tile0 + x.lt == t, "tile in segment");
        REQUIRE(x.nblk == 0 || (uint64_t) (x.ngrp - 1) * g.IG < x.nblk, "group/block counts");
        if (x.lt == 0)
            seg_seen[x.s] = 1;
        blocks_seen += x.nblk;
        groups_seen += x.ngrp;
        // block lengths and the worst-case tile size
        uint64_t tile_bytes = 0;
        for (uint32_t j = 0; j < x.nblk; j++) {
            uint64_t e = x.e0 + (uint64_t) j * g.L;
            REQUIRE(e >= x.seg_a + 1 && e < x.seg_end, "block start");
            uint32_t nj = tile_block_len(g, x, j);
            REQUIRE(nj >= 1 && nj <= g.L && e + nj <= x.seg_end, "block length");
            REQUIRE(nj == g.L || j == x.nblk - 1, "only the last block may be partial");
            tile_bytes += block_bytes(32, nj);
        }
        REQUIRE(tile_bytes <= tl.tile_bytes_max, "tile bytes");
        // units of every thread (every tile of small streams; the tiles at
        // segment edges and the launch ends of large ones)
        const bool edge = tl.ntiles <= 64 || t < 2 || t + 2 >= tl.ntiles || x.lt == 0 || x.lt + 1 == x.seg_tiles;
        const uint32_t nunits = edge ? x.nblk * g.U : 0;
        for (uint32_t tid = 0; tid < tl.threads; tid++) {
            for (uint32_t i = 0; i < kUnitsPerThread; i++) {
                uint32_t v = i * tl.threads + tid;
                if (v >= nunits)
                    continue;
                uint32_t j = v / g.U, u = v - j * g.U;
                uint32_t nj = tile_block_len(g, x, j);
                uint32_t cnt = 4 * u < nj ? (nj - 4 * u < 4 ? nj - 4 * u : 4) : 0;
                REQUIRE(j < 512 && u < 2048 && cnt <= 4, "unit packing");
                REQUIRE(j < tl.G, "unit block");
                if (!cnt)
                    continue;
                uint64_t e = x.e0 + (uint64_t) j * g.L + 4 * u;
                REQUIRE(e + cnt <= x.seg_end && e + cnt <= g.n, "unit values");
                REQUIRE(e >= 1 && e - 1 >= x.seg_a, "previous value");
                // LDS bytes an encoder or decoder touches for this unit, every
                // width: the decoder reads 4 bytes at each of the unit's 4
                // high-byte fields (masked to B bytes) and 4 bytes of low bits
                uint64_t lds = op == 0 ? compress_lds(tl) : stage_bytes(tl, op);
                for (uint32_t w = 0; w <= 32; w++) {
                    uint32_t B = w >> 3, r = w & 7, sb = w ? (nj + 7) >> 3 : 0;
                    uint64_t bo = tile_bytes - block_bytes(w, nj);  // a block of width w ending the tile
                    uint64_t hp = bo + 1 + sb;
                    uint64_t hi_end = hp + 4ull * u * B + 3ull * B + 4;
                    uint64_t lo_end = hp + (uint64_t) B * nj + ((4ull * u * r) >> 3) + 4;
                    uint64_t nib_end = bo + (w ? 1 + (u >> 1) : 0) + 1;
                    uint64_t reach = hi_end > lo_end ? hi_end : lo_end;
                    reach = reach > nib_end ? reach : nib_end;
                    // decoder: within the 16 bytes of slack after the tile's bytes
                    REQUIRE(reach <= tile_bytes + 16, "decoder reach %lu tile %lu w %u", (unsigned long) reach,
                        (unsigned long) tile_bytes, w);
                    // encoder (16-byte OR of the high bytes, 4-byte OR of the low bits)
                    uint64_t enc = w ? hp + 4ull * u * B + 16 : 0;
                    REQUIRE(enc + 16 <= lds && reach + 16 <= lds + 16, "LDS reach %lu of %lu", (unsigned long) enc,
                        (unsigned long) lds);
                }
            }
        }
        if (op == 2) {
            uint32_t nw = tl.threads / 64;
            REQUIRE(kUnitsPerThread * nw <= 64, "decompress wave scan");
        }
        // homomorphic-sum operand tables for k = 32: per-operand need fits the stage
        if (op == 1) {
            uint64_t need = align_up(tile_bytes + 16, 16);
            REQUIRE(need <= stage_bytes(tl, op), "stage");
            REQUIRE(tl.G <= 256, "per-operand scan covers 4 blocks per lane");
        }
    }
    REQUIRE(blocks_seen == g.nblocks, "every block in exactly one tile");
    REQUIRE(groups_seen == g.ngroups, "every group in exactly one tile");
    for (uint32_t s = 0; s < g.S; s++)
        REQUIRE(seg_seen[s], "segment %u has a first tile", s);
    g_cases++;
}

int main()
{
    const uint64_t ns[] = {1, 2, 3, 7, 8, 9, 15, 16, 17, 31, 33, 63, 64, 65, 100, 257, 1000, 2039, 4097,
        9217, 65536, 100003, 262144, 1000003, 16777216, 16777217};
    const uint32_t Ls[] = {1, 2, 3, 4, 5, 7, 8, 9, 16, 31, 32, 33, 36, 37, 63, 64, 65, 100, 127, 128, 129,
        255, 256, 1000, 2048, 4095, 4096};
    const uint32_t Ts[] = {1, 2, 3, 4, 5, 7, 8, 13, 16, 17, 32, 64};
    for (uint64_t n : ns)
        for (uint32_t L : Ls)
            for (uint32_t T : Ts) {
                Geom g;
                if (!make_geom(n, L, T, &g))
                    continue;
                // index layout fits and is ordered
                REQUIRE(g.ix_group >= g.nblocks && g.ix_seed >= g.ix_group + 8 * g.ngroups &&
                            g.ix_meta >= g.ix_seed + 8ull * g.S && g.ix_bytes >= g.ix_meta + 8,
                    "index layout");
                // the parse kernel's group offsets: segment s group q = s*gT + jb/IG
                REQUIRE((uint64_t) g.T * g.gT + g.gR == g.ngroups, "group count");
                check_tiling(g, 512, 0);  // compress
                check_tiling(g, 256, 1);  // homomorphic sum
                check_tiling(g, 512, 2);  // decompress
                const Tiling th = make_tiling(g, 256);
                for (int k : {1, 2, 3, 8, 17, 32})
                    check_stage_plans(th, k, (uint32_t) (n * 131 + L * 7 + T + k));
            }
    std::printf("ok %ld\n", g_cases);
    return 0;
}
