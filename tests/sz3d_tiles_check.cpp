// sz3d_tiles_check.cpp -- host replay of the tiled 3-D kernels' work split.
//
// For many volumes, block sides 4 and 8 and team sizes T, take the sub-volume
// table of make_geom3 and walk every chunk of the tiled compress and every
// block column of the tiled decoder through part_of_unit and chunk_of3, as
// zccl_sz3d.hip does.  Checks: the chunk and column counts match the table's
// prefixes; every block of every sub-volume is covered by exactly one chunk
// (and one column's chunk walk) at the block index the stream order gives it;
// the chunk's extents are the block's; the halo tile and the decoder's z rows
// stay inside their LDS arrays and the volume.  Built and run by
// tests/test_bounds.py with g++; prints "ok <cases>" or the first violation.
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "zccl_sz3d.h"

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

static void check(uint32_t d1, uint32_t d2, uint32_t d3, uint32_t Bs, int T)
{
    Geom3 g;
    std::vector<Part3> parts(T);
    if (!make_geom3(d1, d2, d3, (int) Bs, T, &g, parts.data()))
        return;
    const uint32_t nz = tile_nz3(Bs), wzmax = nz * Bs;
    REQUIRE(nz > 0, "Bs %u has a tiled path", Bs);
    // LDS arrays of the kernels (zccl_sz3d.hip Tile3): halo tile (Bs+1)^2 x (wzmax+1), z rows Bs^2 x wzmax
    const uint64_t qtile = (uint64_t) (Bs + 1) * (Bs + 1) * (wzmax + 1), rows = (uint64_t) Bs * Bs * wzmax;
    uint32_t nu = 0, nc = 0;
    for (int t = 0; t < T; t++) {
        REQUIRE(parts[t].u0 == nu && parts[t].c0 == nc, "prefixes of part %d", t);
        nu += parts[t].b1 * parts[t].b2 * ((parts[t].b3 + nz - 1) / nz);
        nc += parts[t].b1 * parts[t].b2;
    }
    REQUIRE(nu == g.nunits && nc == g.ncols, "counts %u %u vs %u %u", nu, nc, g.nunits, g.ncols);
    std::vector<int> seen(g.nblocks, 0), seen_cols(g.nblocks, 0);
    for (uint32_t u = 0; u < g.nunits; u++) {
        const uint32_t t = part_of_unit(parts.data(), g.T, u, false);
        const Part3& p = parts[t];
        REQUIRE(u >= p.u0 && (t + 1 == g.T || u < parts[t + 1].u0), "unit %u in part %u", u, t);
        const Chunk3 c = chunk_of3(p, u - p.u0, Bs);
        REQUIRE(c.bx < p.b1 && c.by < p.b2 && c.lz0 < p.sz && c.wz >= 1 && c.wz <= wzmax, "chunk of unit %u", u);
        REQUIRE(c.ex == (p.sx - c.bx * Bs < Bs ? p.sx - c.bx * Bs : Bs) && c.ey >= 1 && c.ey <= Bs, "extents");
        // halo tile rows (i, j) <= (ex, ey) and z index <= wz, inside the LDS array and the sub-volume
        REQUIRE((uint64_t) (c.ex * (Bs + 1) + c.ey) * (wzmax + 1) + c.wz < qtile, "halo tile index");
        REQUIRE(c.bx * Bs + c.ex <= p.sx && c.by * Bs + c.ey <= p.sy && c.lz0 + c.wz <= p.sz, "inside the sub-volume");
        for (uint32_t bzl = 0; bzl < c.nbz; bzl++) {
            const uint32_t bz = c.ch * nz + bzl, gb = c.gcol + bzl;
            REQUIRE(bz < p.b3 && gb == p.g0 + (c.bx * p.b2 + c.by) * p.b3 + bz, "block index");
            REQUIRE(gb < g.nblocks, "block %u of %u", gb, g.nblocks);
            const uint32_t ez = c.wz - bzl * Bs < Bs ? c.wz - bzl * Bs : Bs;
            REQUIRE(ez == (p.sz - bz * Bs < Bs ? p.sz - bz * Bs : Bs), "block z extent");
            // decoder z rows: row (i, j) < (ex, ey), z < wz
            REQUIRE((uint64_t) ((c.ex - 1) * Bs + (c.ey - 1)) * wzmax + bzl * Bs + ez - 1 < rows, "z row index");
            seen[gb]++;
        }
    }
    for (uint32_t col = 0; col < g.ncols; col++) {
        const uint32_t t = part_of_unit(parts.data(), g.T, col, true);
        const Part3& p = parts[t];
        const uint32_t nch = (p.b3 + nz - 1) / nz;
        for (uint32_t ch = 0; ch < nch; ch++) {
            const Chunk3 c = chunk_of3(p, (col - p.c0) * nch + ch, Bs);
            REQUIRE(c.ch == ch, "column chunk order");
            for (uint32_t bzl = 0; bzl < c.nbz; bzl++)
                seen_cols[c.gcol + bzl]++;
        }
    }
    for (uint32_t b = 0; b < g.nblocks; b++)
        REQUIRE(seen[b] == 1 && seen_cols[b] == 1, "block %u covered %d / %d times", b, seen[b], seen_cols[b]);
    g_cases++;
}

int main()
{
    const uint32_t dims[] = {1, 2, 3, 5, 8, 13, 31, 64, 65, 100, 257, 300, 517};
    const int Ts[] = {1, 2, 3, 4, 6, 8, 12, 16, 27, 64};
    for (uint32_t Bs : {4u, 8u})
        for (uint32_t a : dims)
            for (uint32_t b : {1u, 7u, 40u, 64u})
                for (uint32_t c : dims)
                    for (int T : Ts)
                        if ((uint64_t) a * b * c <= (1ull << 24))
                            check(a, b, c, Bs, T);
    check(512, 512, 512, 4, 16);
    check(512, 512, 512, 8, 16);
    check(512, 512, 512, 4, 1);
    std::printf("ok %ld\n", g_cases);
    return 0;
}
