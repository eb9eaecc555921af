// zccl_sz3d.h -- geometry of the 3-D SZp codec shared by the host launchers
// and the kernels (zccl_sz3d.hip).
//
// Stream (reference: SZp_compress3D_fast_openmp, ZCCL/src/ZCCL_float.c:
// 1862-2048, and its decoder ZCCL/src/ZCCLd_float.c:2873-3040):
//   [T x u64 payload offsets, relative to the end of this header]
//   [payload of sub-volume 0] ... [payload of sub-volume T-1]
//   payload = one width byte per block (all blocks of the sub-volume, x, y, z
//   order), then for each block of nonzero width w: the MSB-first sign bitmap
//   of its values and their fixed-length bits (w >> 3 little-endian high bytes
//   per value, then the low w & 7 bits MSB-first), values in (i, j, k) order.
// The volume d1 x d2 x d3 (d3 fastest) is cut into T sub-volumes by the
// factorisation of T into three dims (My_Dims_create), every sub-volume into
// cubic blocks of side Bs (smaller at its far edges).  A value's residual is
// the 3-D difference dx dy dz q of q = floor(x * (0.5 / eb) + 0.5) with q = 0
// outside its sub-volume; the decoder is the 3-D prefix sum.
#pragma once

#include <stddef.h>
#include <stdint.h>

#include "zccl_internal.h"

namespace zk {

constexpr int kMaxParts3 = 4096;     // sub-volumes (the reference's OpenMP team size)
constexpr uint32_t kMaxBs3 = 64;     // block side: a payload (< 2^26 bytes) and its width share a word

struct Part3 {
    uint32_t ox, oy, oz;   // first plane along each axis
    uint32_t sx, sy, sz;   // planes along each axis
    uint32_t b1, b2, b3;   // blocks along each axis
    uint32_t g0;           // global index of the sub-volume's first block
    uint32_t nbc;          // blocks of the sub-volumes before it
    uint32_t zr0;          // first z-row (x, y pair) of the sub-volume in the z pass
    uint32_t yc0;          // first y-column (x, z pair) in the y pass
    uint32_t xc0;          // first x-column (y, z pair) in the x pass
    uint32_t u0;           // first chunk of the tiled compress (block sides 4, 8)
    uint32_t c0;           // first block column (bx, by) of the tiled decoder
};

struct Geom3 {
    uint32_t d1, d2, d3;
    uint32_t Bs, T;
    uint32_t nblocks;      // all blocks of all sub-volumes
    uint32_t zrows, ycols, xcols;
    uint32_t nunits;       // chunks of the tiled compress (0: no tiled path)
    uint32_t ncols;        // block columns of the tiled decoder
};

// Blocks per chunk along z of the tiled kernels: a chunk is one (bx, by)
// column of blocks cut to tile_nz3 blocks, whose quantized values with a
// one-value halo fit in LDS.  0 where the tiled path has no instance.
ZK_HD constexpr uint32_t tile_nz3(uint32_t Bs) { return Bs == 4 ? 64u : Bs == 8 ? 8u : 0u; }

// Factorise T over three dims as My_Dims_create (ZCCL/src/ZCCL_multidim.c:
// 100-140): prime factors of T, largest first, each multiplied into the
// currently smallest dim (the first on ties); dims sorted descending.
ZK_HD void dims_create3(int T, int d[3])
{
    d[0] = d[1] = d[2] = 1;
    int f[32], nf = 0, rem = T;
    for (int i = 2; i <= rem && nf < 32; i++)
        while (rem % i == 0 && nf < 32) {
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
            if (d[j] < d[m])
                m = j;
        d[m] *= f[i];
    }
    for (int a = 0; a < 3; a++)
        for (int b = a + 1; b < 3; b++)
            if (d[b] > d[a]) {
                int t = d[a];
                d[a] = d[b];
                d[b] = t;
            }
}

ZK_HD uint32_t nblk3(uint32_t s, uint32_t Bs) { return (s - 1) / Bs + 1; }

// Fill g and parts[T] (sub-volumes in thread order, ZCCL/src/ZCCL_float.c:
// 1901-1935); false when the geometry is out of range or a sub-volume would be
// empty (the reference's DSize3D_init underflows then).
ZK_HD bool make_geom3(uint64_t d1, uint64_t d2, uint64_t d3, int Bs, int T, Geom3* g, Part3* parts)
{
    if (T < 1 || T > kMaxParts3 || Bs < 1 || (uint32_t) Bs > kMaxBs3 || d1 < 1 || d2 < 1 || d3 < 1)
        return false;
    if (d1 * d2 * d3 > kMaxN)
        return false;
    int td[3];
    dims_create3(T, td);
    if (d1 < (uint64_t) td[0] || d2 < (uint64_t) td[1] || d3 < (uint64_t) td[2])
        return false;
    g->d1 = (uint32_t) d1;
    g->d2 = (uint32_t) d2;
    g->d3 = (uint32_t) d3;
    g->Bs = (uint32_t) Bs;
    g->T = (uint32_t) T;
    const uint32_t bx = g->d1 / td[0], rx = g->d1 % td[0], by = g->d2 / td[1], ry = g->d2 % td[1];
    const uint32_t bz = g->d3 / td[2], rz = g->d3 % td[2];
    const uint32_t nz = tile_nz3((uint32_t) Bs);
    uint32_t r = 0, ox = 0, nb = 0, zr = 0, yc = 0, xc = 0, nu = 0, nc = 0;
    for (int px = 0; px < td[0]; px++) {
        const uint32_t sx = bx + ((uint32_t) px < rx ? 1u : 0u);
        uint32_t oy = 0;
        for (int py = 0; py < td[1]; py++) {
            const uint32_t sy = by + ((uint32_t) py < ry ? 1u : 0u);
            uint32_t oz = 0;
            for (int pz = 0; pz < td[2]; pz++) {
                const uint32_t sz = bz + ((uint32_t) pz < rz ? 1u : 0u);
                Part3& p = parts[r++];
                p.ox = ox; p.oy = oy; p.oz = oz;
                p.sx = sx; p.sy = sy; p.sz = sz;
                p.b1 = nblk3(sx, Bs); p.b2 = nblk3(sy, Bs); p.b3 = nblk3(sz, Bs);
                p.g0 = nb;
                p.nbc = nb;
                p.zr0 = zr;
                p.yc0 = yc;
                p.xc0 = xc;
                p.u0 = nu;
                p.c0 = nc;
                nu += nz ? p.b1 * p.b2 * ((p.b3 + nz - 1) / nz) : 0u;
                nc += p.b1 * p.b2;
                nb += p.b1 * p.b2 * p.b3;
                zr += sx * sy;
                yc += sx * sz;
                xc += sy * sz;
                oz += sz;
            }
            oy += sy;
        }
        ox += sx;
    }
    g->nblocks = nb;
    g->zrows = zr;
    g->ycols = yc;
    g->xcols = xc;
    g->nunits = nu;
    g->ncols = nc;
    return true;
}

// Payload bytes of a block of n values at width w (no width byte).
ZK_HD uint32_t payload3(uint32_t w, uint32_t n) { return w ? block_bytes(w, n) - 1 : 0u; }

// Worst case: every block at width 32.
ZK_HD uint64_t worst_case_bytes3(const Geom3& g, const Part3* parts)
{
    uint64_t tot = 8ull * g.T + g.nblocks;
    for (uint32_t t = 0; t < g.T; t++) {
        const Part3& p = parts[t];
        const uint64_t n = (uint64_t) p.sx * p.sy * p.sz;
        tot += 4 * n + n / 8 + (uint64_t) p.b1 * p.b2 * p.b3;
    }
    return tot;
}

// Index of the sub-volume holding global block gb (binary search over g0).
ZK_HD uint32_t part_of_block(const Part3* parts, uint32_t T, uint32_t gb)
{
    uint32_t lo = 0, hi = T;
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) / 2;
        if (parts[mid].g0 <= gb)
            lo = mid;
        else
            hi = mid;
    }
    return lo;
}

// Sub-volume of a tiled work item (first chunk u0, or first block column c0
// with cols set), by binary search over the table.
ZK_HD uint32_t part_of_unit(const Part3* parts, uint32_t T, uint32_t v, bool cols)
{
    uint32_t lo = 0, hi = T;
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) / 2;
        if ((cols ? parts[mid].c0 : parts[mid].u0) <= v)
            lo = mid;
        else
            hi = mid;
    }
    return lo;
}

// A chunk of the tiled kernels: block column (bx, by) of sub-volume p, cut to
// tile_nz3(Bs) blocks along z; rel is the chunk's index inside p (column
// major: rel = (bx * b2 + by) * chunks per column + ch).
struct Chunk3 {
    uint32_t bx, by, ch;   // block column and chunk along z
    uint32_t ex, ey;       // extents of the column's blocks in x and y
    uint32_t lz0, wz, nbz; // chunk's first local z, z values, blocks
    uint32_t gcol;         // global index of the chunk's first block
};

ZK_HD Chunk3 chunk_of3(const Part3& p, uint32_t rel, uint32_t Bs)
{
    const uint32_t nz = tile_nz3(Bs), wzmax = nz * Bs;
    Chunk3 c;
    const uint32_t nch = (p.b3 + nz - 1) / nz;
    const uint32_t col = rel / nch;
    c.ch = rel - col * nch;
    c.bx = col / p.b2;
    c.by = col - c.bx * p.b2;
    c.ex = p.sx - c.bx * Bs < Bs ? p.sx - c.bx * Bs : Bs;
    c.ey = p.sy - c.by * Bs < Bs ? p.sy - c.by * Bs : Bs;
    c.lz0 = c.ch * wzmax;
    c.wz = p.sz - c.lz0 < wzmax ? p.sz - c.lz0 : wzmax;
    c.nbz = (c.wz + Bs - 1) / Bs;
    c.gcol = p.g0 + (c.bx * p.b2 + c.by) * p.b3 + c.ch * nz;
    return c;
}

// Device workspace of one 3-D call, in order: 256 bytes of status words
// (status[0] stream size, status[1] error bits), the per-block size words,
// their exclusive scan (u64), the scan's per-group sums (+ the total) and the
// sub-volume table.
constexpr uint32_t kScan3Group = 1024;  // entries per scan group

struct Ws3 {
    uint64_t words, offs, sums, parts, bytes;
};

ZK_HD Ws3 ws3_layout(uint32_t nblocks, uint32_t T)
{
    Ws3 w;
    const uint64_t groups = ceil_div(nblocks ? nblocks : 1, kScan3Group);
    w.words = 256;
    w.offs = w.words + align_up(4ull * nblocks, 256);
    w.sums = w.offs + align_up(8ull * nblocks + 8, 256);
    w.parts = w.sums + align_up(8ull * groups + 8, 256);
    w.bytes = w.parts + align_up(sizeof(Part3) * T, 256);
    return w;
}

}  // namespace zk
