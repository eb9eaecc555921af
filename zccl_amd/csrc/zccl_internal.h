// zccl_internal.h -- stream geometry, tile mapping and buffer layouts shared by
// the host launchers and the CDNA4 kernels.
//
// Wire format (reference: ZCCL/src/ZCCL_float.c:501-807):
//   [T x u64 chunk offsets][chunk 0 payload]...[chunk T-1 payload]
//   chunk t payload = i32 seed q[t*tb], then ceil((tb-1)/L) blocks of Lorenzo
//   deltas; chunk T-1 additionally carries the remainder segment (seed
//   q[T*tb], then ceil((R-1)/L) blocks) when R = n % T > 0.
//   block = u8 width b; if b > 0: sign bitmap ceil(nb/8) B (MSB first),
//   (b>>3) little-endian high bytes per element of |d|>>(b&7), then the low
//   (b&7) bits of every |d| as one MSB-first bit stream.
//
// Device layout chosen here:
//   * a "segment" is one seeded run of the Lorenzo chain (T regular chunks plus
//     the optional remainder); segments are cut into tiles of G blocks and a
//     tile never straddles two segments;
//   * every segment owns at least one tile (an empty tile carries only the
//     seed when the segment has no blocks);
//   * a device-resident stream travels with an index: widths[nblocks] (one
//     byte per block), tile_off[ntiles] (byte offset of each tile's first
//     block), seed_off[S] (byte offset of each seed) and a meta word holding
//     the stream size.  The index is what lets every kernel find its bytes
//     without a sequential walk of the stream.
#pragma once

#include <stddef.h>
#include <stdint.h>

#if defined(__HIPCC__)
#define ZK_HD __host__ __device__ inline
#else
#define ZK_HD inline
#endif

namespace zk {

constexpr int kMaxOperands = 32;   // operands of one fused homomorphic-sum launch
constexpr uint32_t kMaxL = 4096;   // block length supported by the tile kernels
constexpr uint32_t kMaxThreads = 1024;

struct Geom {
    uint64_t n;
    uint32_t L, T;
    uint64_t tb, R;
    uint32_t S;          // segments: T (+1 when R > 0)
    uint64_t nbT, nbR;   // blocks per regular / remainder segment
    uint64_t nblocks;
    uint32_t U;          // 4-element units per block = ceil(L/4)
    uint32_t G;          // blocks per tile
    uint32_t threads;    // workgroup size: G*U rounded up to a wave
    uint64_t tilesT, tilesR, ntiles;
    uint32_t tile_bytes_max;  // worst-case encoded bytes of one tile
    // index layout (byte offsets inside an index buffer)
    uint64_t ix_tile, ix_seed, ix_meta, ix_bytes;
};

ZK_HD uint64_t ceil_div(uint64_t a, uint64_t b) { return (a + b - 1) / b; }
ZK_HD uint64_t align_up(uint64_t a, uint64_t b) { return ceil_div(a, b) * b; }

// Encoded size of one block of n deltas at width b (width byte included).
// ZCCL/src/ZCCL_TypeManager.c:20-29 plus sign bitmap and width byte.
ZK_HD uint32_t block_bytes(uint32_t b, uint32_t n)
{
    if (b == 0)
        return 1;
    return 1 + ((n + 7) >> 3) + (b >> 3) * n + (((b & 7) * n + 7) >> 3);
}

// Blocks per tile for block length L: 64 while a tile fits 1024 threads.
ZK_HD uint32_t tile_blocks(uint32_t L)
{
    uint32_t U = (L + 3) / 4;
    uint32_t G = kMaxThreads / U;
    return G > 64 ? 64 : (G < 1 ? 1 : G);
}

// Fill g for (n, L, T); returns false when the reference would be undefined
// (n < T: ZCCL/src/ZCCL_float.c:543-544 underflows) or L is out of range.
ZK_HD bool make_geom(uint64_t n, uint32_t L, uint32_t T, Geom* g)
{
    if (L < 1 || L > kMaxL || T < 1 || n < T)
        return false;
    g->n = n;
    g->L = L;
    g->T = T;
    g->tb = n / T;
    g->R = n % T;
    g->S = T + (g->R ? 1u : 0u);
    g->nbT = ceil_div(g->tb - 1, L);
    g->nbR = g->R ? ceil_div(g->R - 1, L) : 0;
    g->nblocks = (uint64_t) T * g->nbT + g->nbR;
    g->U = (L + 3) / 4;
    g->G = tile_blocks(L);
    uint32_t t = g->G * g->U;
    g->threads = (uint32_t) align_up(t < 64 ? 64 : t, 64);
    g->tilesT = g->nbT ? ceil_div(g->nbT, g->G) : 1;
    g->tilesR = g->R ? (g->nbR ? ceil_div(g->nbR, g->G) : 1) : 0;
    g->ntiles = (uint64_t) T * g->tilesT + g->tilesR;
    g->tile_bytes_max = g->G * block_bytes(32, L);
    g->ix_tile = align_up(g->nblocks, 16);
    g->ix_seed = g->ix_tile + 8 * g->ntiles;
    g->ix_meta = align_up(g->ix_seed + 8ull * g->S, 16);
    g->ix_bytes = g->ix_meta + 32;
    return true;
}

// Worst-case stream size: every block at width 32.
ZK_HD uint64_t worst_case_bytes(const Geom& g)
{
    uint64_t tot = 8ull * g.T + 4ull * g.S;
    uint64_t lens[2] = {g.tb, g.R};
    uint64_t cnt[2] = {g.T, g.R ? 1ull : 0ull};
    for (int i = 0; i < 2; i++) {
        if (!cnt[i] || lens[i] < 2)
            continue;
        uint64_t nfull = (lens[i] - 1) / g.L;
        uint32_t tail = (uint32_t) ((lens[i] - 1) % g.L);
        uint64_t per = nfull * block_bytes(32, g.L) + (tail ? block_bytes(32, tail) : 0);
        tot += cnt[i] * per;
    }
    return tot;
}

struct Tile {
    uint32_t s;              // segment
    uint32_t nblk;           // blocks in this tile
    uint64_t lt;             // tile index inside the segment
    uint64_t g0;             // global index of the tile's first block
    uint64_t seg_a, seg_end; // segment element range [seg_a, seg_end)
    uint64_t e0;             // first element of the tile's first block
    uint64_t seg_tile0;      // index of the segment's first tile
};

ZK_HD Tile tile_of(const Geom& g, uint64_t t)
{
    Tile x;
    uint64_t nb_s;
    uint64_t regular = (uint64_t) g.T * g.tilesT;
    if (t < regular) {
        x.s = (uint32_t) (t / g.tilesT);
        x.lt = t % g.tilesT;
        x.seg_a = (uint64_t) x.s * g.tb;
        x.seg_end = x.seg_a + g.tb;
        nb_s = g.nbT;
        x.seg_tile0 = (uint64_t) x.s * g.tilesT;
        x.g0 = (uint64_t) x.s * g.nbT + x.lt * g.G;
    } else {
        x.s = g.T;
        x.lt = t - regular;
        x.seg_a = (uint64_t) g.T * g.tb;
        x.seg_end = g.n;
        nb_s = g.nbR;
        x.seg_tile0 = regular;
        x.g0 = (uint64_t) g.T * g.nbT + x.lt * g.G;
    }
    uint64_t first = x.lt * g.G;
    uint64_t left = nb_s > first ? nb_s - first : 0;
    x.nblk = (uint32_t) (left < g.G ? left : g.G);
    x.e0 = x.seg_a + 1 + first * g.L;
    return x;
}

// Elements in block j of tile x.
ZK_HD uint32_t tile_block_len(const Geom& g, const Tile& x, uint32_t j)
{
    uint64_t e = x.e0 + (uint64_t) j * g.L;
    uint64_t left = x.seg_end - e;
    return (uint32_t) (left < g.L ? left : g.L);
}

// Workspace: decoupled-lookback state.  Two status arrays alternate between
// consecutive launches (parity kept in the ticket word); every launch clears
// the array the next launch will use, so nothing is reset from the host.
struct WsHeader {
    unsigned long long ticket;   // [63:32] parity, [31:0] next tile ticket
    unsigned long long capacity; // status entries per array
    unsigned long long errors;   // sticky error bits (ZK_ERR_*)
    unsigned long long pad[29];
};
static_assert(sizeof(WsHeader) == 256, "workspace header is 256 bytes");

ZK_HD uint64_t workspace_bytes(uint64_t max_tiles) { return sizeof(WsHeader) + 16 * max_tiles; }

enum : unsigned long long {
    ZK_ERR_SPIN = 1ull << 0,      // lookback wait exceeded its bound
    ZK_ERR_NOSPACE = 1ull << 1,   // output capacity too small
    ZK_ERR_WIDTH = 1ull << 2,     // width byte > 32 in a stream
    ZK_ERR_MISMATCH = 1ull << 3,  // index width differs from the stream's width byte
    ZK_ERR_OVERRUN = 1ull << 4,   // stream walk left its stated size
    ZK_ERR_CAPACITY = 1ull << 5,  // workspace too small for the launch
};

// Index meta words (u64 at ix_meta): [0] stream size, [1] block-case counters
// are not kept; [1..3] reserved.

}  // namespace zk
