// zccl_internal.h -- stream geometry, tiling and buffer layouts shared by the
// host launchers and the CDNA4 kernels.
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
//     the optional remainder);
//   * a segment's blocks are cut into index groups of IG blocks (IG = 16 for
//     L <= 144); every segment owns at least one group (an empty one when it
//     has no blocks);
//   * a kernel tile is up to TG consecutive groups of one segment (a tile never
//     straddles two segments);
//   * a device-resident stream travels with an index: widths[nblocks] (one
//     byte per block), group_off[ngroups] (byte offset of each group's first
//     block), seed_off[S] (byte offset of each seed) and a meta word holding
//     the stream size.  With the index every kernel finds its bytes without a
//     sequential walk of the stream.
#pragma once

#include <stddef.h>
#include <stdint.h>
#include <string.h>

#if defined(__HIPCC__)
#define ZK_HD __host__ __device__ inline
#else
#define ZK_HD inline
#endif

namespace zk {

constexpr int kMaxOperands = 32;     // operands of one fused homomorphic-sum launch
constexpr uint32_t kMaxL = 4096;     // block length supported by the tile kernels
constexpr uint32_t kUnitsPerThread = 4;
constexpr uint32_t kMaxTileThreads = 512;                            // 8 waves
constexpr uint32_t kTileUnits = kMaxTileThreads * kUnitsPerThread;  // 4-element units per tile (at most)
constexpr uint32_t kGroupUnits = 576;  // units per index group (at most)
constexpr uint32_t kMaxGroupBlocks = 16;
constexpr uint32_t kMaxTileBlocks = 256;
constexpr uint32_t kStatusStride = 1;  // u64 words between lookback status entries (packed: a 64-wide lookback window reads 512 B)

// Element, block, group and tile counts are 32-bit: streams hold fewer than
// 2^31 values (the reference indexes them with int, ZCCL/src/ZCCL_float.c:553-554).
// Byte offsets inside streams stay 64-bit.
constexpr uint64_t kMaxN = (1ull << 31) - 1;

struct Geom {
    uint32_t n;
    uint32_t L, T;
    uint32_t tb, R;
    uint32_t S;          // segments: T (+1 when R > 0)
    uint32_t nbT, nbR;   // blocks per regular / remainder segment
    uint32_t nblocks;
    uint32_t U;          // 4-element units per block = ceil(L/4)
    uint32_t IG;         // blocks per index group
    uint32_t gT, gR;     // groups per regular / remainder segment
    uint32_t ngroups;
    // index layout (byte offsets inside an index buffer)
    uint32_t ix_group, ix_seed, ix_meta, ix_bytes;
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

// Fill g for (n, L, T); false when the reference would be undefined (n < T:
// ZCCL/src/ZCCL_float.c:543-544 underflows) or L is out of range.
ZK_HD bool make_geom(uint64_t n, uint32_t L, uint32_t T, Geom* g)
{
    if (L < 1 || L > kMaxL || T < 1 || n < T || n > kMaxN)
        return false;
    g->n = (uint32_t) n;
    g->L = L;
    g->T = T;
    g->tb = g->n / T;
    g->R = g->n % T;
    g->S = T + (g->R ? 1u : 0u);
    g->nbT = (uint32_t) ceil_div(g->tb - 1, L);
    g->nbR = g->R ? (uint32_t) ceil_div(g->R - 1, L) : 0u;
    g->nblocks = T * g->nbT + g->nbR;
    g->U = (L + 3) / 4;
    uint32_t ig = kGroupUnits / g->U;
    g->IG = ig > kMaxGroupBlocks ? kMaxGroupBlocks : (ig < 1 ? 1 : ig);
    g->gT = g->nbT ? (uint32_t) ceil_div(g->nbT, g->IG) : 1u;
    g->gR = g->R ? (g->nbR ? (uint32_t) ceil_div(g->nbR, g->IG) : 1u) : 0u;
    g->ngroups = T * g->gT + g->gR;
    g->ix_group = (uint32_t) align_up(g->nblocks, 16);
    g->ix_seed = g->ix_group + 8 * g->ngroups;
    g->ix_meta = (uint32_t) align_up(g->ix_seed + 8ull * g->S, 16);
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

// Per-launch tiling: TG index groups per tile, as many as fill a workgroup of
// (at most) target_threads threads at kUnitsPerThread units each.  Targets
// are multiples of 256 so a workgroup is a whole number of waves per SIMD.
struct Tiling {
    uint32_t TG;         // groups per tile
    uint32_t G;          // blocks per tile (TG * IG)
    uint32_t threads;    // workgroup size
    uint32_t tilesT, tilesR, ntiles;
    uint32_t tile_bytes_max;  // worst-case encoded bytes of one tile
};

ZK_HD Tiling make_tiling(const Geom& g, uint32_t target_threads)
{
    Tiling t;
    uint32_t tg = target_threads * kUnitsPerThread / (g.IG * g.U);
    if (tg < 1)
        tg = 1;
    while (tg > 1 && tg * g.IG > kMaxTileBlocks)
        tg--;
    t.TG = tg;
    t.G = tg * g.IG;
    uint32_t units = t.G * g.U;
    uint32_t th = (uint32_t) ceil_div(units, kUnitsPerThread);
    t.threads = (uint32_t) align_up(th < 64 ? 64 : th, 64);
    t.tilesT = (uint32_t) ceil_div(g.gT, tg);
    t.tilesR = g.gR ? (uint32_t) ceil_div(g.gR, tg) : 0u;
    t.ntiles = g.T * t.tilesT + t.tilesR;
    t.tile_bytes_max = t.G * block_bytes(32, g.L);
    return t;
}

// Lane layout (zccl_lane.hip): one lane per block of the reference's default
// length, one wave per tile of 64 blocks (4 index groups), kLaneWaves tiles
// per workgroup.
constexpr uint32_t kLaneL = 36;
constexpr uint32_t kLaneTileBlocks = 64;
constexpr uint32_t kLaneWaves = 4;

ZK_HD bool lane_layout_ok(const Geom& g) { return g.L == kLaneL && g.IG == 16; }

// Compress in the lane layout: a workgroup of kLaneWaves waves per tile of
// kLaneWaves x 64 blocks (one lookback per workgroup).
ZK_HD Tiling make_lane_wg_tiling(const Geom& g)
{
    Tiling t;
    t.TG = kLaneWaves * kLaneTileBlocks / g.IG;
    t.G = kLaneWaves * kLaneTileBlocks;
    t.threads = 64 * kLaneWaves;
    t.tilesT = (uint32_t) ceil_div(g.gT, t.TG);
    t.tilesR = g.gR ? (uint32_t) ceil_div(g.gR, t.TG) : 0u;
    t.ntiles = g.T * t.tilesT + t.tilesR;
    t.tile_bytes_max = t.G * block_bytes(32, g.L);
    return t;
}

ZK_HD Tiling make_lane_tiling(const Geom& g)
{
    Tiling t;
    t.TG = kLaneTileBlocks / g.IG;
    t.G = kLaneTileBlocks;
    t.threads = 64;
    t.tilesT = (uint32_t) ceil_div(g.gT, t.TG);
    t.tilesR = g.gR ? (uint32_t) ceil_div(g.gR, t.TG) : 0u;
    t.ntiles = g.T * t.tilesT + t.tilesR;
    t.tile_bytes_max = t.G * block_bytes(32, g.L);
    return t;
}

struct Tile {
    uint32_t s;              // segment
    uint32_t nblk;           // blocks in this tile
    uint32_t ngrp;           // index groups in this tile
    uint32_t lt;             // tile index inside the segment
    uint32_t g0;             // global index of the tile's first block
    uint32_t grp0;           // global index of the tile's first group
    uint32_t seg_a, seg_end; // segment element range [seg_a, seg_end)
    uint32_t e0;             // first element of the tile's first block
    uint32_t seg_tile0;      // index of the segment's first tile
    uint32_t seg_tiles;      // tiles in the segment
};

ZK_HD Tile tile_of(const Geom& g, const Tiling& tl, uint32_t t)
{
    Tile x;
    uint32_t nb_s, ng_s;
    const uint32_t regular = g.T * tl.tilesT;
    if (t < regular) {
        x.s = t / tl.tilesT;
        x.lt = t - x.s * tl.tilesT;
        x.seg_a = x.s * g.tb;
        x.seg_end = x.seg_a + g.tb;
        nb_s = g.nbT;
        ng_s = g.gT;
        x.seg_tile0 = x.s * tl.tilesT;
        x.seg_tiles = tl.tilesT;
        x.g0 = x.s * g.nbT;
        x.grp0 = x.s * g.gT;
    } else {
        x.s = g.T;
        x.lt = t - regular;
        x.seg_a = g.T * g.tb;
        x.seg_end = g.n;
        nb_s = g.nbR;
        ng_s = g.gR;
        x.seg_tile0 = regular;
        x.seg_tiles = tl.tilesR;
        x.g0 = g.T * g.nbT;
        x.grp0 = g.T * g.gT;
    }
    const uint32_t first = x.lt * tl.G;
    const uint32_t left = nb_s > first ? nb_s - first : 0u;
    x.nblk = left < tl.G ? left : tl.G;
    const uint32_t gfirst = x.lt * tl.TG;
    const uint32_t gleft = ng_s > gfirst ? ng_s - gfirst : 0u;
    x.ngrp = gleft < tl.TG ? gleft : tl.TG;
    x.g0 += first;
    x.grp0 += gfirst;
    x.e0 = x.seg_a + 1 + first * g.L;
    return x;
}

// Elements in block j of tile x.
ZK_HD uint32_t tile_block_len(const Geom& g, const Tile& x, uint32_t j)
{
    const uint32_t e = x.e0 + j * g.L;
    const uint32_t left = x.seg_end - e;
    return left < g.L ? left : g.L;
}

// Staging plan of the homomorphic sum's operands.  Operand r's tile bytes
// [base, base + tot) are fetched as nch 16-byte chunks starting at the
// 16-byte aligned stream address at or below base (shift = address mod 16),
// so the operand needs 16*nch + 16 LDS bytes (16 of decoder read slack).
// Operands are packed greedily into rounds; consecutive rounds alternate
// between the two halves of the stage, so the next round can be fetched while
// the current one is decoded.  An operand larger than a half forms a round of
// its own over the whole stage (fetched without overlap).
struct StagePlan {
    uint32_t lds[kMaxOperands];   // stage byte offset of the operand's first chunk
    uint32_t nch[kMaxOperands];   // chunks of the operand
    int gend[kMaxOperands];       // round g holds operands [gend[g-1], gend[g])
    uint32_t whole[kMaxOperands]; // round g uses the whole stage
    int nrounds;
};

ZK_HD uint32_t stage_chunks(uint32_t shift, uint32_t tot) { return tot ? (shift + tot + 15) / 16 : 0u; }
ZK_HD uint32_t stage_need(uint32_t shift, uint32_t tot) { return 16 * stage_chunks(shift, tot) + 16; }

// stage = 2 * half bytes; an operand's tot is 0 when it is skipped.  op(r)
// yields operand r's tile bytes and shift (so a device caller can read them
// from registers); sp.lds / sp.nch / sp.gend / sp.whole are only written.
template <typename Op>
ZK_HD void plan_stage_by(Op op, int k, uint32_t half, StagePlan& sp)
{
    int g = 0, r = 0;
    while (r < k) {
        uint32_t tot, shift;
        op(r, tot, shift);
        const uint32_t need0 = stage_need(shift, tot);
        if (need0 > half) {
            sp.lds[r] = 0;
            sp.nch[r] = stage_chunks(shift, tot);
            sp.whole[g] = 1;
            sp.gend[g++] = ++r;
            continue;
        }
        const uint32_t base = (uint32_t) (g & 1) * half;
        uint32_t pos = 0;
        while (r < k) {
            op(r, tot, shift);
            const uint32_t need = stage_need(shift, tot);
            if (need > half || pos + need > half)
                break;
            sp.lds[r] = base + pos;
            sp.nch[r] = stage_chunks(shift, tot);
            pos += need;
            r++;
        }
        sp.whole[g] = 0;
        sp.gend[g++] = r;
    }
    sp.nrounds = g;
}

ZK_HD void plan_stage(const uint32_t* tot, const uint32_t* shift, int k, uint32_t half, StagePlan& sp)
{
    plan_stage_by([&](int r, uint32_t& t, uint32_t& s) { t = tot[r]; s = shift[r]; }, k, half, sp);
}

// Encoder of one 4-value unit in the slot layout (units of a block in
// neighbouring lanes).  unit_fields computes, for deltas d[] (two's
// complement, zero past the block's end) at width w: the magnitudes, the
// MSB-first sign nibble and the unit's low bits as one 4r-bit field with value
// 0 on top (r = w & 7).  emit_unit writes the bytes unit u of a block owns:
// the block's fields start at buf[fo] (its first sign byte; the width byte, if
// the format has one inline, is the caller's), it has nj values, cnt of them
// in this unit, and U units; pnib / plv are the fields of unit u+1 of the same
// block (ignored when u+1 == U).  Ownership: an even unit writes sign byte u/2
// and the r bytes of low bits it shares with unit u+1 (ceil of the bits
// actually present, zero padded); every unit writes its own high bytes.  Every
// byte of the block then has exactly one writer, so the buffer needs neither
// clearing nor atomics.
// ZCCL/src/ZCCL_TypeManager.c:12-91 (fixed-length bits), :824-848 (signs).
struct UnitFields {
    uint32_t m[4];
    uint32_t nib;
    uint32_t lv;
};

ZK_HD uint32_t mag32(uint32_t d) { return (int32_t) d < 0 ? 0u - d : d; }

ZK_HD UnitFields unit_fields(uint32_t w, const uint32_t d[4])
{
    UnitFields f;
    const uint32_t r = w & 7, mask = (1u << r) - 1u;
    f.nib = 0;
    f.lv = 0;
    for (int c = 0; c < 4; c++) {
        f.m[c] = mag32(d[c]);
        f.nib |= (d[c] >> 31) << (3 - c);
        f.lv |= (f.m[c] & mask) << (r * (3 - c));
    }
    return f;
}

// Store the n (< 8) low bytes of v at p, any alignment.
ZK_HD void put_bytes(uint8_t* p, uint64_t v, uint32_t n)
{
    // byte stores: a misaligned multi-byte LDS store costs more than the bytes
    for (uint32_t i = 0; i < 7; i++)
        if (i < n)
            p[i] = (uint8_t) (v >> (8 * i));
}

ZK_HD void emit_unit(uint8_t* buf, uint64_t fo, uint32_t w, uint32_t nj, uint32_t u, uint32_t U, uint32_t cnt,
    const UnitFields& f, uint32_t pnib, uint32_t plv)
{
    if (!cnt || !w)
        return;
    const uint32_t B = w >> 3, r = w & 7;
    const uint64_t hi0 = fo + ((nj + 7) >> 3);
    if ((u & 1) == 0) {
        const bool has_p = u + 1 < U;
        const uint32_t up = 4 * (u + 1);
        const uint32_t pc = has_p && up < nj ? (nj - up < 4 ? nj - up : 4u) : 0u;
        buf[fo + (u >> 1)] = (uint8_t) ((f.nib << 4) | (has_p ? pnib : 0u));
        if (r) {
            const uint64_t cv = ((uint64_t) f.lv << (4 * r)) | (has_p ? plv : 0u);
            put_bytes(buf + hi0 + (uint64_t) B * nj + (u >> 1) * r, __builtin_bswap64(cv << (64 - 8 * r)),
                ((cnt + pc) * r + 7) >> 3);
        }
    }
    if (B == 1 && cnt == 4) {
        const uint32_t h = (f.m[0] >> r) | ((f.m[1] >> r) << 8) | ((f.m[2] >> r) << 16) | ((f.m[3] >> r) << 24);
        uint8_t* q = buf + hi0 + 4 * u;
        q[0] = (uint8_t) h;
        q[1] = (uint8_t) (h >> 8);
        q[2] = (uint8_t) (h >> 16);
        q[3] = (uint8_t) (h >> 24);
    } else if (B) {
        // constant indexes into f.m (a runtime bound would put f in scratch memory)
#pragma unroll
        for (uint32_t c = 0; c < 4; c++) {
            if (c < cnt) {
                const uint32_t h = f.m[c] >> r;
                for (uint32_t b = 0; b < B; b++)
                    buf[hi0 + (uint64_t) (4 * u + c) * B + b] = (uint8_t) (h >> (8 * b));
            }
        }
    }
}

// Workspace: decoupled-lookback state.  Two status arrays alternate between
// consecutive launches.  `seq` holds the parity of the array in use (bit 32)
// and the next tile ticket (low 32 bits): every workgroup takes its tile
// index from it, so tiles are numbered in the order they start and a started
// tile's predecessors have all started, whatever else runs on the device.
// When the last lookback chain of a launch completes, seq is reset to ticket
// 0 of the other parity; every launch clears the array the next one uses, so
// nothing is reset from the host between launches (graph replays included).
struct WsHeader {
    unsigned long long seq;      // parity << 32 | next tile ticket
    unsigned int chains_done;    // chains of the current launch that completed
    unsigned int pad0;
    unsigned long long capacity; // status entries per array
    unsigned long long errors;   // sticky error bits (ZK_ERR_*)
    unsigned long long pad[28];
};
static_assert(sizeof(WsHeader) == 256, "workspace header is 256 bytes");

// Bytes of lookback state per tile: two parity arrays of one kStatusStride
// slot each.
constexpr uint64_t kWsBytesPerTile = 2ull * 8 * kStatusStride;

ZK_HD uint64_t workspace_bytes(uint64_t max_tiles) { return sizeof(WsHeader) + kWsBytesPerTile * max_tiles; }

// Tiles a workspace of `bytes` bytes can serve (inverse of workspace_bytes).
ZK_HD uint64_t workspace_capacity(uint64_t bytes)
{
    return bytes < sizeof(WsHeader) ? 0 : (bytes - sizeof(WsHeader)) / kWsBytesPerTile;
}

enum : unsigned long long {
    ZK_ERR_SPIN = 1ull << 0,      // lookback wait exceeded its bound
    ZK_ERR_NOSPACE = 1ull << 1,   // output capacity too small
    ZK_ERR_WIDTH = 1ull << 2,     // width byte > 32 in a stream
    ZK_ERR_MISMATCH = 1ull << 3,  // index width differs from the stream's width byte
    ZK_ERR_OVERRUN = 1ull << 4,   // stream walk left its stated size
    ZK_ERR_CAPACITY = 1ull << 5,  // workspace too small for the launch
};

// Exact size of a host stream: its last header entry locates chunk T-1,
// whose blocks (and the remainder segment's) are walked by their width bytes.
// Nothing past the bytes the walk proves to be in the stream is read, and the
// walk never leaves the geometry's worst-case size; 0 when the header or a
// width byte is malformed.  Host code (the drop-in layer sizes host copies
// with it).  ZCCL/src/ZCCLd_float.c:597-745 (the same walk).
inline uint64_t host_stream_bytes(const unsigned char* in, const Geom& g)
{
    const uint64_t worst = worst_case_bytes(g);
    uint64_t h = 0;
    memcpy(&h, in + 8ull * (g.T - 1), 8);
    if (h > worst)
        return 0;
    uint64_t pos = 8ull * g.T + h;
    const uint32_t nseg = g.R ? 2u : 1u;
    for (uint32_t sidx = 0; sidx < nseg; sidx++) {
        const uint32_t s = g.T - 1 + sidx;
        const uint64_t nb = s < g.T ? g.nbT : g.nbR;
        const uint64_t len = s < g.T ? g.tb : g.R;
        pos += 4;  // seed
        for (uint64_t jb = 0; jb < nb; jb++) {
            if (pos >= worst)
                return 0;
            const uint32_t w = in[pos];
            if (w > 32)
                return 0;
            const uint64_t left = len - 1 - jb * g.L;
            pos += block_bytes(w, (uint32_t) (left < g.L ? left : g.L));
        }
    }
    return pos <= worst ? pos : 0;
}

}  // namespace zk
