// zccl_kernels.hip -- CDNA4 (gfx950) kernels for the SZp fastmode1 codec and
// the hZCCL homomorphic sum.
//
// One workgroup encodes or decodes one tile of up to 256 blocks (9216 values
// at L = 36).  Thread tid owns the 4-element units v = i*threads + tid
// (i < 4) of the tile, so each wave instruction touches 256 consecutive
// values.  Stream offsets (compress, sum) and running integer values
// (decompress) come from a single-pass decoupled-lookback scan over tiles in
// workgroup-id order (every XCD dispatches its workgroups in order, so the
// lowest unfinished tile is always resident and every wait ends); the input is
// read once and the output written once.  Bytes are staged in LDS: encoders
// OR each unit's sign nibble, high bytes and MSB-first low-bit field into a
// zeroed tile buffer and write the tile out with 16-byte stores; decoders
// pull the tile's bytes in with 16-byte loads and extract fields with
// alignbyte / byte swap / shifts.
//
// Arithmetic follows the reference bit for bit:
//   quantize   q = (int)((double)x * (double)(1.0f/eb))      ZCCL/src/ZCCL_float.c:540,563,580
//              (the host passes (double)(1.0f/eb); out-of-range and NaN give
//              INT_MIN as x86 cvttsd2si does)
//   deltas     d_i = q_i - q_{i-1} in int32, chain restarts at seeds  :576-604
//   width      (int)log2f((float)max|d|)+1 via clz + threshold table  :606
//   dequantize x' = (float)q * eb in FP32                   ZCCL/src/ZCCLd_float.c:588-628
//   sum        seeds and deltas add in int32; blocks re-encoded     :1072, 1099-1219
#include <hip/hip_runtime.h>

#include <atomic>

#include "zccl_device.h"
#include "zccl_internal.h"
#include "zccl_kernels.h"
#include "zccl_unit.h"

namespace zk {

// ------------------------------------------------------------ LDS helpers

__device__ __forceinline__ uint32_t lds_u8(const uint32_t* buf, uint32_t addr)
{
    return ((const uint8_t*) buf)[addr];
}

// 4 little-endian bytes starting at an arbitrary LDS byte address.
__device__ __forceinline__ uint32_t lds_rd4(const uint32_t* buf, uint32_t addr)
{
    uint32_t w = addr >> 2, sh = addr & 3;
    return __builtin_amdgcn_alignbyte(buf[w + 1], buf[w], sh);
}

__device__ __forceinline__ void lds_rd16(const uint32_t* buf, uint32_t addr, uint32_t o[4])
{
    uint32_t w = addr >> 2, sh = addr & 3;
    uint32_t a = buf[w], b = buf[w + 1], c = buf[w + 2], d = buf[w + 3], e = buf[w + 4];
    o[0] = __builtin_amdgcn_alignbyte(b, a, sh);
    o[1] = __builtin_amdgcn_alignbyte(c, b, sh);
    o[2] = __builtin_amdgcn_alignbyte(d, c, sh);
    o[3] = __builtin_amdgcn_alignbyte(e, d, sh);
}

// OR 4 little-endian bytes into LDS at an arbitrary byte address.
__device__ __forceinline__ void lds_or4(uint32_t* buf, uint32_t addr, uint32_t v)
{
    if (!v)
        return;
    uint32_t w = addr >> 2, sh = (addr & 3) << 3;
    atomicOr(&buf[w], v << sh);
    if (sh) {
        uint32_t hi = v >> (32 - sh);
        if (hi)
            atomicOr(&buf[w + 1], hi);
    }
}

// OR up to 16 little-endian bytes (v[0..3]) into LDS at a byte address.
__device__ __forceinline__ void lds_or16(uint32_t* buf, uint32_t addr, const uint32_t v[4])
{
    uint32_t w = addr >> 2, sh = (addr & 3) << 3;
    uint32_t o[5];
    if (sh == 0) {
        o[0] = v[0]; o[1] = v[1]; o[2] = v[2]; o[3] = v[3]; o[4] = 0;
    } else {
        o[0] = v[0] << sh;
        o[1] = (v[1] << sh) | (v[0] >> (32 - sh));
        o[2] = (v[2] << sh) | (v[1] >> (32 - sh));
        o[3] = (v[3] << sh) | (v[2] >> (32 - sh));
        o[4] = v[3] >> (32 - sh);
    }
#pragma unroll
    for (int i = 0; i < 5; i++)
        if (o[i])
            atomicOr(&buf[w + i], o[i]);
}

// ------------------------------------------------------------ unit codec

// OR the encoding of unit u (cnt deltas d[], two's complement) of a block of
// nj elements at width w whose width byte sits at LDS byte bo.
// ZCCL/src/ZCCL_TypeManager.c:12-91 and :824-848.
__device__ __forceinline__ void encode_unit(uint32_t* buf, uint32_t bo, uint32_t w, uint32_t nj,
    uint32_t u, uint32_t cnt, const uint32_t d[4])
{
    const uint32_t B = w >> 3, r = w & 7, sb = (nj + 7) >> 3;
    uint32_t m[4], sg = 0;
#pragma unroll
    for (int c = 0; c < 4; c++) {
        bool live = (uint32_t) c < cnt;
        m[c] = live ? umag((int32_t) d[c]) : 0u;
        sg |= (live && (int32_t) d[c] < 0 ? 1u : 0u) << (3 - c);
    }
    lds_or4(buf, bo + 1 + (u >> 1), (u & 1) ? sg : (sg << 4));
    if (B) {
        uint32_t h[4];
#pragma unroll
        for (int c = 0; c < 4; c++)
            h[c] = m[c] >> r;
        uint32_t W[4] = {0, 0, 0, 0};
        switch (B) {
        case 1:
            W[0] = h[0] | (h[1] << 8) | (h[2] << 16) | (h[3] << 24);
            break;
        case 2:
            W[0] = h[0] | (h[1] << 16);
            W[1] = h[2] | (h[3] << 16);
            break;
        case 3:
            W[0] = h[0] | (h[1] << 24);
            W[1] = (h[1] >> 8) | (h[2] << 16);
            W[2] = (h[2] >> 16) | (h[3] << 8);
            break;
        default:
            W[0] = h[0]; W[1] = h[1]; W[2] = h[2]; W[3] = h[3];
            break;
        }
        lds_or16(buf, bo + 1 + sb + 4 * u * B, W);
    }
    if (r) {
        uint32_t bitpos = 4 * u * r, s = bitpos & 7, mask = (1u << r) - 1, V = 0;
#pragma unroll
        for (int c = 0; c < 4; c++)
            if ((uint32_t) c < cnt)
                V |= (m[c] & mask) << (32 - s - (c + 1) * r);
        lds_or4(buf, bo + 1 + sb + B * nj + (bitpos >> 3), __builtin_bswap32(V));
    }
}

// Write nbytes of LDS (from LDS byte 0) to dst.  16-byte aligned interior
// chunks go out as whole stores; the partial first and last chunk are written
// byte by byte so neighbouring tiles' bytes are never touched.
__device__ __forceinline__ void store_tile(const uint32_t* buf, uint8_t* dst, uint32_t nbytes)
{
    uintptr_t d0 = (uintptr_t) dst, d1 = d0 + nbytes;
    uintptr_t c0 = d0 & ~(uintptr_t) 15;
    uint32_t nch = nbytes ? (uint32_t) ((d1 - c0 + 15) >> 4) : 0u;
    for (uint32_t c = threadIdx.x; c < nch; c += blockDim.x) {
        uintptr_t ca = c0 + 16ull * c;
        if (ca >= d0 && ca + 16 <= d1) {
            uint32_t o[4];
            lds_rd16(buf, (uint32_t) (ca - d0), o);
            *(uint4*) ca = make_uint4(o[0], o[1], o[2], o[3]);
        } else {
            uintptr_t s = ca > d0 ? ca : d0, e = ca + 16 < d1 ? ca + 16 : d1;
            for (uintptr_t p = s; p < e; p++)
                *(uint8_t*) p = (uint8_t) lds_u8(buf, (uint32_t) (p - d0));
        }
    }
}

// ------------------------------------------------------------ tile state

extern __shared__ uint32_t zk_lds[];

struct TileShared {
    unsigned long long bc[4];
    uint32_t wave[64];
    uint32_t bmax[kMaxTileBlocks];
    uint32_t boff[kMaxTileBlocks];
    uint8_t width[kMaxTileBlocks];
};

// Per-thread unit coordinates: unit v = i*threads + tid of the tile, packed
// as block j (9 bits) | unit-in-block u (11 bits) | value count (3 bits); a
// zero count means the thread has no unit i.
struct Units {
    uint32_t pk[kUnitsPerThread];
};

__device__ __forceinline__ uint32_t unit_j(uint32_t p) { return p & 511u; }
__device__ __forceinline__ uint32_t unit_u(uint32_t p) { return (p >> 9) & 2047u; }
__device__ __forceinline__ uint32_t unit_cnt(uint32_t p) { return p >> 20; }

__device__ __forceinline__ Units tile_units(const Geom& g, const Tile& x)
{
    Units un;
    const uint32_t nunits = x.nblk * g.U;
#pragma unroll
    for (uint32_t i = 0; i < kUnitsPerThread; i++) {
        uint32_t v = i * blockDim.x + threadIdx.x;
        un.pk[i] = 0;
        if (v < nunits) {
            uint32_t j = v / g.U;
            uint32_t u = v - j * g.U;
            uint32_t nj = tile_block_len(g, x, j);
            uint32_t cnt = 4 * u < nj ? min(4u, nj - 4 * u) : 0u;
            un.pk[i] = j | (u << 9) | (cnt << 20);
        }
    }
    return un;
}

// Fold the magnitudes of the thread's units into the per-block maxima.
__device__ __forceinline__ void fold_block_max(uint32_t* bmax, const Units& un, const uint32_t (&d)[kUnitsPerThread][4])
{
#pragma unroll
    for (uint32_t i = 0; i < kUnitsPerThread; i++) {
        const uint32_t cnt = unit_cnt(un.pk[i]);
        uint32_t umax = 0;
#pragma unroll
        for (int c = 0; c < 4; c++)
            if ((uint32_t) c < cnt)
                umax = max(umax, umag((int32_t) d[i][c]));
        if (umax)
            atomicMax(&bmax[unit_j(un.pk[i])], umax);
    }
}

// Tail shared by every encoder, run by the whole workgroup once the tile's
// bytes sit in LDS at buf[0, total) and its aggregate is published: wave 0
// resolves the tile's stream offset (lookback), then the tile is stored, the
// seed and chunk header written by a segment's first tile, and the output
// index (widths, group offsets, seed offsets, size) filled in.  SH provides
// bc[], width[] and boff[] of the tile's blocks.
template <typename SH>
__device__ __forceinline__ void finish_tile(const Geom& g, const Tile& x, uint64_t t, SH& sh, const uint32_t* buf,
    uint32_t total, unsigned long long* status, WsHeader* ws, uint32_t seed, uint8_t* out, uint64_t out_cap,
    uint8_t* index, unsigned long long* size_out, bool last_tile, uint32_t nchains)
{
    const uint32_t tid = threadIdx.x;
    if (tid < 64) {
        unsigned long long ex = lookback(status, t, 0, total, 0, ws, true);
        if (tid == 0)
            sh.bc[2] = ex;
    }
    __syncthreads();
    const uint64_t P = 8ull * g.T + 4ull * (x.s + 1) + sh.bc[2];
    const bool fits = P + total <= out_cap;
    if (fits)
        store_tile(buf, out + P, total);
    else if (tid == 0)
        raise_error(ws, ZK_ERR_NOSPACE);
    IndexView ix = index_view(g, index);
    for (uint32_t j = tid; j < x.nblk; j += blockDim.x)
        ix.widths[x.g0 + j] = sh.width[j];
    if (tid < x.ngrp)
        ix.group_off[x.grp0 + tid] = P + (x.nblk ? sh.boff[tid * g.IG] : 0u);
    if (tid == 0) {
        if (x.lt == 0) {
            uint64_t so = P - 4;
            if (so + 4 <= out_cap) {
                store_bytes(out + so, &seed, 4);
                if (x.s < g.T) {
                    unsigned long long h = so - 8ull * g.T;
                    store_bytes(out + 8ull * x.s, &h, 8);
                }
            }
            ix.seed_off[x.s] = so;
        }
        if (last_tile) {
            // the last tile ends the stream: if it fits, every tile fits.  A
            // stream that did not fit records size 0 so no consumer reads it.
            unsigned long long sz = fits ? P + total : 0ull;
            ix.meta[0] = sz;
            if (size_out)
                *size_out = sz;
            chain_finished(ws, nchains);
        }
    }
}

// Output side shared by compress and the homomorphic sum: per-block widths
// from sh.bmax, block offsets, the lookback for the tile's stream offset,
// encoding into LDS, the tile store, seed / header entry and output index.
__device__ __forceinline__ void encode_tile(const Geom& g, const Tiling& tl, const Tile& x, uint64_t t,
    TileShared& sh, uint32_t* buf, unsigned long long* status, WsHeader* ws, const Units& un,
    const uint32_t (&d)[kUnitsPerThread][4], uint32_t seed,
    uint8_t* out, uint64_t out_cap, uint8_t* index, unsigned long long* size_out, bool last_tile,
    uint32_t nchains)
{
    const uint32_t tid = threadIdx.x;
    if (tid < 64) {
        // wave 0: widths, sizes and offsets of the tile's blocks (up to 4 per
        // lane, contiguous runs) with one wave scan
        const uint32_t cpt = (uint32_t) ceil_div(tl.G, 64);
        uint32_t loc[4] = {0, 0, 0, 0};
        uint32_t run = 0;
        for (uint32_t c = 0; c < cpt; c++) {
            uint32_t j = tid * cpt + c;
            if (j < x.nblk) {
                uint32_t w = dwidth(sh.bmax[j]);
                sh.width[j] = (uint8_t) w;
                loc[c] = run;
                run += block_bytes(w, tile_block_len(g, x, j));
            }
        }
        const uint32_t incl = wave_incl_scan(run);
        for (uint32_t c = 0; c < cpt; c++) {
            uint32_t j = tid * cpt + c;
            if (j < x.nblk)
                sh.boff[j] = incl - run + loc[c];
        }
        if (tid == 63) {
            sh.bc[0] = incl;
            lookback_publish(status, t, 0, incl, 0);  // successors see the aggregate now
        }
    }
    {
        // the other waves (wave 0 too when it is the only one) clear the
        // encode buffer for the largest tile
        const uint32_t nvec = (tl.tile_bytes_max + 32 + 15) / 16;
        const uint32_t first = blockDim.x > 64 ? 64u : 0u;
        if (tid >= first)
            for (uint32_t i = tid - first; i < nvec; i += blockDim.x - first)
                reinterpret_cast<uint4*>(buf)[i] = make_uint4(0, 0, 0, 0);
    }
    __syncthreads();
    const uint32_t total = (uint32_t) sh.bc[0];
#pragma unroll
    for (uint32_t i = 0; i < kUnitsPerThread; i++) {
        const uint32_t pk = un.pk[i], cnt = unit_cnt(pk);
        if (cnt) {
            const uint32_t j = unit_j(pk);
            uint32_t w = sh.width[j];
            if (w)
                encode_unit(buf, sh.boff[j], w, tile_block_len(g, x, j), unit_u(pk), cnt, d[i]);
        }
    }
    for (uint32_t j = tid; j < x.nblk; j += blockDim.x)
        lds_or4(buf, sh.boff[j], sh.width[j]);
    // the encoding needs only offsets inside the tile; wave 0 resolves the
    // tile's stream offset after it, when predecessors have had longer to
    // publish
    finish_tile(g, x, t, sh, buf, total, status, ws, seed, out, out_cap, index, size_out, last_tile, nchains);
}

// ------------------------------------------------------------ compress

// Fetch `bytes` bytes from src (a global address with any 4-byte alignment)
// into LDS at dst + (src mod 16), as aligned 16-byte chunks moved by
// direct-to-LDS loads in runs of 64 per wave.  Chunks not wholly inside
// [lo, hi) are read byte by byte within it.  Returns src mod 16.
__device__ __forceinline__ uint32_t fetch_range(const uint8_t* src, uint32_t bytes, const uint8_t* lo, const uint8_t* hi,
    uint8_t* dst)
{
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6, nw = blockDim.x >> 6;
    const uint32_t shift = (uint32_t) ((uintptr_t) src & 15u);
    const uint8_t* a = src - shift;
    const uint32_t nch = (shift + bytes + 15) / 16;
    const uint32_t nruns = (nch + 63) / 64;
    for (uint32_t q = wid; q < nruns; q += nw) {
        const uint32_t c = q * 64 + lane;
        uint8_t* lbase = dst + q * 1024u;
        if (c >= nch)
            continue;
        const uint8_t* g = a + 16ull * c;
        if (g >= lo && g + 16 <= hi) {
            __builtin_amdgcn_global_load_lds((const void*) g, (__attribute__((address_space(3))) void*) lbase, 16, 0, 0);
        } else {
            uint32_t t[4] = {0, 0, 0, 0};
            for (int i = 0; i < 16; i++)
                if (g + i >= lo && g + i < hi)
                    t[i >> 2] |= (uint32_t) g[i] << (8 * (i & 3));
            __builtin_memcpy(lbase + 16u * lane, t, 16);
        }
    }
    return shift;
}

__device__ __forceinline__ void compress_tile(const Geom& g, const Tiling& tl, const OpParams& p, const CompressIO& io,
    uint64_t t, unsigned long long* status, WsHeader* ws, uint32_t nchains)
{
    __shared__ TileShared sh;
    uint32_t* buf = zk_lds;
    const Tile x = tile_of(g, tl, t);
    const uint32_t tid = threadIdx.x;
    for (uint32_t j = tid; j < tl.G; j += blockDim.x)
        sh.bmax[j] = 0;
    const Units un = tile_units(g, x);
    // the tile's values and the one before them, staged in the encode buffer
    const uint32_t tile_end = x.nblk ? x.e0 + (x.nblk - 1) * g.L + tile_block_len(g, x, x.nblk - 1) : x.e0;
    const uint32_t shift = fetch_range((const uint8_t*) (io.x + x.e0 - 1), 4u * (tile_end - x.e0 + 1),
        (const uint8_t*) io.x, (const uint8_t*) (io.x + g.n), (uint8_t*) buf);
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
    const uint8_t* xs = (const uint8_t*) buf + shift;  // float i of the tile at xs + 4 * (e - (e0 - 1))
    uint32_t d[kUnitsPerThread][4];
#pragma unroll
    for (uint32_t i = 0; i < kUnitsPerThread; i++) {
        const uint32_t pk = un.pk[i], cnt = unit_cnt(pk);
        const uint32_t rel = 1 + unit_j(pk) * g.L + 4 * unit_u(pk);  // index of the unit's first value
        // dword reads: xs is 4-byte aligned, a 16-byte read from it is not
        const float* xf = (const float*) (xs + 4u * rel);
        const float v[4] = {xf[0], xf[1], xf[2], xf[3]};
        const float pv = xf[-1];
        const uint32_t prev = cnt ? (uint32_t) quant(pv, p.inv) : 0u;
#pragma unroll
        for (int c = 0; c < 4; c++)
            d[i][c] = (uint32_t) c < cnt ? (uint32_t) quant(v[c], p.inv) : 0u;
        // deltas in place (int32 wrap)
#pragma unroll
        for (int c = 3; c >= 1; c--)
            d[i][c] -= d[i][c - 1];
        d[i][0] -= prev;
    }
    __syncthreads();
    fold_block_max(sh.bmax, un, d);
    __syncthreads();
    uint32_t seed = 0;
    if (x.lt == 0 && tid == 0)
        seed = (uint32_t) quant(io.x[x.seg_a], p.inv);
    encode_tile(g, tl, x, t, sh, buf, status, ws, un, d, seed, io.out, io.out_cap, io.index, io.size_out,
        t == tl.ntiles - 1, nchains);
}

// ------------------------------------------------------------ operand staging

struct OperandShared {
    unsigned long long base[kMaxOperands];  // stream offset of the tile's first block
    unsigned long long size[kMaxOperands];  // stream size
    unsigned long long seedoff[kMaxOperands];
    uint32_t tot[kMaxOperands];             // tile bytes
    uint32_t lds[kMaxOperands];             // LDS position of the operand's bytes
    uint32_t bad[kMaxOperands];
    uint32_t shift[kMaxOperands];           // (stream address of the tile bytes) mod 16
};

// Read the k operands' widths and offsets for tile x; per-operand scans give
// each block's byte offset inside its operand's tile bytes.  Table s_wo[k][G]
// holds offset | width << 16 (a width above 32 raises ZK_ERR_WIDTH and reads
// as 0).
__device__ __forceinline__ void operand_tables(const Geom& g, const Tiling& tl, const Tile& x, int k, const Operand* in,
    OperandShared& os, uint32_t* s_wo, WsHeader* ws)
{
    const uint32_t G = tl.G;
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6, nw = blockDim.x >> 6;
    // one wave per operand: its offsets, and the widths of blocks 4*lane ..
    // 4*lane+3 as one (unaligned) dword; bytes past the tile's blocks are
    // masked, and stay inside the index (the group offsets follow the widths)
    for (int r = wid; r < k; r += nw) {
        const IndexView ix = index_view(g, in[r].index);
        const unsigned long long base = ix.group_off[x.grp0];
        const unsigned long long size = ix.meta[0];
        uint32_t wv = 0;
        if (4u * lane < x.nblk)
            __builtin_memcpy(&wv, ix.widths + x.g0 + 4u * lane, 4);
        uint32_t sz[4], wj[4], run = 0;
#pragma unroll
        for (int c = 0; c < 4; c++) {
            const uint32_t j = 4 * lane + c;
            uint32_t w = j < x.nblk ? (wv >> (8 * c)) & 255u : 0u;
            if (w > 32) {
                raise_error(ws, ZK_ERR_WIDTH);
                w = 0;
            }
            wj[c] = w;
            sz[c] = j < x.nblk ? block_bytes(w, tile_block_len(g, x, j)) : 0u;
            run += sz[c];
        }
        const uint32_t incl = wave_incl_scan(run);
        uint32_t off = incl - run;
#pragma unroll
        for (int c = 0; c < 4; c++) {
            const uint32_t j = 4 * lane + c;
            if (j < G)
                s_wo[r * G + j] = off | (wj[c] << 16);
            off += sz[c];
        }
        if (lane == 63) {
            os.base[r] = base;
            os.shift[r] = (uint32_t) ((uintptr_t) (in[r].bytes + base) & 15u);
            os.size[r] = size;
            os.seedoff[r] = x.lt == 0 ? ix.seed_off[x.s] : 0ull;
            os.tot[r] = incl;
            const bool bad = base + incl > size;
            os.bad[r] = bad;
            if (bad)
                raise_error(ws, ZK_ERR_OVERRUN);
        }
    }
    __syncthreads();
}

// Every staged block's width byte must equal the index width.
__device__ __forceinline__ void check_width_bytes(const uint8_t* sbytes, const OperandShared& os, const uint32_t* s_wo,
    uint32_t G, uint32_t nblk, int r0, int r1, WsHeader* ws)
{
    bool bad = false;
    for (uint32_t idx = threadIdx.x; idx < (uint32_t) (r1 - r0) * G; idx += blockDim.x) {
        const uint32_t r = r0 + idx / G, j = idx % G;
        if (j < nblk && !os.bad[r]) {
            const uint32_t e = s_wo[r * G + j];
            bad |= sbytes[os.lds[r] + (e & 0xFFFFu)] != (e >> 16);
        }
    }
    if (bad)
        raise_error(ws, ZK_ERR_MISMATCH);
}

// Group operands [r0, r) so their tile bytes (plus 16 bytes of read slack
// each) fit the stage area; fills os.lds for the group and returns its end.
// Executed by one thread; the caller synchronises.
__device__ __forceinline__ int plan_group(OperandShared& os, int r0, int k, uint32_t stage_bytes)
{
    uint32_t pos = 0;
    int r = r0;
    while (r < k) {
        uint32_t n = os.bad[r] ? 0u : os.tot[r];
        uint32_t need = (uint32_t) align_up(n + 16, 16);
        if (r > r0 && pos + need > stage_bytes)
            break;
        os.lds[r] = pos;
        pos += need;
        r++;
    }
    return r;
}

// Load operands [r0, r1) of a group into LDS: LDS byte os.lds[r] + i holds
// stream byte os.base[r] + i.  The workgroup's threads are split evenly over
// the group's operands; each thread issues its operand's 16-byte chunks in
// batches before writing any of them, so a group costs about one memory round
// trip.  Only the operand's last chunk can cross the stream's end; it is read
// byte by byte when it does, so no byte past the stream is read.
constexpr int kLoadBatch = 8;

__device__ __forceinline__ void load_group(const Operand* in, const OperandShared& os, int r0, int r1, uint32_t* stage)
{
    const uint32_t ng = (uint32_t) (r1 - r0);
    const uint32_t tpo = blockDim.x / ng;  // threads per operand (blockDim >= 64 >= ng)
    const uint32_t r = r0 + threadIdx.x / tpo;
    const uint32_t q = threadIdx.x % tpo;
    if (r >= (uint32_t) r1 || os.bad[r])
        return;
    const uint32_t nch = (os.tot[r] + 15) / 16;
    if (!nch)
        return;
    const uint64_t base = os.base[r], size = os.size[r];
    const uint8_t* src = in[r].bytes + base;
    uint4* dst = (uint4*) ((uint8_t*) stage + os.lds[r]);
    const uint32_t nfast = nch - 1;  // chunks that end inside the stream's tile bytes
    for (uint32_t c0 = q; c0 < nfast; c0 += kLoadBatch * tpo) {
        uint4 v[kLoadBatch];
#pragma unroll
        for (int b = 0; b < kLoadBatch; b++) {
            const uint32_t c = c0 + b * tpo;
            if (c < nfast)
                v[b] = ld16(src + 16ull * c);
        }
#pragma unroll
        for (int b = 0; b < kLoadBatch; b++) {
            const uint32_t c = c0 + b * tpo;
            if (c < nfast)
                dst[c] = v[b];
        }
    }
    if (q == (nfast % tpo)) {
        const uint64_t off = base + 16ull * nfast;
        uint4 v;
        if (off + 16 <= size) {
            v = ld16(src + 16ull * nfast);
        } else {
            uint32_t t[4] = {0, 0, 0, 0};
            for (uint32_t i = 0; i < 16 && off + i < size; i++)
                t[i >> 2] |= (uint32_t) src[16ull * nfast + i] << (8 * (i & 3));
            v = make_uint4(t[0], t[1], t[2], t[3]);
        }
        dst[nfast] = v;
    }
}

// 32-bit LDS address of a pointer into the workgroup's shared memory.
__device__ __forceinline__ uint32_t lds_addr(const void* p)
{
    return (uint32_t) (uintptr_t) (const __attribute__((address_space(3))) void*) p;
}

// Direct-to-LDS load of 16 bytes per lane: lane i's bytes from g land at LDS
// byte lds_base + 16 * i (lds_base is the same for every lane).  Issued from
// inline assembly on purpose: the compiler models the builtin form as an
// asynchronous write to all of LDS and puts a full vector-memory wait in
// front of every later LDS read, which would make each round's decode wait
// for the next round's prefetch.  The instruction writes no register; the
// caller waits for the vector-memory counter (s_waitcnt) before it reads
// the bytes.
__device__ __forceinline__ void dma16(const void* g, uint32_t lds_base)
{
    asm volatile(
        "s_nop 0\n\t"
        "global_load_lds_dwordx4 %0, off"
        :
        : "v"(g), "{m0}"(__builtin_amdgcn_readfirstlane(lds_base))
        : "memory");
}

// Fetch round g of a staging plan into the stage: every wave takes runs of 64
// consecutive 16-byte chunks of one operand and moves them from HBM to LDS
// with direct-to-LDS loads (no registers, completion tracked by the wave's
// vector memory counter).  Chunks not wholly inside the stream (the first
// chunk when the aligned start precedes the stream, the last one past the
// stream's end) are read byte by byte instead, so no byte outside the stream
// is read.  The caller waits for the counter and synchronises before use.
__device__ __forceinline__ void issue_round(const Operand* in, const OperandShared& os, const StagePlan& sp, int g,
    uint8_t* stage)
{
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6, nw = blockDim.x >> 6;
    const int r0 = g ? sp.gend[g - 1] : 0, r1 = sp.gend[g];
    uint32_t run = 0;
    for (int r = r0; r < r1; r++) {
        const uint32_t nch = sp.nch[r];
        const uint32_t nruns = (nch + 63) / 64;
        const int64_t start = (int64_t) os.base[r] - (int64_t) os.shift[r];  // stream offset of chunk 0
        const int64_t size = (int64_t) os.size[r];
        const uint8_t* bytes = in[r].bytes;
        for (uint32_t q = 0; q < nruns; q++, run++) {
            if ((int) (run % (uint32_t) nw) != wid)
                continue;
            const uint32_t c = q * 64 + lane;
            uint8_t* lbase = stage + sp.lds[r] + q * 1024u;
            if (c >= nch)
                continue;
            const int64_t off = start + 16ll * c;
            if (off >= 0 && off + 16 <= size) {
                dma16(bytes + off, lds_addr(lbase));
            } else {
                uint32_t t[4] = {0, 0, 0, 0};
                for (int i = 0; i < 16; i++)
                    if (off + i >= 0 && off + i < size)
                        t[i >> 2] |= (uint32_t) bytes[off + i] << (8 * (i & 3));
                __builtin_memcpy(lbase + 16u * lane, t, 16);
            }
        }
    }
}

// ------------------------------------------------------------ homomorphic sum

__device__ __forceinline__ void homsum_tile(const Geom& g, const Tiling& tl, const OpParams& p, const HomsumIO& io,
    uint64_t t, unsigned long long* status, WsHeader* ws, uint32_t nchains)
{
    __shared__ TileShared sh;
    __shared__ OperandShared os;
    const Tile x = tile_of(g, tl, t);
    const uint32_t tid = threadIdx.x;
    uint32_t* stage = zk_lds;
    uint32_t* s_wo = zk_lds + p.stage_bytes / 4;
    const Units un = tile_units(g, x);
    for (uint32_t j = tid; j < tl.G; j += blockDim.x)
        sh.bmax[j] = 0;  // visible to the block-max atomics after the staging barriers
    const uint32_t last_j = x.nblk ? x.nblk - 1 : 0u;
    const uint32_t last_len = x.nblk ? tile_block_len(g, x, last_j) : 0u;
    uint32_t nj[kUnitsPerThread], sbj[kUnitsPerThread];
#pragma unroll
    for (uint32_t i = 0; i < kUnitsPerThread; i++) {
        nj[i] = unit_j(un.pk[i]) == last_j ? last_len : g.L;
        sbj[i] = (nj[i] + 7) >> 3;
    }
    operand_tables(g, tl, x, p.k, io.in, os, s_wo, ws);
    uint32_t acc[kUnitsPerThread][4];
#pragma unroll
    for (uint32_t i = 0; i < kUnitsPerThread; i++)
#pragma unroll
        for (int c = 0; c < 4; c++)
            acc[i][c] = 0;
    __shared__ StagePlan sp;
    if (tid == 0) {
        uint32_t tot[kMaxOperands];
        for (int r = 0; r < p.k; r++)
            tot[r] = os.bad[r] ? 0u : os.tot[r];
        plan_stage(tot, os.shift, p.k, p.stage_bytes / 2, sp);
        for (int r = 0; r < p.k; r++)
            os.lds[r] = sp.lds[r] + os.shift[r];
    }
    __syncthreads();
    uint8_t* stage8 = (uint8_t*) stage;
    issue_round(io.in, os, sp, 0, stage8);
    for (int round = 0; round < sp.nrounds; round++) {
        __builtin_amdgcn_s_waitcnt(0);
        __syncthreads();  // round `round` is in LDS
        const bool overlap = round + 1 < sp.nrounds && !sp.whole[round] && !sp.whole[round + 1];
        if (overlap)
            issue_round(io.in, os, sp, round + 1, stage8);  // lands in the other half during the decode
        const int r0 = round ? sp.gend[round - 1] : 0, r1 = sp.gend[round];
        check_width_bytes(stage8, os, s_wo, tl.G, x.nblk, r0, r1, ws);
        for (int r = r0; r < r1; r++) {
            if (__builtin_amdgcn_readfirstlane(os.bad[r]))
                continue;
            const uint8_t* ob = stage8 + __builtin_amdgcn_readfirstlane(os.lds[r]);
            const uint32_t* tab = s_wo + r * tl.G;
#pragma unroll
            for (uint32_t i = 0; i < kUnitsPerThread; i++) {
                const uint32_t e = tab[unit_j(un.pk[i])];
                if (e >> 16)
                    accumulate_unit(ob + (e & 0xFFFFu), e >> 16, nj[i], sbj[i], unit_u(un.pk[i]), acc[i]);
            }
        }
        __syncthreads();  // this round's half is free
        if (round + 1 < sp.nrounds && !overlap)
            issue_round(io.in, os, sp, round + 1, stage8);
    }
    uint32_t seed = 0;
    if (x.lt == 0 && tid == 0) {
        for (int r = 0; r < p.k; r++) {
            uint64_t so = os.seedoff[r];
            if (so + 4 <= os.size[r])
                seed += load_u32_bytes(io.in[r].bytes + so);
            else
                raise_error(ws, ZK_ERR_OVERRUN);
        }
    }
#pragma unroll
    for (uint32_t i = 0; i < kUnitsPerThread; i++)
#pragma unroll
        for (int c = 0; c < 4; c++)
            if ((uint32_t) c >= unit_cnt(un.pk[i]))
                acc[i][c] = 0;
    fold_block_max(sh.bmax, un, acc);
    __syncthreads();
    encode_tile(g, tl, x, t, sh, stage, status, ws, un, acc, seed, io.out, io.out_cap, io.index, io.size_out,
        t == tl.ntiles - 1, nchains);
}

// ------------------------------------------------------------ homomorphic sum, block length fixed at compile time

// Slot layout: a wave slot is BPS whole blocks with one 4-value unit per
// lane (lane = U * block-in-slot + unit), so no block straddles two waves
// and the two units that share a sign byte or a byte of low bits sit in
// neighbouring lanes.  Each of the W waves owns 4 consecutive slots, so a
// tile holds G = 4 * W * BPS blocks; with IG = 16 this is the tile that
// make_tiling(g, 64 W) cuts (TG = W * 16 / U groups), so the layout in HBM
// and the lookback chains are those of the generic kernels.  The sum uses
// 4 waves, compress 8.
template <uint32_t L, uint32_t W = 4>
struct SlotMap {
    static constexpr uint32_t U = (L + 3) / 4;
    static constexpr uint32_t BPS = 64 / U;
    static constexpr uint32_t G = 4 * W * BPS;
    static constexpr uint32_t SB = (L + 7) / 8;
    static constexpr uint32_t ACTIVE = BPS * U;
    static_assert(U >= 4 && U <= 36 && G <= kMaxTileBlocks, "slot layout needs 13 <= L <= 144");
};

template <uint32_t L, uint32_t W = 4>
struct SlotShared {
    unsigned long long bc[4];
    uint32_t bmax[SlotMap<L, W>::G];
    uint32_t boff[SlotMap<L, W>::G];
    uint8_t width[SlotMap<L, W>::G];
};

// Decoder of the slot-layout sum: add the up-to-4 values of unit u of one
// operand block (width byte at bp, width w > 0, nj values, sb sign bytes) to
// acc[] as magnitude XOR sign mask; the +1 that completes the two's
// complement of each negative value is counted in negs instead, one byte per
// value (byte 3 - c counts value c; at most 32 operands, so no byte
// overflows), and added once after all operands.  u4 = 4u, uh = u >> 1 and
// nsh = (u & 1) ? 0 : 4 are the lane's unit constants.  The sign byte, the
// unit's first 4 high bytes and the dword holding its low bits are read
// together, so a slot of widths below 16 costs one LDS round trip; the reads
// reach at most 20 bytes past the block's last byte.  Same fields as accumulate_unit:
// ZCCL/src/ZCCL_TypeManager.c:423-517 and :906-950.
__device__ __forceinline__ void accumulate_slot(const uint8_t* bp, uint32_t w, uint32_t nj, uint32_t sb,
    uint32_t u4, uint32_t uh, uint32_t nsh, uint32_t acc[4], uint32_t& negs)
{
    const uint32_t B = w >> 3, r = w & 7;
    const uint8_t* hp = bp + 1 + sb;
    const uint32_t bitpos = __umul24(u4, r);  // a multiple of 4
    const uint8_t* hu = hp + __umul24(u4, B);
    const uint32_t sbyte = bp[1 + uh];
    const uint32_t H0 = lds_ld4(hu);
    const uint32_t Xr = lds_ld4(hp + __umul24(B, nj) + (bitpos >> 3));
    const uint32_t nib = __builtin_amdgcn_ubfe(sbyte, nsh, 4);
    negs += __umul24(nib, 0x00204081u) & 0x01010101u;  // bit 3 - c of nib -> byte 3 - c
    uint32_t hs[4];
    if (B <= 1) {
#pragma unroll
        for (int c = 0; c < 4; c++)
            hs[c] = __builtin_amdgcn_ubfe(H0, 8 * c, 8 * B);
    } else if (B == 2) {
        // widths of 16 and more: rare, a second read
        const uint2 H = lds_ld8(hu);
        hs[0] = H.x & 0xFFFFu;
        hs[1] = H.x >> 16;
        hs[2] = H.y & 0xFFFFu;
        hs[3] = H.y >> 16;
    } else {
        const uint4 W = lds_ld16(hu);
        if (B == 3) {
            hs[0] = W.x & 0xFFFFFFu;
            hs[1] = __builtin_amdgcn_alignbyte(W.y, W.x, 3) & 0xFFFFFFu;
            hs[2] = __builtin_amdgcn_alignbyte(W.z, W.y, 2) & 0xFFFFFFu;
            hs[3] = W.z >> 8;
        } else {
            hs[0] = W.x; hs[1] = W.y; hs[2] = W.z; hs[3] = W.w;
        }
    }
    // the unit's 4r low bits, MSB-first from bit (bitpos & 7) of the
    // big-endian dword; shifted to the top of lo, value c's field is lo's top
    // r bits.  One 64-bit shift of {hs[c], lo} by r gives the magnitude
    // (high dword) and the next value's window (low dword), r == 0 included
    // (4.8 % faster launches than a bit-field extract and a position update
    // per value).  4r + (bitpos & 4) <= 32, so the dword holds every field.
    uint32_t lo = __builtin_bswap32(Xr) << (bitpos & 4u);
#pragma unroll
    for (int c = 0; c < 4; c++) {
        const uint64_t v = ((uint64_t) hs[c] << 32 | lo) << r;
        const uint32_t m = (uint32_t) (v >> 32);
        lo = (uint32_t) v;
        const uint32_t sm = (uint32_t) __builtin_amdgcn_sbfe((int32_t) nib, 3 - c, 1);  // 0 or ~0
        acc[c] += m ^ sm;
    }
}

// Add one operand's values to the lane's four slots: the four table entries
// are read together, the data of each slot one slot at a time (batching the
// data reads of all four slots as well was measured slower: it takes
// registers that cost workgroups per CU, and residency hides more latency).
// tab points at the table entry of the lane's block in slot 0; slot i's
// block is BPS entries further.  Every block's width byte in the stream must
// equal its index width: the byte is read with the slot's other reads and a
// mismatch sets `bad`.
template <uint32_t BPS>
__device__ __forceinline__ void decode_operand(const uint8_t* ob, const uint32_t* tab, const uint32_t (&nj)[4],
    uint32_t u4, uint32_t uh, uint32_t nsh, uint32_t (&acc)[4][4], uint32_t (&negs)[4], bool& bad)
{
    uint32_t e[4];
#pragma unroll
    for (uint32_t i = 0; i < 4; i++)
        e[i] = tab[i * BPS];
#pragma unroll
    for (uint32_t i = 0; i < 4; i++) {
        const bool has = u4 < nj[i];
        const uint32_t w = has ? e[i] >> 16 : 0u;
        const uint8_t* bp = ob + (e[i] & 0xFFFFu);  // inside the operand's staged bytes for any table entry
        const uint32_t wb = bp[0];
        if (w)
            accumulate_slot(bp, w, nj[i], (nj[i] + 7) >> 3, u4, uh, nsh, acc[i], negs[i]);
        bad |= has && wb != (e[i] >> 16);
    }
}

// Encode this lane's unit in the slot layout (emit_unit in zccl_internal.h
// holds the byte ownership rules; the unit-0 lane writes the width byte); the
// next lane's fields arrive by a lane shuffle, so every lane of the wave
// calls this.
template <uint32_t U>
__device__ __forceinline__ void encode_unit_pair(uint8_t* buf, uint32_t bo, uint32_t w, uint32_t nj, uint32_t u,
    uint32_t cnt, const uint32_t d[4])
{
    const UnitFields f = unit_fields(w, d);
    const uint32_t pnib = __shfl_down(f.nib, 1, 64);
    const uint32_t plv = __shfl_down(f.lv, 1, 64);
    if (u == 0 && cnt)
        buf[bo] = (uint8_t) w;
    emit_unit(buf, bo + 1, w, nj, u, U, cnt, f, pnib, plv);
}

// Output side of the slot-layout kernels, once sh.bmax holds every block's
// largest magnitude and d[] the lane's four units (two's complement): wave 0
// computes widths (dwidth), block sizes and offsets with one wave scan and
// publishes the tile's aggregate; every lane encodes its units into buf (one
// writer per byte); finish_tile resolves the tile's offset and stores it.
template <uint32_t L, uint32_t W = 4>
__device__ __forceinline__ void slot_output(const Geom& g, const Tiling& tl, const Tile& x, uint64_t t,
    SlotShared<L, W>& sh, uint32_t* buf, uint32_t jb0, const uint32_t (&nj)[4], uint32_t u, uint32_t last_len,
    const uint32_t (&d)[4][4], uint32_t seed, uint8_t* out, uint64_t out_cap, uint8_t* index,
    unsigned long long* size_out, unsigned long long* status, WsHeader* ws, uint32_t nchains)
{
    using M = SlotMap<L, W>;
    constexpr uint32_t G = M::G;
    const uint32_t tid = threadIdx.x;
    const uint32_t u4 = 4 * u;
    __syncthreads();
    if (tid < 64) {
        // wave 0: widths, sizes and offsets of the tile's blocks with one wave scan
        constexpr uint32_t CPT = (G + 63) / 64;
        uint32_t loc[CPT];
        uint32_t run = 0;
#pragma unroll
        for (uint32_t c = 0; c < CPT; c++) {
            const uint32_t j = tid * CPT + c;
            loc[c] = run;
            if (j < x.nblk) {
                const uint32_t w = dwidth(sh.bmax[j]);
                sh.width[j] = (uint8_t) w;
                run += block_bytes(w, j == x.nblk - 1 ? last_len : L);
            }
        }
        const uint32_t incl = wave_incl_scan(run);
#pragma unroll
        for (uint32_t c = 0; c < CPT; c++) {
            const uint32_t j = tid * CPT + c;
            if (j < x.nblk)
                sh.boff[j] = incl - run + loc[c];
        }
        if (tid == 63) {
            sh.bc[0] = incl;
            lookback_publish(status, t, 0, incl, 0);  // successors see the aggregate now
        }
    }
    __syncthreads();
    const uint32_t total = (uint32_t) sh.bc[0];
#pragma unroll
    for (uint32_t i = 0; i < 4; i++) {
        const uint32_t j = jb0 + i * M::BPS, cnt = u4 < nj[i] ? min(4u, nj[i] - u4) : 0u;
        encode_unit_pair<M::U>((uint8_t*) buf, sh.boff[j], cnt ? sh.width[j] : 0u, nj[i], u, cnt, d[i]);
    }
    finish_tile(g, x, t, sh, buf, total, status, ws, seed, out, out_cap, index, size_out, t == tl.ntiles - 1, nchains);
}

template <uint32_t L>
__device__ __forceinline__ void homsum_tile_slots(const Geom& g, const Tiling& tl, const OpParams& p, const HomsumIO& io,
    uint64_t t, unsigned long long* status, WsHeader* ws, uint32_t nchains)
{
    using M = SlotMap<L>;
    constexpr uint32_t G = M::G;
    __shared__ SlotShared<L> sh;
    __shared__ OperandShared os;
    __shared__ StagePlan sp;
    const Tile x = tile_of(g, tl, t);
    const uint32_t tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    uint32_t* stage = zk_lds;
    uint32_t* s_wo = zk_lds + p.stage_bytes / 4;
    uint8_t* stage8 = (uint8_t*) stage;
    // unit of this lane inside its block, the same in all four slots
    // (the idle lanes past the wave's BPS blocks take unit 0 of the slots'
    // first blocks with no values)
    const bool on = lane < M::ACTIVE;
    const uint32_t jj = on ? lane / M::U : 0u, u = on ? lane - jj * M::U : 0u;
    const uint32_t u4 = 4 * u, uh = u >> 1, nsh = (u & 1) ? 0u : 4u;
    // (a block holds at most L values; the bound lets the compiler use 24-bit products)
    const uint32_t last_len = x.nblk ? min(tile_block_len(g, x, x.nblk - 1), L) : 0u;
    // slot i's block is jb0 + i * BPS; nj[i] its values (0: no block), and
    // the lane's value count min(4, nj - 4u) is derived where needed
    const uint32_t jb0 = wv * 4 * M::BPS + jj;
    uint32_t nj[4];
#pragma unroll
    for (uint32_t i = 0; i < 4; i++) {
        const uint32_t j = jb0 + i * M::BPS;
        nj[i] = on && j < x.nblk ? (j == x.nblk - 1 ? last_len : L) : 0u;
    }
    for (uint32_t j = tid; j < G; j += blockDim.x)
        sh.bmax[j] = 0;
    operand_tables(g, tl, x, p.k, io.in, os, s_wo, ws);
    uint32_t acc[4][4], negs[4];
#pragma unroll
    for (uint32_t i = 0; i < 4; i++) {
        negs[i] = 0;
#pragma unroll
        for (int c = 0; c < 4; c++)
            acc[i][c] = 0;
    }
    // operand r's tile bytes and shift in lane r of every wave: the plan (wave
    // 0) and the decode read them with readlane instead of LDS round trips
    const uint32_t op_tot = lane < (uint32_t) p.k && !os.bad[lane] ? os.tot[lane] : 0u;
    const uint32_t op_shift = lane < (uint32_t) p.k ? os.shift[lane] : 0u;
    if (tid == 0)
        plan_stage_by([&](int r, uint32_t& tt, uint32_t& sh) {
            tt = __builtin_amdgcn_readlane(op_tot, r);
            sh = __builtin_amdgcn_readlane(op_shift, r);
        }, p.k, p.stage_bytes / 2, sp);
    __syncthreads();
    if (lane < (uint32_t) p.k)
        os.lds[lane] = sp.lds[lane] + op_shift;
    // every wave wrote the same values; the staging reads them after this barrier
    __syncthreads();
    const uint32_t op_lds = lane < (uint32_t) p.k ? sp.lds[lane] + op_shift : 0u;
    const uint32_t op_bad = lane < (uint32_t) p.k ? os.bad[lane] : 1u;
    bool bad = false;  // a staged width byte differs from its index width
    issue_round(io.in, os, sp, 0, stage8);
    for (int round = 0; round < sp.nrounds; round++) {
        __builtin_amdgcn_s_waitcnt(0);
        __syncthreads();  // round `round` is in LDS
        const bool overlap = round + 1 < sp.nrounds && !sp.whole[round] && !sp.whole[round + 1];
        if (overlap)
            issue_round(io.in, os, sp, round + 1, stage8);  // lands in the other half during the decode
        const int r0 = round ? sp.gend[round - 1] : 0, r1 = sp.gend[round];
        for (int r = r0; r < r1; r++) {
            if (__builtin_amdgcn_readlane(op_bad, r))
                continue;
            decode_operand<M::BPS>(stage8 + __builtin_amdgcn_readlane(op_lds, r), s_wo + r * G + jb0, nj, u4, uh,
                nsh, acc, negs, bad);
        }
        __syncthreads();  // this round's half is free
        if (round + 1 < sp.nrounds && !overlap)
            issue_round(io.in, os, sp, round + 1, stage8);
    }
    if (bad)
        raise_error(ws, ZK_ERR_MISMATCH);
    uint32_t seed = 0;
    if (x.lt == 0 && tid == 0) {
        for (int r = 0; r < p.k; r++) {
            uint64_t so = os.seedoff[r];
            if (so + 4 <= os.size[r])
                seed += load_u32_bytes(io.in[r].bytes + so);
            else
                raise_error(ws, ZK_ERR_OVERRUN);
        }
    }
    // complete the negative values' two's complement; values past the block's
    // end decoded garbage: clear them; then block maxima
#pragma unroll
    for (uint32_t i = 0; i < 4; i++) {
        const uint32_t cnt = u4 < nj[i] ? min(4u, nj[i] - u4) : 0u;
        uint32_t umax = 0;
#pragma unroll
        for (int c = 0; c < 4; c++) {
            acc[i][c] += (negs[i] >> (8 * (3 - c))) & 0xFFu;
            if ((uint32_t) c >= cnt)
                acc[i][c] = 0;
            umax = max(umax, umag((int32_t) acc[i][c]));
        }
        if (umax)
            atomicMax(&sh.bmax[jb0 + i * M::BPS], umax);
    }
    slot_output<L>(g, tl, x, t, sh, stage, jb0, nj, u, last_len, acc, seed, io.out, io.out_cap, io.index, io.size_out,
        status, ws, nchains);
}

// Compress with the slot layout (36-value blocks, 8 waves): the tile's
// values and the one before them are staged in LDS (direct-to-LDS loads);
// each lane quantizes its four units (FP64 product, truncation,
// ZCCL/src/ZCCL_float.c:563,577-604) and takes their deltas, the chain
// continuing across blocks; block maxima by LDS atomics; then the output side
// shared with the sum (one writer per encoded byte, no buffer clearing, no
// LDS atomics in the encoder).
template <uint32_t L, uint32_t W>
__device__ __forceinline__ void compress_tile_slots(const Geom& g, const Tiling& tl, const OpParams& p,
    const CompressIO& io, uint64_t t, unsigned long long* status, WsHeader* ws, uint32_t nchains)
{
    using M = SlotMap<L, W>;
    constexpr uint32_t G = M::G;
    __shared__ SlotShared<L, W> sh;
    uint32_t* buf = zk_lds;
    const Tile x = tile_of(g, tl, t);
    const uint32_t tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const bool on = lane < M::ACTIVE;
    const uint32_t jj = on ? lane / M::U : 0u, u = on ? lane - jj * M::U : 0u;
    const uint32_t u4 = 4 * u;
    // (a block holds at most L values; the bound lets the compiler use 24-bit products)
    const uint32_t last_len = x.nblk ? min(tile_block_len(g, x, x.nblk - 1), L) : 0u;
    const uint32_t jb0 = wv * 4 * M::BPS + jj;
    uint32_t nj[4];
#pragma unroll
    for (uint32_t i = 0; i < 4; i++) {
        const uint32_t j = jb0 + i * M::BPS;
        nj[i] = on && j < x.nblk ? (j == x.nblk - 1 ? last_len : L) : 0u;
    }
    for (uint32_t j = tid; j < G; j += blockDim.x)
        sh.bmax[j] = 0;
    const uint32_t tile_end = x.nblk ? x.e0 + (x.nblk - 1) * g.L + last_len : x.e0;
    const uint32_t shift = fetch_range((const uint8_t*) (io.x + x.e0 - 1), 4u * (tile_end - x.e0 + 1),
        (const uint8_t*) io.x, (const uint8_t*) (io.x + g.n), (uint8_t*) buf);
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
    // value e of the tile at xs[e - (e0 - 1)]; every lane reads its units'
    // positions whether or not they hold values (all inside the staged
    // range's LDS allocation) and masks the rest
    const float* xs = (const float*) ((const uint8_t*) buf + shift);
    uint32_t d[4][4];
#pragma unroll
    for (uint32_t i = 0; i < 4; i++) {
        const uint32_t cnt = u4 < nj[i] ? min(4u, nj[i] - u4) : 0u;
        const float* xf = xs + 1 + (jb0 + i * M::BPS) * L + u4;
        const float pv = xf[-1];
        const float v[4] = {xf[0], xf[1], xf[2], xf[3]};
        const uint32_t prev = cnt ? (uint32_t) quant(pv, p.inv) : 0u;
#pragma unroll
        for (uint32_t c = 0; c < 4; c++)
            d[i][c] = c < cnt ? (uint32_t) quant(v[c], p.inv) : 0u;
#pragma unroll
        for (int c = 3; c >= 1; c--)
            d[i][c] -= d[i][c - 1];
        d[i][0] -= prev;
        uint32_t umax = 0;
#pragma unroll
        for (uint32_t c = 0; c < 4; c++) {
            if (c >= cnt)
                d[i][c] = 0;
            umax = max(umax, umag((int32_t) d[i][c]));
        }
        if (umax)
            atomicMax(&sh.bmax[jb0 + i * M::BPS], umax);
    }
    uint32_t seed = 0;
    if (x.lt == 0 && tid == 0)
        seed = (uint32_t) quant(io.x[x.seg_a], p.inv);
    // (slot_output's first barrier also ends every read of the staged values,
    // so the encode may reuse the buffer)
    slot_output<L, W>(g, tl, x, t, sh, buf, jb0, nj, u, last_len, d, seed, io.out, io.out_cap, io.index,
        io.size_out, status, ws, nchains);
}

// ------------------------------------------------------------ decompress

__device__ __forceinline__ void decompress_tile(const Geom& g, const Tiling& tl, const OpParams& p, const DecompressIO& io,
    uint64_t t, unsigned long long* status, WsHeader* ws, uint32_t nchains)
{
    __shared__ TileShared sh;
    __shared__ OperandShared os;
    const Tile x = tile_of(g, tl, t);
    const uint32_t tid = threadIdx.x;
    uint32_t* stage = zk_lds;
    uint32_t* s_wo = zk_lds + p.stage_bytes / 4;
    const Units un = tile_units(g, x);
    operand_tables(g, tl, x, 1, io.in, os, s_wo, ws);
    if (tid == 0)
        plan_group(os, 0, 1, p.stage_bytes);
    __syncthreads();
    load_group(io.in, os, 0, 1, stage);
    if (tid == 0) {
        uint32_t seed = 0;
        if (x.lt == 0) {
            uint64_t so = os.seedoff[0];
            if (so + 4 <= os.size[0])
                seed = load_u32_bytes(io.in[0].bytes + so);
            else
                raise_error(ws, ZK_ERR_OVERRUN);
        }
        sh.bc[3] = seed;
    }
    __syncthreads();
    uint32_t pre[kUnitsPerThread][4], us[kUnitsPerThread];
    const uint32_t last_j = x.nblk ? x.nblk - 1 : 0u;
    const uint32_t last_len = x.nblk ? tile_block_len(g, x, last_j) : 0u;
    const uint8_t* base = (const uint8_t*) stage + os.lds[0];
    const bool bad0 = __builtin_amdgcn_readfirstlane(os.bad[0]);
    // the units' table entries read together; every unit decoded without a
    // branch on its width (a zero width adds nothing, a lane without values
    // reads block 0 and masks); the width byte of every block in the stream
    // must equal its index width
    uint32_t e[kUnitsPerThread];
#pragma unroll
    for (uint32_t i = 0; i < kUnitsPerThread; i++)
        e[i] = s_wo[unit_j(un.pk[i])];
    bool badw = false;
#pragma unroll
    for (uint32_t i = 0; i < kUnitsPerThread; i++) {
        uint32_t d[4] = {0, 0, 0, 0};
        const uint32_t pk = un.pk[i], cnt = unit_cnt(pk);
        const uint32_t j = unit_j(pk);
        if (!bad0) {
            const uint32_t nj = j == last_j ? last_len : g.L;
            const uint8_t* bp = base + (e[i] & 0xFFFFu);
            const uint32_t wb = bp[0];
            accumulate_unit(bp, e[i] >> 16, nj, (nj + 7) >> 3, unit_u(pk), d);
            badw |= cnt && wb != (e[i] >> 16);
        }
        uint32_t run = 0;
#pragma unroll
        for (int c = 0; c < 4; c++) {
            run += (uint32_t) c < cnt ? d[c] : 0u;
            pre[i][c] = run;
        }
        us[i] = run;
    }
    if (badw)
        raise_error(ws, ZK_ERR_MISMATCH);
    // scan of unit sums in unit order v = i*threads + tid
    const int lane = tid & 63, wid = tid >> 6, nw = blockDim.x >> 6;
    uint32_t winc[kUnitsPerThread];
#pragma unroll
    for (uint32_t i = 0; i < kUnitsPerThread; i++) {
        winc[i] = wave_incl_scan(us[i]);
        if (lane == 63)
            sh.wave[i * nw + wid] = winc[i];
    }
    __syncthreads();
    if (wid == 0) {
        const int nv = kUnitsPerThread * nw;
        uint32_t v = lane < nv ? sh.wave[lane] : 0u;
        uint32_t inc = wave_incl_scan(v);
        if (lane < nv)
            sh.wave[lane] = inc - v;
        if (lane == nv - 1)
            sh.bc[1] = inc;
    }
    __syncthreads();
    const uint32_t tile_sum = (uint32_t) sh.bc[1];
    const uint32_t seed = (uint32_t) sh.bc[3];
    if (tid < 64) {
        unsigned long long ex = lookback(status, t, x.seg_tile0, tile_sum, x.lt == 0 ? seed : 0u, ws);
        if (tid == 0)
            sh.bc[2] = ex;
    }
    __syncthreads();
    const uint32_t carry = (uint32_t) sh.bc[2];
#pragma unroll
    for (uint32_t i = 0; i < kUnitsPerThread; i++) {
        const uint32_t pk = un.pk[i], cnt = unit_cnt(pk);
        if (!cnt)
            continue;
        const uint32_t c0 = carry + sh.wave[i * nw + wid] + winc[i] - us[i];
        float* op = io.out + x.e0 + (uint64_t) unit_j(pk) * g.L + 4 * unit_u(pk);
        if (cnt == 4) {
            float4 v;
            v.x = dequant(c0 + pre[i][0], p.eb);
            v.y = dequant(c0 + pre[i][1], p.eb);
            v.z = dequant(c0 + pre[i][2], p.eb);
            v.w = dequant(c0 + pre[i][3], p.eb);
            __builtin_memcpy(op, &v, 16);
        } else {
#pragma unroll
            for (uint32_t c = 0; c < 3; c++)
                if (c < cnt)
                    op[c] = dequant(c0 + pre[i][c], p.eb);
        }
    }
    if (tid == 0) {
        if (x.lt == 0)
            io.out[x.seg_a] = dequant(seed, p.eb);
        if (x.lt == x.seg_tiles - 1)
            chain_finished(ws, nchains);
    }
}

// ------------------------------------------------------------ kernels

__global__ void __launch_bounds__(kCompressThreads) k_compress(CompressArgs a)
{
    uint64_t t;
    unsigned long long* status;
    if (!launch_status(a.ws, a.job.tl.ntiles, t, status))
        return;
    compress_tile(a.job.g, a.job.tl, a.job.p, a.job.io, t, status, a.ws, 1);
}

__global__ void __launch_bounds__(kHomsumThreads, 4) k_homsum(HomsumArgs a)
{
    uint64_t t;
    unsigned long long* status;
    if (!launch_status(a.ws, a.job.tl.ntiles, t, status))
        return;
    homsum_tile(a.job.g, a.job.tl, a.job.p, a.job.io, t, status, a.ws, 1);
}

// Sums whose blocks hold kSlotL values and whose tiles match the slot layout.
constexpr uint32_t kSlotL = 36;

__global__ void __launch_bounds__(kHomsumThreads, 6) k_homsum_slots(HomsumArgs a)
{
    uint64_t t;
    unsigned long long* status;
    if (!launch_status(a.ws, a.job.tl.ntiles, t, status))
        return;
    homsum_tile_slots<kSlotL>(a.job.g, a.job.tl, a.job.p, a.job.io, t, status, a.ws, 1);
}

__global__ void __launch_bounds__(kCompressThreads) k_compress_slots(CompressArgs a)
{
    uint64_t t;
    unsigned long long* status;
    if (!launch_status(a.ws, a.job.tl.ntiles, t, status))
        return;
    compress_tile_slots<kSlotL, kCompressThreads / 64>(a.job.g, a.job.tl, a.job.p, a.job.io, t, status, a.ws, 1);
}

__global__ void __launch_bounds__(kDecompressThreads) k_decompress(DecompressArgs a)
{
    uint64_t t;
    unsigned long long* status;
    if (!launch_status(a.ws, a.job.tl.ntiles, t, status))
        return;
    decompress_tile(a.job.g, a.job.tl, a.job.p, a.job.io, t, status, a.ws, a.job.g.S);
}

// Batched launches (see BatchArgs).
__device__ __forceinline__ void job_of(const BatchArgs& b, uint64_t gb, uint32_t& jidx, uint64_t& t)
{
    // 32-bit arithmetic, results forced into scalar registers (every lane of
    // the workgroup agrees)
    const uint32_t g32 = (uint32_t) gb;
    uint32_t j, tt;
    if (b.interleave) {
        j = g32 % b.njobs;
        tt = g32 / b.njobs;
    } else {
        uint32_t lo = 0, hi = b.njobs;
        while (hi - lo > 1) {
            uint32_t mid = (lo + hi) / 2;
            if (b.job_tile0[mid] <= gb)
                lo = mid;
            else
                hi = mid;
        }
        j = lo;
        tt = g32 - (uint32_t) b.job_tile0[lo];
    }
    jidx = __builtin_amdgcn_readfirstlane(j);
    t = __builtin_amdgcn_readfirstlane(tt);
}

// Copy one descriptor into LDS (the workgroup reads it many times).
template <typename D>
__device__ __forceinline__ void stage_desc(const void* base, uint32_t jidx, D* dst)
{
    static_assert(sizeof(D) % 8 == 0, "descriptors are 8-byte multiples");
    const unsigned long long* src = (const unsigned long long*) ((const D*) base + jidx);
    unsigned long long* d = (unsigned long long*) dst;
    for (uint32_t i = threadIdx.x; i < sizeof(D) / 8; i += blockDim.x)
        d[i] = src[i];
    __syncthreads();
}

// Uniform batches: geometry in kernel arguments, IO blocks per job.
__global__ void __launch_bounds__(kHomsumThreads, 4) k_homsum_batch(BatchArgs b)
{
    __shared__ HomsumIO sio;
    unsigned long long* status;
    uint64_t tg;
    if (!launch_status(b.ws, b.total_tiles, tg, status))
        return;
    uint32_t jidx;
    uint64_t t;
    job_of(b, tg, jidx, t);
    stage_desc(b.jobs, jidx, &sio);
    homsum_tile(b.g, b.tl, b.p, sio, t, status + b.job_tile0[jidx] * kStatusStride, b.ws, b.total_chains);
}

__global__ void __launch_bounds__(kHomsumThreads, 6) k_homsum_batch_slots(BatchArgs b)
{
    __shared__ HomsumIO sio;
    unsigned long long* status;
    uint64_t tg;
    if (!launch_status(b.ws, b.total_tiles, tg, status))
        return;
    uint32_t jidx;
    uint64_t t;
    job_of(b, tg, jidx, t);
    stage_desc(b.jobs, jidx, &sio);
    // a uniform batch's jobs have equal tile counts: job j's tiles start at j * ntiles
    homsum_tile_slots<kSlotL>(b.g, b.tl, b.p, sio, t, status + (uint64_t) jidx * b.tl.ntiles * kStatusStride, b.ws,
        b.total_chains);
}

__global__ void __launch_bounds__(kCompressThreads) k_compress_batch(BatchArgs b)
{
    __shared__ CompressIO sio;
    unsigned long long* status;
    uint64_t tg;
    if (!launch_status(b.ws, b.total_tiles, tg, status))
        return;
    uint32_t jidx;
    uint64_t t;
    job_of(b, tg, jidx, t);
    stage_desc(b.jobs, jidx, &sio);
    compress_tile(b.g, b.tl, b.p, sio, t, status + b.job_tile0[jidx] * kStatusStride, b.ws, b.total_chains);
}

__global__ void __launch_bounds__(kCompressThreads) k_compress_batch_slots(BatchArgs b)
{
    __shared__ CompressIO sio;
    unsigned long long* status;
    uint64_t tg;
    if (!launch_status(b.ws, b.total_tiles, tg, status))
        return;
    uint32_t jidx;
    uint64_t t;
    job_of(b, tg, jidx, t);
    stage_desc(b.jobs, jidx, &sio);
    // a uniform batch's jobs have equal tile counts: job j's tiles start at j * ntiles
    compress_tile_slots<kSlotL, kCompressThreads / 64>(b.g, b.tl, b.p, sio, t,
        status + (uint64_t) jidx * b.tl.ntiles * kStatusStride, b.ws, b.total_chains);
}

__global__ void __launch_bounds__(kDecompressThreads) k_decompress_batch(BatchArgs b)
{
    __shared__ DecompressIO sio;
    unsigned long long* status;
    uint64_t tg;
    if (!launch_status(b.ws, b.total_tiles, tg, status))
        return;
    uint32_t jidx;
    uint64_t t;
    job_of(b, tg, jidx, t);
    stage_desc(b.jobs, jidx, &sio);
    decompress_tile(b.g, b.tl, b.p, sio, t, status + b.job_tile0[jidx] * kStatusStride, b.ws, b.total_chains);
}

// Mixed-geometry batches: whole descriptors per job.
__global__ void __launch_bounds__(kHomsumThreads, 4) k_homsum_mixed(BatchArgs b)
{
    __shared__ HomsumJob sj;
    unsigned long long* status;
    uint64_t tg;
    if (!launch_status(b.ws, b.total_tiles, tg, status))
        return;
    uint32_t jidx;
    uint64_t t;
    job_of(b, tg, jidx, t);
    stage_desc(b.jobs, jidx, &sj);
    homsum_tile(sj.g, sj.tl, sj.p, sj.io, t, status + b.job_tile0[jidx] * kStatusStride, b.ws, b.total_chains);
}

__global__ void __launch_bounds__(kCompressThreads) k_compress_mixed(BatchArgs b)
{
    __shared__ CompressJob sj;
    unsigned long long* status;
    uint64_t tg;
    if (!launch_status(b.ws, b.total_tiles, tg, status))
        return;
    uint32_t jidx;
    uint64_t t;
    job_of(b, tg, jidx, t);
    stage_desc(b.jobs, jidx, &sj);
    compress_tile(sj.g, sj.tl, sj.p, sj.io, t, status + b.job_tile0[jidx] * kStatusStride, b.ws, b.total_chains);
}

__global__ void __launch_bounds__(kDecompressThreads) k_decompress_mixed(BatchArgs b)
{
    __shared__ DecompressJob sj;
    unsigned long long* status;
    uint64_t tg;
    if (!launch_status(b.ws, b.total_tiles, tg, status))
        return;
    uint32_t jidx;
    uint64_t t;
    job_of(b, tg, jidx, t);
    stage_desc(b.jobs, jidx, &sj);
    decompress_tile(sj.g, sj.tl, sj.p, sj.io, t, status + b.job_tile0[jidx] * kStatusStride, b.ws, b.total_chains);
}

// ------------------------------------------------------------ parse

// Rebuild the index of a stream that carries none (for example one written
// by the CPU library).  One wave walks one chunk: the wave pulls a window of
// the chunk into LDS and follows the width bytes through it; the remainder
// segment continues chunk T-1's walk (ZCCL/src/ZCCLd_float.c:597-745).
constexpr uint32_t kParseWindow = 8192;

__global__ void __launch_bounds__(64) k_parse(ParseArgs a)
{
    __shared__ uint32_t win[kParseWindow / 4 + 8];
    __shared__ unsigned long long s_pos[2];
    const Geom& g = a.g;
    const uint32_t chunk = blockIdx.x;
    const int lane = threadIdx.x;
    if (chunk >= g.T)
        return;
    IndexView ix = index_view(g, a.index);
    const uint8_t* in = a.bytes;
    const uint64_t cap = a.in_cap;
    if (lane == 0) {
        unsigned long long pos = 0;
        bool ok = 8ull * g.T <= cap;
        if (ok)
            pos = 8ull * g.T + load_u64_bytes(in + 8ull * chunk);
        s_pos[0] = ok ? pos : ~0ull;
    }
    __syncthreads();
    unsigned long long pos = s_pos[0];
    if (pos == ~0ull) {
        if (lane == 0)
            raise_error(a.ws, ZK_ERR_OVERRUN);
        return;
    }
    uint64_t win_lo = 1, win_hi = 0;  // empty window
    bool bad = false;
    const int nseg = (chunk == g.T - 1 && g.R) ? 2 : 1;
    for (int sidx = 0; sidx < nseg && !bad; sidx++) {
        const uint32_t s = chunk + (uint32_t) sidx;
        const uint64_t nb = s < g.T ? g.nbT : g.nbR;
        const uint64_t len = s < g.T ? g.tb : g.R;
        const uint64_t b0 = (uint64_t) s * g.nbT;
        const uint64_t grp0 = (uint64_t) s * g.gT;
        if (lane == 0) {
            ix.seed_off[s] = pos;
            if (nb == 0)
                ix.group_off[grp0] = pos + 4;
        }
        pos += 4;
        for (uint64_t jb = 0; jb < nb; jb++) {
            if (pos < win_lo || pos >= win_hi) {
                uint64_t lo = pos & ~(uint64_t) 3;
                uint64_t hi = lo + kParseWindow;
                if (hi > cap)
                    hi = cap;
                if (lo >= hi) {
                    bad = true;
                    break;
                }
                __syncthreads();
                for (uint64_t q = lo + 4ull * lane; q < hi; q += 256) {
                    uint32_t v = 0;
                    for (int b = 0; b < 4 && q + b < hi; b++)
                        v |= (uint32_t) in[q + b] << (8 * b);
                    win[(q - lo) >> 2] = v;
                }
                __syncthreads();
                win_lo = lo;
                win_hi = hi;
            }
            uint32_t w = lds_u8(win, (uint32_t) (pos - win_lo));
            if (w > 32) {
                bad = true;
                break;
            }
            uint64_t left = len - 1 - jb * g.L;
            uint32_t nbl = (uint32_t) (left < g.L ? left : g.L);
            if (lane == 0) {
                ix.widths[b0 + jb] = (uint8_t) w;
                if (jb % g.IG == 0)
                    ix.group_off[grp0 + jb / g.IG] = pos;
            }
            pos += block_bytes(w, nbl);
            if (pos > cap) {
                bad = true;
                break;
            }
        }
    }
    if (bad) {
        if (lane == 0)
            raise_error(a.ws, ZK_ERR_OVERRUN);
        return;
    }
    if (lane == 0 && chunk == g.T - 1) {
        ix.meta[0] = pos;
        if (a.size_out)
            *a.size_out = pos;
    }
}

// ------------------------------------------------------------ parallel parse
//
// The serial walk above runs one wave per chunk: a 2^24-value stream keeps
// 16 waves busy for milliseconds.  Here a chunk's block region (the bytes
// after its seed) is cut into tiles of kParseTile bytes.  The first block
// starting inside a tile lies within its first maxb = block_bytes(32, L)
// bytes, so pass one walks full-size blocks from every such candidate offset
// to the tile's end (exit offset past the tile's end, blocks walked; dead
// when a width byte exceeds 32).  Walks from different offsets merge within a
// few blocks, so a tile's surviving candidates usually share one exit and
// the next tile's entry is known without this tile's.  Pass two (a
// workgroup per chunk) chains entries from the previous tile's common exit
// and scans the block counts; pass three walks each tile from its entry and writes the
// widths, group offsets, seeds, the chunk's last (possibly short) block, the
// remainder segment and the stream size.  Same walk as the serial parse
// (ZCCL/src/ZCCLd_float.c:597-745), same index.

// Block region of chunk c: [cs, ce), cs just past the chunk's seed; ce the
// next chunk's seed (from the header) or, for the last chunk, its worst-case
// end.  Both clamped to the readable bytes.
__device__ __forceinline__ bool parse_region(const ParseArgs& a, const ParseLayout& pl, uint32_t c, uint64_t& cs,
    uint64_t& ce)
{
    const Geom& g = a.g;
    if (8ull * g.T > a.in_cap)
        return false;
    cs = 8ull * g.T + load_u64_bytes(a.bytes + 8ull * c) + 4;
    ce = c + 1 < g.T ? 8ull * g.T + load_u64_bytes(a.bytes + 8ull * (c + 1)) : cs + (uint64_t) g.nbT * pl.maxb;
    if (ce > a.in_cap)
        ce = a.in_cap;
    return cs <= ce && ce - cs <= (uint64_t) pl.tpc * kParseTile;
}

// Bytes [ts, te) of the stream into LDS by NT threads: buf[0] holds the
// 16-byte aligned bytes from ts & ~15 (16-byte loads, all issued before the
// first LDS store, where the stream pointer is 16-byte aligned; dword loads
// otherwise), bytes past the readable end read as zero.  Returns the staged
// dwords.  Every thread of the group calls this.
template <uint32_t NT>
__device__ __forceinline__ uint32_t parse_stage(const ParseArgs& a, uint64_t ts, uint64_t te, uint32_t* buf,
    uint32_t tid)
{
    constexpr uint32_t K = (kParseTile / 16 + 1 + NT - 1) / NT;
    const uint64_t base = ts & ~15ull;
    const uint32_t n16 = (uint32_t) ((te - base + 15) >> 4);
    const bool aligned = ((uintptr_t) a.bytes & 15) == 0;
    uint4 v[K];
#pragma unroll
    for (uint32_t k = 0; k < K; k++) {
        const uint32_t i = tid + k * NT;
        const uint64_t q = base + 16ull * i;
        v[k] = make_uint4(0, 0, 0, 0);
        if (i < n16) {
            if (aligned && q + 16 <= a.in_cap) {
                v[k] = *(const uint4*) (a.bytes + q);
            } else {
                uint32_t d[4] = {0, 0, 0, 0};
                for (uint32_t b = 0; b < 16 && q + b < a.in_cap; b++)
                    d[b >> 2] |= (uint32_t) a.bytes[q + b] << (8 * (b & 3));
                v[k] = make_uint4(d[0], d[1], d[2], d[3]);
            }
        }
    }
#pragma unroll
    for (uint32_t k = 0; k < K; k++) {
        const uint32_t i = tid + k * NT;
        if (i < n16)
            reinterpret_cast<uint4*>(buf)[i] = v[k];
    }
    return 4 * n16;
}

// A corrupt stream: the sticky workspace error and the parse's own flag
// (which makes k_parse_finish leave no stream size).
__device__ __forceinline__ void parse_fail(const ParseArgs& a, const ParseLayout& pl)
{
    raise_error(a.ws, ZK_ERR_OVERRUN);
    __hip_atomic_store((unsigned int*) ((char*) a.ws + pl.flag_off), 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ uint32_t parse_block_bytes(uint32_t w, uint32_t L)
{
    return w ? 1u + ((L + 7u) >> 3) + ((w * L + 7u) >> 3) : 1u;  // (w >> 3) L + ceil((w & 7) L / 8) = ceil(w L / 8)
}

// One bit per nonzero dword of a wave's staged bytes (ndw dwords), built with
// ballots: the walks below cross runs of zero bytes (width-0 blocks of one
// byte each) a bitmap word at a time.
constexpr uint32_t kParseNzWords = 2 * ((kParseTile / 4 + 4 + 63) / 64);

__device__ __forceinline__ void parse_nonzero_map(const uint32_t* buf, uint32_t ndw, uint32_t* nzb)
{
    const uint32_t lane = threadIdx.x & 63;
    for (uint32_t i = 0; i < kParseNzWords / 2; i++) {
        const uint32_t d = 64 * i + lane;
        const unsigned long long m = __ballot(d < ndw && buf[d] != 0);
        if (lane == 0) {
            nzb[2 * i] = (uint32_t) m;
            nzb[2 * i + 1] = (uint32_t) (m >> 32);
        }
    }
}

// Index of the first nonzero staged byte at or after byte q, or lim when every
// byte in [q, lim) is zero (byte indices count from the staged buffer's start).
__device__ __forceinline__ uint32_t parse_next_nonzero(const uint32_t* buf, const uint32_t* nzb, uint32_t q, uint32_t lim)
{
    if (q >= lim)
        return lim;
    const uint32_t d = q >> 2;
    const uint32_t v = buf[d] >> (8 * (q & 3));
    if (v)
        return min(q + ((uint32_t) __builtin_ctz(v) >> 3), lim);
    const uint32_t kend = (lim + 127) >> 7;  // bitmap words of the dwords below ceil(lim / 4)
    uint32_t k = (d + 1) >> 5;
    if (k >= kend)
        return lim;
    uint32_t m = nzb[k] & (~0u << ((d + 1) & 31));
    while (!m) {
        if (++k >= kend)
            return lim;
        m = nzb[k];
    }
    const uint32_t d2 = 32 * k + (uint32_t) __builtin_ctz(m);
    return min(4 * d2 + ((uint32_t) __builtin_ctz(buf[d2]) >> 3), lim);
}

// Pass one: one workgroup per tile.  Wave 0 walks candidates 0..63 and
// records, for every offset one of its walks steps on, the lane and the
// blocks it had walked there; then the four waves walk the remaining
// candidates, each stopping at the first recorded offset: from there its
// walk is that lane's, so its exit is that lane's and its count its own
// blocks so far plus what that lane walked from the offset on.  Walks merge
// within a few blocks, so the later candidates take a few steps instead of
// the whole tile.
// Two waves per tile: wave 0's recorded walks take the first 64 candidates,
// and both waves the remaining ones (128 per round, one round for blocks of
// up to 36 values), so more tiles fit on a CU (the walks are chains of
// dependent LDS reads: tiles in flight, not instructions, set the pace).
constexpr uint32_t kParseSpecThreads = 128;
constexpr uint32_t kParseSpecWaves = kParseSpecThreads / 64;

__global__ void __launch_bounds__(kParseSpecThreads) k_parse_spec(ParseArgs a, ParseLayout pl)
{
    __shared__ uint32_t buf[kParseTile / 4 + 5];  // one dword past the staged bytes for the 8-byte window
    __shared__ uint32_t nzb[kParseNzWords];
    __shared__ uint32_t seen[kParseTile];          // offset -> wave-0 lane << 16 | blocks before it, or ~0
    __shared__ uint32_t s_cnt[64], s_ex[64];       // wave 0's walks: blocks, exit (kParseNoExit: dead)
    __shared__ uint32_t s_lo[kParseSpecWaves], s_hi[kParseSpecWaves];
    const uint32_t tid = threadIdx.x, wv = tid >> 6, lane = tid & 63;
    const uint64_t gt = blockIdx.x;
    const uint32_t c = (uint32_t) (gt / pl.tpc), t = (uint32_t) (gt % pl.tpc);
    uint64_t cs, ce;
    if (!parse_region(a, pl, c, cs, ce))
        return;  // pass two reports the chunk
    const uint64_t ts = cs + (uint64_t) t * kParseTile;
    if (ts >= ce)
        return;  // past the chunk's region
    const uint64_t te = ts + kParseTile < ce ? ts + kParseTile : ce;
    const uint32_t ndw = parse_stage<kParseSpecThreads>(a, ts, te, buf, tid);
    for (uint32_t i = tid; i < kParseTile; i += kParseSpecThreads)
        seen[i] = ~0u;
    __syncthreads();
    if (wv == 0)
        parse_nonzero_map(buf, ndw, nzb);
    __syncthreads();
    const uint32_t off = (uint32_t) (ts & 15);
    const uint32_t len = (uint32_t) (te - ts);
    const uint32_t hb = 1u + ((a.g.L + 7u) >> 3);
    uint32_t* table = (uint32_t*) ((char*) a.ws + pl.table_off) + gt * pl.maxb;
    uint32_t lo = 0xFFFFFFFFu, hi = 0;
    // Walk candidate e to the tile's end, one step per loop trip for every
    // lane still inside the tile: the byte at pos and the 8-byte window from
    // it; a width byte above 32 ends the walk (dead), a zero byte is a
    // one-byte block and the window's run of zero bytes is taken at once; a
    // run that fills the window goes on through the nonzero-dword bitmap.
    // record: note every offset stepped on; else stop at a noted offset.
    // Returns the exit past the tile's end or kParseNoExit, and the blocks.
    auto walk = [&](uint32_t e, bool record, uint32_t& cnt) -> uint32_t {
        uint32_t pos = e < pl.maxb ? e : len, ex = kParseNoExit;
        bool dead = false, merged = false;
        cnt = 0;
        while (true) {
            bool act = pos < len;
            if (!__any(act))
                break;
            if (record) {
                if (act)
                    seen[pos] = (lane << 16) | cnt;
            } else if (act) {
                const uint32_t m = seen[pos];
                if (m != ~0u) {
                    const uint32_t l0 = m >> 16;
                    ex = s_ex[l0];
                    cnt += s_cnt[l0] - (m & 0xFFFFu);
                    merged = true;
                    pos = len;
                    act = false;
                }
            }
            const uint32_t q = off + (act ? pos : 0u);
            const uint64_t x = ((uint64_t) buf[(q >> 2) + 1] << 32 | buf[q >> 2]) >> (8 * (q & 3));
            const uint32_t w = (uint32_t) x & 0xFFu;
            const uint32_t span = 8u - (q & 3);
            const uint32_t zr = x ? (uint32_t) __builtin_ctzll(x) >> 3 : span;
            const uint32_t bsz = hb + ((__umul24(w, a.g.L) + 7u) >> 3);  // (used only for 0 < w <= 32)
            const uint32_t z = min(zr, len - min(pos, len));
            const bool bad = act && w > 32;
            dead |= bad;
            cnt += act && !bad ? (w ? 1u : z) : 0u;
            pos = act ? (bad ? len : pos + (w ? bsz : z)) : pos;
            const bool longz = act && w == 0 && zr == span && pos < len;
            if (__any(longz)) {
                if (longz) {
                    const uint32_t nz = parse_next_nonzero(buf, nzb, off + pos, off + len) - off;
                    cnt += nz - pos;
                    pos = nz;
                }
            }
        }
        return merged ? ex : (dead ? kParseNoExit : pos - len);
    };
    // blocks walked and the exit past the tile's end (kParseNoExit when the
    // walk met a width byte above 32: past the last chunk's blocks that is
    // the next segment's seed, not an error)
    auto keep = [&](uint32_t e, uint32_t ex, uint32_t cnt) {
        if (e < pl.maxb) {
            table[e] = (cnt << 16) | ex;
            if (ex != kParseNoExit) {
                lo = min(lo, ex);
                hi = max(hi, ex);
            }
        }
    };
    if (wv == 0) {
        uint32_t cnt;
        const uint32_t ex = walk(lane, true, cnt);
        s_cnt[lane] = cnt;
        s_ex[lane] = ex;
        keep(lane, ex, cnt);
    }
    __syncthreads();
    for (uint32_t e0 = 64 + 64 * wv; e0 < pl.maxb; e0 += kParseSpecThreads) {
        uint32_t cnt;
        const uint32_t ex = walk(e0 + lane, false, cnt);
        keep(e0 + lane, ex, cnt);
    }
    // one exit for every surviving candidate: the next tile's entry
    for (int o = 32; o >= 1; o >>= 1) {
        lo = min(lo, (uint32_t) __shfl_xor((int) lo, o));
        hi = max(hi, (uint32_t) __shfl_xor((int) hi, o));
    }
    if (lane == 0) {
        s_lo[wv] = lo;
        s_hi[wv] = hi;
    }
    __syncthreads();
    if (tid == 0) {
        for (uint32_t w = 1; w < kParseSpecWaves; w++) {
            s_lo[0] = min(s_lo[0], s_lo[w]);
            s_hi[0] = max(s_hi[0], s_hi[w]);
        }
        lo = s_lo[0];
        hi = s_hi[0];
        ((uint32_t*) ((char*) a.ws + pl.conv_off))[gt] = lo == hi ? lo : kParseDead;
    }
}

// Exclusive scan over a workgroup of kParseResolveThreads threads.
constexpr uint32_t kParseResolveThreads = 1024;

__device__ __forceinline__ uint32_t parse_block_excl(uint32_t v, uint32_t* s_w)
{
    const uint32_t lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const uint32_t incl = wave_scan_dpp(v);
    __syncthreads();  // s_w free
    if (lane == 63)
        s_w[wv] = incl;
    __syncthreads();
    uint32_t before = incl - v;
    for (uint32_t i = 0; i < wv; i++)
        before += s_w[i];
    return before;
}

// Pass two: one workgroup per chunk.  Thread i takes a run of consecutive
// tiles; its first tile's entry is the previous tile's common exit (tile 0:
// the region's start), and the tables carry it through the run.  A run whose
// previous tile has no common exit waits: when the resolved runs before it
// hold fewer blocks than the chunk, one thread follows the tables from the
// last known exit through the runs in order until the chunk's blocks are
// covered.  The runs' block counts are then scanned into each tile's first
// block index.
__global__ void __launch_bounds__(kParseResolveThreads) k_parse_resolve(ParseArgs a, ParseLayout pl)
{
    __shared__ uint32_t s_w[kParseResolveThreads / 64];
    __shared__ uint32_t s_exit[kParseResolveThreads];
    __shared__ uint32_t s_sum[kParseResolveThreads];
    __shared__ uint32_t s_first, s_fix;
    const uint32_t c = blockIdx.x, tid = threadIdx.x;
    uint64_t cs, ce;
    const bool ok = parse_region(a, pl, c, cs, ce);
    if (!ok && tid == 0)
        parse_fail(a, pl);
    const uint32_t ntc = ok ? (uint32_t) ceil_div(ce - cs, kParseTile) : 0u;
    if (ok && ntc == 0 && tid == 0)
        parse_fail(a, pl);  // no bytes for the chunk's blocks
    const uint32_t per = (uint32_t) ceil_div(ntc, kParseResolveThreads);
    const uint32_t nb = a.g.nbT;
    const uint64_t g0 = (uint64_t) c * pl.tpc;
    const uint32_t* table = (const uint32_t*) ((const char*) a.ws + pl.table_off);
    const uint32_t* conv = (const uint32_t*) ((const char*) a.ws + pl.conv_off);
    uint32_t* entry = (uint32_t*) ((char*) a.ws + pl.entry_off);
    uint32_t* base = (uint32_t*) ((char*) a.ws + pl.base_off);
    // follow the tables from entry e through tiles [r0, r1): entries and
    // counts written, the block total returned, e left at the exit
    // (kParseNoExit once a tile had no exit: the later tiles get no entry)
    auto chain = [&](uint32_t r0, uint32_t r1, uint32_t& e) {
        uint32_t s = 0;
        for (uint32_t t = r0; t < r1; t++) {
            uint32_t cnt = 0;
            if (e < pl.maxb) {
                const uint32_t x = table[(g0 + t) * pl.maxb + e];
                cnt = x >> 16;
                entry[g0 + t] = e;
                e = x & 0xFFFFu;
            } else {
                entry[g0 + t] = kParseDead;
            }
            base[g0 + t] = cnt;
            s += cnt;
        }
        return s;
    };
    if (tid == 0) {
        s_first = 0xFFFFFFFFu;
        s_fix = 0;
    }
    __syncthreads();
    const uint32_t t0 = min(tid * per, ntc), t1 = min(t0 + per, ntc);
    uint32_t sum = 0, e = 0;
    if (t0 < t1) {
        if (t0 > 0)
            e = conv[g0 + t0 - 1];
        if (e == kParseDead) {
            for (uint32_t t = t0; t < t1; t++) {
                entry[g0 + t] = kParseDead;
                base[g0 + t] = 0;
            }
            atomicMin(&s_first, tid);
            e = kParseNoExit;
        } else {
            sum = chain(t0, t1, e);
        }
    }
    s_exit[tid] = e;
    s_sum[tid] = sum;
    uint32_t before = parse_block_excl(sum, s_w);
    const uint32_t first = s_first;
    if (tid == first && before < nb)
        s_fix = 1;
    __syncthreads();
    if (s_fix) {
        // rare in valid streams (a tile whose candidate walks kept more than
        // one exit): serial, up to the end of the chunk's blocks
        if (tid == 0) {
            uint32_t x = s_exit[first - 1], run = 0;
            for (uint32_t i = 0; i < first; i++)
                run += s_sum[i];
            for (uint32_t r = first; r < kParseResolveThreads && run < nb && x < pl.maxb; r++) {
                const uint32_t r0 = min(r * per, ntc), r1 = min(r0 + per, ntc);
                s_sum[r] = chain(r0, r1, x);
                run += s_sum[r];
            }
        }
        __syncthreads();
        before = parse_block_excl(s_sum[tid], s_w);
    }
    for (uint32_t t = t0; t < t1; t++) {
        const uint32_t cnt = base[g0 + t];
        base[g0 + t] = before;
        before += cnt;
    }
}

// Pass three: one wave per tile, walking the tile from its entry.  The walk is
// the same in every lane; each block's position and width go to the lane of
// its place in a batch of 64, and a full batch (or the end) writes its widths
// and group offsets at once.  A run of zero bytes (one-byte width-0 blocks) is
// written by the whole wave.
__global__ void __launch_bounds__(256) k_parse_emit(ParseArgs a, ParseLayout pl)
{
    __shared__ uint32_t buf[4][kParseTile / 4 + 4];
    __shared__ uint32_t nzb[4][kParseNzWords];
    // (the wave index read as uniform: the walk below then runs on scalar registers)
    const uint32_t wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
    const uint64_t gt = (uint64_t) blockIdx.x * 4 + wv;
    if (gt >= pl.ntiles)
        return;
    const Geom& g = a.g;
    const uint32_t c = (uint32_t) (gt / pl.tpc), t = (uint32_t) (gt % pl.tpc);
    uint64_t cs, ce;
    if (!parse_region(a, pl, c, cs, ce))
        return;
    const uint64_t ts = cs + (uint64_t) t * kParseTile;
    if (ts >= ce)
        return;
    const uint64_t te = ts + kParseTile < ce ? ts + kParseTile : ce;
    const uint32_t b = ((const uint32_t*) ((const char*) a.ws + pl.base_off))[gt];
    const uint32_t nb = g.nbT;
    if (b >= nb)
        return;  // past the chunk's blocks
    IndexView ix = index_view(g, a.index);
    const uint32_t e = ((const uint32_t*) ((const char*) a.ws + pl.entry_off))[gt];
    if (e == kParseDead) {
        if (lane == 0)
            parse_fail(a, pl);
        return;
    }
    const uint32_t ndw = parse_stage<64>(a, ts, te, buf[wv], lane);
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
    parse_nonzero_map(buf[wv], ndw, nzb[wv]);
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
    const uint32_t off = (uint32_t) (ts & 15);
    const uint8_t* tb = (const uint8_t*) buf[wv] + off;
    const uint64_t b0 = (uint64_t) c * g.nbT, grp0 = (uint64_t) c * g.gT;
    if (t == 0 && lane == 0)
        ix.seed_off[c] = cs - 4;
    const uint32_t len = (uint32_t) (te - ts);
    const uint32_t hb = 1u + ((g.L + 7u) >> 3);
    const uint32_t last_len = g.tb - 1 - (nb - 1) * g.L;  // values in the chunk's last block
    uint32_t pos = e, j = b, k = 0;
    uint32_t my_pos = 0, my_j = 0, my_w = 0;
    bool bad = false;
    auto flush = [&]() {
        if (lane < k) {
            ix.widths[b0 + my_j] = (uint8_t) my_w;
            if (my_j % g.IG == 0)
                ix.group_off[grp0 + my_j / g.IG] = ts + my_pos;
        }
        k = 0;
    };
    // every block but the chunk's last (a run of zero bytes may end the chunk)
    while (pos < len && j + 1 < nb) {
        const uint32_t w = __builtin_amdgcn_readfirstlane(tb[pos]);
        if (w > 32) {
            bad = true;
            break;
        }
        if (w == 0) {
            const uint32_t nz = parse_next_nonzero(buf[wv], nzb[wv], off + pos + 1, off + len) - off;
            const uint32_t r = min(nz - pos, nb - j);
            flush();
            for (uint32_t i = lane; i < r; i += 64) {
                const uint32_t jj = j + i;
                ix.widths[b0 + jj] = 0;
                if (jj % g.IG == 0)
                    ix.group_off[grp0 + jj / g.IG] = ts + pos + i;
            }
            pos += r;
            j += r;
            continue;
        }
        if (lane == k) {
            my_pos = pos;
            my_j = j;
            my_w = w;
        }
        pos += hb + ((__umul24(w, g.L) + 7u) >> 3);
        j++;
        if (++k == 64)
            flush();
    }
    // the chunk's last block (last_len values), if it starts in this tile
    if (!bad && pos < len && j + 1 == nb) {
        const uint32_t w = __builtin_amdgcn_readfirstlane(tb[pos]);
        if (w > 32) {
            bad = true;
        } else {
            if (lane == k) {
                my_pos = pos;
                my_j = j;
                my_w = w;
            }
            pos += block_bytes(w, last_len);
            j++;
            k++;  // (at most 64: the loop flushed at 64)
        }
    }
    flush();
    if (lane != 0)
        return;
    if (bad) {
        parse_fail(a, pl);
        return;
    }
    if (j < nb) {
        if (te == ce)
            parse_fail(a, pl);  // the region ended before the chunk's blocks did
        return;                 // the chunk goes on in the next tile
    }
    // the chunk's last block ends here
    uint64_t end = ts + pos;
    if (c + 1 < g.T) {
        if (end != ce)
            parse_fail(a, pl);
        return;
    }
    if (g.R) {
        // the remainder segment (fewer than T values): seed, then its blocks
        ix.seed_off[g.T] = end;
        end += 4;
        const uint64_t rb0 = (uint64_t) g.T * g.nbT, rgrp0 = (uint64_t) g.T * g.gT;
        if (g.nbR == 0)
            ix.group_off[rgrp0] = end;
        for (uint32_t jb = 0; jb < g.nbR; jb++) {
            if (end >= a.in_cap) {
                parse_fail(a, pl);
                return;
            }
            const uint32_t w = a.bytes[end];
            if (w > 32) {
                parse_fail(a, pl);
                return;
            }
            ix.widths[rb0 + jb] = (uint8_t) w;
            if (jb % g.IG == 0)
                ix.group_off[rgrp0 + jb / g.IG] = end;
            const uint32_t left = g.R - 1 - jb * g.L;
            end += block_bytes(w, left < g.L ? left : g.L);
        }
    }
    if (end > a.in_cap) {
        parse_fail(a, pl);
        return;
    }
    ix.meta[0] = end;
    if (a.size_out)
        *a.size_out = end;
}

// Before pass one: no stream size yet, and the flag clear (the flag and the
// tables share the workspace with lookback status words of earlier launches).
__global__ void k_parse_begin(ParseArgs a, ParseLayout pl)
{
    if (threadIdx.x == 0) {
        *(unsigned long long*) (a.index + a.g.ix_meta) = 0;
        if (a.size_out)
            *a.size_out = 0;
        *(unsigned long long*) ((char*) a.ws + pl.flag_off) = 0;
    }
}

// After pass three: a stream found corrupt anywhere keeps no size; then the
// flag and tables are zeroed (the lookback area is zero between launches).
// Workgroup 0 reads the flag before clearing it; the others clear the rest.
__global__ void k_parse_finish(ParseArgs a, ParseLayout pl)
{
    unsigned long long* w = (unsigned long long*) ((char*) a.ws + pl.flag_off);
    const uint64_t words = (pl.bytes - pl.flag_off) / 8;
    if (blockIdx.x == 0) {
        if (threadIdx.x == 0) {
            if (*(const unsigned int*) w) {
                *(unsigned long long*) (a.index + a.g.ix_meta) = 0;
                if (a.size_out)
                    *a.size_out = 0;
            }
            w[0] = 0;
            w[1] = 0;
        }
        return;
    }
    for (uint64_t i = 2 + (uint64_t) (blockIdx.x - 1) * blockDim.x + threadIdx.x; i < words;
         i += (uint64_t) (gridDim.x - 1) * blockDim.x)
        w[i] = 0;
}

// ------------------------------------------------------------ case counts

// Block-case histogram of a pairwise add, the statistics the reference's
// _record variant keeps (ZCCL/src/ZCCLd_float.c:2175-2246, 2761-2764):
// [0] both widths zero, [1] first zero only, [2] second zero only, [3] neither.
__global__ void k_case_count(const uint8_t* w1, const uint8_t* w2, uint64_t nblocks,
    unsigned long long* counts)
{
    // (four counters in registers: an array indexed by the case would be
    // placed in memory, which the compiler maps to LDS)
    uint32_t c0 = 0, c1 = 0, c2 = 0, c3 = 0;
    for (uint64_t i = blockIdx.x * (uint64_t) blockDim.x + threadIdx.x; i < nblocks;
         i += (uint64_t) gridDim.x * blockDim.x) {
        const uint32_t k = 2u * (w1[i] != 0) + (w2[i] != 0);
        c0 += k == 0;
        c1 += k == 1;
        c2 += k == 2;
        c3 += k == 3;
    }
#pragma unroll
    for (int k = 0; k < 4; k++) {
        uint32_t v = k == 0 ? c0 : k == 1 ? c1 : k == 2 ? c2 : c3;
#pragma unroll
        for (int d = 32; d >= 1; d >>= 1)
            v += __shfl_xor(v, d, 64);
        if ((threadIdx.x & 63) == 0 && v)
            atomicAdd(&counts[k], (unsigned long long) v);
    }
}

// ------------------------------------------------------------ utilities

__global__ void k_fill_u64(unsigned long long* p, uint64_t n, unsigned long long v)
{
    for (uint64_t i = blockIdx.x * (uint64_t) blockDim.x + threadIdx.x; i < n;
         i += (uint64_t) gridDim.x * blockDim.x)
        p[i] = v;
}

__global__ void k_ws_init(WsHeader* ws, uint64_t capacity)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        ws->seq = 0;
        ws->chains_done = 0;
        ws->capacity = capacity;
        ws->errors = 0;
    }
}

// ------------------------------------------------------------ launchers

size_t compress_lds_bytes(const Tiling& tl) { return (size_t) align_up(tl.tile_bytes_max + 64, 16); }

// Decompress (k = 1): one worst-case tile.  Homomorphic sum: two halves of
// at least kStageMin / 2 bytes each; the whole stage holds one worst-case
// operand (aligned chunks plus slack).
static std::atomic<uint32_t> g_stage_min{kStageMin};

bool set_stage_min_bytes(uint32_t bytes)
{
    if (bytes == 0)
        bytes = kStageMin;
    if (bytes % 32 || bytes > 65536)
        return false;
    g_stage_min.store(bytes);
    return true;
}

uint32_t get_stage_min_bytes() { return g_stage_min.load(); }

uint32_t stage_bytes_for(const Tiling& tl, int k)
{
    if (k <= 1)
        return (uint32_t) align_up(tl.tile_bytes_max + 64, 16);
    uint64_t need = align_up(tl.tile_bytes_max + 48, 32);
    const uint32_t m = g_stage_min.load();
    return (uint32_t) (need < m ? m : need);
}

size_t operand_lds_bytes(const Tiling& tl, int k, uint32_t stage)
{
    return (size_t) (stage + 4ull * k * tl.G);
}

// The slot-layout compress: the same block length and index groups at the
// compress workgroup size (8 waves).
static bool slot_layout_compress(const Geom& g, const Tiling& tl)
{
    return g.L == kSlotL && g.IG == 16 && tl.threads == kCompressThreads &&
        tl.G == SlotMap<kSlotL, kCompressThreads / 64>::G;
}

hipError_t launch_compress(const CompressArgs& a, hipStream_t s)
{
    if (slot_layout_compress(a.job.g, a.job.tl))
        hipLaunchKernelGGL(k_compress_slots, dim3((uint32_t) a.job.tl.ntiles), dim3(a.job.tl.threads),
            compress_lds_bytes(a.job.tl), s, a);
    else
        hipLaunchKernelGGL(k_compress, dim3((uint32_t) a.job.tl.ntiles), dim3(a.job.tl.threads),
            compress_lds_bytes(a.job.tl), s, a);
    return hipGetLastError();
}

// The slot-layout sum serves the reference's default block length when the
// tiling is the one its layout assumes (256 threads, 16-block index groups).
static bool slot_layout(const Geom& g, const Tiling& tl)
{
    return g.L == kSlotL && g.IG == 16 && tl.threads == kHomsumThreads && tl.G == SlotMap<kSlotL>::G;
}

hipError_t launch_homsum(const HomsumArgs& a, hipStream_t s)
{
    const size_t lds = operand_lds_bytes(a.job.tl, a.job.p.k, a.job.p.stage_bytes);
    if (slot_layout(a.job.g, a.job.tl))
        hipLaunchKernelGGL(k_homsum_slots, dim3((uint32_t) a.job.tl.ntiles), dim3(a.job.tl.threads), lds, s, a);
    else
        hipLaunchKernelGGL(k_homsum, dim3((uint32_t) a.job.tl.ntiles), dim3(a.job.tl.threads), lds, s, a);
    return hipGetLastError();
}

hipError_t launch_decompress(const DecompressArgs& a, hipStream_t s)
{
    hipLaunchKernelGGL(k_decompress, dim3((uint32_t) a.job.tl.ntiles), dim3(a.job.tl.threads),
        operand_lds_bytes(a.job.tl, 1, a.job.p.stage_bytes), s, a);
    return hipGetLastError();
}

hipError_t launch_batch(BatchKind kind, const BatchArgs& b, uint32_t threads, size_t lds, hipStream_t s)
{
    dim3 grid((uint32_t) b.total_tiles), block(threads);
    switch (kind) {
    case BatchKind::Compress:
        if (b.uniform && slot_layout_compress(b.g, b.tl) && threads == kCompressThreads)
            hipLaunchKernelGGL(k_compress_batch_slots, grid, block, lds, s, b);
        else if (b.uniform)
            hipLaunchKernelGGL(k_compress_batch, grid, block, lds, s, b);
        else
            hipLaunchKernelGGL(k_compress_mixed, grid, block, lds, s, b);
        break;
    case BatchKind::Homsum:
        if (b.uniform && slot_layout(b.g, b.tl) && threads == kHomsumThreads)
            hipLaunchKernelGGL(k_homsum_batch_slots, grid, block, lds, s, b);
        else if (b.uniform)
            hipLaunchKernelGGL(k_homsum_batch, grid, block, lds, s, b);
        else
            hipLaunchKernelGGL(k_homsum_mixed, grid, block, lds, s, b);
        break;
    case BatchKind::Decompress:
        if (b.uniform)
            hipLaunchKernelGGL(k_decompress_batch, grid, block, lds, s, b);
        else
            hipLaunchKernelGGL(k_decompress_mixed, grid, block, lds, s, b);
        break;
    }
    return hipGetLastError();
}

hipError_t launch_parse(const ParseArgs& a, hipStream_t s)
{
    const ParseLayout pl = parse_layout(a.g);
    if (parse_parallel_ok(a.g) && a.ws_bytes >= pl.bytes) {
        const uint32_t groups = (uint32_t) ceil_div(pl.ntiles, 4);
        const uint64_t words = (pl.bytes - pl.flag_off) / 8;
        hipLaunchKernelGGL(k_parse_begin, dim3(1), dim3(64), 0, s, a, pl);
        hipLaunchKernelGGL(k_parse_spec, dim3((uint32_t) pl.ntiles), dim3(kParseSpecThreads), 0, s, a, pl);
        hipLaunchKernelGGL(k_parse_resolve, dim3(a.g.T), dim3(kParseResolveThreads), 0, s, a, pl);
        hipLaunchKernelGGL(k_parse_emit, dim3(groups), dim3(256), 0, s, a, pl);
        hipLaunchKernelGGL(k_parse_finish, dim3(1 + (uint32_t) std::min<uint64_t>(ceil_div(words, 256), 4096)), dim3(256),
            0, s, a, pl);
        return hipGetLastError();
    }
    // the stream size starts at 0 (a fill kernel, safe under graph capture):
    // only a parse that reaches the stream's end writes it, so a failed parse
    // leaves no stale size for later kernels to read up to
    hipLaunchKernelGGL(k_fill_u64, dim3(1), dim3(64), 0, s, (unsigned long long*) (a.index + a.g.ix_meta), 1ull, 0ull);
    if (a.size_out)
        hipLaunchKernelGGL(k_fill_u64, dim3(1), dim3(64), 0, s, a.size_out, 1ull, 0ull);
    hipLaunchKernelGGL(k_parse, dim3(a.g.T), dim3(64), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_case_count(const uint8_t* w1, const uint8_t* w2, uint64_t nblocks,
    unsigned long long* counts, hipStream_t s)
{
    hipLaunchKernelGGL(k_fill_u64, dim3(1), dim3(64), 0, s, counts, 4, 0ull);
    uint64_t blocks = ceil_div(nblocks, 256);
    blocks = blocks < 1 ? 1 : (blocks > 2048 ? 2048 : blocks);
    hipLaunchKernelGGL(k_case_count, dim3((uint32_t) blocks), dim3(256), 0, s, w1, w2, nblocks, counts);
    return hipGetLastError();
}

hipError_t launch_ws_init(WsHeader* ws, uint64_t capacity, hipStream_t s)
{
    uint64_t n = 2 * capacity * kStatusStride;
    unsigned long long* st = (unsigned long long*) ((char*) ws + sizeof(WsHeader));
    if (n) {
        uint32_t blocks = (uint32_t) (ceil_div(n, 256) < 1024 ? ceil_div(n, 256) : 1024);
        hipLaunchKernelGGL(k_fill_u64, dim3(blocks), dim3(256), 0, s, st, n, 0ull);
    }
    hipLaunchKernelGGL(k_ws_init, dim3(1), dim3(64), 0, s, ws, capacity);
    return hipGetLastError();
}

}  // namespace zk
