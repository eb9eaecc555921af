// zccl_kernels.hip -- CDNA4 (gfx950) kernels for the SZp fastmode1 codec and
// the hZCCL homomorphic sum.
//
// One workgroup encodes or decodes one tile of G blocks (G = 64 for L <= 64).
// Thread (j, u) owns the 4-element unit u of block j, so a wave64 covers 256
// consecutive elements and the tile's elements are contiguous in thread order.
// Stream offsets come from a single-pass decoupled-lookback scan over tiles
// (tickets drawn in launch order), so the input is read once and the output
// written once.  Bytes are staged in LDS: encoders OR each unit's sign nibble,
// high bytes and MSB-first low-bit field into a zeroed tile buffer, then the
// tile leaves with aligned 16-byte stores (byte stores only at its two ends);
// decoders pull the tile's bytes in with aligned 16-byte loads and extract
// fields with alignbyte / byte-swap / shifts.
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

#include "zccl_internal.h"
#include "zccl_kernels.h"

namespace zk {

// ------------------------------------------------------------ arithmetic

// Width thresholds c_w for w = 21..32: (int)log2f((float)m)+1 equals w+1
// once m >= 2^w - c_w (glibc 2.35 log2f, measured over every m < 2^32).
__constant__ uint32_t kBumpC[12] = {1, 2, 5, 11, 22, 45, 91, 183, 367, 735, 1471, 2943};

__device__ __forceinline__ uint32_t dwidth(uint32_t m)
{
    if (m == 0)
        return 0;
    uint32_t w = 32u - (uint32_t) __clz((int) m);
    if (w >= 21) {
        uint64_t thr = (1ull << w) - kBumpC[w - 21];
        if ((uint64_t) m >= thr)
            w++;
    }
    return w;
}

__device__ __forceinline__ int32_t quant(float x, double inv)
{
    double p = (double) x * inv;
    return (p > -2147483649.0 && p < 2147483648.0) ? (int32_t) p : INT32_MIN;
}

__device__ __forceinline__ float dequant(uint32_t q, float eb)
{
    return __fmul_rn(__int2float_rn((int32_t) q), eb);
}

__device__ __forceinline__ uint32_t umag(int32_t d) { return d < 0 ? 0u - (uint32_t) d : (uint32_t) d; }

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

// ------------------------------------------------------------ scans

__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t x)
{
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        uint32_t y = __shfl_up(x, d, 64);
        if (lane >= d)
            x += y;
    }
    return x;
}

__device__ __forceinline__ unsigned long long wave_sum64(unsigned long long x)
{
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1)
        x += __shfl_xor(x, d, 64);
    return x;
}

// Exclusive scan of one uint32 per thread over the workgroup (blockDim a
// multiple of 64, at most 1024).  total receives the workgroup sum.
__device__ uint32_t block_excl_scan(uint32_t v, uint32_t* s_wave, uint32_t& total)
{
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6, nw = blockDim.x >> 6;
    uint32_t x = wave_incl_scan(v);
    if (lane == 63)
        s_wave[wid] = x;
    __syncthreads();
    if (wid == 0) {
        uint32_t t = lane < nw ? s_wave[lane] : 0u;
        t = wave_incl_scan(t);
        if (lane < nw)
            s_wave[lane] = t;
    }
    __syncthreads();
    uint32_t base = wid ? s_wave[wid - 1] : 0u;
    total = s_wave[nw - 1];
    __syncthreads();
    return base + x - v;
}

// ------------------------------------------------------------ lookback

constexpr unsigned long long kFlagAgg = 1ull << 62;
constexpr unsigned long long kFlagInc = 2ull << 62;
constexpr unsigned long long kValMask = (1ull << 62) - 1;
constexpr uint32_t kSpinLimit = 1u << 22;

__device__ __forceinline__ void st_publish(unsigned long long* p, unsigned long long v)
{
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ unsigned long long st_poll(unsigned long long* p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ void raise_error(WsHeader* ws, unsigned long long bit)
{
    atomicOr(&ws->errors, bit);
}

// Decoupled lookback, executed by wave 0 of the workgroup.  Tiles t0..t form
// one scan; tile t0 starts from `init`.  Returns init + sum of the aggregates
// of tiles t0..t-1 (modulo 2^62; callers mask further) and publishes tile t's
// inclusive value.  Every wait is bounded: past kSpinLimit polls the tile
// records ZK_ERR_SPIN and proceeds, so the grid always drains.
__device__ unsigned long long lookback(unsigned long long* status, uint64_t t, uint64_t t0,
    unsigned long long agg, unsigned long long init, WsHeader* ws)
{
    const int lane = threadIdx.x & 63;
    if (t == t0) {
        if (lane == 0)
            st_publish(&status[t], kFlagInc | ((init + agg) & kValMask));
        return init;
    }
    if (lane == 0)
        st_publish(&status[t], kFlagAgg | (agg & kValMask));
    unsigned long long excl = 0;
    int64_t end = (int64_t) t;
    uint32_t spins = 0;
    bool timed_out = false;
    while (true) {
        int64_t idx = end - 1 - lane;
        bool valid = idx >= (int64_t) t0;
        unsigned long long st = 0;
        while (true) {
            if (valid)
                st = st_poll(&status[idx]);
            bool waiting = valid && (st >> 62) == 0;
            if (!__any(waiting))
                break;
            if (++spins > kSpinLimit) {
                if (waiting)
                    st = kFlagInc;
                timed_out = true;
                break;
            }
            __builtin_amdgcn_s_sleep(1);
        }
        unsigned long long pm = __ballot(valid && (st >> 62) == 2);
        unsigned long long val = valid ? (st & kValMask) : 0ull;
        if (pm) {
            int first = __ffsll((long long) pm) - 1;
            if (lane > first)
                val = 0;
            excl += wave_sum64(val);
            break;
        }
        excl += wave_sum64(val);
        end -= 64;
    }
    if (timed_out && lane == 0)
        raise_error(ws, ZK_ERR_SPIN);
    excl += init;
    if (lane == 0)
        st_publish(&status[t], kFlagInc | ((excl + agg) & kValMask));
    return excl;
}

// Ticket in launch order plus the duty of clearing the status array the next
// launch will use.  Returns false when this workgroup has nothing to do.
__device__ bool take_ticket(WsHeader* ws, uint64_t ntiles, uint64_t* s_bc, uint64_t& t,
    unsigned long long*& status)
{
    if (threadIdx.x == 0) {
        unsigned long long old = atomicAdd(&ws->ticket, 1ull);
        uint32_t tk = (uint32_t) old;
        uint32_t par = (uint32_t) (old >> 32) & 1u;
        if ((uint64_t) tk == ntiles - 1)
            atomicExch(&ws->ticket, (unsigned long long) (par ^ 1u) << 32);
        s_bc[0] = old;
        s_bc[1] = ws->capacity;
    }
    __syncthreads();
    unsigned long long old = s_bc[0];
    uint64_t cap = s_bc[1];
    t = (uint32_t) old;
    uint32_t par = (uint32_t) (old >> 32) & 1u;
    unsigned long long* base = (unsigned long long*) ((char*) ws + sizeof(WsHeader));
    if (cap < ntiles || t >= ntiles) {
        if (threadIdx.x == 0)
            raise_error(ws, ZK_ERR_CAPACITY);
        return false;
    }
    status = base + (uint64_t) par * cap;
    unsigned long long* other = base + (uint64_t) (par ^ 1u) * cap;
    uint64_t per = ceil_div(cap, ntiles);
    uint64_t lo = t * per, hi = lo + per < cap ? lo + per : cap;
    for (uint64_t i = lo + threadIdx.x; i < hi; i += blockDim.x)
        other[i] = 0;
    return true;
}

// ------------------------------------------------------------ global <-> LDS

// Copy src[lo, hi) into LDS so that LDS byte i holds byte (a0 + i) where
// a0 = (src + lo) rounded down to 16 bytes; returns (src + lo) - a0.  Whole
// 16-byte loads are used only when they lie inside [src, src + size); the
// rest is read byte by byte, so nothing outside the stream is touched.
__device__ uint32_t load_tile(uint32_t* buf, const uint8_t* src, uint64_t lo, uint64_t hi,
    uint64_t size)
{
    uintptr_t b0 = (uintptr_t) (src + lo), b1 = (uintptr_t) (src + hi);
    uintptr_t a0 = b0 & ~(uintptr_t) 15;
    uintptr_t v0 = (uintptr_t) src, v1 = v0 + size;
    uint32_t nch = (uint32_t) ((b1 - a0 + 15) >> 4);
    for (uint32_t c = threadIdx.x; c < nch; c += blockDim.x) {
        uintptr_t ca = a0 + 16ull * c;
        uint4 v;
        if (ca >= v0 && ca + 16 <= v1) {
            v = *(const uint4*) ca;
        } else {
            uint32_t w[4] = {0, 0, 0, 0};
            uintptr_t s = ca > b0 ? ca : b0, e = ca + 16 < b1 ? ca + 16 : b1;
            for (uintptr_t p = s; p < e; p++)
                w[(p - ca) >> 2] |= (uint32_t) (*(const uint8_t*) p) << (8 * ((p - ca) & 3));
            v = make_uint4(w[0], w[1], w[2], w[3]);
        }
        *(uint4*) &buf[4 * c] = v;
    }
    return (uint32_t) (b0 - a0);
}

// Write nbytes of LDS (starting at LDS byte 0) to dst.  Interior 16-byte
// aligned chunks go out as whole stores; the partial first and last chunk are
// written byte by byte so neighbouring tiles' bytes are never touched.
__device__ void store_tile(const uint32_t* buf, uint8_t* dst, uint32_t nbytes)
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

__device__ __forceinline__ void store_bytes(uint8_t* dst, const void* src, int n)
{
    const uint8_t* s = (const uint8_t*) src;
    for (int i = 0; i < n; i++)
        dst[i] = s[i];
}

__device__ __forceinline__ uint32_t load_u32_bytes(const uint8_t* p)
{
    return (uint32_t) p[0] | ((uint32_t) p[1] << 8) | ((uint32_t) p[2] << 16) | ((uint32_t) p[3] << 24);
}

__device__ __forceinline__ uint64_t load_u64_bytes(const uint8_t* p)
{
    return (uint64_t) load_u32_bytes(p) | ((uint64_t) load_u32_bytes(p + 4) << 32);
}

// ------------------------------------------------------------ unit codec

// Decode the up-to-4 deltas of unit u of a block whose width byte sits at LDS
// byte bo.  Layout: ZCCL/src/ZCCL_TypeManager.c:423-517 (extract) and
// :906-950 (sign bitmap).
__device__ __forceinline__ void decode_unit(const uint32_t* buf, uint32_t bo, uint32_t w,
    uint32_t nj, uint32_t u, uint32_t d[4])
{
    const uint32_t B = w >> 3, r = w & 7, sb = (nj + 7) >> 3;
    uint32_t sbyte = lds_u8(buf, bo + 1 + (u >> 1));
    uint32_t nib = (u & 1) ? (sbyte & 15u) : (sbyte >> 4);
    uint32_t mag[4] = {0, 0, 0, 0};
    if (B) {
        uint32_t W[4];
        lds_rd16(buf, bo + 1 + sb + 4 * u * B, W);
        switch (B) {
        case 1:
            mag[0] = W[0] & 0xFF; mag[1] = (W[0] >> 8) & 0xFF;
            mag[2] = (W[0] >> 16) & 0xFF; mag[3] = W[0] >> 24;
            break;
        case 2:
            mag[0] = W[0] & 0xFFFF; mag[1] = W[0] >> 16;
            mag[2] = W[1] & 0xFFFF; mag[3] = W[1] >> 16;
            break;
        case 3:
            mag[0] = W[0] & 0xFFFFFF;
            mag[1] = (W[0] >> 24) | ((W[1] & 0xFFFF) << 8);
            mag[2] = (W[1] >> 16) | ((W[2] & 0xFF) << 16);
            mag[3] = W[2] >> 8;
            break;
        default:
            mag[0] = W[0]; mag[1] = W[1]; mag[2] = W[2]; mag[3] = W[3];
            break;
        }
#pragma unroll
        for (int c = 0; c < 4; c++)
            mag[c] <<= r;
    }
    if (r) {
        uint32_t bitpos = 4 * u * r;
        uint32_t X = __builtin_bswap32(lds_rd4(buf, bo + 1 + sb + B * nj + (bitpos >> 3))) << (bitpos & 7);
#pragma unroll
        for (int c = 0; c < 4; c++)
            mag[c] |= (X << (c * r)) >> (32 - r);
    }
#pragma unroll
    for (int c = 0; c < 4; c++)
        d[c] = ((nib >> (3 - c)) & 1u) ? 0u - mag[c] : mag[c];
}

// OR the encoding of unit u (cnt deltas, magnitudes m[], sign nibble sg with
// element 0 in bit 3) of a block of nj elements at width w whose width byte
// sits at LDS byte bo.  ZCCL/src/ZCCL_TypeManager.c:12-91 and :824-848.
__device__ __forceinline__ void encode_unit(uint32_t* buf, uint32_t bo, uint32_t w, uint32_t nj,
    uint32_t u, uint32_t cnt, const uint32_t m[4], uint32_t sg)
{
    const uint32_t B = w >> 3, r = w & 7, sb = (nj + 7) >> 3;
    lds_or4(buf, bo + 1 + (u >> 1), (u & 1) ? sg : (sg << 4));
    if (B) {
        uint32_t h[4];
#pragma unroll
        for (int c = 0; c < 4; c++)
            h[c] = (uint32_t) c < cnt ? (m[c] >> r) : 0u;
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

// ------------------------------------------------------------ index access

struct IndexView {
    uint8_t* widths;
    unsigned long long* tile_off;
    unsigned long long* seed_off;
    unsigned long long* meta;
};

__device__ __forceinline__ IndexView index_view(const Geom& g, uint8_t* ix)
{
    IndexView v;
    v.widths = ix;
    v.tile_off = (unsigned long long*) (ix + g.ix_tile);
    v.seed_off = (unsigned long long*) (ix + g.ix_seed);
    v.meta = (unsigned long long*) (ix + g.ix_meta);
    return v;
}

// ------------------------------------------------------------ shared LDS

extern __shared__ uint32_t zk_lds[];

struct TileShared {
    uint64_t bc[4];
    uint32_t wave[16];
    uint32_t bmax[64];
    uint32_t boff[64];
    uint32_t width[64];
};

// Width, size and tile offsets of the blocks whose magnitudes were folded into
// sh.bmax; encodes the thread's unit; writes the tile, its seed and header
// entry, and the output index.  Shared tail of compress and homomorphic sum.
__device__ void encode_tile(const Geom& g, const Tile& x, uint64_t t, TileShared& sh,
    unsigned long long* status, WsHeader* ws, uint32_t j, uint32_t u, uint32_t nj, uint32_t cnt,
    const uint32_t m[4], uint32_t sg, uint32_t seed, uint8_t* out, uint64_t out_cap,
    uint8_t* index, unsigned long long* size_out)
{
    const uint32_t tid = threadIdx.x;
    uint32_t* buf = zk_lds;
    uint32_t bsz = 0;
    if (tid < x.nblk) {
        uint32_t w = dwidth(sh.bmax[tid]);
        sh.width[tid] = w;
        bsz = block_bytes(w, tile_block_len(g, x, tid));
    }
    uint32_t total;
    uint32_t boff = block_excl_scan(bsz, sh.wave, total);
    if (tid < x.nblk)
        sh.boff[tid] = boff;
    if (tid < 64) {
        unsigned long long ex = lookback(status, t, 0, total, 0, ws);
        if (tid == 0)
            sh.bc[2] = ex;
    }
    uint32_t nwords = (total + 3) / 4 + 8;
    for (uint32_t i = tid; i < nwords; i += blockDim.x)
        buf[i] = 0;
    __syncthreads();
    const uint64_t P = 8ull * g.T + 4ull * (x.s + 1) + sh.bc[2];
    if (cnt) {
        uint32_t w = sh.width[j];
        if (w)
            encode_unit(buf, sh.boff[j], w, nj, u, cnt, m, sg);
    }
    if (tid < x.nblk)
        lds_or4(buf, sh.boff[tid], sh.width[tid]);
    __syncthreads();
    const bool fits = P + total <= out_cap;
    if (fits)
        store_tile(buf, out + P, total);
    else if (tid == 0)
        raise_error(ws, ZK_ERR_NOSPACE);
    IndexView ix = index_view(g, index);
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
        ix.tile_off[t] = P;
        if (t == g.ntiles - 1) {
            // the last tile ends the stream: if it fits, every tile fits.  A
            // stream that did not fit records size 0 so no consumer reads it.
            unsigned long long sz = fits ? P + total : 0ull;
            ix.meta[0] = sz;
            if (size_out)
                *size_out = sz;
        }
    }
    if (tid < x.nblk)
        ix.widths[x.g0 + tid] = (uint8_t) sh.width[tid];
}

// ------------------------------------------------------------ compress

__global__ void __launch_bounds__(1024) k_compress(CompressArgs a)
{
    __shared__ TileShared sh;
    const Geom& g = a.g;
    uint64_t t;
    unsigned long long* status;
    if (!take_ticket(a.ws, g.ntiles, sh.bc, t, status))
        return;
    const Tile x = tile_of(g, t);
    const uint32_t tid = threadIdx.x;
    const uint32_t j = tid / g.U, u = tid - j * g.U;
    if (tid < 64)
        sh.bmax[tid] = 0;
    const uint32_t nj = j < x.nblk ? tile_block_len(g, x, j) : 0u;
    const uint32_t cnt = 4 * u < nj ? min(4u, nj - 4 * u) : 0u;
    const uint64_t e = x.e0 + (uint64_t) j * g.L + 4 * u;
    int32_t q[4] = {0, 0, 0, 0};
    if (cnt == 4 && (((uintptr_t) (a.x + e)) & 15) == 0) {
        float4 v = *(const float4*) (a.x + e);
        q[0] = quant(v.x, a.inv); q[1] = quant(v.y, a.inv);
        q[2] = quant(v.z, a.inv); q[3] = quant(v.w, a.inv);
    } else {
#pragma unroll
        for (int c = 0; c < 4; c++)
            if ((uint32_t) c < cnt)
                q[c] = quant(a.x[e + c], a.inv);
    }
    // previous element: the last element of lane-1's unit (the units of a
    // tile are contiguous in thread order); lane 0 of each wave reads it
    int32_t last = cnt ? q[cnt - 1] : 0;
    int32_t prev = __shfl_up(last, 1, 64);
    if ((tid & 63) == 0 && cnt)
        prev = quant(a.x[e - 1], a.inv);
    uint32_t m[4] = {0, 0, 0, 0}, sg = 0, umax = 0;
    int32_t p = prev;
#pragma unroll
    for (int c = 0; c < 4; c++) {
        if ((uint32_t) c < cnt) {
            int32_t d = (int32_t) ((uint32_t) q[c] - (uint32_t) p);
            p = q[c];
            m[c] = umag(d);
            sg |= (d < 0 ? 1u : 0u) << (3 - c);
            umax = max(umax, m[c]);
        }
    }
    __syncthreads();
    if (umax)
        atomicMax(&sh.bmax[j], umax);
    __syncthreads();
    uint32_t seed = 0;
    if (x.lt == 0 && tid == 0)
        seed = (uint32_t) quant(a.x[x.seg_a], a.inv);
    encode_tile(g, x, t, sh, status, a.ws, j, u, nj, cnt, m, sg, seed, a.out, a.out_cap, a.index,
        a.size_out);
}

// ------------------------------------------------------------ operand decode

// Stage operand tile bytes in LDS and return each block's LDS offset in
// sh.boff, width in sh.width.  Widths come from the operand's index; a width
// byte in the stream that disagrees with the index raises ZK_ERR_MISMATCH.
__device__ void stage_operand(const Geom& g, const Tile& x, uint64_t t, TileShared& sh,
    const uint8_t* bytes, const uint8_t* index, WsHeader* ws)
{
    const uint32_t tid = threadIdx.x;
    IndexView ix = index_view(g, (uint8_t*) index);
    uint32_t bsz = 0, w = 0;
    if (tid < x.nblk) {
        w = ix.widths[x.g0 + tid];
        if (w > 32) {
            raise_error(ws, ZK_ERR_WIDTH);
            w = 0;
        }
        bsz = block_bytes(w, tile_block_len(g, x, tid));
    }
    uint32_t total;
    uint32_t boff = block_excl_scan(bsz, sh.wave, total);
    const uint64_t P = ix.tile_off[t];
    const uint64_t size = ix.meta[0];
    uint32_t shift = 0;
    if (P + total <= size) {
        shift = load_tile(zk_lds, bytes, P, P + total, size);
    } else {
        if (tid == 0)
            raise_error(ws, ZK_ERR_OVERRUN);
        total = 0;
    }
    if (tid < x.nblk) {
        sh.boff[tid] = shift + boff;
        sh.width[tid] = total ? w : 0u;
    }
    __syncthreads();
    if (tid < x.nblk && total && lds_u8(zk_lds, shift + boff) != w)
        raise_error(ws, ZK_ERR_MISMATCH);
}

// Seed of the tile's segment, read only when it lies inside the stream.
__device__ uint32_t load_seed(const Geom& g, const Tile& x, const uint8_t* bytes, const uint8_t* index,
    WsHeader* ws)
{
    IndexView ix = index_view(g, (uint8_t*) index);
    uint64_t so = ix.seed_off[x.s];
    if (so + 4 <= ix.meta[0])
        return load_u32_bytes(bytes + so);
    raise_error(ws, ZK_ERR_OVERRUN);
    return 0;
}

// ------------------------------------------------------------ homomorphic sum

__global__ void __launch_bounds__(1024) k_homsum(HomsumArgs a)
{
    __shared__ TileShared sh;
    const Geom& g = a.g;
    uint64_t t;
    unsigned long long* status;
    if (!take_ticket(a.ws, g.ntiles, sh.bc, t, status))
        return;
    const Tile x = tile_of(g, t);
    const uint32_t tid = threadIdx.x;
    const uint32_t j = tid / g.U, u = tid - j * g.U;
    const uint32_t nj = j < x.nblk ? tile_block_len(g, x, j) : 0u;
    const uint32_t cnt = 4 * u < nj ? min(4u, nj - 4 * u) : 0u;
    uint32_t acc[4] = {0, 0, 0, 0};
    uint32_t seed = 0;
    for (int r = 0; r < a.k; r++) {
        stage_operand(g, x, t, sh, a.in[r].bytes, a.in[r].index, a.ws);
        if (cnt) {
            uint32_t w = sh.width[j];
            if (w) {
                uint32_t d[4];
                decode_unit(zk_lds, sh.boff[j], w, nj, u, d);
#pragma unroll
                for (int c = 0; c < 4; c++)
                    acc[c] += d[c];
            }
        }
        if (x.lt == 0 && tid == 0)
            seed += load_seed(g, x, a.in[r].bytes, a.in[r].index, a.ws);
        __syncthreads();
    }
    if (tid < 64)
        sh.bmax[tid] = 0;
    uint32_t m[4] = {0, 0, 0, 0}, sg = 0, umax = 0;
#pragma unroll
    for (int c = 0; c < 4; c++) {
        if ((uint32_t) c < cnt) {
            int32_t d = (int32_t) acc[c];
            m[c] = umag(d);
            sg |= (d < 0 ? 1u : 0u) << (3 - c);
            umax = max(umax, m[c]);
        }
    }
    __syncthreads();
    if (umax)
        atomicMax(&sh.bmax[j], umax);
    __syncthreads();
    encode_tile(g, x, t, sh, status, a.ws, j, u, nj, cnt, m, sg, seed, a.out, a.out_cap, a.index,
        a.size_out);
}

// ------------------------------------------------------------ decompress

__global__ void __launch_bounds__(1024) k_decompress(DecompressArgs a)
{
    __shared__ TileShared sh;
    const Geom& g = a.g;
    uint64_t t;
    unsigned long long* status;
    if (!take_ticket(a.ws, g.ntiles, sh.bc, t, status))
        return;
    const Tile x = tile_of(g, t);
    const uint32_t tid = threadIdx.x;
    const uint32_t j = tid / g.U, u = tid - j * g.U;
    const uint32_t nj = j < x.nblk ? tile_block_len(g, x, j) : 0u;
    const uint32_t cnt = 4 * u < nj ? min(4u, nj - 4 * u) : 0u;
    stage_operand(g, x, t, sh, a.bytes, a.index, a.ws);
    uint32_t d[4] = {0, 0, 0, 0};
    if (cnt) {
        uint32_t w = sh.width[j];
        if (w)
            decode_unit(zk_lds, sh.boff[j], w, nj, u, d);
    }
    // running sum inside the unit, then across the tile in element order
    uint32_t pre[4];
    uint32_t run = 0;
#pragma unroll
    for (int c = 0; c < 4; c++) {
        run += (uint32_t) c < cnt ? d[c] : 0u;
        pre[c] = run;
    }
    uint32_t tile_sum;
    uint32_t texcl = block_excl_scan(run, sh.wave, tile_sum);
    if (x.lt == 0 && tid == 0)
        sh.bc[3] = load_seed(g, x, a.bytes, a.index, a.ws);
    __syncthreads();
    const uint32_t seed = (uint32_t) sh.bc[3];
    if (tid < 64) {
        unsigned long long ex = lookback(status, t, x.seg_tile0, tile_sum, x.lt == 0 ? seed : 0u, a.ws);
        if (tid == 0)
            sh.bc[2] = ex;
    }
    __syncthreads();
    const uint32_t carry = (uint32_t) sh.bc[2] + texcl;
    const uint64_t e = x.e0 + (uint64_t) j * g.L + 4 * u;
    if (cnt == 4 && (((uintptr_t) (a.out + e)) & 15) == 0) {
        float4 v;
        v.x = dequant(carry + pre[0], a.eb);
        v.y = dequant(carry + pre[1], a.eb);
        v.z = dequant(carry + pre[2], a.eb);
        v.w = dequant(carry + pre[3], a.eb);
        *(float4*) (a.out + e) = v;
    } else {
#pragma unroll
        for (int c = 0; c < 4; c++)
            if ((uint32_t) c < cnt)
                a.out[e + c] = dequant(carry + pre[c], a.eb);
    }
    if (x.lt == 0 && tid == 0)
        a.out[x.seg_a] = dequant(seed, a.eb);
}

// ------------------------------------------------------------ parse

// Rebuild the index of a stream that carries none (for example one written
// by the CPU library).  One wave walks one chunk: the wave pulls a window of
// the chunk into LDS and lane 0 follows the width bytes through it; the
// remainder segment continues chunk T-1's walk (ZCCL/src/ZCCLd_float.c:597-745).
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
        const uint64_t g0 = (uint64_t) s * g.nbT;
        const uint64_t tile0 = (uint64_t) s * g.tilesT;
        if (lane == 0) {
            ix.seed_off[s] = pos;
            if (nb == 0)
                ix.tile_off[tile0] = pos + 4;
        }
        pos += 4;
        for (uint64_t jb = 0; jb < nb; jb++) {
            // refill the window when the width byte is outside it
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
                ix.widths[g0 + jb] = (uint8_t) w;
                if (jb % g.G == 0)
                    ix.tile_off[tile0 + jb / g.G] = pos;
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

// ------------------------------------------------------------ case counts

// Block-case histogram of a pairwise add, the statistics the reference's
// _record variant keeps (ZCCL/src/ZCCLd_float.c:2175-2246, 2761-2764):
// [0] both widths zero, [1] first zero only, [2] second zero only, [3] neither.
__global__ void k_case_count(const uint8_t* w1, const uint8_t* w2, uint64_t nblocks,
    unsigned long long* counts)
{
    uint32_t c[4] = {0, 0, 0, 0};
    for (uint64_t i = blockIdx.x * (uint64_t) blockDim.x + threadIdx.x; i < nblocks;
         i += (uint64_t) gridDim.x * blockDim.x) {
        uint32_t a = w1[i] != 0, b = w2[i] != 0;
        c[a * 2 + b] += 1;
    }
    // c index: (a,b) = (0,0)->0, (0,1)->1, (1,0)->2, (1,1)->3
#pragma unroll
    for (int k = 0; k < 4; k++) {
        uint32_t v = c[k];
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
        ws->ticket = 0;
        ws->capacity = capacity;
        ws->errors = 0;
    }
}

// ------------------------------------------------------------ launchers

static size_t lds_bytes(const Geom& g) { return (size_t) align_up(g.tile_bytes_max + 64, 16); }

hipError_t launch_compress(const CompressArgs& a, hipStream_t s)
{
    hipLaunchKernelGGL(k_compress, dim3((uint32_t) a.g.ntiles), dim3(a.g.threads), lds_bytes(a.g), s, a);
    return hipGetLastError();
}

hipError_t launch_homsum(const HomsumArgs& a, hipStream_t s)
{
    hipLaunchKernelGGL(k_homsum, dim3((uint32_t) a.g.ntiles), dim3(a.g.threads), lds_bytes(a.g), s, a);
    return hipGetLastError();
}

hipError_t launch_decompress(const DecompressArgs& a, hipStream_t s)
{
    hipLaunchKernelGGL(k_decompress, dim3((uint32_t) a.g.ntiles), dim3(a.g.threads), lds_bytes(a.g), s, a);
    return hipGetLastError();
}

hipError_t launch_parse(const ParseArgs& a, hipStream_t s)
{
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
    uint64_t n = 2 * capacity;
    unsigned long long* st = (unsigned long long*) ((char*) ws + sizeof(WsHeader));
    if (n) {
        uint32_t blocks = (uint32_t) (ceil_div(n, 256) < 1024 ? ceil_div(n, 256) : 1024);
        hipLaunchKernelGGL(k_fill_u64, dim3(blocks), dim3(256), 0, s, st, n, 0ull);
    }
    hipLaunchKernelGGL(k_ws_init, dim3(1), dim3(64), 0, s, ws, capacity);
    return hipGetLastError();
}

}  // namespace zk
