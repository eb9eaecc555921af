// zccl_lane.hip -- block-per-lane kernels for 36-value blocks, the
// reference's default block length (examples/compression_allreduce.c,
// jobs/*.sh): the fused homomorphic sum.
//
// Layout.  A wave owns one tile of 64 consecutive blocks of one segment and
// every lane owns one whole block.  For each operand a lane reads its block's
// bytes straight from HBM into registers with range-checked buffer loads (a
// read past the operand's stream returns zeros, so no read ever leaves the
// stream), decodes its 36 deltas and adds them in int32.  No LDS staging and
// no workgroup barrier sit on this path: the only per-operand cross-lane step
// is one DPP scan of block sizes that places each lane's block.  The next
// operand's loads are in flight while the current one is decoded.
//
// Each wave is a tile of the decoupled lookback; a workgroup of four waves
// takes one ticket for four consecutive tiles.  The summed block is
// re-encoded into the wave's own LDS buffer with OR-writes of dwords (the
// buffer is zeroed first, and a lane ORs zeros wherever its dwords overlap a
// neighbour's bytes, so overlapping writes need no ordering) and leaves with
// 16-byte stores.
//
// Blocks of width 16 and more, and a segment's last (short) block, take a
// per-unit path with the same arithmetic.
//
// Wire format and arithmetic (reference):
//   block      width byte, sign bitmap MSB first, (w>>3) high bytes per
//              value, MSB-first low (w&7) bits     ZCCL/src/ZCCL_TypeManager.c:12-91, 824-848
//   sum        deltas add in int32, block re-encoded at the width of the
//              largest magnitude                   ZCCL/src/ZCCLd_float.c:1099-1219
//   width      (int)log2f((float)max)+1             ZCCL/src/ZCCLd_float.c:1205
#include <hip/hip_runtime.h>

#include <atomic>

#include "zccl_device.h"
#include "zccl_internal.h"
#include "zccl_kernels.h"

namespace zk {

// Dynamic LDS: one output buffer per wave, 16-byte aligned (the buffers are
// zeroed with 16-byte stores).
extern __shared__ __attribute__((aligned(16))) uint32_t zk_lds[];

namespace {

constexpr uint32_t L = kLaneL;          // values per block
constexpr uint32_t SB = (L + 7) / 8;    // sign bytes of a whole block
constexpr uint32_t HEADW = 11;          // dwords of bytes [0, 44): width, signs, 36 high bytes
constexpr uint32_t NGRP = (L + 7) / 8;  // groups of (up to) 8 values in the low bits
constexpr uint32_t kWorstBlock = 1 + SB + 4 * L;   // 150 bytes at width 32
// Output buffer of one wave: 8 KB, so that five workgroups of four waves
// share a CU's 160 KB of LDS.  A tile's bytes plus the reach of the
// OR-writes past a block's end (at most 48 bytes) fit unless the tile is
// wide (the worst tile is 64 blocks of kWorstBlock bytes); such a tile is
// encoded half a wave at a time, and a half always fits.
constexpr uint32_t kOutBuf = 8192;
static_assert(kLaneTileBlocks / 2 * kWorstBlock + 64 <= kOutBuf, "half a worst-case tile fits a wave buffer");
static_assert(kOutBuf % 16 == 0, "wave buffers stay 16-byte aligned");

static_assert(L % 4 == 0 && L > 32 && L <= 40, "the lane decoder is laid out for 36-value blocks");

// Encoded size of a whole block at width w (w = 0..32): 1 + SB + 36 B +
// ceil(36 r / 8) = 6 + ceil(9 w / 2).
// a * b + c for a, b below 2^24: one full-rate instruction (the compiler
// picks a quarter-rate 32-bit multiply where it cannot prove the range).
__device__ __forceinline__ uint32_t mad_u24(uint32_t a, uint32_t b, uint32_t c)
{
    uint32_t d;
    asm("v_mad_u32_u24 %0, %1, %2, %3" : "=v"(d) : "v"(a), "v"(b), "v"(c));
    return d;
}

__device__ __forceinline__ uint32_t full_block_bytes(uint32_t w) { return w ? 6u + (mad_u24(w, 9u, 1u) >> 1) : 1u; }

// block_bytes for n <= L values at width w <= 32.
__device__ __forceinline__ uint32_t lane_block_bytes(uint32_t w, uint32_t n)
{
    return w ? 1u + ((n + 7u) >> 3) + mad_u24(w >> 3, n, (mad_u24(w & 7u, n, 7u) >> 3)) : 1u;
}

// (w >> 3) * L, the high bytes of a block, for w <= 32.
__device__ __forceinline__ uint32_t high_bytes(uint32_t w) { return mad_u24(w >> 3, L, 0u); }

using View = __amdgpu_buffer_rsrc_t;
// Four floats at 4-byte alignment (one 16-byte store where the address is
// only dword aligned).
typedef float f32x4u __attribute__((ext_vector_type(4), aligned(4)));

// Range-checked view of [base, base + bytes) (base 4-byte aligned, bytes a
// multiple of 4).  A buffer load returns zeros for every dword that does not
// lie wholly inside the view, whatever its alignment: a dword load at an
// unaligned offset that reaches past the view's last dword reads as zero even
// where its first bytes are inside.
__device__ __forceinline__ View make_view(const uint8_t* base, uint64_t bytes)
{
    const uint32_t n = bytes > 0xFFFFFFFCull ? 0xFFFFFFFCu : (uint32_t) bytes;
    return __builtin_amdgcn_make_buffer_rsrc((void*) base, (short) 0, (int) n, 0x00020000);
}

__device__ __forceinline__ uint32_t bl4(View v, uint32_t off) { return __builtin_amdgcn_raw_buffer_load_b32(v, off, 0, 0); }

__device__ __forceinline__ uint2 bl8(View v, uint32_t off)
{
    const auto x = __builtin_amdgcn_raw_buffer_load_b64(v, off, 0, 0);
    return make_uint2(x[0], x[1]);
}

__device__ __forceinline__ uint3 bl12(View v, uint32_t off)
{
    const auto x = __builtin_amdgcn_raw_buffer_load_b96(v, off, 0, 0);
    return make_uint3(x[0], x[1], x[2]);
}

__device__ __forceinline__ uint4 bl16(View v, uint32_t off)
{
    const auto x = __builtin_amdgcn_raw_buffer_load_b128(v, off, 0, 0);
    return make_uint4(x[0], x[1], x[2], x[3]);
}

// 4 / 8 bytes from any offset through aligned dword loads only, so every
// byte inside the view reads as itself (up to its very end).
__device__ __forceinline__ uint32_t bl4a(View v, uint32_t off)
{
    const uint2 x = bl8(v, off & ~3u);
    return __builtin_amdgcn_alignbyte(x.y, x.x, off);  // uses off & 3
}

__device__ __forceinline__ uint2 bl8a(View v, uint32_t off)
{
    const uint3 x = bl12(v, off & ~3u);
    return make_uint2(__builtin_amdgcn_alignbyte(x.y, x.x, off), __builtin_amdgcn_alignbyte(x.z, x.y, off));
}

// Read-only data of the launch (descriptors, input indices) through the
// scalar cache when the address is the same for the whole wave.
template <typename T>
__device__ __forceinline__ T ldc(const T* p)
{
    return *(const __attribute__((address_space(4))) T*) (uintptr_t) p;
}

// |d| of an int32 delta as unsigned (|INT_MIN| = 2^31).
__device__ __forceinline__ uint32_t mag(uint32_t d) { return (uint32_t) max((int32_t) d, (int32_t) (0u - d)); }

// One operand's block of this lane, loaded: bytes [0, 44) of the block (width
// byte, sign bytes, the high bytes of a width below 16) and, for each group
// of 8 values, the 8 bytes from the group's first byte of low bits as two
// big-endian words.
struct Block {
    uint32_t h[HEADW];
    uint32_t ge[2 * NGRP];
    uint32_t w;   // index width (0..32)
    uint32_t o;   // offset of the block in the operand's view
    uint32_t lim; // length of the operand's view (the same for the wave)
    bool fast;    // decode_fast serves this lane (else the general decoder, if w > 0)
};

// Add a whole block of width w < 16 to acc: magnitude = (high byte << r) |
// low r bits, times the value's sign as +1 or -1.  A zero width adds zeros
// whatever the bytes.  wide: some lane of the wave has w >= 8 (otherwise no
// lane has high bytes and they are not extracted).
// ZCCL/src/ZCCL_TypeManager.c:423-517 (extract), :906-950 (signs).
// a * b + c for a, b of 24 bits (signed) and any c, in one instruction.
__device__ __forceinline__ uint32_t mad_i24(uint32_t a, uint32_t b, uint32_t c)
{
    uint32_t d;
    asm("v_mad_i32_i24 %0, %1, %2, %3" : "=v"(d) : "v"(a), "v"(b), "v"(c));
    return d;
}

// N < L: only the sum of the block's values is wanted; value j is added to
// acc[j % N] (N short chains of adds instead of L values to add up).
template <bool wide, uint32_t N = L>
__device__ __forceinline__ void decode_fast(const Block& q, uint32_t (&acc)[N])
{
    const uint32_t r = q.w & 7, hb = q.w & 8;  // hb: bits per high byte (0 or 8)
    const uint32_t s0 = __builtin_amdgcn_alignbyte(q.h[1], q.h[0], 1);  // sign bytes 0..3, memory order
    const uint32_t s1 = q.h[1] >> 8;                                   // sign byte 4 in bits 7..0
    // Sign words: a value's sign bit moved to an odd bit with a one below it,
    // so the 2-bit field there reads -1 (sign set) or +1 as a signed number.
    // so0/so1 hold the values whose bit is odd in s0/s1, se0/se1 (shifted up
    // one) the values whose bit is even.
    uint32_t so0 = (s0 & 0xAAAAAAAAu) | 0x55555555u, se0 = ((s0 << 1) & 0xAAAAAAAAu) | 0x55555555u;
    uint32_t so1 = (s1 & 0xAAu) | 0x55u, se1 = ((s1 << 1) & 0xAAu) | 0x55u;
    // (opaque to the optimiser, which would otherwise extract the sign bit
    // alone and OR in the one: two instructions per value instead of one)
    asm volatile("" : "+v"(so0), "+v"(se0), "+v"(so1), "+v"(se1));
#pragma unroll
    for (uint32_t p = 0; p < NGRP; p++) {
        // the group's low bits big-endian: value 8p+c's field is r bits from bit 64 - (c+1) r
        const uint32_t e0 = q.ge[2 * p], e1 = q.ge[2 * p + 1];
        uint32_t gh = e0;
#pragma unroll
        for (uint32_t c = 0; c < 8; c++) {
            const uint32_t j = 8 * p + c;
            if (j >= L)
                break;
            if (c == 4)  // the window of value 4: bits 4r.. of e0:e1 (any value when r = 0)
                gh = __builtin_amdgcn_alignbit(e0, e1, 32 - 4 * r);
            const uint32_t hbyte = wide ? __builtin_amdgcn_ubfe(q.h[(6 + j) >> 2], 8 * ((6 + j) & 3), hb) : 0u;
            // {high byte, window} << r: the magnitude (< 2^15) in the upper
            // dword, the next value's window in the lower one
            const uint64_t v = (((uint64_t) hbyte << 32) | gh) << r;
            const uint32_t m = (uint32_t) (v >> 32);
            gh = (uint32_t) v;
            const uint32_t bit = j < 32 ? 8 * (j >> 3) + 7 - (j & 7) : 7 - (j - 32);
            const uint32_t sw = j < 32 ? ((bit & 1) ? so0 : se0) : ((bit & 1) ? so1 : se1);
            const int32_t f = __builtin_amdgcn_sbfe((int32_t) sw, bit & ~1u, 2);
            acc[j % N] = mad_i24(m, (uint32_t) f, acc[j % N]);
        }
    }
}

// Add a block of any width (nj values, nj <= L) unit by unit; values past nj
// add garbage that the caller clears.  Same fields as decode_fast; aligned
// loads only (this path also serves blocks that end in the view's last
// dword), each value's high bytes with a load of their own.  Branch-free and
// one unit's loads at a time: this path is rare and must not raise the
// kernel's register count.  Returns whether the block's width byte differs
// from w.
__device__ __forceinline__ bool decode_general(View v, uint32_t o, uint32_t w, uint32_t nj, uint32_t (&acc)[L])
{
    const uint32_t B = w >> 3, r = w & 7, nsb = (nj + 7) >> 3;
    const uint32_t hmask = B >= 4 ? ~0u : (1u << (8 * B)) - 1u;
    const uint32_t hp = o + 1 + nsb, lp = hp + B * nj;
    const bool bad = (bl4a(v, o) & 0xFFu) != w;
    const uint2 sg = bl8a(v, o + 1);
#pragma unroll
    for (uint32_t u = 0; u < L / 4; u++) {
        __builtin_amdgcn_sched_barrier(0);
        uint32_t H[4];
#pragma unroll
        for (uint32_t c = 0; c < 4; c++)
            H[c] = bl4a(v, hp + (4 * u + c) * B);
        const uint32_t X = bl4a(v, lp + ((4 * u * r) >> 3));
        const uint32_t sw = (u >> 1) < 4 ? sg.x : sg.y;
        const uint32_t nib = __builtin_amdgcn_ubfe(sw, 8 * ((u >> 1) & 3) + ((u & 1) ? 0 : 4), 4);
        uint32_t lo = __builtin_bswap32(X) << ((4 * u * r) & 4u);
#pragma unroll
        for (uint32_t c = 0; c < 4; c++) {
            const uint32_t hs = H[c] & hmask;
            const uint64_t x = (((uint64_t) hs << 32) | lo) << r;
            const uint32_t m = (uint32_t) (x >> 32);
            lo = (uint32_t) x;
            const uint32_t sm = (uint32_t) __builtin_amdgcn_sbfe((int32_t) nib, 3 - c, 1);
            acc[4 * u + c] += (m ^ sm) - sm;
        }
    }
    return bad;
}

// OR bytes into the wave's zeroed LDS buffer.
__device__ __forceinline__ void lds_or(uint32_t* p, uint32_t v)
{
    __hip_atomic_fetch_or(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

// Byte-shift selector for v_perm: perm(x_i, x_{i-1}, sel) is the dword at
// 4i of a byte array placed sh bytes into the aligned grid.
__device__ __forceinline__ uint32_t shift_sel(uint32_t sh) { return 0x07060504u - sh * 0x01010101u; }

// Encode a whole block of width 0 < w < 16 at byte o of the buffer: width
// byte, sign bitmap MSB first, one high byte per value when w >= 8, then the
// low r bits of every magnitude as an MSB-first stream (groups of 8 values
// fill r whole bytes).  wide: some lane of the wave has w >= 8 (otherwise
// the high bytes are known to be absent).
// ZCCL/src/ZCCL_TypeManager.c:12-91 (fixed-length bits), :824-848 (signs).
// s: signs of values 0..31 (value j at bit 31 - j), s2: signs of values
// 32..35 (value 32 at bit 3); m: magnitudes.
__device__ __forceinline__ void encode_fast(uint32_t* ob, uint32_t o, uint32_t w, uint32_t s, uint32_t s2,
    const uint32_t (&m)[L], bool wide)
{
    const uint32_t r = w & 7, B = w >> 3;
    const uint32_t sb = __builtin_bswap32(s);  // sign bytes 0..3 in memory order
    // high bytes (zero when w < 8: then m < 2^r)
    uint32_t H[L];
#pragma unroll
    for (uint32_t j = 0; j < L; j++)
        H[j] = m[j] >> r;
    uint32_t x[HEADW + 1];
    x[0] = w | (sb << 8);
    x[1] = (sb >> 24) | (s2 << (8 + 40 - L)) | (H[0] << 16) | (H[1] << 24);  // sign byte 4: value 32 at bit 7
#pragma unroll
    for (uint32_t i = 2; i < HEADW; i++) {
        const uint32_t j = 4 * i - 6;
        x[i] = H[j] | (j + 1 < L ? H[j + 1] << 8 : 0u) | (j + 2 < L ? H[j + 2] << 16 : 0u) |
               (j + 3 < L ? H[j + 3] << 24 : 0u);
    }
    x[HEADW] = 0;
    const uint32_t sel = shift_sel(o & 3);
    uint32_t* p = ob + (o >> 2);
    uint32_t prev = 0;
    const uint32_t nd = wide ? HEADW + 1 : 3;
#pragma unroll
    for (uint32_t i = 0; i < HEADW + 1; i++) {
        if (i < nd)
            lds_or(p + i, __builtin_amdgcn_perm(x[i], prev, sel));
        prev = x[i];
    }
    if (r) {
        const uint32_t mask = (1u << r) - 1, lo0 = o + 1 + SB + B * L;
#pragma unroll
        for (uint32_t g = 0; g < NGRP; g++) {
            uint32_t ua = 0, ub = 0;
#pragma unroll
            for (uint32_t c = 0; c < 4; c++) {
                const uint32_t j = 8 * g + c;
                ua = (ua << r) | (m[j] & mask);
                if (j + 4 < L)
                    ub = (ub << r) | (m[j + 4] & mask);
            }
            // 8 fields (the last group: 4 and zeros) = r bytes, top-aligned in Y
            const uint64_t G = ((uint64_t) ua << (4 * r)) | ub;
            const uint64_t Y = G << (64 - 8 * r);
            const uint32_t z0 = __builtin_bswap32((uint32_t) (Y >> 32)), z1 = __builtin_bswap32((uint32_t) Y);
            const uint32_t a = lo0 + g * r;
            const uint32_t gs = shift_sel(a & 3);
            uint32_t* q = ob + (a >> 2);
            lds_or(q, __builtin_amdgcn_perm(z0, 0u, gs));
            lds_or(q + 1, __builtin_amdgcn_perm(z1, z0, gs));
            lds_or(q + 2, __builtin_amdgcn_perm(0u, z1, gs));
        }
    }
}

// OR the n <= 4 little-endian bytes of v at LDS byte address a (two dwords).
__device__ __forceinline__ void or_bytes(uint32_t* ob, uint32_t a, uint32_t v)
{
    const uint64_t x = (uint64_t) v << (8 * (a & 3));
    uint32_t* q = ob + (a >> 2);
    lds_or(q, (uint32_t) x);
    lds_or(q + 1, (uint32_t) (x >> 32));
}

// Encode a block of any width w > 0 with 0 < nj <= L values (values past nj
// are zero) by OR-writes, value by value and without branches, for the rare
// blocks encode_fast does not take (widths from 16, a segment's short last
// block).  Writes of the zero values past nj go to the last value's bytes
// or at most 30 bytes past the block.  Same format and arguments as
// encode_fast.
__device__ __forceinline__ void encode_general(uint32_t* ob, uint32_t o, uint32_t w, uint32_t nj, uint32_t s,
    uint32_t s2, const uint32_t (&m)[L])
{
    const uint32_t B = w >> 3, r = w & 7, nsb = (nj + 7) >> 3;
    const uint32_t hmask = B >= 4 ? ~0u : (1u << (8 * B)) - 1u;
    const uint32_t fmask = (1u << r) - 1u;
    const uint32_t hp = o + 1 + nsb, lo0 = hp + B * nj;
    // width byte and sign bytes (the bits of values past nj are zero, and so
    // are the sign bytes past nsb that these dwords also cover)
    const uint32_t sb = __builtin_bswap32(s);
    const uint32_t x0 = w | (sb << 8), x1 = (sb >> 24) | (s2 << (8 + 40 - L));
    const uint32_t sel = shift_sel(o & 3);
    uint32_t* p = ob + (o >> 2);
    lds_or(p, __builtin_amdgcn_perm(x0, 0u, sel));
    lds_or(p + 1, __builtin_amdgcn_perm(x1, x0, sel));
    lds_or(p + 2, __builtin_amdgcn_perm(0u, x1, sel));
#pragma unroll
    for (uint32_t j = 0; j < L; j++) {
        __builtin_amdgcn_sched_barrier(0);  // value by value: few registers live
        // high bytes: B bytes of m >> r at hp + j B (m >> r < 2^32, so width 32 has 4)
        or_bytes(ob, hp + min(j, nj - 1) * B, (m[j] >> r) & hmask);
        // low bits: r bits at bit j r of the MSB-first stream (a 16-bit window)
        const uint32_t bit = j * r;
        const uint32_t be = (m[j] & fmask) << (16 - r - (bit & 7));
        or_bytes(ob, lo0 + (bit >> 3), ((be >> 8) & 0xFFu) | ((be & 0xFFu) << 8));
    }
}

// Write bytes [0, nbytes) of the wave's LDS buffer to dst with 16-byte stores
// on dst's 16-byte grid; the partial first and last chunk go out byte by
// byte, so bytes outside [dst, dst + nbytes) are never written.
__device__ __forceinline__ void store_wave(const uint32_t* ob, uint8_t* dst, uint32_t nbytes)
{
    const uint32_t lane = threadIdx.x & 63;
    const uintptr_t d0 = (uintptr_t) dst, d1 = d0 + nbytes;
    const uintptr_t c0 = d0 & ~(uintptr_t) 15;
    const uint32_t nch = nbytes ? (uint32_t) ((d1 - c0 + 15) >> 4) : 0u;
    for (uint32_t c = lane; c < nch; c += 64) {
        const uintptr_t ca = c0 + 16ull * c;
        if (ca >= d0 && ca + 16 <= d1) {
            const uint32_t o = (uint32_t) (ca - d0), wi = o >> 2, sh = o & 3;
            const uint32_t a = ob[wi], b = ob[wi + 1], cc = ob[wi + 2], dd = ob[wi + 3], e = ob[wi + 4];
            *(uint4*) ca = make_uint4(__builtin_amdgcn_alignbyte(b, a, sh), __builtin_amdgcn_alignbyte(cc, b, sh),
                __builtin_amdgcn_alignbyte(dd, cc, sh), __builtin_amdgcn_alignbyte(e, dd, sh));
        } else {
            const uintptr_t s = ca > d0 ? ca : d0, e = ca + 16 < d1 ? ca + 16 : d1;
            for (uintptr_t p = s; p < e; p++)
                *(uint8_t*) p = ((const uint8_t*) ob)[p - d0];
        }
    }
}

// The operands of one sum: in device memory (read through the scalar cache)
// or in the kernel arguments.
struct SumIO {
    const HomsumIO* io;
    __device__ const uint8_t* bytes(int r) const { return ldc(&io->in[r].bytes); }
    __device__ const uint8_t* index(int r) const { return ldc(&io->in[r].index); }
};

struct ArgIO {
    const HomsumIO& io;
    __device__ const uint8_t* bytes(int r) const { return io.in[r].bytes; }
    __device__ const uint8_t* index(int r) const { return io.in[r].index; }
};

// Staging area of one operand's tile bytes: half of the wave's output buffer
// (the buffer is free until the encode), less the reach of the block reads
// past the staged bytes (up to 47 bytes).
constexpr uint32_t kStageHalf = kOutBuf / 2;
constexpr uint32_t kStageCap = kStageHalf - 64;
static_assert(kStageHalf % 16 == 0, "stage halves stay 16-byte aligned");

// A lane's block from tile bytes staged in LDS at st (offset o, index width
// w): aligned dword reads (a misaligned LDS read is legal but slow).
__device__ __forceinline__ void read_block_staged(const uint32_t* st, uint32_t o, uint32_t w, Block& q)
{
    const uint32_t* p = st + (o >> 2);
    uint32_t D[HEADW + 1];
#pragma unroll
    for (uint32_t i = 0; i < HEADW + 1; i++)
        D[i] = p[i];
#pragma unroll
    for (uint32_t i = 0; i < HEADW; i++)
        q.h[i] = __builtin_amdgcn_alignbyte(D[i + 1], D[i], o);  // uses o & 3
    const uint32_t r = w & 7, lo = o + 1 + SB + high_bytes(w);
    uint32_t a = lo;
#pragma unroll
    for (uint32_t gp = 0; gp < NGRP; gp++, a += r) {
        const uint32_t* pg = st + (a >> 2);
        const uint32_t E0 = pg[0], E1 = pg[1], E2 = pg[2];
        // the big-endian dwords at byte a & 3 (alignbyte uses a & 3)
        q.ge[2 * gp] = __builtin_bswap32(__builtin_amdgcn_alignbyte(E1, E0, a));
        q.ge[2 * gp + 1] = __builtin_bswap32(__builtin_amdgcn_alignbyte(E2, E1, a));
    }
}

// The same straight from a view, per lane (tiles too large to stage).
__device__ __forceinline__ void read_block_direct(View v, uint32_t o, uint32_t w, Block& q)
{
    const uint4 h0 = bl16(v, o), h1 = bl16(v, o + 16);
    const uint3 h2 = bl12(v, o + 32);
    q.h[0] = h0.x; q.h[1] = h0.y; q.h[2] = h0.z; q.h[3] = h0.w;
    q.h[4] = h1.x; q.h[5] = h1.y; q.h[6] = h1.z; q.h[7] = h1.w;
    q.h[8] = h2.x; q.h[9] = h2.y; q.h[10] = h2.z;
    const uint32_t r = w & 7, lo = o + 1 + SB + high_bytes(w);
#pragma unroll
    for (uint32_t gp = 0; gp < NGRP; gp++) {
        const uint2 e = bl8(v, lo + gp * r);
        q.ge[2 * gp] = __builtin_bswap32(e.x);
        q.ge[2 * gp + 1] = __builtin_bswap32(e.y);
    }
}

// Shared state of a workgroup of lane waves that owns one lookback tile:
// each wave's bytes, and the tile's stream offset.
struct WgShared {
    uint32_t wtot[kLaneWaves];
    unsigned long long ex;
};

// Forms of the lane sum's output side.
enum SumMode : int {
    kSumWave = 0,    // single pass, a lookback tile per wave
    kSumStaged = 1,  // pass one of the two-pass sum: no lookback; the tile's bytes go to
                     // slot sidx of the stage area `area` of R tiles, its sizes and seed to
                     // record sidx, and only the widths to the output index
    kSumWg = 2,      // single pass, a lookback tile per workgroup of kLaneWaves waves
};

// One tile of a fused k-operand sum, executed by one wave.  obuf: the wave's
// LDS buffer (kOutBuf bytes, 16-byte aligned).  kSumWg: t is the wave's tile
// of the lane tiling, t_wg the workgroup's lookback tile (the last of the
// launch when wg_last), wsh the workgroup's shared state; every wave of the
// workgroup calls this (a wave past its segment's blocks with an empty tile).
template <int Mode, typename Src>
__device__ __forceinline__ void homsum_tile_lane(const Geom& g, const Tiling& tl, int k, const Src& src,
    uint8_t* out, uint64_t out_cap, uint8_t* oindex, unsigned long long* size_out, uint64_t t,
    unsigned long long* status, WsHeader* ws, uint32_t nchains, uint32_t* obuf, uint8_t* area = nullptr,
    uint64_t R = 0, uint64_t sidx = 0, uint64_t t_wg = 0, bool wg_last = false, WgShared* wsh = nullptr,
    const Tile* xe = nullptr)
{
    const uint32_t lane = threadIdx.x & 63;
    const Tile x = xe ? *xe : tile_of(g, tl, (uint32_t) t);
    const bool has = lane < x.nblk;
    const uint32_t nj = has ? min(tile_block_len(g, x, lane), L) : 0u;
    bool bad_byte = false, bad_width = false, bad_over = false;

    // ---- operand table: lane r < k holds operand r's stream, index, tile
    // offset and stream size, read once per tile; each operand's values come
    // out of it with readlane, so no operand waits on a scalar load chain
    const uint8_t* tb_bytes = nullptr;
    const uint8_t* tb_index = nullptr;
    uint64_t tb_goff = 0, tb_size = 0;
    if (lane < (uint32_t) k) {
        tb_bytes = src.bytes((int) lane);
        tb_index = src.index((int) lane);
        const IndexView iv = index_view(g, tb_index);
        tb_goff = iv.group_off[x.grp0];
        tb_size = iv.meta[0];
    }
    // (readlane returns int: widen through uint32_t, never by sign extension)
    auto rl64 = [](uint64_t v, int r) -> uint64_t {
        const uint32_t hi = (uint32_t) __builtin_amdgcn_readlane((uint32_t) (v >> 32), r);
        const uint32_t lo = (uint32_t) __builtin_amdgcn_readlane((uint32_t) v, r);
        return ((uint64_t) hi << 32) | lo;
    };
    struct Op {
        const uint8_t* bytes;
        const uint8_t* index;
        uint64_t goff, size;
    };
    auto op_of = [&](int r) {
        Op o;
        o.bytes = (const uint8_t*) rl64((uint64_t) tb_bytes, r);
        o.index = (const uint8_t*) rl64((uint64_t) tb_index, r);
        o.goff = rl64(tb_goff, r);
        o.size = rl64(tb_size, r);
        return o;
    };
    auto widths = [&](const Op& o) -> uint32_t { return has ? (uint32_t) o.index[x.g0 + lane] : 0u; };

    // this lane's block of an operand: width, size, offset in the operand's
    // view of the tile bytes (one scan of the block sizes), and whether the
    // tile lies inside its stream.  The view starts at the 4-byte aligned
    // address at or below the tile's bytes and ends with the stream's last
    // dword (lim: its length).
    struct Loc {
        View v;
        uint32_t w, bs, o, lim, tot;
        bool skip;
    };
    auto locate = [&](const Op& p, uint32_t wq) {
        Loc c;
        c.w = wq;
        if (c.w > 32) {
            bad_width = true;
            c.w = 0;
        }
        c.bs = has ? (nj == L ? full_block_bytes(c.w) : lane_block_bytes(c.w, nj)) : 0u;
        const uint32_t incl = wave_scan_dpp(c.bs);
        c.tot = __builtin_amdgcn_readlane(incl, 63);
        c.skip = p.goff > p.size || c.tot > p.size - p.goff;
        const uint8_t* base = p.bytes + p.goff;
        const uint32_t delta = (uint32_t) ((uintptr_t) base & 3u);
        const uint64_t len = c.skip ? 0ull : align_up(delta + (p.size - p.goff), 4);
        c.lim = len > 0xFFFFFFFCull ? 0xFFFFFFFCu : (uint32_t) len;
        c.v = make_view(base - delta, len);
        c.o = delta + incl - c.bs;
        c.tot += delta;  // bytes of the view that hold the tile
        return c;
    };
    // move an operand's tile bytes (view bytes [0, tot)) into LDS half `half`
    // with 16-byte direct-to-LDS buffer loads, 64 lanes per instruction; the
    // view's range check zero-fills past the stream.  False when they do not
    // fit (then the blocks are read with per-lane loads).
    auto stage = [&](const Loc& c, uint32_t half) -> bool {
        const uint32_t nch = (c.tot + 15) >> 4;
        if (c.skip || 16 * nch > kStageCap)
            return false;
        uint8_t* st = (uint8_t*) obuf + half * kStageHalf;
        for (uint32_t run = 0; run * 64 < nch; run++) {
            const uint32_t ch = run * 64 + lane;
            if (ch < nch)
                __builtin_amdgcn_raw_ptr_buffer_load_lds(c.v, (__attribute__((address_space(3))) void*) (st + run * 1024),
                    16, 16 * ch, 0, 0, 0);
        }
        return true;
    };
    auto read_staged = [&](const Loc& c, uint32_t half, Block& q) {
        read_block_staged(obuf + half * (kStageHalf / 4), c.o, c.w, q);
    };
    auto read_direct = [&](const Loc& c, Block& q) { read_block_direct(c.v, c.o, c.w, q); };

    uint32_t acc[L];
#pragma unroll
    for (uint32_t j = 0; j < L; j++)
        acc[j] = 0;
    // bit r: this lane's block of operand r needs the general decoder, which
    // runs after the pipelined loop (fewer registers live there)
    uint32_t gen = 0;

    // ---- operand pipeline: operand r+1's tile bytes move into the other LDS
    // half while operand r is decoded, operand r+2's widths are on their way
    {
        Op on = op_of(0);
        Loc cn = locate(on, widths(on));
        bool sn = stage(cn, 0);
        uint32_t wq2 = 0;
        Op o2;
        if (k > 1) {
            o2 = op_of(1);
            wq2 = widths(o2);
        }
        for (int r = 0; r < k; r++) {
            const Loc c = cn;
            const bool staged = sn;
            __builtin_amdgcn_s_waitcnt(0);  // operand r's bytes are in LDS, operand r+1's widths in registers
            Block q;
            q.w = c.w;
            q.o = c.o;
            q.lim = c.lim;
            if (staged)
                read_staged(c, (uint32_t) r & 1u, q);
            else
                read_direct(c, q);
            // (phase fences keep the scheduler from interleaving the block
            // reads, the next operand's staging and the decode, which would
            // hold all their values at once)
            __builtin_amdgcn_sched_barrier(0);
            if (r + 1 < k) {
                cn = locate(o2, wq2);
                sn = stage(cn, (uint32_t) (r + 1) & 1u);
            }
            if (r + 2 < k) {
                o2 = op_of(r + 2);
                wq2 = widths(o2);
            }
            __builtin_amdgcn_sched_barrier(0);
            // decode_fast reads a block correctly unless the per-lane loads
            // read it and it ends in the view's last three bytes (a dword
            // straddling the end reads as zero); such a block, a wide one or
            // a short one goes to the general decoder.  Zero-width and empty
            // lanes add nothing either way.  Branch-free: a lane that does
            // not use the fast decoder runs it at width 0, which adds zeros.
            bad_over |= c.skip;
            const bool fast = !has || q.w == 0 || (nj == L && q.w < 16 && (staged || q.o + c.bs + 3 <= q.lim));
            const bool use = fast && !c.skip;
            if (!fast && !c.skip)
                gen |= 1u << r;
            // (a zero-width block in the view's last three bytes is not
            // checked when read per lane: its first dword reads as zero)
            bad_byte |= use && has && (q.w || staged || q.o + 4 <= q.lim) && (q.h[0] & 0xFFu) != q.w;
            q.w = use ? q.w : 0u;
            decode_fast<true>(q, acc);
        }
    }
    if (__any(gen)) {
        for (int r = 0; r < k; r++) {
            if (!__any((gen >> r) & 1u))
                continue;
            const Op o = op_of(r);
            const Loc c = locate(o, widths(o));
            if ((gen >> r) & 1u)
                bad_byte |= decode_general(c.v, c.o, c.w, nj, acc);
        }
    }
    if (__any(bad_width) && lane == 0)
        raise_error(ws, ZK_ERR_WIDTH);
    if (__any(bad_byte) && lane == 0)
        raise_error(ws, ZK_ERR_MISMATCH);
    if (__any(bad_over) && lane == 0)
        raise_error(ws, ZK_ERR_OVERRUN);

    // seed of the segment (its first tile): the operands' seeds add in int32
    uint32_t seed = 0;
    if (x.lt == 0) {
        uint32_t sv = 0;
        if (lane < (uint32_t) k) {
            const IndexView iv = index_view(g, tb_index);
            const uint64_t so = iv.seed_off[x.s];
            if (so + 4 <= tb_size)
                sv = load_u32_bytes(tb_bytes + so);
            else
                raise_error(ws, ZK_ERR_OVERRUN);
        }
        seed = (uint32_t) wave_sum64(sv);
    }

    // ---- output: widths, sizes, offsets
    if (nj < L) {
#pragma unroll
        for (uint32_t j = 0; j < L; j++)
            if (j >= nj)
                acc[j] = 0;
    }
    __builtin_amdgcn_sched_barrier(0);
    // signs, then the magnitudes in place of the deltas
    uint32_t s = 0, s2 = 0;
#pragma unroll
    for (uint32_t j = 0; j < 32; j++)
        s = __builtin_amdgcn_alignbit(s, acc[j], 31);  // (s << 1) | sign of value j
#pragma unroll
    for (uint32_t j = 32; j < L; j++)
        s2 = __builtin_amdgcn_alignbit(s2, acc[j], 31);
    // (phase fences: without them the scheduler computes the magnitudes
    // before the sign words and keeps both arrays live)
    __builtin_amdgcn_sched_barrier(0);
    uint32_t mx = 0;
#pragma unroll
    for (uint32_t j = 0; j < L; j++) {
        acc[j] = mag(acc[j]);
        mx = max(mx, acc[j]);
    }
    __builtin_amdgcn_sched_barrier(0);
    const uint32_t w = has ? dwidth(mx) : 0u;
    const uint32_t bs = has ? (nj == L ? full_block_bytes(w) : lane_block_bytes(w, nj)) : 0u;
    const uint32_t incl = wave_scan_dpp(bs);
    const uint32_t excl = incl - bs;
    const uint32_t total = __builtin_amdgcn_readlane(incl, 63);
    if (Mode == kSumWave && lane == 0)
        lookback_publish(status, t, 0, total, 0);  // successors see the aggregate now
    const bool fast = nj == L && w < 16;
    const bool wide = __any(has && w >= 8);

    if constexpr (Mode == kSumWg) {
        // the workgroup's tile: its waves' bytes in wave order, one lookback
        const uint32_t wv = threadIdx.x >> 6;
        if (lane == 0)
            wsh->wtot[wv] = total;
        __syncthreads();
        uint32_t woff = 0, gtot = 0;
#pragma unroll
        for (uint32_t i = 0; i < kLaneWaves; i++) {
            const uint32_t v = wsh->wtot[i];
            woff += i < wv ? v : 0u;
            gtot += v;
        }
        if (threadIdx.x == 0)
            lookback_publish(status, t_wg, 0, gtot, 0);  // successors see the aggregate now
        const bool split = total + 64 > kOutBuf;
        if (!split) {
            const uint32_t nz = (total + 64 + 15) / 16;
            for (uint32_t i = lane; i < nz; i += 64)
                reinterpret_cast<uint4*>(obuf)[i] = make_uint4(0, 0, 0, 0);
            if (has && w && fast)
                encode_fast(obuf, excl, w, s, s2, acc, wide);
            if (has && w && !fast)
                encode_general(obuf, excl, w, nj, s, s2, acc);
        }
        if (wv == 0) {
            const unsigned long long ex = lookback(status, t_wg, 0, gtot, 0, ws, true);
            if (lane == 0)
                wsh->ex = ex;
        }
        __syncthreads();
        const uint64_t hdr = 8ull * g.T + 4ull * (x.s + 1);
        const uint64_t P0 = hdr + wsh->ex;  // the workgroup tile's stream offset
        const uint64_t P = P0 + woff;       // this wave's
        const bool fits = P0 + gtot <= out_cap;
        if (!split) {
            if (fits)
                store_wave(obuf, out + P, total);
        } else {
            // a wave too wide for the buffer: half a wave at a time, value by value
            const uint32_t mid = __builtin_amdgcn_readlane(excl, 32);
            for (uint32_t i = lane; i < (mid + 64 + 15) / 16; i += 64)
                reinterpret_cast<uint4*>(obuf)[i] = make_uint4(0, 0, 0, 0);
            if (lane < 32 && has && w)
                encode_general(obuf, excl, w, nj, s, s2, acc);
            if (fits)
                store_wave(obuf, out + P, mid);
            for (uint32_t i = lane; i < (total - mid + 64 + 15) / 16; i += 64)
                reinterpret_cast<uint4*>(obuf)[i] = make_uint4(0, 0, 0, 0);
            if (lane >= 32 && has && w)
                encode_general(obuf, excl - mid, w, nj, s, s2, acc);
            if (fits)
                store_wave(obuf, out + P + mid, total - mid);
        }
        if (!fits && threadIdx.x == 0)
            raise_error(ws, ZK_ERR_NOSPACE);
        IndexView ox = index_view(g, oindex);
        if (has)
            ox.widths[x.g0 + lane] = (uint8_t) w;
        if ((lane & 15) == 0 && lane / 16 < x.ngrp)
            ox.group_off[x.grp0 + lane / 16] = P + excl;
        if (lane == 0 && x.lt == 0) {
            const uint64_t so = P0 - 4;
            if (so + 4 <= out_cap) {
                store_bytes(out + so, &seed, 4);
                if (x.s < g.T) {
                    const unsigned long long h = so - 8ull * g.T;
                    store_bytes(out + 8ull * x.s, &h, 8);
                }
            }
            ox.seed_off[x.s] = so;
        }
        if (threadIdx.x == 0 && wg_last) {
            // the last tile ends the stream: if it fits, every tile fits
            const unsigned long long sz = fits ? P0 + gtot : 0ull;
            ox.meta[0] = sz;
            if (size_out)
                *size_out = sz;
            chain_finished(ws, nchains);
        }
        return;
    }

    if constexpr (Mode == kSumStaged) {
        uint8_t* slot = area + (8 + sizeof(StageRec)) * R + kStageSlot * sidx;
        if (total + 64 <= kOutBuf) {
            const uint32_t nz = (total + 64 + 15) / 16;
            for (uint32_t i = lane; i < nz; i += 64)
                reinterpret_cast<uint4*>(obuf)[i] = make_uint4(0, 0, 0, 0);
            if (has && w && fast)
                encode_fast(obuf, excl, w, s, s2, acc, wide);
            if (has && w && !fast)
                encode_general(obuf, excl, w, nj, s, s2, acc);
            // whole 16-byte chunks: the slot is the tile's own (total <= kStageSlot, a multiple of 16)
            for (uint32_t i = lane; i < (total + 15) / 16; i += 64)
                reinterpret_cast<uint4*>(slot)[i] = reinterpret_cast<const uint4*>(obuf)[i];
        } else {
            // too wide for the buffer: half a wave at a time, value by value
            const uint32_t mid = __builtin_amdgcn_readlane(excl, 32);
            for (uint32_t i = lane; i < (mid + 64 + 15) / 16; i += 64)
                reinterpret_cast<uint4*>(obuf)[i] = make_uint4(0, 0, 0, 0);
            if (lane < 32 && has && w)
                encode_general(obuf, excl, w, nj, s, s2, acc);
            store_wave(obuf, slot, mid);
            for (uint32_t i = lane; i < (total - mid + 64 + 15) / 16; i += 64)
                reinterpret_cast<uint4*>(obuf)[i] = make_uint4(0, 0, 0, 0);
            if (lane >= 32 && has && w)
                encode_general(obuf, excl - mid, w, nj, s, s2, acc);
            store_wave(obuf, slot + mid, total - mid);
        }
        if (has)
            index_view(g, oindex).widths[x.g0 + lane] = (uint8_t) w;
        const uint32_t e16 = __builtin_amdgcn_readlane(excl, 16), e32 = __builtin_amdgcn_readlane(excl, 32),
                       e48 = __builtin_amdgcn_readlane(excl, 48);
        if (lane == 0) {
            StageRec rec;
            rec.total = total;
            rec.e16 = (uint16_t) e16;
            rec.e32 = (uint16_t) e32;
            rec.e48 = (uint16_t) e48;
            rec.pad = 0;
            rec.seed = seed;
            reinterpret_cast<StageRec*>(area + 8 * R)[sidx] = rec;
        }
        return;
    }

    // ---- encode into the LDS buffer and store
    const uint64_t hdr = 8ull * g.T + 4ull * (x.s + 1);
    unsigned long long ex;
    if (total + 64 <= kOutBuf) {
        const uint32_t nz = (total + 64 + 15) / 16;
        for (uint32_t i = lane; i < nz; i += 64)
            reinterpret_cast<uint4*>(obuf)[i] = make_uint4(0, 0, 0, 0);
        // (two separate branches rather than if/else: the compiler would hoist
        // what the two encoders share above an if/else and keep it live)
        if (has && w && fast)
            encode_fast(obuf, excl, w, s, s2, acc, wide);
        if (has && w && !fast)
            encode_general(obuf, excl, w, nj, s, s2, acc);
        // the tile's stream offset: after its encoding, so predecessors had longer to publish
        ex = lookback(status, t, 0, total, 0, ws, true);
        if (hdr + ex + total <= out_cap)
            store_wave(obuf, out + hdr + ex, total);
    } else {
        // a tile too wide for the buffer (its blocks average over 127 bytes):
        // half a wave at a time, each half encoded value by value after the
        // offset is known.  Straight-line: a loop here would keep the
        // accumulators in scratch memory.
        ex = lookback(status, t, 0, total, 0, ws, true);
        const uint32_t mid = __builtin_amdgcn_readlane(excl, 32);  // lane 32's block offset
        const bool ok = hdr + ex + total <= out_cap;
        for (uint32_t i = lane; i < (mid + 64 + 15) / 16; i += 64)
            reinterpret_cast<uint4*>(obuf)[i] = make_uint4(0, 0, 0, 0);
        if (lane < 32 && has && w)
            encode_general(obuf, excl, w, nj, s, s2, acc);
        if (ok)
            store_wave(obuf, out + hdr + ex, mid);
        for (uint32_t i = lane; i < (total - mid + 64 + 15) / 16; i += 64)
            reinterpret_cast<uint4*>(obuf)[i] = make_uint4(0, 0, 0, 0);
        if (lane >= 32 && has && w)
            encode_general(obuf, excl - mid, w, nj, s, s2, acc);
        if (ok)
            store_wave(obuf, out + hdr + ex + mid, total - mid);
    }
    const uint64_t P = hdr + ex;
    const bool fits = P + total <= out_cap;
    if (!fits && lane == 0)
        raise_error(ws, ZK_ERR_NOSPACE);

    // ---- output index, seed and chunk header
    IndexView ox = index_view(g, oindex);
    if (has)
        ox.widths[x.g0 + lane] = (uint8_t) w;
    if ((lane & 15) == 0 && lane / 16 < x.ngrp)
        ox.group_off[x.grp0 + lane / 16] = P + excl;
    if (lane == 0) {
        if (x.lt == 0) {
            const uint64_t so = P - 4;
            if (so + 4 <= out_cap) {
                store_bytes(out + so, &seed, 4);
                if (x.s < g.T) {
                    const unsigned long long h = so - 8ull * g.T;
                    store_bytes(out + 8ull * x.s, &h, 8);
                }
            }
            ox.seed_off[x.s] = so;
        }
        if (t == tl.ntiles - 1) {
            // the last tile ends the stream: if it fits, every tile fits
            const unsigned long long sz = fits ? P + total : 0ull;
            ox.meta[0] = sz;
            if (size_out)
                *size_out = sz;
            chain_finished(ws, nchains);
        }
    }
}

// ---------------------------------------------------------------- decompress
//
// Two passes, no wait between tiles.  Pass one: lane j decodes block j of the
// wave's tile (the sum's loaders and decoders, one operand) and the wave
// writes the tile's sum of deltas.  A scan per segment turns the sums into
// each tile's carry, starting from the segment's seed.  Pass two decodes the
// tile again, takes the running sums of the deltas in the lane, places the
// blocks with one scan of the block sums across the wave, dequantizes, and
// stores the values through the wave's LDS buffer, half a tile at a time,
// with 16-byte stores.  Arithmetic: prefix sums of the deltas in int32 from
// the segment's seed, x' = (float) q * eb (ZCCL/src/ZCCLd_float.c:588-628).
//
// The tile sums live in the workspace's lookback status array of the parity
// that the next lookback launch clears (the one other than its own), so the
// two passes leave no state that a later launch must see cleared.
__device__ __forceinline__ uint32_t* dec_sums(WsHeader* ws)
{
    const unsigned long long par = (__hip_atomic_load(&ws->seq, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) >> 32) & 1ull;
    return (uint32_t*) ((char*) ws + sizeof(WsHeader) + (par ^ 1ull) * ldc(&ws->capacity) * kStatusStride * 8);
}

// The segment's seed (0 with an error when its offset is past the stream).
__device__ __forceinline__ uint32_t segment_seed(const Geom& g, const uint8_t* bytes, const uint8_t* index,
    uint32_t s, WsHeader* ws)
{
    const IndexView iv = index_view(g, index);
    const uint64_t so = iv.seed_off[s], size = iv.meta[0];
    if (so + 4 <= size)
        return load_u32_bytes(bytes + so);
    raise_error(ws, ZK_ERR_OVERRUN);
    return 0;
}

// Launch tile tg of a decompress in the lane layout: one stream
// (DecompressArgs), or a batch (uniform: shared geometry and bound, per-job
// pointers; mixed: per-job descriptors).
struct DecJob {
    Geom g;
    Tiling tl;
    float eb;
    const uint8_t* bytes;
    const uint8_t* index;
    float* out;
    uint32_t t;         // tile inside the job
    uint64_t tile0;     // the job's first launch tile
};

__device__ __forceinline__ bool dec_job(const DecompressArgs& a, uint64_t tg, DecJob& d)
{
    if (tg >= a.job.tl.ntiles)
        return false;
    d.g = a.job.g;
    d.tl = a.job.tl;
    d.eb = a.job.p.eb;
    d.bytes = a.job.io.in[0].bytes;
    d.index = a.job.io.in[0].index;
    d.out = a.job.io.out;
    d.t = (uint32_t) tg;
    d.tile0 = 0;
    return true;
}

__device__ __forceinline__ bool dec_job(const BatchArgs& b, uint64_t tg, DecJob& d)
{
    if (tg >= b.total_tiles)
        return false;
    if (b.uniform) {
        const uint32_t j = __builtin_amdgcn_readfirstlane((uint32_t) (tg / b.tl.ntiles));
        const DecompressIO* io = (const DecompressIO*) b.jobs + j;
        d.g = b.g;
        d.tl = b.tl;
        d.eb = b.p.eb;
        d.bytes = ldc(&io->in[0].bytes);
        d.index = ldc(&io->in[0].index);
        d.out = ldc(&io->out);
        d.tile0 = (uint64_t) j * b.tl.ntiles;
    } else {
        uint32_t lo = 0, hi = b.njobs;
        while (hi - lo > 1) {
            const uint32_t mid = (lo + hi) / 2;
            if (ldc(&b.job_tile0[mid]) <= tg)
                lo = mid;
            else
                hi = mid;
        }
        const uint32_t j = __builtin_amdgcn_readfirstlane(lo);
        const DecompressJob* job = (const DecompressJob*) b.jobs + j;
        d.g = job->g;
        d.tl = job->tl;
        d.eb = ldc(&job->p.eb);
        d.bytes = ldc(&job->io.in[0].bytes);
        d.index = ldc(&job->io.in[0].index);
        d.out = ldc(&job->io.out);
        d.tile0 = ldc(&b.job_tile0[j]);
    }
    d.t = __builtin_amdgcn_readfirstlane((uint32_t) (tg - d.tile0));
    return true;
}

// One tile of a decompress pass as it moves through the pipeline: its launch
// index, the lane's width byte and the tile's offset and stream size from the
// index (job, geometry and tile are derived again where needed, with scalar
// work, so that three tiles in flight hold few registers); then the lane's
// block located in the tile bytes, and the bytes staged.
struct DecPre {
    bool ok;
    uint64_t idx;
    uint64_t goff, size;
    uint32_t wq;
    uint32_t carry;    // pass two: the tile's carry
};

struct DecLoc {
    View v;
    uint32_t w, bs, o, lim, nj;
    bool has, skip, staged;
};

// Issue the index reads of launch tile tg (and its carry in pass two).
template <int Pass, typename Args>
__device__ __forceinline__ DecPre dec_prefetch(const Args& a, uint64_t tg, const uint32_t* sums)
{
    DecPre p;
    DecJob d;
    p.ok = dec_job(a, tg, d);
    if (!p.ok)
        return p;
    const Geom& g = d.g;
    const Tile x = tile_of(g, d.tl, d.t);
    p.idx = tg;
    const IndexView iv = index_view(g, d.index);
    p.goff = iv.group_off[x.grp0];
    p.size = iv.meta[0];
    const uint32_t lane = threadIdx.x & 63;
    p.wq = lane < x.nblk ? (uint32_t) iv.widths[x.g0 + lane] : 0u;
    p.carry = Pass == 2 ? sums[tg] : 0u;
    return p;
}

// Locate the lane's block of tile p (one scan of the block sizes) and stage
// the tile's bytes into LDS half `half` of the wave buffer (16-byte
// direct-to-LDS loads, zeros past the stream); tiles too large to stage are
// read per lane.
template <typename Args>
__device__ __forceinline__ DecLoc dec_locate(const Args& a, const DecPre& p, uint32_t* obuf, uint32_t half,
    bool& bad_width)
{
    DecJob d;
    dec_job(a, p.idx, d);
    const Geom& g = d.g;
    const Tile x = tile_of(g, d.tl, d.t);
    DecLoc c;
    const uint32_t lane = threadIdx.x & 63;
    c.has = lane < x.nblk;
    c.nj = c.has ? min(tile_block_len(g, x, lane), L) : 0u;
    c.w = p.wq;
    if (c.w > 32) {
        bad_width = true;
        c.w = 0;
    }
    c.bs = c.has ? (c.nj == L ? full_block_bytes(c.w) : lane_block_bytes(c.w, c.nj)) : 0u;
    const uint32_t binc = wave_scan_dpp(c.bs);
    const uint32_t tot0 = __builtin_amdgcn_readlane(binc, 63);
    c.skip = p.goff > p.size || tot0 > p.size - p.goff;
    const uint8_t* base = d.bytes + p.goff;
    const uint32_t delta = (uint32_t) ((uintptr_t) base & 3u);
    const uint64_t vlen = c.skip ? 0ull : align_up(delta + (p.size - p.goff), 4);
    c.lim = vlen > 0xFFFFFFFCull ? 0xFFFFFFFCu : (uint32_t) vlen;
    c.v = make_view(base - delta, vlen);
    c.o = delta + binc - c.bs;
    const uint32_t nch = (tot0 + delta + 15) >> 4;
    c.staged = !c.skip && 16 * nch <= kStageCap;
    if (c.staged) {
        uint8_t* st = (uint8_t*) obuf + half * kStageHalf;
        for (uint32_t run = 0; run * 64 < nch; run++) {
            const uint32_t ch = run * 64 + lane;
            if (ch < nch)
                __builtin_amdgcn_raw_ptr_buffer_load_lds(c.v, (__attribute__((address_space(3))) void*) (st + run * 1024),
                    16, 16 * ch, 0, 0, 0);
        }
    }
    return c;
}

// Store the wave's blocks of values (lane l: v[0..nj), at base + 36 l; the
// lanes below nblk hold one, all whole but perhaps the last) so that every
// store instruction writes contiguous, 16-byte aligned chunks.  The blocks
// start D dwords past the 16-byte grid; chunk c of the wave covers dwords
// [4c - D, 4c - D + 4) from base.  A quarter of the wave (16 blocks, 144
// chunks) at a time: its lanes write their values to LDS (ob) at their
// offsets in the quarter's chunks, the lane before the quarter its last D
// values, and the wave stores the chunks with streaming (non-temporal)
// stores.  Lane 0's first chunk is partly the previous tile's and the last
// whole block's last D values start the next tile's (both by dword stores),
// and a segment's short last block goes value by value.  (Unaligned HBM
// writes, and 16-byte stores strided 144 bytes across the lanes, both ran at
// under half the rate of aligned contiguous ones.)
__device__ __forceinline__ void store_blocks(float* base, const float (&v)[L], uint32_t nj, uint32_t nblk,
    uint32_t* ob)
{
    typedef unsigned int u32x4v __attribute__((ext_vector_type(4)));
    const uint32_t lane = threadIdx.x & 63;
    const bool full = lane < nblk && nj == L;
    const unsigned long long fm = __ballot(full);
    const uint32_t nfull = (uint32_t) __popcll(fm);  // whole blocks: a prefix of the lanes
    const bool next_full = lane < 63 && ((fm >> (lane + 1)) & 1ull);
    const uint32_t D = __builtin_amdgcn_readfirstlane((uint32_t) (((uintptr_t) base >> 2) & 3u));
    float* a0 = base - D;  // 16-byte aligned
    const uint32_t nch = 9 * nfull;
    float* of = reinterpret_cast<float*>(ob);
#pragma unroll
    for (uint32_t q = 0; q < 4; q++) {
        const uint32_t c0 = 144 * q;
        if (c0 >= nch)
            break;
        const uint32_t c1 = min(c0 + 144, nch);
        if (full && (lane >> 4) == q) {
            const uint32_t at = D + L * (lane & 15);
#pragma unroll
            for (uint32_t j = 0; j < L; j++)
                if (j < L - 3 || at + j < 576)
                    of[at + j] = v[j];
        }
        if (full && q > 0 && lane == 16 * q - 1) {
#pragma unroll
            for (uint32_t j = L - 3; j < L; j++)
                if (j >= L - D)
                    of[j - (L - D)] = v[j];
        }
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
        for (uint32_t c = c0 + lane; c < c1; c += 64) {
            const uint4 x = reinterpret_cast<const uint4*>(ob)[c - c0];
            float* g = a0 + 4 * c;
            if (c == 0 && D) {
                if (D <= 1)
                    g[1] = __uint_as_float(x.y);
                if (D <= 2)
                    g[2] = __uint_as_float(x.z);
                g[3] = __uint_as_float(x.w);
            } else {
                const u32x4v y = {x.x, x.y, x.z, x.w};
                __builtin_nontemporal_store(y, reinterpret_cast<u32x4v*>(g));
            }
        }
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
    }
    float* p = base + lane * L;
    if (full && D && !next_full) {
#pragma unroll
        for (uint32_t m = L - 3; m < L; m++)
            if (m >= L - D)
                p[m] = v[m];
    }
    if (lane < nblk && !full) {
#pragma unroll
        for (uint32_t j = 0; j < L; j++)
            if (j < nj)
                p[j] = v[j];
    }
}

// A decompress pass over launch tiles first, first + stride, ... of one
// wave.  Pass 1 writes each tile's sum of deltas to sums[idx]; pass 2 reads
// the tile's carry from sums[idx] and writes its values.  Per step, tile i
// is decoded while tile i+1's bytes move into the other LDS half and tile
// i+2's index entries are on their way.  The lane decoders are the sum's
// (whole blocks below width 16 that the reads covered take the fast one, the
// rest the general one; a lane that does not use the fast decoder runs it at
// width 0, which adds zeros).
template <int Pass, typename Args>
__device__ __forceinline__ void dec_run(const Args& a, WsHeader* ws, uint64_t first, uint64_t stride, uint32_t* obuf)
{
    const uint32_t lane = threadIdx.x & 63;
    uint32_t* sums = dec_sums(ws);
    bool bad_width = false, bad_byte = false, bad_over = false;
    DecPre pc = dec_prefetch<Pass>(a, first, sums);
    if (!pc.ok)
        return;
    __builtin_amdgcn_s_waitcnt(0);
    DecLoc cc = dec_locate(a, pc, obuf, 0, bad_width);
    DecPre pn = dec_prefetch<Pass>(a, first + stride, sums);
    for (uint32_t half = 0;; half ^= 1u) {
        __builtin_amdgcn_s_waitcnt(0);  // tile i's bytes in LDS, tile i+1's index entries in registers
        Block q;
        q.w = cc.w;
        q.o = cc.o;
        q.lim = cc.lim;
        if (cc.staged)
            read_block_staged(obuf + half * (kStageHalf / 4), cc.o, cc.w, q);
        else
            read_block_direct(cc.v, cc.o, cc.w, q);
        __builtin_amdgcn_sched_barrier(0);
        DecLoc cn;
        DecPre pnn;
        pnn.ok = false;
        if (pn.ok) {
            cn = dec_locate(a, pn, obuf, half ^ 1u, bad_width);
            pnn = dec_prefetch<Pass>(a, pn.idx + stride, sums);
        }
        __builtin_amdgcn_sched_barrier(0);
        const bool fast = !cc.has || cc.w == 0 || (cc.nj == L && cc.w < 16 && (cc.staged || cc.o + cc.bs + 3 <= cc.lim));
        const bool use = fast && !cc.skip;
        bad_byte |= use && cc.has && (cc.w || cc.staged || cc.o + 4 <= cc.lim) && (q.h[0] & 0xFFu) != cc.w;
        bad_over |= cc.skip;
        q.w = use ? cc.w : 0u;
        const bool gen = !fast && !cc.skip;
        if constexpr (Pass == 1) {
            // the block's sum only (a lane on the fast path has nj = L or
            // adds zeros; the general decoder adds garbage past nj)
            uint32_t part[2] = {0u, 0u};
            decode_fast<true>(q, part);
            uint32_t b = part[0] + part[1];
            if (__any(gen)) {
                uint32_t acc[L];
#pragma unroll
                for (uint32_t j = 0; j < L; j++)
                    acc[j] = 0;
                if (gen)
                    bad_byte |= decode_general(cc.v, cc.o, cc.w, cc.nj, acc);
#pragma unroll
                for (uint32_t j = 0; j < L; j++)
                    if (j < cc.nj)
                        b += acc[j];
            }
            const uint32_t inc = wave_scan_dpp(b);
            if (lane == 63)
                sums[pc.idx] = inc;
        } else {
            uint32_t acc[L];
#pragma unroll
            for (uint32_t j = 0; j < L; j++)
                acc[j] = 0;
            decode_fast<true>(q, acc);
            if (__any(gen)) {
                if (gen)
                    bad_byte |= decode_general(cc.v, cc.o, cc.w, cc.nj, acc);
            }
            if (cc.nj < L) {
#pragma unroll
                for (uint32_t j = 0; j < L; j++)
                    if (j >= cc.nj)
                        acc[j] = 0;
            }
            DecJob d;
            dec_job(a, pc.idx, d);
            const Geom& g = d.g;
            const Tile x = tile_of(g, d.tl, d.t);
#pragma unroll
            for (uint32_t j = 1; j < L; j++)
                acc[j] += acc[j - 1];
            const uint32_t bsum = acc[L - 1];
            const uint32_t carry = pc.carry + wave_scan_dpp(bsum) - bsum;
            const float eb = d.eb;
            float vals[L];
#pragma unroll
            for (uint32_t jj = 0; jj < L; jj++)
                vals[jj] = dequant(carry + acc[jj], eb);
            // (this tile's LDS half is free: its bytes are in registers)
            store_blocks(d.out + x.e0, vals, cc.nj, x.nblk, obuf + half * (kStageHalf / 4));
            if (lane == 0 && x.lt == 0)
                d.out[x.seg_a] = dequant(segment_seed(g, d.bytes, d.index, x.s, ws), eb);
        }
        if (!pn.ok)
            break;
        pc = pn;
        cc = cn;
        pn = pnn;
    }
    if (__any(bad_width) && lane == 0)
        raise_error(ws, ZK_ERR_WIDTH);
    if (__any(bad_byte) && lane == 0)
        raise_error(ws, ZK_ERR_MISMATCH);
    if (__any(bad_over) && lane == 0)
        raise_error(ws, ZK_ERR_OVERRUN);
}

// Scan of one segment's tile sums into carries (in place): the seed plus the
// sums of the segment's earlier tiles, mod 2^32.  One workgroup.
constexpr uint32_t kDecScanThreads = 1024;

__device__ __forceinline__ void decompress_scan_lane(const Geom& g, const Tiling& tl, const uint8_t* bytes,
    const uint8_t* index, uint32_t s, uint32_t* sums, uint64_t tile0, WsHeader* ws)
{
    __shared__ uint32_t s_w[kDecScanThreads / 64];
    __shared__ uint32_t s_seed;
    const uint32_t tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const uint32_t first = s < g.T ? s * tl.tilesT : g.T * tl.tilesT;
    const uint32_t cnt = s < g.T ? tl.tilesT : tl.tilesR;
    if (tid == 0)
        s_seed = segment_seed(g, bytes, index, s, ws);
    uint32_t* p = sums + tile0 + first;
    const uint32_t per = (cnt + kDecScanThreads - 1) / kDecScanThreads;
    const uint32_t a = min(tid * per, cnt), b = min(a + per, cnt);
    uint32_t run = 0;
    for (uint32_t i = a; i < b; i++)
        run += p[i];
    const uint32_t inc = wave_scan_dpp(run);
    if (lane == 63)
        s_w[wv] = inc;
    __syncthreads();
    uint32_t before = s_seed + inc - run;
    for (uint32_t i = 0; i < wv; i++)
        before += s_w[i];
    for (uint32_t i = a; i < b; i++) {
        const uint32_t v = p[i];
        p[i] = before;
        before += v;
    }
}

// ---------------------------------------------------------------- compress
//
// A workgroup of four waves owns a tile of 256 blocks (make_lane_wg_tiling),
// wave w blocks 64w .. 64w + 63, one block per lane.  Each wave stages its
// blocks' values and the one before them into its LDS buffer (16-byte
// direct-to-LDS loads through a range-checked view of the input; reads past
// the input return zeros), each lane reads its block's 36 values and the one
// before, quantizes them (FP64 product, truncation), takes the deltas, the
// sign words, the magnitudes and the width, and the wave scans the block
// sizes.  The workgroup adds its waves' sizes, publishes the tile's
// aggregate, encodes every block into the wave buffers (the lane encoders of
// the sum), resolves the tile's stream offset by the decoupled lookback
// (one chain per stream, as the slot kernel), and each wave stores its bytes.
// Arithmetic: ZCCL/src/ZCCL_float.c:563,577-606 (quantize, delta, width),
// ZCCL/src/ZCCL_TypeManager.c:12-91, 824-848 (packing, signs).
constexpr uint32_t kCompBuf = 9728;  // a wave's values (<= 2305 floats) or a worst-case 64-block encode + reach
static_assert(kCompBuf >= kLaneTileBlocks * kWorstBlock + 64 && kCompBuf >= 4 * (kLaneTileBlocks * L + 1) + 16,
    "a compress wave buffer holds the staged values and the worst-case encode");
static_assert(kCompBuf % 16 == 0, "compress wave buffers stay 16-byte aligned");

struct CompShared {
    uint32_t wtot[kLaneWaves];
    unsigned long long ex;
};

__device__ __forceinline__ void compress_tile_lane(const Geom& g, const Tiling& tl, double inv, const float* xin,
    uint8_t* out, uint64_t out_cap, uint8_t* oindex, unsigned long long* size_out, uint64_t t,
    unsigned long long* status, WsHeader* ws, uint32_t nchains, uint32_t* obuf, CompShared& sh)
{
    const uint32_t lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const Tile x = tile_of(g, tl, (uint32_t) t);
    const uint32_t jb = kLaneTileBlocks * wv + lane;  // block of the tile
    const bool has = jb < x.nblk;
    const uint32_t nj = has ? min(tile_block_len(g, x, jb), L) : 0u;
    // the wave's values and the one before them (the segment's first value,
    // the seed, before its first block)
    const uint32_t nbw = x.nblk > kLaneTileBlocks * wv ? min(x.nblk - kLaneTileBlocks * wv, kLaneTileBlocks) : 0u;
    const uint32_t nvals = nbw ? (nbw - 1) * L + min(tile_block_len(g, x, kLaneTileBlocks * wv + nbw - 1), L) : 0u;
    const uint32_t v0 = x.e0 + kLaneTileBlocks * L * wv - 1;  // input index of LDS float 0
    if (nbw) {
        // (a view of the wave's values alone: its range check ends them)
        const View vx = make_view((const uint8_t*) (xin + v0), 4ull * (nvals + 1));
        const uint32_t nch = (4 * (nvals + 1) + 15) >> 4;
        for (uint32_t run = 0; run * 64 < nch; run++) {
            const uint32_t ch = run * 64 + lane;
            if (ch < nch)
                __builtin_amdgcn_raw_ptr_buffer_load_lds(vx,
                    (__attribute__((address_space(3))) void*) ((uint8_t*) obuf + run * 1024), 16, 16 * ch, 0, 0, 0);
        }
    }
    __builtin_amdgcn_s_waitcnt(0);
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
    // quantize, deltas, signs and magnitudes (values past nj are zero)
    const float* xv = reinterpret_cast<const float*>(obuf) + lane * L;
    uint32_t m[L];
    uint32_t s = 0, s2 = 0, mx = 0;
    {
        uint32_t prev = (uint32_t) quant(xv[0], inv);
#pragma unroll
        for (uint32_t j = 0; j < L; j++) {
            const uint32_t q = (uint32_t) quant(xv[j + 1], inv);
            const uint32_t d = j < nj ? q - prev : 0u;
            prev = q;
            if (j < 32)
                s = __builtin_amdgcn_alignbit(s, d, 31);  // (s << 1) | sign of value j
            else
                s2 = __builtin_amdgcn_alignbit(s2, d, 31);
            m[j] = mag(d);
            mx = max(mx, m[j]);
        }
    }
    const uint32_t w = has ? dwidth(mx) : 0u;
    const uint32_t bs = has ? (nj == L ? full_block_bytes(w) : lane_block_bytes(w, nj)) : 0u;
    const uint32_t incl = wave_scan_dpp(bs);
    const uint32_t excl = incl - bs;
    const uint32_t wtot = __builtin_amdgcn_readlane(incl, 63);
    if (lane == 0)
        sh.wtot[wv] = wtot;
    __syncthreads();  // every wave's values are read: the buffers are free for the encode
    uint32_t woff = 0, total = 0;
#pragma unroll
    for (uint32_t i = 0; i < kLaneWaves; i++) {
        const uint32_t v = sh.wtot[i];
        woff += i < wv ? v : 0u;
        total += v;
    }
    if (threadIdx.x == 0)
        lookback_publish(status, t, 0, total, 0);  // successors see the aggregate now
    const bool fast = nj == L && w < 16;
    const bool wide = __any(has && w >= 8);
    const uint32_t nz = (wtot + 64 + 15) / 16;
    for (uint32_t i = lane; i < nz; i += 64)
        reinterpret_cast<uint4*>(obuf)[i] = make_uint4(0, 0, 0, 0);
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
    // (two separate branches rather than if/else: the compiler would hoist
    // what the two encoders share above an if/else and keep it live)
    if (has && w && fast)
        encode_fast(obuf, excl, w, s, s2, m, wide);
    if (has && w && !fast)
        encode_general(obuf, excl, w, nj, s, s2, m);
    if (wv == 0) {
        const unsigned long long ex = lookback(status, t, 0, total, 0, ws, true);
        if (lane == 0)
            sh.ex = ex;
    }
    __syncthreads();
    const uint64_t P = 8ull * g.T + 4ull * (x.s + 1) + sh.ex;
    const bool fits = P + total <= out_cap;
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
    if (fits)
        store_wave(obuf, out + P + woff, wtot);
    else if (threadIdx.x == 0)
        raise_error(ws, ZK_ERR_NOSPACE);
    IndexView ox = index_view(g, oindex);
    if (has)
        ox.widths[x.g0 + jb] = (uint8_t) w;
    if (has && (jb % g.IG) == 0)
        ox.group_off[x.grp0 + jb / g.IG] = P + woff + excl;
    if (threadIdx.x == 0) {
        if (x.nblk == 0 && x.ngrp)
            ox.group_off[x.grp0] = P;  // a segment of its seed alone keeps one group entry
        if (x.lt == 0) {
            const uint32_t seed = (uint32_t) quant(xin[x.seg_a], inv);
            const uint64_t so = P - 4;
            if (so + 4 <= out_cap) {
                store_bytes(out + so, &seed, 4);
                if (x.s < g.T) {
                    const unsigned long long h = so - 8ull * g.T;
                    store_bytes(out + 8ull * x.s, &h, 8);
                }
            }
            ox.seed_off[x.s] = so;
        }
        if (t == tl.ntiles - 1) {
            // the last tile ends the stream: if it fits, every tile fits
            const unsigned long long sz = fits ? P + total : 0ull;
            ox.meta[0] = sz;
            if (size_out)
                *size_out = sz;
            chain_finished(ws, nchains);
        }
    }
}

}  // namespace

// One fused sum (kernel arguments hold its geometry and pointers).
__global__ void __launch_bounds__(64 * kLaneWaves, 4) k_homsum_lane(HomsumArgs a)
{
    uint64_t t0;
    unsigned long long* status;
    if (!launch_status(a.ws, a.job.tl.ntiles, t0, status, kLaneWaves))
        return;
    const uint32_t wv = threadIdx.x >> 6;
    const uint64_t t = t0 + wv;
    if (t >= a.job.tl.ntiles)
        return;
    const ArgIO src{a.job.io};
    homsum_tile_lane<kSumWave>(a.job.g, a.job.tl, a.job.p.k, src, a.job.io.out, a.job.io.out_cap, a.job.io.index,
        a.job.io.size_out, t, status, a.ws, 1, zk_lds + wv * (kOutBuf / 4));
}

// Uniform batch: shared geometry, per-job pointers in device memory; tiles of
// the jobs interleave (tile b is tile b / njobs of job b % njobs).
__global__ void __launch_bounds__(64 * kLaneWaves, 5) k_homsum_lane_batch(BatchArgs b)
{
    uint64_t t0;
    unsigned long long* status;
    if (!launch_status(b.ws, b.total_tiles, t0, status, kLaneWaves))
        return;
    const uint32_t wv = threadIdx.x >> 6;
    const uint64_t tg = t0 + wv;
    if (tg >= b.total_tiles)
        return;
    uint32_t jidx, t;
    if (b.interleave) {
        jidx = (uint32_t) tg % b.njobs;
        t = (uint32_t) tg / b.njobs;
    } else {
        jidx = 0;
        t = (uint32_t) tg;
    }
    jidx = __builtin_amdgcn_readfirstlane(jidx);
    t = __builtin_amdgcn_readfirstlane(t);
    const HomsumIO* io = (const HomsumIO*) b.jobs + jidx;
    const SumIO src{io};
    homsum_tile_lane<kSumWave>(b.g, b.tl, b.p.k, src, ldc(&io->out), ldc(&io->out_cap), ldc(&io->index),
        ldc(&io->size_out), t, status + (uint64_t) jidx * b.tl.ntiles * kStatusStride, b.ws, b.total_chains,
        zk_lds + wv * (kOutBuf / 4));
}

// Mixed batch: every job carries its own geometry; job j owns launch tiles
// [job_tile0[j], job_tile0[j+1]).
__global__ void __launch_bounds__(64 * kLaneWaves, 4) k_homsum_lane_mixed(BatchArgs b)
{
    uint64_t t0;
    unsigned long long* status;
    if (!launch_status(b.ws, b.total_tiles, t0, status, kLaneWaves))
        return;
    const uint32_t wv = threadIdx.x >> 6;
    const uint64_t tg = t0 + wv;
    if (tg >= b.total_tiles)
        return;
    uint32_t lo = 0, hi = b.njobs;
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) / 2;
        if (ldc(&b.job_tile0[mid]) <= tg)
            lo = mid;
        else
            hi = mid;
    }
    const uint32_t jidx = __builtin_amdgcn_readfirstlane(lo);
    const uint64_t tile0 = ldc(&b.job_tile0[jidx]);
    const uint32_t t = __builtin_amdgcn_readfirstlane((uint32_t) (tg - tile0));
    const HomsumJob* job = (const HomsumJob*) b.jobs + jidx;
    const Geom g = job->g;
    const Tiling tl = job->tl;
    const SumIO src{&job->io};
    homsum_tile_lane<kSumWave>(g, tl, ldc(&job->p.k), src, ldc(&job->io.out), ldc(&job->io.out_cap), ldc(&job->io.index),
        ldc(&job->io.size_out), t, status + tile0 * kStatusStride, b.ws, b.total_chains, zk_lds + wv * (kOutBuf / 4));
}

// The two passes: a persistent grid, wave w of workgroup b starting at launch
// tile 4b + w and striding by the grid's waves.
template <int Pass, typename Args>
__device__ __forceinline__ void dec_pass(const Args& a, WsHeader* ws)
{
    const uint32_t wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    dec_run<Pass>(a, ws, (uint64_t) blockIdx.x * kLaneWaves + wv, (uint64_t) gridDim.x * kLaneWaves,
        zk_lds + wv * (kOutBuf / 4));
}

// five workgroups per CU: the LDS buffers' limit
constexpr uint32_t kDecGroupsPerCU = 5;
static_assert(kDecGroupsPerCU == (uint32_t) kDecGroupsMax, "the launch bound is the grid's limit");

__global__ void __launch_bounds__(64 * kLaneWaves, kDecGroupsPerCU) k_decompress_lane_sums(DecompressArgs a)
{
    dec_pass<1>(a, a.ws);
}
__global__ void __launch_bounds__(64 * kLaneWaves, kDecGroupsPerCU) k_decompress_lane_out(DecompressArgs a)
{
    dec_pass<2>(a, a.ws);
}
__global__ void __launch_bounds__(64 * kLaneWaves, kDecGroupsPerCU) k_decompress_lane_batch_sums(BatchArgs b)
{
    dec_pass<1>(b, b.ws);
}
__global__ void __launch_bounds__(64 * kLaneWaves, kDecGroupsPerCU) k_decompress_lane_batch_out(BatchArgs b)
{
    dec_pass<2>(b, b.ws);
}

// Compress in the lane layout: one stream, a uniform batch (shared geometry
// and bound, per-job pointers, tiles of the jobs interleaved) or a mixed one.
__global__ void __launch_bounds__(64 * kLaneWaves, 4) k_compress_lane(CompressArgs a)
{
    __shared__ CompShared sh;
    uint64_t t;
    unsigned long long* status;
    if (!launch_status(a.ws, a.job.tl.ntiles, t, status))
        return;
    compress_tile_lane(a.job.g, a.job.tl, a.job.p.inv, a.job.io.x, a.job.io.out, a.job.io.out_cap, a.job.io.index,
        a.job.io.size_out, t, status, a.ws, 1, zk_lds + (threadIdx.x >> 6) * (kCompBuf / 4), sh);
}

__global__ void __launch_bounds__(64 * kLaneWaves, 4) k_compress_lane_batch(BatchArgs b)
{
    __shared__ CompShared sh;
    uint64_t tg;
    unsigned long long* status;
    if (!launch_status(b.ws, b.total_tiles, tg, status))
        return;
    uint32_t jidx, t;
    if (b.interleave) {
        jidx = (uint32_t) tg % b.njobs;
        t = (uint32_t) tg / b.njobs;
    } else {
        jidx = 0;
        t = (uint32_t) tg;
    }
    jidx = __builtin_amdgcn_readfirstlane(jidx);
    t = __builtin_amdgcn_readfirstlane(t);
    const CompressIO* io = (const CompressIO*) b.jobs + jidx;
    compress_tile_lane(b.g, b.tl, b.p.inv, ldc(&io->x), ldc(&io->out), ldc(&io->out_cap), ldc(&io->index),
        ldc(&io->size_out), t, status + (uint64_t) jidx * b.tl.ntiles * kStatusStride, b.ws, b.total_chains,
        zk_lds + (threadIdx.x >> 6) * (kCompBuf / 4), sh);
}

__global__ void __launch_bounds__(64 * kLaneWaves, 4) k_compress_lane_mixed(BatchArgs b)
{
    __shared__ CompShared sh;
    uint64_t tg;
    unsigned long long* status;
    if (!launch_status(b.ws, b.total_tiles, tg, status))
        return;
    uint32_t lo = 0, hi = b.njobs;
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) / 2;
        if (ldc(&b.job_tile0[mid]) <= tg)
            lo = mid;
        else
            hi = mid;
    }
    const uint32_t jidx = __builtin_amdgcn_readfirstlane(lo);
    const uint64_t tile0 = ldc(&b.job_tile0[jidx]);
    const uint32_t t = __builtin_amdgcn_readfirstlane((uint32_t) (tg - tile0));
    const CompressJob* job = (const CompressJob*) b.jobs + jidx;
    const Geom g = job->g;
    const Tiling tl = job->tl;
    compress_tile_lane(g, tl, ldc(&job->p.inv), ldc(&job->io.x), ldc(&job->io.out), ldc(&job->io.out_cap),
        ldc(&job->io.index), ldc(&job->io.size_out), t, status + tile0 * kStatusStride, b.ws, b.total_chains,
        zk_lds + (threadIdx.x >> 6) * (kCompBuf / 4), sh);
}

// Scans: workgroup (s, j) takes segment s of job j.
__global__ void __launch_bounds__(kDecScanThreads) k_decompress_lane_scan(DecompressArgs a)
{
    if (blockIdx.x >= a.job.g.S)
        return;
    decompress_scan_lane(a.job.g, a.job.tl, a.job.io.in[0].bytes, a.job.io.in[0].index, blockIdx.x, dec_sums(a.ws), 0,
        a.ws);
}

__global__ void __launch_bounds__(kDecScanThreads) k_decompress_lane_batch_scan(BatchArgs b)
{
    const uint32_t j = blockIdx.y;
    if (b.uniform) {
        if (blockIdx.x >= b.g.S)
            return;
        const DecompressIO* io = (const DecompressIO*) b.jobs + j;
        decompress_scan_lane(b.g, b.tl, ldc(&io->in[0].bytes), ldc(&io->in[0].index), blockIdx.x, dec_sums(b.ws),
            (uint64_t) j * b.tl.ntiles, b.ws);
    } else {
        const DecompressJob* job = (const DecompressJob*) b.jobs + j;
        const Geom g = job->g;
        if (blockIdx.x >= g.S)
            return;
        const Tiling tl = job->tl;
        decompress_scan_lane(g, tl, ldc(&job->io.in[0].bytes), ldc(&job->io.in[0].index), blockIdx.x, dec_sums(b.ws),
            ldc(&b.job_tile0[j]), b.ws);
    }
}

// Single-pass sum with a lookback tile per workgroup (layout 2): the
// workgroup's tile comes from make_lane_wg_tiling, each wave takes one tile
// of the lane tiling inside it.  The wave's lane tile (its index in t_lane),
// or an empty tile past its segment's blocks.
__device__ __forceinline__ Tile wg_wave_tile(const Geom& g, const Tiling& tl, const Tiling& tw, uint32_t t_wg,
    uint32_t wv, uint32_t& t_lane)
{
    const Tile xw = tile_of(g, tw, t_wg);
    const uint32_t base = xw.s < g.T ? xw.s * tl.tilesT : g.T * tl.tilesT;
    const uint32_t seg_tiles = xw.s < g.T ? tl.tilesT : tl.tilesR;
    const uint32_t lt = kLaneWaves * xw.lt + wv;
    t_lane = base + min(lt, seg_tiles - 1);
    Tile x = tile_of(g, tl, t_lane);
    if (lt >= seg_tiles) {
        x.nblk = 0;
        x.ngrp = 0;
        x.lt = lt;
    }
    return x;
}

__global__ void __launch_bounds__(64 * kLaneWaves, 4) k_homsum_lane_wg(HomsumArgs a)
{
    __shared__ WgShared sh;
    uint64_t t_wg;
    unsigned long long* status;
    if (!launch_status(a.ws, a.job.tl.ntiles, t_wg, status))
        return;
    const uint32_t wv = threadIdx.x >> 6;
    const Tiling tl = make_lane_tiling(a.job.g);
    uint32_t t_lane;
    const Tile x = wg_wave_tile(a.job.g, tl, a.job.tl, (uint32_t) t_wg, wv, t_lane);
    const ArgIO src{a.job.io};
    homsum_tile_lane<kSumWg>(a.job.g, tl, a.job.p.k, src, a.job.io.out, a.job.io.out_cap, a.job.io.index,
        a.job.io.size_out, t_lane, status, a.ws, 1, zk_lds + wv * (kOutBuf / 4), nullptr, 0, 0, t_wg,
        t_wg == a.job.tl.ntiles - 1, &sh, &x);
}

// (four workgroups per CU: at five the workgroup state spills)
__global__ void __launch_bounds__(64 * kLaneWaves, 4) k_homsum_lane_wg_batch(BatchArgs b)
{
    __shared__ WgShared sh;
    uint64_t tg;
    unsigned long long* status;
    if (!launch_status(b.ws, b.total_tiles, tg, status))
        return;
    uint32_t jidx, t;
    if (b.interleave) {
        jidx = (uint32_t) tg % b.njobs;
        t = (uint32_t) tg / b.njobs;
    } else {
        jidx = 0;
        t = (uint32_t) tg;
    }
    jidx = __builtin_amdgcn_readfirstlane(jidx);
    t = __builtin_amdgcn_readfirstlane(t);
    const uint32_t wv = threadIdx.x >> 6;
    const Tiling tl = make_lane_tiling(b.g);
    uint32_t t_lane;
    const Tile x = wg_wave_tile(b.g, tl, b.tl, t, wv, t_lane);
    const HomsumIO* io = (const HomsumIO*) b.jobs + jidx;
    const SumIO src{io};
    homsum_tile_lane<kSumWg>(b.g, tl, b.p.k, src, ldc(&io->out), ldc(&io->out_cap), ldc(&io->index),
        ldc(&io->size_out), t_lane, status + (uint64_t) jidx * b.tl.ntiles * kStatusStride, b.ws, b.total_chains,
        zk_lds + wv * (kOutBuf / 4), nullptr, 0, 0, t, t == b.tl.ntiles - 1, &sh, &x);
}

__global__ void __launch_bounds__(64 * kLaneWaves, 4) k_homsum_lane_wg_mixed(BatchArgs b)
{
    __shared__ WgShared sh;
    uint64_t tg;
    unsigned long long* status;
    if (!launch_status(b.ws, b.total_tiles, tg, status))
        return;
    uint32_t lo = 0, hi = b.njobs;
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) / 2;
        if (ldc(&b.job_tile0[mid]) <= tg)
            lo = mid;
        else
            hi = mid;
    }
    const uint32_t jidx = __builtin_amdgcn_readfirstlane(lo);
    const uint64_t tile0 = ldc(&b.job_tile0[jidx]);
    const uint32_t t = __builtin_amdgcn_readfirstlane((uint32_t) (tg - tile0));
    const HomsumJob* job = (const HomsumJob*) b.jobs + jidx;
    const Geom g = job->g;
    const Tiling tw = job->tl;
    const uint32_t wv = threadIdx.x >> 6;
    const Tiling tl = make_lane_tiling(g);
    uint32_t t_lane;
    const Tile x = wg_wave_tile(g, tl, tw, t, wv, t_lane);
    const SumIO src{&job->io};
    homsum_tile_lane<kSumWg>(g, tl, ldc(&job->p.k), src, ldc(&job->io.out), ldc(&job->io.out_cap),
        ldc(&job->io.index), ldc(&job->io.size_out), t_lane, status + tile0 * kStatusStride, b.ws, b.total_chains,
        zk_lds + wv * (kOutBuf / 4), nullptr, 0, 0, t, t == tw.ntiles - 1, &sh, &x);
}

// ---------------------------------------------------------------- two-pass sum

// The hardware hands workgroup b of a launch to XCD b % 8.  The index this
// returns gives each XCD a contiguous run of workgroups (XCD x: runs of
// about nwg / 8 starting at x * nwg / 8), so tiles that share cache lines at
// their edges, and their index entries, meet in one L2.  A bijection on
// [0, nwg).
__device__ __forceinline__ uint32_t xcd_block(uint32_t b, uint32_t nwg)
{
    const uint32_t q = nwg / 8, rem = nwg % 8, x = b % 8, s = b / 8;
    return x * q + min(x, rem) + s;
}

// Pass one of a single sum: tile t of workgroup-wave t / kLaneWaves, slot t.
__global__ void __launch_bounds__(64 * kLaneWaves, 4) k_homsum_lane_staged(HomsumArgs a, uint8_t* stage)
{
    const uint32_t wv = threadIdx.x >> 6;
    const uint64_t t = (uint64_t) xcd_block(blockIdx.x, gridDim.x) * kLaneWaves + wv;
    if (t >= a.job.tl.ntiles)
        return;
    const ArgIO src{a.job.io};
    homsum_tile_lane<kSumStaged>(a.job.g, a.job.tl, a.job.p.k, src, a.job.io.out, a.job.io.out_cap, a.job.io.index,
        a.job.io.size_out, t, nullptr, a.ws, 1, zk_lds + wv * (kOutBuf / 4), stage, a.job.tl.ntiles, t);
}

// Pass one of a uniform batch: a workgroup takes kLaneWaves consecutive
// tiles of one job (ceil(ntiles / kLaneWaves) workgroups per job, jobs in
// order of the XCD-contiguous index), slot job * ntiles + tile.
__global__ void __launch_bounds__(64 * kLaneWaves, 5) k_homsum_lane_batch_staged(BatchArgs b, uint8_t* stage)
{
    const uint32_t wv = threadIdx.x >> 6;
    const uint32_t wpj = (uint32_t) ceil_div(b.tl.ntiles, kLaneWaves);
    const uint32_t g = xcd_block(blockIdx.x, gridDim.x);
    const uint32_t jidx = __builtin_amdgcn_readfirstlane(g / wpj);
    const uint32_t t = __builtin_amdgcn_readfirstlane((g % wpj) * kLaneWaves + wv);
    if (jidx >= b.njobs || t >= b.tl.ntiles)
        return;
    const HomsumIO* io = (const HomsumIO*) b.jobs + jidx;
    const SumIO src{io};
    homsum_tile_lane<kSumStaged>(b.g, b.tl, b.p.k, src, ldc(&io->out), ldc(&io->out_cap), ldc(&io->index),
        ldc(&io->size_out), t, nullptr, b.ws, 1, zk_lds + wv * (kOutBuf / 4), stage, b.total_tiles,
        (uint64_t) jidx * b.tl.ntiles + t);
}

// Byte j (0..15) of a 16-byte chunk held in registers.
__device__ __forceinline__ uint32_t chunk_byte(const uint4& x, uint32_t j)
{
    const uint32_t w = j < 8 ? (j < 4 ? x.x : x.y) : (j < 12 ? x.z : x.w);
    return (w >> (8 * (j & 3))) & 0xFFu;
}

// Copy n bytes from src (16-byte aligned, a stage slot) to dst, one wave:
// 16-byte stores on dst's 16-byte grid, the partial first and last chunk
// byte by byte from registers, so no byte outside [dst, dst + n) is
// written.  The source is read through a range-checked view, four chunks
// per lane in flight; a partial first chunk reads the slot's first 16 bytes.
__device__ __forceinline__ void copy_wave(uint8_t* dst, const uint8_t* src, uint32_t n)
{
    const uint32_t lane = threadIdx.x & 63;
    const View v = make_view(src, kStageSlot);
    const uintptr_t d0 = (uintptr_t) dst, d1 = d0 + n;
    const uintptr_t c0 = d0 & ~(uintptr_t) 15;
    const uint32_t nch = n ? (uint32_t) ((d1 - c0 + 15) >> 4) : 0u;
    for (uint32_t base = 0; base < nch; base += 256) {
        uint4 x[4];
#pragma unroll
        for (uint32_t u = 0; u < 4; u++) {
            const uint32_t c = base + 64 * u + lane;
            const uintptr_t ca = c0 + 16ull * c;
            x[u] = c < nch ? bl16(v, ca >= d0 ? (uint32_t) (ca - d0) : 0u) : make_uint4(0, 0, 0, 0);
        }
#pragma unroll
        for (uint32_t u = 0; u < 4; u++) {
            const uint32_t c = base + 64 * u + lane;
            const uintptr_t ca = c0 + 16ull * c;
            if (c >= nch)
                continue;
            if (ca >= d0 && ca + 16 <= d1) {
                *(uint4*) ca = x[u];
            } else {
                // x[u] holds source bytes from destination byte sb on (byte
                // indices fixed at compile time: a variable one would put
                // the chunks in memory, which the compiler maps to LDS)
                const uintptr_t sb = ca > d0 ? ca : d0, e = ca + 16 < d1 ? ca + 16 : d1;
#pragma unroll
                for (uint32_t i = 0; i < 16; i++)
                    if (sb + i < e)
                        *(uint8_t*) (sb + i) = (uint8_t) chunk_byte(x[u], i);
            }
        }
    }
}

// Stream offset of every tile of one sum (pass two, first kernel): one
// workgroup of 1024 threads per sum; thread i adds the sizes of tiles
// [8i, 8i + 8) of each round of 8192 tiles, then a workgroup scan.  A tile's
// offset counts the stream header, one seed per segment started and the
// sizes of all earlier tiles.
constexpr uint32_t kScanThreads = 1024;
constexpr uint32_t kScanPer = 8;

__device__ __forceinline__ void scan_sum(const Geom& g, const Tiling& tl, const uint8_t* stage, uint64_t R,
    uint64_t sidx0)
{
    __shared__ unsigned long long s_w[kScanThreads / 64];
    __shared__ unsigned long long s_run;
    const uint32_t tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    unsigned long long* pos = (unsigned long long*) stage;
    const StageRec* recs = reinterpret_cast<const StageRec*>(stage + 8 * R);
    if (tid == 0)
        s_run = 0;
    for (uint32_t base = 0; base < tl.ntiles; base += kScanThreads * kScanPer) {
        uint32_t tot[kScanPer];
        uint64_t sum = 0;
#pragma unroll
        for (uint32_t u = 0; u < kScanPer; u++) {
            const uint32_t t = base + tid * kScanPer + u;
            tot[u] = t < tl.ntiles ? recs[sidx0 + t].total : 0u;
        }
#pragma unroll
        for (uint32_t u = 0; u < kScanPer; u++)
            sum += tot[u];
        // workgroup exclusive scan of the per-thread sums (values < 2^32 per
        // wave: 64 x 8 tiles of at most 9600 bytes)
        const uint32_t incl = wave_scan_dpp((uint32_t) sum);
        if (lane == 63)
            s_w[wv] = incl;
        __syncthreads();
        unsigned long long before = s_run;
        for (uint32_t i = 0; i < wv; i++)
            before += s_w[i];
        uint64_t at = before + (incl - (uint32_t) sum);
#pragma unroll
        for (uint32_t u = 0; u < kScanPer; u++) {
            const uint32_t t = base + tid * kScanPer + u;
            if (t < tl.ntiles) {
                const Tile x = tile_of(g, tl, t);
                pos[sidx0 + t] = 8ull * g.T + 4ull * (x.s + 1) + at;
            }
            at += tot[u];
        }
        __syncthreads();
        if (tid == kScanThreads - 1)
            s_run = at;
        __syncthreads();
    }
}

// One tile's placement (pass two, second kernel), one wave: its bytes, the
// output index group offsets, for a segment's first tile the seed and chunk
// header, for the sum's last tile the stream size.
__device__ __forceinline__ void place_tile(const Geom& g, const Tiling& tl, uint32_t t, uint8_t* out,
    uint64_t out_cap, uint8_t* oindex, unsigned long long* size_out, const uint8_t* stage, uint64_t R,
    uint64_t sidx, WsHeader* ws)
{
    const uint32_t lane = threadIdx.x & 63;
    const StageRec rec = reinterpret_cast<const StageRec*>(stage + 8 * R)[sidx];
    const uint64_t P = ((const unsigned long long*) stage)[sidx];
    const Tile x = tile_of(g, tl, t);
    const bool fits = P + rec.total <= out_cap;
    if (!fits && lane == 0)
        raise_error(ws, ZK_ERR_NOSPACE);
    if (fits)
        copy_wave(out + P, stage + (8 + sizeof(StageRec)) * R + kStageSlot * sidx, rec.total);
    IndexView ox = index_view(g, oindex);
    if (lane < x.ngrp) {
        const uint32_t e = lane == 0 ? 0u : lane == 1 ? rec.e16 : lane == 2 ? rec.e32 : rec.e48;
        ox.group_off[x.grp0 + lane] = P + e;
    }
    if (lane == 0) {
        if (x.lt == 0) {
            const uint64_t so = P - 4;
            if (so + 4 <= out_cap) {
                store_bytes(out + so, &rec.seed, 4);
                if (x.s < g.T) {
                    const unsigned long long h = so - 8ull * g.T;
                    store_bytes(out + 8ull * x.s, &h, 8);
                }
            }
            ox.seed_off[x.s] = so;
        }
        if (t == tl.ntiles - 1) {
            // the last tile ends the stream: if it fits, every tile fits
            const unsigned long long sz = fits ? P + rec.total : 0ull;
            ox.meta[0] = sz;
            if (size_out)
                *size_out = sz;
        }
    }
}

__global__ void __launch_bounds__(kScanThreads) k_scan_lane(HomsumArgs a, const uint8_t* stage)
{
    scan_sum(a.job.g, a.job.tl, stage, a.job.tl.ntiles, 0);
}

__global__ void __launch_bounds__(kScanThreads) k_scan_lane_batch(BatchArgs b, const uint8_t* stage)
{
    scan_sum(b.g, b.tl, stage, b.total_tiles, (uint64_t) blockIdx.x * b.tl.ntiles);
}

__global__ void __launch_bounds__(64 * kLaneWaves) k_place_lane(HomsumArgs a, const uint8_t* stage)
{
    const uint64_t t = (uint64_t) xcd_block(blockIdx.x, gridDim.x) * kLaneWaves + (threadIdx.x >> 6);
    if (t >= a.job.tl.ntiles)
        return;
    place_tile(a.job.g, a.job.tl, (uint32_t) t, a.job.io.out, a.job.io.out_cap, a.job.io.index, a.job.io.size_out,
        stage, a.job.tl.ntiles, t, a.ws);
}

// Uniform batch: workgroups as in pass one.
__global__ void __launch_bounds__(64 * kLaneWaves) k_place_lane_batch(BatchArgs b, const uint8_t* stage)
{
    const uint32_t wpj = (uint32_t) ceil_div(b.tl.ntiles, kLaneWaves);
    const uint32_t g = xcd_block(blockIdx.x, gridDim.x);
    const uint32_t jidx = __builtin_amdgcn_readfirstlane(g / wpj);
    const uint32_t t = __builtin_amdgcn_readfirstlane((g % wpj) * kLaneWaves + (threadIdx.x >> 6));
    if (jidx >= b.njobs || t >= b.tl.ntiles)
        return;
    const HomsumIO* io = (const HomsumIO*) b.jobs + jidx;
    place_tile(b.g, b.tl, t, ldc(&io->out), ldc(&io->out_cap), ldc(&io->index), ldc(&io->size_out), stage,
        b.total_tiles, (uint64_t) jidx * b.tl.ntiles + t, b.ws);
}

size_t lane_lds_bytes() { return (size_t) kLaneWaves * kOutBuf; }

hipError_t launch_homsum_lane(const HomsumArgs& a, hipStream_t s)
{
    const uint32_t groups = (uint32_t) ceil_div(a.job.tl.ntiles, kLaneWaves);
    hipLaunchKernelGGL(k_homsum_lane, dim3(groups), dim3(64 * kLaneWaves), lane_lds_bytes(), s, a);
    return hipGetLastError();
}

hipError_t launch_homsum_lane_batch(const BatchArgs& b, hipStream_t s)
{
    const uint32_t groups = (uint32_t) ceil_div(b.total_tiles, kLaneWaves);
    if (b.uniform)
        hipLaunchKernelGGL(k_homsum_lane_batch, dim3(groups), dim3(64 * kLaneWaves), lane_lds_bytes(), s, b);
    else
        hipLaunchKernelGGL(k_homsum_lane_mixed, dim3(groups), dim3(64 * kLaneWaves), lane_lds_bytes(), s, b);
    return hipGetLastError();
}

// (four measured fastest on the 8 x 2^24 batch, the grid's limit being five)
constexpr int kDecGroupsDefault = 4;
static std::atomic<int> g_dec_groups{kDecGroupsDefault};

bool set_decompress_groups_per_cu(int groups)
{
    if (groups == 0)
        groups = kDecGroupsDefault;
    if (groups < 1 || groups > kDecGroupsMax)
        return false;
    g_dec_groups.store(groups);
    return true;
}

int get_decompress_groups_per_cu() { return g_dec_groups.load(); }

uint32_t device_cu_count()
{
    static std::atomic<int> cus[64];
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64)
        dev = 0;
    int n = cus[dev].load();
    if (n <= 0) {
        if (hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || n <= 0)
            n = 256;
        cus[dev].store(n);
    }
    return (uint32_t) n;
}

// Workgroups of a persistent decompress pass: the tuned number per CU (at
// most as many as the device holds at once), or one per four tiles when that
// is fewer.
static uint32_t dec_groups(uint64_t tiles)
{
    return (uint32_t) std::min<uint64_t>(ceil_div(tiles, kLaneWaves), (uint64_t) device_cu_count() * g_dec_groups.load());
}

hipError_t launch_decompress_lane(const DecompressArgs& a, hipStream_t s)
{
    const dim3 grid(dec_groups(a.job.tl.ntiles)), block(64 * kLaneWaves);
    hipLaunchKernelGGL(k_decompress_lane_sums, grid, block, lane_lds_bytes(), s, a);
    hipLaunchKernelGGL(k_decompress_lane_scan, dim3(a.job.g.S), dim3(kDecScanThreads), 0, s, a);
    hipLaunchKernelGGL(k_decompress_lane_out, grid, block, lane_lds_bytes(), s, a);
    return hipGetLastError();
}

hipError_t launch_decompress_lane_batch(const BatchArgs& b, hipStream_t s)
{
    const dim3 grid(dec_groups(b.total_tiles)), block(64 * kLaneWaves);
    hipLaunchKernelGGL(k_decompress_lane_batch_sums, grid, block, lane_lds_bytes(), s, b);
    hipLaunchKernelGGL(k_decompress_lane_batch_scan, dim3(b.max_segments, b.njobs), dim3(kDecScanThreads), 0, s, b);
    hipLaunchKernelGGL(k_decompress_lane_batch_out, grid, block, lane_lds_bytes(), s, b);
    return hipGetLastError();
}

hipError_t launch_homsum_lane_wg(const HomsumArgs& a, hipStream_t s)
{
    hipLaunchKernelGGL(k_homsum_lane_wg, dim3((uint32_t) a.job.tl.ntiles), dim3(64 * kLaneWaves), lane_lds_bytes(),
        s, a);
    return hipGetLastError();
}

hipError_t launch_homsum_lane_wg_batch(const BatchArgs& b, hipStream_t s)
{
    if (b.uniform)
        hipLaunchKernelGGL(k_homsum_lane_wg_batch, dim3((uint32_t) b.total_tiles), dim3(64 * kLaneWaves),
            lane_lds_bytes(), s, b);
    else
        hipLaunchKernelGGL(k_homsum_lane_wg_mixed, dim3((uint32_t) b.total_tiles), dim3(64 * kLaneWaves),
            lane_lds_bytes(), s, b);
    return hipGetLastError();
}

size_t compress_lane_lds_bytes() { return (size_t) kLaneWaves * kCompBuf; }

hipError_t launch_compress_lane(const CompressArgs& a, hipStream_t s)
{
    hipLaunchKernelGGL(k_compress_lane, dim3((uint32_t) a.job.tl.ntiles), dim3(64 * kLaneWaves),
        compress_lane_lds_bytes(), s, a);
    return hipGetLastError();
}

hipError_t launch_compress_lane_batch(const BatchArgs& b, hipStream_t s)
{
    if (b.uniform)
        hipLaunchKernelGGL(k_compress_lane_batch, dim3((uint32_t) b.total_tiles), dim3(64 * kLaneWaves),
            compress_lane_lds_bytes(), s, b);
    else
        hipLaunchKernelGGL(k_compress_lane_mixed, dim3((uint32_t) b.total_tiles), dim3(64 * kLaneWaves),
            compress_lane_lds_bytes(), s, b);
    return hipGetLastError();
}

hipError_t launch_homsum_lane_staged(const HomsumArgs& a, uint8_t* stage, hipStream_t s)
{
    const uint64_t tiles = a.job.tl.ntiles;
    hipLaunchKernelGGL(k_homsum_lane_staged, dim3((uint32_t) ceil_div(tiles, kLaneWaves)), dim3(64 * kLaneWaves),
        lane_lds_bytes(), s, a, stage);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess)
        return e;
    hipLaunchKernelGGL(k_scan_lane, dim3(1), dim3(kScanThreads), 0, s, a, (const uint8_t*) stage);
    hipLaunchKernelGGL(k_place_lane, dim3((uint32_t) ceil_div(tiles, kLaneWaves)), dim3(64 * kLaneWaves), 0, s, a,
        (const uint8_t*) stage);
    return hipGetLastError();
}

hipError_t launch_homsum_lane_batch_staged(const BatchArgs& b, uint8_t* stage, hipStream_t s)
{
    if (!b.uniform)
        return hipErrorInvalidValue;
    const uint32_t groups = (uint32_t) (ceil_div(b.tl.ntiles, kLaneWaves) * b.njobs);
    hipLaunchKernelGGL(k_homsum_lane_batch_staged, dim3(groups), dim3(64 * kLaneWaves), lane_lds_bytes(), s, b, stage);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess)
        return e;
    hipLaunchKernelGGL(k_scan_lane_batch, dim3(b.njobs), dim3(kScanThreads), 0, s, b, (const uint8_t*) stage);
    hipLaunchKernelGGL(k_place_lane_batch, dim3(groups), dim3(64 * kLaneWaves), 0, s, b, (const uint8_t*) stage);
    return hipGetLastError();
}

}  // namespace zk
