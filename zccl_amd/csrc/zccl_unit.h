// zccl_unit.h -- device-side decoder of one 4-value unit of an encoded block,
// shared by the 1-D kernels (zccl_kernels.hip) and the 3-D codec
// (zccl_sz3d.hip).  Reads are from LDS (any byte alignment).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace zk {

// LDS reads at any byte address, built from 4-byte aligned dword reads and
// byte shifts (v_alignbyte).  A ds_read at a misaligned address is legal on
// gfx950 but slow: the fused sum ran 1.47x faster with the decoder's reads
// aligned this way.  The byte shift is the address itself: v_alignbyte uses
// only bits 1:0 of it, so no mask is needed (0.9 % faster launches).  Each
// reads up to 4 bytes past the last byte it returns.
__device__ __forceinline__ uint32_t lds_ld4(const uint8_t* p)
{
    const uint32_t a = (uint32_t) (uintptr_t) p;  // v_alignbyte reads bits 1:0 only
    const uint32_t* q = (const uint32_t*) __builtin_align_down(p, 4);
    return __builtin_amdgcn_alignbyte(q[1], q[0], a);
}

__device__ __forceinline__ uint2 lds_ld8(const uint8_t* p)
{
    const uint32_t a = (uint32_t) (uintptr_t) p;  // v_alignbyte reads bits 1:0 only
    const uint32_t* q = (const uint32_t*) __builtin_align_down(p, 4);
    const uint32_t d0 = q[0], d1 = q[1], d2 = q[2];
    return make_uint2(__builtin_amdgcn_alignbyte(d1, d0, a), __builtin_amdgcn_alignbyte(d2, d1, a));
}

__device__ __forceinline__ uint4 lds_ld16(const uint8_t* p)
{
    const uint32_t a = (uint32_t) (uintptr_t) p;  // v_alignbyte reads bits 1:0 only
    const uint32_t* q = (const uint32_t*) __builtin_align_down(p, 4);
    const uint32_t d0 = q[0], d1 = q[1], d2 = q[2], d3 = q[3], d4 = q[4];
    return make_uint4(__builtin_amdgcn_alignbyte(d1, d0, a), __builtin_amdgcn_alignbyte(d2, d1, a),
        __builtin_amdgcn_alignbyte(d3, d2, a), __builtin_amdgcn_alignbyte(d4, d3, a));
}

// Add the up-to-4 deltas of unit u of a block (width byte at bp, width w,
// nj elements, sb = ceil(nj/8) sign bytes) to acc[].  Magnitude = (high
// bytes << r) | low r bits; the sign bit s (MSB-first bitmap) is applied as
// (m ^ -s) + s.  The sign byte, the unit's first 4 high bytes and the dword
// holding its low bits are read together (one LDS round trip below width 16);
// a zero width adds nothing.  Reads reach at most 20 bytes past the block's
// last byte.  ZCCL/src/ZCCL_TypeManager.c:423-517 (extract) and :906-950
// (sign bitmap).
__device__ __forceinline__ void accumulate_unit(const uint8_t* bp, uint32_t w, uint32_t nj, uint32_t sb,
    uint32_t u, uint32_t acc[4])
{
    const uint32_t B = w >> 3, r = w & 7;
    const uint8_t* hp = bp + 1 + sb;
    const uint8_t* hu = hp + 4 * u * B;
    const uint32_t bitpos = 4 * u * r;  // a multiple of 4
    const uint32_t sbyte = bp[1 + (u >> 1)];
    const uint32_t H0 = lds_ld4(hu);
    const uint32_t Xr = lds_ld4(hp + B * nj + (bitpos >> 3));
    const uint32_t nib = w ? __builtin_amdgcn_ubfe(sbyte, (~u & 1u) << 2, 4) : 0u;
    uint32_t m[4];
    if (B <= 1) {
        // widths below 16, the common case; B == 0 masks the bytes to nothing
#pragma unroll
        for (int c = 0; c < 4; c++)
            m[c] = __builtin_amdgcn_ubfe(H0, 8 * c, 8 * B);
    } else if (B == 2) {
        const uint2 H = lds_ld8(hu);
        m[0] = H.x & 0xFFFFu;
        m[1] = H.x >> 16;
        m[2] = H.y & 0xFFFFu;
        m[3] = H.y >> 16;
    } else {
        const uint4 H = lds_ld16(hu);
        if (B == 3) {
            m[0] = H.x & 0xFFFFFFu;
            m[1] = __builtin_amdgcn_alignbyte(H.y, H.x, 3) & 0xFFFFFFu;
            m[2] = __builtin_amdgcn_alignbyte(H.z, H.y, 2) & 0xFFFFFFu;
            m[3] = H.z >> 8;
        } else {
            m[0] = H.x; m[1] = H.y; m[2] = H.z; m[3] = H.w;
        }
    }
    // low bits for every width: the unit's fields start at bit (bitpos & 7)
    // of the big-endian dword; shifted to the top of lo, one 64-bit shift of
    // {m[c], lo} by r gives value c's magnitude (high dword) and the next
    // window (low dword), r == 0 included
    uint32_t lo = __builtin_bswap32(Xr) << (bitpos & 4u);
#pragma unroll
    for (int c = 0; c < 4; c++) {
        const uint64_t v = ((uint64_t) m[c] << 32 | lo) << r;
        const uint32_t mc = (uint32_t) (v >> 32);
        lo = (uint32_t) v;
        const uint32_t s = (uint32_t) __builtin_amdgcn_sbfe((int32_t) nib, 3 - c, 1);  // 0 or ~0
        acc[c] += (mc ^ s) - s;
    }
}

}  // namespace zk
