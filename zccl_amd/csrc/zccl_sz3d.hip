// zccl_sz3d.hip -- CDNA4 (gfx950) kernels of the 3-D SZp codec
// (SZp_compress3D_fast_openmp, ZCCL/src/ZCCL_float.c:1862-2048, and
// SZp_decompress3D_fast_openmp, ZCCL/src/ZCCLd_float.c:2873-3040).
//
// The reference walks each sub-volume with one CPU thread and running
// buffers; its residual is the exact 3-D difference dx dy dz q (zero outside
// the sub-volume) and its decoder the exact 3-D prefix sum, so both are done
// here with full parallelism:
//   compress   k3_table (sub-volume table) -> k3_widths (one wave per block:
//              residuals from x, block max, width, payload size) -> scan of the
//              payload sizes -> k3_encode (one wave per block: width header
//              byte, sub-volume offset, sign bitmap and fixed-length bits with
//              the one-writer-per-byte unit encoder of the 1-D sum)
//   decompress k3_table -> k3_dsizes (widths read from the stream's own
//              header) -> scan -> k3_decode (residuals into the output buffer
//              as int32) -> k3_zpass / k3_ypass (inclusive prefix sums along z
//              and y) -> k3_xpass (along x, then x' = (float)(q * 2 eb) in place)
// Arithmetic: q = (int)floor((double)x * (0.5 / eb) + 0.5) with the product and
// the sum rounded separately (the reference is built without FMA; this file is
// built with -ffp-contract=off); residuals and sums wrap in 32 bits.
#include <hip/hip_runtime.h>

#include "zccl_internal.h"
#include "zccl_kernels.h"
#include "zccl_sz3d.h"
#include "zccl_unit.h"

namespace zk {

namespace {

__device__ __forceinline__ int32_t quant3(float v, double inv)
{
    const double p = floor(__dadd_rn(__dmul_rn((double) v, inv), 0.5));
    return (p >= -2147483648.0 && p < 2147483648.0) ? (int32_t) p : INT32_MIN;
}

__device__ __forceinline__ void raise3(unsigned long long* status, unsigned long long bit)
{
    atomicOr(&status[1], bit);
}

// Block of a sub-volume: position, extents and value count.
struct Blk3 {
    uint32_t bx, by, bz;
    uint32_t ex, ey, ez;
    uint32_t cnt;
};

__device__ __forceinline__ Blk3 block_of(const Part3& p, uint32_t lb, uint32_t Bs)
{
    Blk3 b;
    const uint32_t yz = p.b2 * p.b3;
    b.bx = lb / yz;
    const uint32_t rest = lb - b.bx * yz;
    b.by = rest / p.b3;
    b.bz = rest - b.by * p.b3;
    b.ex = min(Bs, p.sx - b.bx * Bs);
    b.ey = min(Bs, p.sy - b.by * Bs);
    b.ez = min(Bs, p.sz - b.bz * Bs);
    b.cnt = b.ex * b.ey * b.ez;
    return b;
}

// Residual of the value at local (lx, ly, lz) of sub-volume p: dx dy dz q with
// q = 0 outside the sub-volume (the reference's err_1 / err_2 / err_3 chain,
// ZCCL/src/ZCCL_float.c:1983-1990).
__device__ __forceinline__ uint32_t resid3(const float* x, const Geom3& g, const Part3& p, double inv,
    uint32_t lx, uint32_t ly, uint32_t lz)
{
    const uint64_t sx = (uint64_t) g.d2 * g.d3, sy = g.d3;
    const uint64_t i0 = ((uint64_t) (p.ox + lx) * g.d2 + (p.oy + ly)) * g.d3 + p.oz + lz;
    uint32_t acc = (uint32_t) quant3(x[i0], inv);
    if (lz)
        acc -= (uint32_t) quant3(x[i0 - 1], inv);
    if (ly) {
        acc -= (uint32_t) quant3(x[i0 - sy], inv);
        if (lz)
            acc += (uint32_t) quant3(x[i0 - sy - 1], inv);
    }
    if (lx) {
        acc -= (uint32_t) quant3(x[i0 - sx], inv);
        if (lz)
            acc += (uint32_t) quant3(x[i0 - sx - 1], inv);
        if (ly) {
            acc += (uint32_t) quant3(x[i0 - sx - sy], inv);
            if (lz)
                acc -= (uint32_t) quant3(x[i0 - sx - sy - 1], inv);
        }
    }
    return acc;
}

// Local coordinates of value e of block b (values in i, j, k order).
__device__ __forceinline__ void value_pos(const Blk3& b, uint32_t Bs, uint32_t e, uint32_t& lx, uint32_t& ly,
    uint32_t& lz)
{
    const uint32_t yz = b.ey * b.ez;
    const uint32_t i = e / yz, rest = e - i * yz, j = rest / b.ez, k = rest - j * b.ez;
    lx = b.bx * Bs + i;
    ly = b.by * Bs + j;
    lz = b.bz * Bs + k;
}

__device__ __forceinline__ uint32_t wave_max(uint32_t v)
{
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1)
        v = max(v, (uint32_t) __shfl_xor(v, d, 64));
    return v;
}

__device__ __forceinline__ unsigned long long wave_incl_scan64(unsigned long long v)
{
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const unsigned long long y = __shfl_up(v, d, 64);
        if (lane >= d)
            v += y;
    }
    return v;
}

__device__ __forceinline__ uint32_t wave_incl_scan32(uint32_t v)
{
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const uint32_t y = __shfl_up(v, d, 64);
        if (lane >= d)
            v += y;
    }
    return v;
}

constexpr uint32_t kSizeMask = (1u << 26) - 1;

// Sub-volume table, built on the device by one thread (no host copy on the
// stream the caller passes), and the status words cleared.
__global__ void k3_table(uint32_t d1, uint32_t d2, uint32_t d3, int Bs, int T, Part3* parts,
    unsigned long long* status)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        Geom3 g;
        make_geom3(d1, d2, d3, Bs, T, &g, parts);
        status[0] = 0;
        status[1] = 0;
    }
}

// Width and payload size of every block: one wave per block.
__global__ void __launch_bounds__(256) k3_widths(Sz3Args a)
{
    const uint32_t lane = threadIdx.x & 63;
    const uint32_t gb = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (gb >= a.g.nblocks)
        return;
    const uint32_t t = __builtin_amdgcn_readfirstlane(part_of_block(a.parts, a.g.T, gb));
    const Part3 p = a.parts[t];
    const Blk3 b = block_of(p, gb - p.g0, a.g.Bs);
    uint32_t m = 0;
    for (uint32_t e = lane; e < b.cnt; e += 64) {
        uint32_t lx, ly, lz;
        value_pos(b, a.g.Bs, e, lx, ly, lz);
        m = max(m, mag32(resid3(a.x, a.g, p, a.inv, lx, ly, lz)));
    }
    m = wave_max(m);
    if (lane == 0) {
        const uint32_t w = m ? 32u - (uint32_t) __clz((int) m) : 0u;
        a.words[gb] = payload3(w, b.cnt) | (w << 26);
    }
}

// Payload sizes of a stream's blocks from its own header: one thread per block.
__global__ void __launch_bounds__(256) k3_dsizes(Sz3Args a)
{
    const uint32_t gb = blockIdx.x * blockDim.x + threadIdx.x;
    if (gb >= a.g.nblocks)
        return;
    const uint32_t t = part_of_block(a.parts, a.g.T, gb);
    const Part3 p = a.parts[t];
    const Blk3 b = block_of(p, gb - p.g0, a.g.Bs);
    uint64_t h;
    __builtin_memcpy(&h, a.in + 8ull * t, 8);
    const uint64_t wpos = 8ull * a.g.T + h + (gb - p.g0);
    uint32_t w = 0;
    if (wpos < a.in_size) {
        w = a.in[wpos];
        if (w > 32) {
            raise3(a.status, ZK_ERR_WIDTH);
            w = 0;
        }
    } else {
        raise3(a.status, ZK_ERR_OVERRUN);
    }
    a.words[gb] = payload3(w, b.cnt) | (w << 26);
}

// Exclusive scan of the payload sizes (low 26 bits of words[]) into offs[]:
// group sums, a scan of the group sums by one workgroup, then each group's
// local scan.  256 threads, kScan3Group = 1024 entries per group.
__global__ void __launch_bounds__(256) k3_scan_groups(Sz3Args a)
{
    __shared__ unsigned long long s_w[4];
    const uint32_t base = blockIdx.x * kScan3Group;
    unsigned long long v = 0;
    for (uint32_t i = threadIdx.x; i < kScan3Group; i += 256)
        if (base + i < a.g.nblocks)
            v += a.words[base + i] & kSizeMask;
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1)
        v += __shfl_xor(v, d, 64);
    if ((threadIdx.x & 63) == 0)
        s_w[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0)
        a.part_sums[blockIdx.x] = s_w[0] + s_w[1] + s_w[2] + s_w[3];
}

// One workgroup: exclusive scan of the group sums in place; the stream size
// (header, width bytes, payloads) goes to status[0] and the optional size
// word; a compress whose stream exceeds the output capacity raises NOSPACE.
__global__ void __launch_bounds__(256) k3_scan_top(Sz3Args a, uint32_t ngroups, int check_cap)
{
    __shared__ unsigned long long s_w[4];
    __shared__ unsigned long long s_carry;
    if (threadIdx.x == 0)
        s_carry = 0;
    __syncthreads();
    for (uint32_t c0 = 0; c0 < ngroups; c0 += 256) {
        const uint32_t i = c0 + threadIdx.x;
        const unsigned long long v = i < ngroups ? a.part_sums[i] : 0ull;
        const unsigned long long inc = wave_incl_scan64(v);
        if ((threadIdx.x & 63) == 63)
            s_w[threadIdx.x >> 6] = inc;
        __syncthreads();
        unsigned long long pre = s_carry;
        for (uint32_t w = 0; w < (threadIdx.x >> 6); w++)
            pre += s_w[w];
        if (i < ngroups)
            a.part_sums[i] = pre + inc - v;
        __syncthreads();
        if (threadIdx.x == 255)
            s_carry = pre + inc;
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        const unsigned long long total = 8ull * a.g.T + a.g.nblocks + s_carry;
        a.status[0] = total;
        a.part_sums[ngroups] = s_carry;
        if (check_cap && total > a.out_cap)
            raise3(a.status, ZK_ERR_NOSPACE);
        if (a.size_out)
            *a.size_out = (check_cap && total > a.out_cap) ? 0ull : total;
    }
}

__global__ void __launch_bounds__(256) k3_scan_apply(Sz3Args a)
{
    __shared__ unsigned long long s_w[4];
    const uint32_t base = blockIdx.x * kScan3Group + threadIdx.x * 4;
    uint32_t v[4], run = 0;
#pragma unroll
    for (int c = 0; c < 4; c++) {
        v[c] = base + c < a.g.nblocks ? a.words[base + c] & kSizeMask : 0u;
        run += v[c];
    }
    const unsigned long long inc = wave_incl_scan64(run);
    if ((threadIdx.x & 63) == 63)
        s_w[threadIdx.x >> 6] = inc;
    __syncthreads();
    unsigned long long pre = a.part_sums[blockIdx.x] + inc - run;
    for (uint32_t w = 0; w < (threadIdx.x >> 6); w++)
        pre += s_w[w];
#pragma unroll
    for (int c = 0; c < 4; c++) {
        if (base + c < a.g.nblocks)
            a.offs[base + c] = pre;
        pre += v[c];
    }
}

// LDS bytes per wave for one block's payload: blocks whose payload fits are
// encoded (decoded) in LDS and moved with dword copies; larger ones work on
// global memory byte by byte.  Side 9 (729 values, at most 3008 bytes) fits.
constexpr uint32_t kStage3 = 4096;

// Make a wave's LDS writes visible to its other lanes.
__device__ __forceinline__ void wave_sync_lds()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Header byte, sub-volume offset and payload of every block: one wave per
// block, 4-value units in consecutive lanes (so unit pairs share a wave).
__global__ void __launch_bounds__(256) k3_encode(Sz3Args a)
{
    __shared__ uint32_t s_stage[4][kStage3 / 4];
    __shared__ uint32_t s_res[4][256];
    const uint32_t lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const uint32_t gb = blockIdx.x * 4 + wv;
    if (gb >= a.g.nblocks || (a.status[1] & ZK_ERR_NOSPACE))
        return;
    const uint32_t t = __builtin_amdgcn_readfirstlane(part_of_block(a.parts, a.g.T, gb));
    const Part3 p = a.parts[t];
    const uint32_t lb = gb - p.g0, Bs = a.g.Bs;
    const Blk3 b = block_of(p, lb, Bs);
    const uint32_t word = a.words[gb], w = word >> 26, size = word & kSizeMask;
    const uint64_t sub = (uint64_t) p.nbc + a.offs[p.g0];  // sub-volume start after the header
    const uint64_t hdr = 8ull * a.g.T;
    if (lane == 0) {
        a.out[hdr + sub + lb] = (uint8_t) w;
        if (lb == 0)
            __builtin_memcpy(a.out + 8ull * t, &sub, 8);
    }
    if (!w)
        return;
    const uint64_t fo = hdr + p.nbc + (uint64_t) p.b1 * p.b2 * p.b3 + a.offs[gb];
    const bool staged = size <= kStage3;
    uint8_t* st = (uint8_t*) s_stage[wv];
    const uint32_t U = (b.cnt + 3) / 4;
    uint32_t* res = s_res[wv];
    for (uint32_t u0 = 0; u0 < U; u0 += 64) {
        // residuals of the next 256 values, one per lane per step, then each
        // unit lane takes its four
#pragma unroll
        for (uint32_t c = 0; c < 4; c++) {
            const uint32_t e = 4 * u0 + 64 * c + lane;
            uint32_t r = 0;
            if (e < b.cnt) {
                uint32_t lx, ly, lz;
                value_pos(b, Bs, e, lx, ly, lz);
                r = resid3(a.x, a.g, p, a.inv, lx, ly, lz);
            }
            res[64 * c + lane] = r;
        }
        wave_sync_lds();
        const uint32_t u = u0 + lane;
        const uint32_t cnt = u < U ? min(4u, b.cnt - 4 * u) : 0u;
        uint32_t d[4];
#pragma unroll
        for (uint32_t c = 0; c < 4; c++)
            d[c] = c < cnt ? res[4 * lane + c] : 0u;
        wave_sync_lds();
        const UnitFields f = unit_fields(w, d);
        const uint32_t pnib = __shfl_down(f.nib, 1, 64);
        const uint32_t plv = __shfl_down(f.lv, 1, 64);
        if (staged)
            emit_unit(st, 0, w, b.cnt, u, U, cnt, f, pnib, plv);
        else
            emit_unit(a.out, fo, w, b.cnt, u, U, cnt, f, pnib, plv);
    }
    if (staged) {
        // every payload byte has one writer; dword stores cover only whole
        // dwords of this payload, the tail goes out byte by byte
        wave_sync_lds();
        uint8_t* g = a.out + fo;
        const uint32_t nw = size / 4;
        for (uint32_t i = lane; i < nw; i += 64) {
            const uint32_t v = ((const uint32_t*) st)[i];
            __builtin_memcpy(g + 4 * i, &v, 4);
        }
        for (uint32_t i = 4 * nw + lane; i < size; i += 64)
            g[i] = st[i];
    }
}

// Byte j of a payload (0 past its end, which no well-formed block reads).
__device__ __forceinline__ uint32_t pbyte(const uint8_t* p, uint32_t size, uint32_t j)
{
    return j < size ? p[j] : 0u;
}

// Residuals of every block into the output buffer as int32 at their positions
// in the volume: one wave per block, one 4-value unit per lane.  A payload of
// at most kStage3 bytes is copied into LDS (dword loads, 16 zero bytes of
// slack) and decoded with the 1-D unit decoder; a larger one is read byte by
// byte from global memory.
// ZCCL/src/ZCCL_TypeManager.c:423-517 (fixed-length bits), :906-950 (signs).
__global__ void __launch_bounds__(256) k3_decode(Sz3Args a)
{
    __shared__ uint32_t s_stage[4][kStage3 / 4 + 8];
    __shared__ uint32_t s_res[4][256];
    const uint32_t lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const uint32_t gb = blockIdx.x * 4 + wv;
    if (gb >= a.g.nblocks)
        return;
    const uint32_t t = __builtin_amdgcn_readfirstlane(part_of_block(a.parts, a.g.T, gb));
    const Part3 p = a.parts[t];
    const uint32_t lb = gb - p.g0, Bs = a.g.Bs;
    const Blk3 b = block_of(p, lb, Bs);
    const uint32_t word = a.words[gb], w = word >> 26, size = word & kSizeMask;
    uint64_t h;
    __builtin_memcpy(&h, a.in + 8ull * t, 8);
    const uint64_t fo = 8ull * a.g.T + h + (uint64_t) p.b1 * p.b2 * p.b3 + (a.offs[gb] - a.offs[p.g0]);
    const bool inside = fo + size <= a.in_size;
    if (w && !inside && lane == 0)
        raise3(a.status, ZK_ERR_OVERRUN);
    const uint8_t* pl = a.in + fo;
    const uint32_t B = w >> 3, r = w & 7, sb = (b.cnt + 7) / 8;
    const bool staged = w && inside && size <= kStage3;
    uint8_t* st = (uint8_t*) s_stage[wv];
    if (staged) {
        const uint32_t nw = size / 4;
        for (uint32_t i = lane; i < nw; i += 64) {
            uint32_t v;
            __builtin_memcpy(&v, pl + 4 * i, 4);
            ((uint32_t*) st)[i] = v;
        }
        for (uint32_t i = 4 * nw + lane; i < size; i += 64)
            st[i] = pl[i];
        if (lane < 16)
            st[size + lane] = 0;
        wave_sync_lds();
    }
    int32_t* q = (int32_t*) a.y;
    uint32_t* res = s_res[wv];
    const uint32_t U = (b.cnt + 3) / 4;
    for (uint32_t u0 = 0; u0 < U; u0 += 64) {
        const uint32_t u = u0 + lane;
        const uint32_t cnt = u < U ? min(4u, b.cnt - 4 * u) : 0u;
        uint32_t v[4] = {0, 0, 0, 0};
        if (!cnt) {
        } else if (staged) {
            accumulate_unit(st - 1, w, b.cnt, sb, u, v);  // fields start right after the (absent) width byte
        } else if (w && inside) {
            for (uint32_t c = 0; c < cnt; c++) {
                const uint32_t e = 4 * u + c;
                uint32_t m = 0;
                for (uint32_t k = 0; k < B; k++)
                    m |= pbyte(pl, size, sb + e * B + k) << (8 * k);
                uint32_t lo = 0;
                for (uint32_t bit = e * r, s2 = 0; s2 < r; s2++, bit++)
                    lo = (lo << 1) | ((pbyte(pl, size, sb + B * b.cnt + (bit >> 3)) >> (7 - (bit & 7))) & 1u);
                m = (m << r) | lo;
                v[c] = ((pbyte(pl, size, e >> 3) >> (7 - (e & 7))) & 1u) ? 0u - m : m;
            }
        }
        // the unit lanes' values go through LDS so all lanes write them out,
        // one value each per step
#pragma unroll
        for (uint32_t c = 0; c < 4; c++)
            res[4 * lane + c] = v[c];
        wave_sync_lds();
#pragma unroll
        for (uint32_t c = 0; c < 4; c++) {
            const uint32_t e = 4 * u0 + 64 * c + lane;
            if (e < b.cnt) {
                uint32_t lx, ly, lz;
                value_pos(b, Bs, e, lx, ly, lz);
                q[((uint64_t) (p.ox + lx) * a.g.d2 + (p.oy + ly)) * a.g.d3 + p.oz + lz] = (int32_t) res[64 * c + lane];
            }
        }
        wave_sync_lds();
    }
}

// Inclusive prefix sums along z: one wave per (sub-volume, x, y) row.
__global__ void __launch_bounds__(256) k3_zpass(Sz3Args a)
{
    const uint32_t lane = threadIdx.x & 63;
    const uint32_t row = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (row >= a.g.zrows)
        return;
    uint32_t lo = 0, hi = a.g.T;
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) / 2;
        if (a.parts[mid].zr0 <= row)
            lo = mid;
        else
            hi = mid;
    }
    const Part3 p = a.parts[__builtin_amdgcn_readfirstlane(lo)];
    const uint32_t r = row - p.zr0, lx = r / p.sy, ly = r - lx * p.sy;
    uint32_t* q = (uint32_t*) a.y + ((uint64_t) (p.ox + lx) * a.g.d2 + (p.oy + ly)) * a.g.d3 + p.oz;
    uint32_t carry = 0;
    for (uint32_t z0 = 0; z0 < p.sz; z0 += 64) {
        const uint32_t z = z0 + lane;
        const uint32_t v = z < p.sz ? q[z] : 0u;
        const uint32_t inc = wave_incl_scan32(v) + carry;
        if (z < p.sz)
            q[z] = inc;
        carry = __shfl(inc, 63, 64);
    }
}

// Along y: one thread per (sub-volume, x, z) column.
__global__ void __launch_bounds__(256) k3_ypass(Sz3Args a)
{
    const uint32_t col = blockIdx.x * blockDim.x + threadIdx.x;
    if (col >= a.g.ycols)
        return;
    uint32_t lo = 0, hi = a.g.T;
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) / 2;
        if (a.parts[mid].yc0 <= col)
            lo = mid;
        else
            hi = mid;
    }
    const Part3 p = a.parts[lo];
    const uint32_t c = col - p.yc0, lx = c / p.sz, lz = c - lx * p.sz;
    uint32_t* q = (uint32_t*) a.y + ((uint64_t) (p.ox + lx) * a.g.d2 + p.oy) * a.g.d3 + p.oz + lz;
    uint32_t run = q[0];
    for (uint32_t y = 1; y < p.sy; y++) {
        run += q[(uint64_t) y * a.g.d3];
        q[(uint64_t) y * a.g.d3] = run;
    }
}

// Along x, then dequantize in place: one thread per (sub-volume, y, z) column.
// x' = (float)((double)q * 2 eb), ZCCL/src/ZCCLd_float.c:3030.
__global__ void __launch_bounds__(256) k3_xpass(Sz3Args a)
{
    const uint32_t col = blockIdx.x * blockDim.x + threadIdx.x;
    if (col >= a.g.xcols)
        return;
    uint32_t lo = 0, hi = a.g.T;
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) / 2;
        if (a.parts[mid].xc0 <= col)
            lo = mid;
        else
            hi = mid;
    }
    const Part3 p = a.parts[lo];
    const uint32_t c = col - p.xc0, ly = c / p.sz, lz = c - ly * p.sz;
    const uint64_t sx = (uint64_t) a.g.d2 * a.g.d3;
    uint32_t* q = (uint32_t*) a.y + ((uint64_t) p.ox * a.g.d2 + (p.oy + ly)) * a.g.d3 + p.oz + lz;
    uint32_t run = 0;
    for (uint32_t x = 0; x < p.sx; x++) {
        run += q[x * sx];
        a.y[(q - (uint32_t*) a.y) + x * sx] = (float) __dmul_rn((double) (int32_t) run, a.tw);
    }
}

// ------------------------------------------------------------ tiled path
// Block sides 4 and 8.  A workgroup takes one chunk: a (bx, by) column of
// blocks of one sub-volume, cut to Tile3::NZ blocks along z.  Compress
// quantizes the chunk's values and a one-value halo (x - 1, y - 1, z - 1; zero
// outside the sub-volume) once into LDS and takes every residual from there;
// the decoder decodes a chunk's blocks into an LDS tile of z rows and scans
// them along z there (carrying across the column's chunks), so only the y and
// x passes remain.

template <uint32_t BS>
struct Tile3 {
    static constexpr uint32_t NZ = tile_nz3(BS);             // blocks per chunk
    static constexpr uint32_t WZ = NZ * BS;                  // z values per chunk
    static constexpr uint32_t SZ = WZ + 1;                   // with the z halo
    static constexpr uint32_t S1 = BS + 1;                   // x and y with the halo
    static constexpr uint32_t U = BS * BS * BS / 4;          // units of a whole block
    static constexpr uint32_t BPW = U >= 64 ? 1u : 64u / U;  // blocks per wave step
    static constexpr uint32_t STEPS = U >= 64 ? U / 64 : 1u; // wave steps per block
    static constexpr uint32_t UPR = BS / 4;                  // units per z row of a block
    // worst-case payload (signs + 32-bit values), 16-byte multiple, + slack
    static constexpr uint32_t PAY = ((BS * BS * BS + 7) / 8 + 4 * BS * BS * BS + 15) / 16 * 16 + 32;
};

// Quantized values of a chunk and its halo into qt[(i * S1 + j) * SZ + k]:
// tile (i, j, k) holds q at local (bx Bs + i - 1, by Bs + j - 1, lz0 + k - 1).
// A wave takes RG rows at a time and issues all their loads before it
// quantizes any, so each wave keeps up to RG * KS loads in flight.
template <uint32_t BS>
__device__ __forceinline__ void load_qtile(const Sz3Args& a, const Part3& p, uint32_t bx, uint32_t by, uint32_t lz0,
    uint32_t wz, uint32_t ex, uint32_t ey, int32_t* qt)
{
    using C = Tile3<BS>;
    constexpr uint32_t KS = (C::SZ + 63) / 64, RG = 4;
    const uint32_t lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    for (uint32_t r0 = RG * wv; r0 < C::S1 * C::S1; r0 += 4 * RG) {
        float v[RG][KS];
        bool live[RG];
#pragma unroll
        for (uint32_t h = 0; h < RG; h++) {
            const uint32_t row = r0 + h, i = row / C::S1, j = row - i * C::S1;
            live[h] = row < C::S1 * C::S1 && i <= ex && j <= ey && (i > 0 || bx > 0) && (j > 0 || by > 0);
            // element of (i - 1, j - 1, lz0) in the volume; every load below is
            // unconditional (clamped into the row, or element 0 for a dead row)
            // so the compiler keeps them all in flight
            const uint64_t rb = live[h] ?
                ((uint64_t) (p.ox + bx * BS + i - 1) * a.g.d2 + (p.oy + by * BS + j - 1)) * a.g.d3 + p.oz + lz0 : 0ull;
#pragma unroll
            for (uint32_t s = 0; s < KS; s++) {
                const uint32_t kk = min(lane + 64 * s, wz);
                const uint64_t e = rb + kk - ((kk > 0 || lz0 > 0) ? 1u : 0u);
                v[h][s] = a.x[live[h] ? e : 0ull];
            }
        }
#pragma unroll
        for (uint32_t h = 0; h < RG; h++) {
            const uint32_t row = r0 + h;
            if (row >= C::S1 * C::S1)
                continue;
            int32_t* dst = qt + row * C::SZ;
#pragma unroll
            for (uint32_t s = 0; s < KS; s++) {
                const uint32_t k = lane + 64 * s;
                if (k <= wz)
                    dst[k] = live[h] && (k > 0 || lz0 > 0) ? quant3(v[h][s], a.inv) : 0;
            }
        }
    }
}

// Column term of the residual at tile z index tk: the x-y difference of the
// four rows around (i, j) (value coordinates).
template <uint32_t BS>
__device__ __forceinline__ uint32_t col_term(const int32_t* qt, uint32_t i, uint32_t j, uint32_t tk)
{
    using C = Tile3<BS>;
    const int32_t* a0 = qt + ((i + 1) * C::S1 + (j + 1)) * C::SZ + tk;
    const int32_t* a2 = a0 - C::S1 * C::SZ;
    return (uint32_t) a0[0] - (uint32_t) a0[-(int) C::SZ] - (uint32_t) a2[0] + (uint32_t) a2[-(int) C::SZ];
}

// Residuals of unit u of a block whose first value has chunk-local z kb:
// returns the unit's value count (0..4) and its residuals in r[].
template <uint32_t BS>
__device__ __forceinline__ uint32_t unit_resid(const int32_t* qt, uint32_t ex, uint32_t ey, uint32_t ez, uint32_t cnt,
    uint32_t kb, uint32_t u, uint32_t r[4])
{
    using C = Tile3<BS>;
    const uint32_t n = 4 * u < cnt ? min(4u, cnt - 4 * u) : 0u;
    if (!n)
        return 0;
    if (cnt == BS * BS * BS) {
        // whole block: the unit is 4 consecutive values of one z row
        const uint32_t i = u / (BS * C::UPR), j = (u / C::UPR) % BS, k0 = kb + (u % C::UPR) * 4;
        uint32_t prev = col_term<BS>(qt, i, j, k0);
#pragma unroll
        for (uint32_t c = 0; c < 4; c++) {
            const uint32_t cur = col_term<BS>(qt, i, j, k0 + c + 1);
            r[c] = cur - prev;
            prev = cur;
        }
    } else {
        const uint32_t yz = ey * ez;
#pragma unroll
        for (uint32_t c = 0; c < 4; c++) {
            r[c] = 0;
            if (c < n) {
                const uint32_t e = 4 * u + c, i = e / yz, rest = e - i * yz, j = rest / ez, k = rest - j * ez;
                r[c] = col_term<BS>(qt, i, j, kb + k + 1) - col_term<BS>(qt, i, j, kb + k);
            }
        }
    }
    return n;
}

template <uint32_t BS>
__device__ __forceinline__ Chunk3 chunk_of(const Part3& p, uint32_t rel)
{
    return chunk_of3(p, rel, BS);
}

// Width and payload size of every block of a chunk.
template <uint32_t BS>
__global__ void __launch_bounds__(256) k3t_widths(Sz3Args a)
{
    using C = Tile3<BS>;
    __shared__ int32_t qt[C::S1 * C::S1 * C::SZ];
    const uint32_t lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const uint32_t t = part_of_unit(a.parts, a.g.T, blockIdx.x, false);
    const Part3 p = a.parts[t];
    const Chunk3 c = chunk_of<BS>(p, blockIdx.x - p.u0);
    load_qtile<BS>(a, p, c.bx, c.by, c.lz0, c.wz, c.ex, c.ey, qt);
    __syncthreads();
    const uint32_t ul = C::BPW > 1 ? lane % C::U : lane;
    for (uint32_t base = wv * C::BPW; base < c.nbz; base += 4 * C::BPW) {
        const uint32_t bzl = base + (C::BPW > 1 ? lane / C::U : 0u);
        const uint32_t ez = bzl < c.nbz ? min(BS, c.wz - bzl * BS) : 0u;
        const uint32_t cnt = c.ex * c.ey * ez;
        uint32_t m = 0;
        for (uint32_t s = 0; s < C::STEPS; s++) {
            uint32_t r[4];
            const uint32_t n = unit_resid<BS>(qt, c.ex, c.ey, ez, cnt, bzl * BS, C::BPW > 1 ? ul : s * 64 + lane, r);
#pragma unroll
            for (uint32_t k = 0; k < 4; k++)
                if (k < n)
                    m = max(m, mag32(r[k]));
        }
        if (C::BPW > 1) {
#pragma unroll
            for (uint32_t d = C::U / 2; d >= 1; d >>= 1)
                m = max(m, (uint32_t) __shfl_xor(m, d, 64));
        } else {
            m = wave_max(m);
        }
        if (ul == 0 && bzl < c.nbz) {
            const uint32_t w = m ? 32u - (uint32_t) __clz((int) m) : 0u;
            a.words[c.gcol + bzl] = payload3(w, cnt) | (w << 26);
        }
    }
}

// Header byte, sub-volume offset and payload of every block of a chunk; the
// payloads are built in LDS by the unit encoder of the 1-D sum and stored with
// dword stores that cover only the block's own bytes.
template <uint32_t BS>
__global__ void __launch_bounds__(256) k3t_encode(Sz3Args a)
{
    using C = Tile3<BS>;
    __shared__ int32_t qt[C::S1 * C::S1 * C::SZ];
    __shared__ uint32_t s_pay[4][C::BPW][C::PAY / 4];
    if (a.status[1] & ZK_ERR_NOSPACE)
        return;
    const uint32_t lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const uint32_t t = part_of_unit(a.parts, a.g.T, blockIdx.x, false);
    const Part3 p = a.parts[t];
    const uint32_t rel = blockIdx.x - p.u0;
    const Chunk3 c = chunk_of<BS>(p, rel);
    const uint64_t hdr = 8ull * a.g.T;
    const uint64_t sub = (uint64_t) p.nbc + a.offs[p.g0];  // sub-volume start after the header
    if (rel == 0 && threadIdx.x == 0)
        __builtin_memcpy(a.out + 8ull * t, &sub, 8);
    load_qtile<BS>(a, p, c.bx, c.by, c.lz0, c.wz, c.ex, c.ey, qt);
    __syncthreads();
    const uint32_t ul = C::BPW > 1 ? lane % C::U : lane;
    const uint32_t step = C::BPW > 1 ? C::U : 64u;
    uint8_t* st = (uint8_t*) s_pay[wv][C::BPW > 1 ? lane / C::U : 0u];
    for (uint32_t base = wv * C::BPW; base < c.nbz; base += 4 * C::BPW) {
        const uint32_t bzl = base + (C::BPW > 1 ? lane / C::U : 0u);
        const bool valid = bzl < c.nbz;
        const uint32_t gb = c.gcol + bzl;
        const uint32_t word = valid ? a.words[gb] : 0u, w = word >> 26, size = word & kSizeMask;
        const uint32_t ez = valid ? min(BS, c.wz - bzl * BS) : 0u;
        const uint32_t cnt = c.ex * c.ey * ez, Ub = (cnt + 3) / 4;
        if (valid && ul == 0)
            a.out[hdr + sub + (gb - p.g0)] = (uint8_t) w;
        for (uint32_t s = 0; s < C::STEPS; s++) {
            const uint32_t u = C::BPW > 1 ? ul : s * 64 + lane;
            uint32_t d[4] = {0, 0, 0, 0};
            const uint32_t n = w ? unit_resid<BS>(qt, c.ex, c.ey, ez, cnt, bzl * BS, u, d) : 0u;
            const UnitFields f = unit_fields(w, d);
            const uint32_t pnib = __shfl_down(f.nib, 1, 64);
            const uint32_t plv = __shfl_down(f.lv, 1, 64);
            emit_unit(st, 0, w, cnt, u, Ub, n, f, pnib, plv);
        }
        wave_sync_lds();
        if (valid && w) {
            uint8_t* g = a.out + hdr + p.nbc + (uint64_t) p.b1 * p.b2 * p.b3 + a.offs[gb];
            const uint32_t nw = size / 4;
            for (uint32_t i = ul; i < nw; i += step) {
                const uint32_t v = ((const uint32_t*) st)[i];
                __builtin_memcpy(g + 4 * i, &v, 4);
            }
            for (uint32_t i = 4 * nw + ul; i < size; i += step)
                g[i] = st[i];
        }
        wave_sync_lds();
    }
}

// Residuals of a block column, chunk by chunk: blocks decoded into an LDS tile
// of z rows, each row scanned along z (carried across chunks) and written to
// the output buffer as int32.  ZCCL/src/ZCCL_TypeManager.c:423-517, :906-950.
template <uint32_t BS>
__global__ void __launch_bounds__(256) k3t_decode(Sz3Args a)
{
    using C = Tile3<BS>;
    __shared__ uint4 rt4[BS * BS * C::WZ / 4];
    __shared__ uint32_t s_pay[4][C::BPW][C::PAY / 4];
    __shared__ uint32_t s_carry[BS * BS];
    uint32_t* rt = (uint32_t*) rt4;
    const uint32_t lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const uint32_t t = part_of_unit(a.parts, a.g.T, blockIdx.x, true);
    const Part3 p = a.parts[t];
    const uint32_t nch = (p.b3 + C::NZ - 1) / C::NZ;
    uint64_t h;
    __builtin_memcpy(&h, a.in + 8ull * t, 8);
    const uint64_t pbase = 8ull * a.g.T + h + (uint64_t) p.b1 * p.b2 * p.b3;
    const unsigned long long o0 = a.offs[p.g0];
    for (uint32_t i = threadIdx.x; i < BS * BS; i += 256)
        s_carry[i] = 0;
    const uint32_t ul = C::BPW > 1 ? lane % C::U : lane;
    const uint32_t step = C::BPW > 1 ? C::U : 64u;
    uint8_t* st = (uint8_t*) s_pay[wv][C::BPW > 1 ? lane / C::U : 0u];
    for (uint32_t ch = 0; ch < nch; ch++) {
        const Chunk3 c = chunk_of<BS>(p, (blockIdx.x - p.c0) * nch + ch);
        for (uint32_t base = wv * C::BPW; base < c.nbz; base += 4 * C::BPW) {
            const uint32_t bzl = base + (C::BPW > 1 ? lane / C::U : 0u);
            const bool valid = bzl < c.nbz;
            const uint32_t gb = c.gcol + bzl;
            const uint32_t word = valid ? a.words[gb] : 0u, w = word >> 26, size = word & kSizeMask;
            const uint32_t ez = valid ? min(BS, c.wz - bzl * BS) : 0u;
            const uint32_t cnt = c.ex * c.ey * ez, sb = (cnt + 7) / 8;
            const uint64_t fo = pbase + ((valid ? a.offs[gb] : o0) - o0);
            const bool inside = fo + size <= a.in_size;
            if (valid && w && !inside && ul == 0)
                raise3(a.status, ZK_ERR_OVERRUN);
            const bool ok = valid && w && inside;
            if (ok) {
                // all of the payload's dword loads in flight before any store
                constexpr uint32_t PS = (C::PAY / 4 + (C::BPW > 1 ? C::U : 64u) - 1) / (C::BPW > 1 ? C::U : 64u);
                const uint8_t* pl = a.in + fo;
                const uint32_t nw = size / 4;
                uint32_t v[PS];
                // unconditional loads (clamped into the payload, or the stream's
                // first bytes when it has no whole dword) stay in flight together
#pragma unroll
                for (uint32_t m = 0; m < PS; m++) {
                    const uint32_t i = min(ul + m * step, nw ? nw - 1 : 0u);
                    __builtin_memcpy(&v[m], nw ? pl + 4 * i : a.in, 4);
                }
#pragma unroll
                for (uint32_t m = 0; m < PS; m++) {
                    const uint32_t i = ul + m * step;
                    if (i < nw)
                        ((uint32_t*) st)[i] = v[m];
                }
                for (uint32_t i = 4 * nw + ul; i < size; i += step)
                    st[i] = pl[i];
            }
            wave_sync_lds();
            for (uint32_t s = 0; s < C::STEPS; s++) {
                const uint32_t u = C::BPW > 1 ? ul : s * 64 + lane;
                const uint32_t n = 4 * u < cnt ? min(4u, cnt - 4 * u) : 0u;
                uint32_t v[4] = {0, 0, 0, 0};
                if (ok && n)
                    accumulate_unit(st - 1, w, cnt, sb, u, v);  // fields start right after the (absent) width byte
                if (n == 4 && cnt == BS * BS * BS) {
                    const uint32_t i = u / (BS * C::UPR), j = (u / C::UPR) % BS, k0 = bzl * BS + (u % C::UPR) * 4;
                    rt4[((i * BS + j) * C::WZ + k0) / 4] = make_uint4(v[0], v[1], v[2], v[3]);
                } else if (n) {
                    const uint32_t yz = c.ey * ez;
#pragma unroll
                    for (uint32_t k = 0; k < 4; k++) {
                        if (k < n) {
                            const uint32_t e = 4 * u + k, i = e / yz, rest = e - i * yz, j = rest / ez;
                            rt[(i * BS + j) * C::WZ + bzl * BS + (rest - j * ez)] = v[k];
                        }
                    }
                }
            }
            wave_sync_lds();
        }
        __syncthreads();
        for (uint32_t row = wv; row < c.ex * c.ey; row += 4) {
            const uint32_t i = row / c.ey, j = row - i * c.ey;
            const uint32_t* src = rt + (i * BS + j) * C::WZ;
            uint32_t carry = s_carry[i * BS + j];
            uint32_t* q = (uint32_t*) a.y + ((uint64_t) (p.ox + c.bx * BS + i) * a.g.d2 + (p.oy + c.by * BS + j)) * a.g.d3 +
                p.oz + c.lz0;
            for (uint32_t k0 = 0; k0 < c.wz; k0 += 64) {
                const uint32_t k = k0 + lane;
                const uint32_t v = k < c.wz ? src[k] : 0u;
                const uint32_t inc = wave_incl_scan32(v) + carry;
                if (k < c.wz)
                    q[k] = inc;
                carry = __shfl(inc, 63, 64);
            }
            if (lane == 0)
                s_carry[i * BS + j] = carry;
        }
        __syncthreads();
    }
}

}  // namespace

static hipError_t scan3(const Sz3Args& a, int check_cap, hipStream_t s)
{
    const uint32_t groups = (uint32_t) ceil_div(a.g.nblocks ? a.g.nblocks : 1, kScan3Group);
    hipLaunchKernelGGL(k3_scan_groups, dim3(groups), dim3(256), 0, s, a);
    hipLaunchKernelGGL(k3_scan_top, dim3(1), dim3(256), 0, s, a, groups, check_cap);
    hipLaunchKernelGGL(k3_scan_apply, dim3(groups), dim3(256), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_sz3_table(const Geom3& g, Part3* parts, unsigned long long* status, hipStream_t s)
{
    hipLaunchKernelGGL(k3_table, dim3(1), dim3(64), 0, s, g.d1, g.d2, g.d3, (int) g.Bs, (int) g.T, parts, status);
    return hipGetLastError();
}

template <uint32_t BS>
static hipError_t compress_tiled(const Sz3Args& a, hipStream_t s)
{
    hipLaunchKernelGGL(k3t_widths<BS>, dim3(a.g.nunits), dim3(256), 0, s, a);
    hipError_t e = scan3(a, 1, s);
    if (e != hipSuccess)
        return e;
    hipLaunchKernelGGL(k3t_encode<BS>, dim3(a.g.nunits), dim3(256), 0, s, a);
    return hipGetLastError();
}

template <uint32_t BS>
static hipError_t decompress_tiled(const Sz3Args& a, hipStream_t s)
{
    hipLaunchKernelGGL(k3_dsizes, dim3((uint32_t) ceil_div(a.g.nblocks, 256)), dim3(256), 0, s, a);
    hipError_t e = scan3(a, 0, s);
    if (e != hipSuccess)
        return e;
    hipLaunchKernelGGL(k3t_decode<BS>, dim3(a.g.ncols), dim3(256), 0, s, a);
    hipLaunchKernelGGL(k3_ypass, dim3((uint32_t) ceil_div(a.g.ycols, 256)), dim3(256), 0, s, a);
    hipLaunchKernelGGL(k3_xpass, dim3((uint32_t) ceil_div(a.g.xcols, 256)), dim3(256), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_sz3_compress(const Sz3Args& a, hipStream_t s)
{
    if (a.g.nunits && a.g.Bs == 4)
        return compress_tiled<4>(a, s);
    if (a.g.nunits && a.g.Bs == 8)
        return compress_tiled<8>(a, s);
    const uint32_t waves = (uint32_t) ceil_div(a.g.nblocks, 4);
    hipLaunchKernelGGL(k3_widths, dim3(waves), dim3(256), 0, s, a);
    hipError_t e = scan3(a, 1, s);
    if (e != hipSuccess)
        return e;
    hipLaunchKernelGGL(k3_encode, dim3(waves), dim3(256), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_sz3_decompress(const Sz3Args& a, hipStream_t s)
{
    if (a.g.nunits && a.g.Bs == 4)
        return decompress_tiled<4>(a, s);
    if (a.g.nunits && a.g.Bs == 8)
        return decompress_tiled<8>(a, s);
    hipLaunchKernelGGL(k3_dsizes, dim3((uint32_t) ceil_div(a.g.nblocks, 256)), dim3(256), 0, s, a);
    hipError_t e = scan3(a, 0, s);
    if (e != hipSuccess)
        return e;
    hipLaunchKernelGGL(k3_decode, dim3((uint32_t) ceil_div(a.g.nblocks, 4)), dim3(256), 0, s, a);
    hipLaunchKernelGGL(k3_zpass, dim3((uint32_t) ceil_div(a.g.zrows, 4)), dim3(256), 0, s, a);
    hipLaunchKernelGGL(k3_ypass, dim3((uint32_t) ceil_div(a.g.ycols, 256)), dim3(256), 0, s, a);
    hipLaunchKernelGGL(k3_xpass, dim3((uint32_t) ceil_div(a.g.xcols, 256)), dim3(256), 0, s, a);
    return hipGetLastError();
}

}  // namespace zk
