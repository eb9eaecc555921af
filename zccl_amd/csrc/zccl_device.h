// zccl_device.h -- device helpers shared by the 1-D kernels
// (zccl_kernels.hip, zccl_lane.hip): the reference's arithmetic, wave scans,
// the decoupled lookback with its tile tickets, and index access.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "zccl_internal.h"

namespace zk {

// ------------------------------------------------------------ arithmetic

// Width thresholds c_w for w = 21..32: (int)log2f((float)m)+1 equals w+1
// once m >= 2^w - c_w (glibc 2.35 log2f, measured over every m < 2^32).
static __constant__ uint32_t kBumpC[12] = {1, 2, 5, 11, 22, 45, 91, 183, 367, 735, 1471, 2943};

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

// 16 bytes from any byte address (gfx950 global loads need no alignment).
__device__ __forceinline__ uint4 ld16(const void* p)
{
    uint4 v;
    __builtin_memcpy(&v, p, 16);
    return v;
}

__device__ __forceinline__ float4 ldf4(const float* p)
{
    float4 v;
    __builtin_memcpy(&v, p, 16);
    return v;
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

// Inclusive scan over the 64 lanes of a wave with DPP adds: shifts of 1, 2, 4
// and 8 lanes inside each row of 16, then the row totals by row broadcasts
// (lane 15 into rows 1 and 3, lane 31 into rows 2 and 3).  Every lane of the
// wave must be active.
__device__ __forceinline__ uint32_t wave_scan_dpp(uint32_t x)
{
    x += (uint32_t) __builtin_amdgcn_update_dpp(0, (int) x, 0x111, 0xF, 0xF, false);  // row_shr:1
    x += (uint32_t) __builtin_amdgcn_update_dpp(0, (int) x, 0x112, 0xF, 0xF, false);  // row_shr:2
    x += (uint32_t) __builtin_amdgcn_update_dpp(0, (int) x, 0x114, 0xF, 0xF, false);  // row_shr:4
    x += (uint32_t) __builtin_amdgcn_update_dpp(0, (int) x, 0x118, 0xF, 0xF, false);  // row_shr:8
    x += (uint32_t) __builtin_amdgcn_update_dpp(0, (int) x, 0x142, 0xA, 0xF, false);  // row_bcast:15
    x += (uint32_t) __builtin_amdgcn_update_dpp(0, (int) x, 0x143, 0xC, 0xF, false);  // row_bcast:31
    return x;
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

// Decoupled lookback, executed by one wave of the tile.  Tiles t0..t form
// one scan (status entry of tile i at status[i * kStatusStride]); tile t0
// starts from `init`.  Returns init + the aggregates of tiles t0..t-1 (mod
// 2^62; callers mask further) and publishes tile t's inclusive value.
// Predecessors are polled 64 at a time (one per lane), and a window waits
// only for those up to the nearest inclusive value.  A waiting wave backs off
// (s_sleep doubling up to 64 x 64 cycles): polls are uncached device-scope
// loads, and thousands of waves polling at full rate take memory bandwidth
// from the tiles they wait for.  Every wait is bounded: past kSpinLimit
// polls the tile records ZK_ERR_SPIN and proceeds, so the grid always drains.
// First half of the lookback, callable early by any one thread: the tile's
// aggregate (the chain's first tile: its inclusive value) becomes visible.
__device__ __forceinline__ void lookback_publish(unsigned long long* status, uint64_t t, uint64_t t0,
    unsigned long long agg, unsigned long long init)
{
    st_publish(&status[t * kStatusStride], t == t0 ? kFlagInc | ((init + agg) & kValMask) : kFlagAgg | (agg & kValMask));
}

// Sleep of the backoff step b (s_sleep takes an immediate: 64 cycles per unit).
__device__ __forceinline__ void backoff_sleep(uint32_t b)
{
    switch (b) {
    case 0: __builtin_amdgcn_s_sleep(1); break;
    case 1: __builtin_amdgcn_s_sleep(2); break;
    case 2: __builtin_amdgcn_s_sleep(4); break;
    case 3: __builtin_amdgcn_s_sleep(8); break;
    case 4: __builtin_amdgcn_s_sleep(16); break;
    case 5: __builtin_amdgcn_s_sleep(32); break;
    default: __builtin_amdgcn_s_sleep(64); break;
    }
}

// `published`: lookback_publish already ran for this tile.
__device__ __forceinline__ unsigned long long lookback(unsigned long long* status, uint64_t t, uint64_t t0,
    unsigned long long agg, unsigned long long init, WsHeader* ws, bool published = false)
{
    const int lane = threadIdx.x & 63;
    if (!published && lane == 0)
        lookback_publish(status, t, t0, agg, init);
    if (t == t0)
        return init;
    unsigned long long excl = 0;
    int64_t end = (int64_t) t;
    uint32_t spins = 0;
    bool timed_out = false;
    while (true) {
        const int64_t idx = end - 1 - lane;
        const bool act = idx >= (int64_t) t0;
        unsigned long long st = 0;
        uint32_t b = 0;
        while (true) {
            if (act)
                st = st_poll(&status[(uint64_t) idx * kStatusStride]);
            // only the predecessors up to the nearest inclusive one matter: a
            // slow tile further back must not hold this one
            const unsigned long long zm = __ballot(act && (st >> 62) == 0);
            const unsigned long long im = __ballot(act && (st >> 62) == 2);
            const unsigned long long need = im ? ((im & (~im + 1)) << 1) - 1 : ~0ull;
            if (!(zm & need))
                break;
            if (++spins > kSpinLimit) {
                if (act && (st >> 62) == 0)
                    st = kFlagInc;
                timed_out = true;
                break;
            }
            backoff_sleep(b);
            b += b < 6 ? 1u : 0u;
        }
        unsigned long long pm = __ballot(act && (st >> 62) == 2);
        unsigned long long val = act ? (st & kValMask) : 0ull;
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
        st_publish(&status[t * kStatusStride], kFlagInc | ((excl + agg) & kValMask));
    return excl;
}

// Tile ticket and status arrays of the current launch.  The workgroup takes
// the next ticket (its tile index t, in start order) and clears its share of
// the array the next launch will use.  Returns false when the workspace is
// too small (then no workgroup of the launch takes a ticket or does anything).
// With tiles_per_group > 1 a ticket stands for that many consecutive tiles
// (one per wave of the workgroup) and t is the first of them.
__device__ __forceinline__ bool launch_status(WsHeader* ws, uint64_t ntiles_total, uint64_t& t,
    unsigned long long*& status, uint32_t tiles_per_group = 1)
{
    __shared__ unsigned long long s_ticket;
    const uint64_t cap = ws->capacity;
    if (cap < ntiles_total) {
        if (threadIdx.x == 0 && blockIdx.x == 0)
            raise_error(ws, ZK_ERR_CAPACITY);
        return false;
    }
    if (threadIdx.x == 0)
        s_ticket = atomicAdd(&ws->seq, 1ull);
    __syncthreads();
    const unsigned long long tk = s_ticket;
    const uint32_t par = (uint32_t) (tk >> 32) & 1u;
    const uint64_t ticket = (uint32_t) tk;
    t = ticket * tiles_per_group;
    unsigned long long* base = (unsigned long long*) ((char*) ws + sizeof(WsHeader));
    status = base + (uint64_t) par * cap * kStatusStride;
    unsigned long long* other = base + (uint64_t) (par ^ 1u) * cap * kStatusStride;
    uint64_t per = ceil_div(cap, ceil_div(ntiles_total, tiles_per_group));
    uint64_t lo = ticket * per, hi = lo + per < cap ? lo + per : cap;
    for (uint64_t i = lo + threadIdx.x; i < hi; i += blockDim.x)
        other[i * kStatusStride] = 0;
    return true;
}

// Called by one thread of the last tile of a lookback chain once that tile
// has published.  A finished chain implies every tile of the chain took its
// ticket; when the last chain of the launch finishes, every ticket of the
// launch has been taken and the ticket word restarts at 0 with the other
// parity for the next launch.
__device__ __forceinline__ void chain_finished(WsHeader* ws, uint32_t nchains)
{
    unsigned int old = atomicAdd(&ws->chains_done, 1u);
    if (old == nchains - 1) {
        ws->chains_done = 0;
        const unsigned long long par = (__hip_atomic_load(&ws->seq, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) >> 32) & 1ull;
        __hip_atomic_store(&ws->seq, (par ^ 1ull) << 32, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
}

// ------------------------------------------------------------ index access

struct IndexView {
    uint8_t* widths;
    unsigned long long* group_off;
    unsigned long long* seed_off;
    unsigned long long* meta;
};

__device__ __forceinline__ IndexView index_view(const Geom& g, const uint8_t* ix)
{
    IndexView v;
    v.widths = (uint8_t*) ix;
    v.group_off = (unsigned long long*) (ix + g.ix_group);
    v.seed_off = (unsigned long long*) (ix + g.ix_seed);
    v.meta = (unsigned long long*) (ix + g.ix_meta);
    return v;
}

__device__ __forceinline__ uint32_t load_u32_bytes(const uint8_t* p)
{
    return (uint32_t) p[0] | ((uint32_t) p[1] << 8) | ((uint32_t) p[2] << 16) | ((uint32_t) p[3] << 24);
}

__device__ __forceinline__ uint64_t load_u64_bytes(const uint8_t* p)
{
    return (uint64_t) load_u32_bytes(p) | ((uint64_t) load_u32_bytes(p + 4) << 32);
}

__device__ __forceinline__ void store_bytes(uint8_t* dst, const void* src, int n)
{
    const uint8_t* s = (const uint8_t*) src;
    for (int i = 0; i < n; i++)
        dst[i] = s[i];
}


}  // namespace zk
