// zccl_calib.hip -- copy kernels of known traffic, for calibrating the
// rocprofv3 byte counters (FETCH_SIZE, WRITE_SIZE) against the load paths
// the sum kernels use.  Not part of the codec.
//
// mode 0: 16-byte global loads into registers, 16-byte global stores.
// mode 1: 16-byte buffer loads straight into LDS (buffer_load_dwordx4 ... lds,
//         the sum's staging path), LDS reads, 16-byte global stores.
// Each reads `bytes` bytes (a multiple of 16) once and writes them once.
#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

namespace {

constexpr uint32_t kThreads = 256;
constexpr uint32_t kChunkPerThread = 4;  // 16-byte chunks per thread per pass

__global__ void __launch_bounds__(kThreads) k_calib_copy_global(const uint4* __restrict__ src, uint4* __restrict__ dst,
    uint64_t nchunks)
{
    const uint64_t base = ((uint64_t) blockIdx.x * kThreads * kChunkPerThread) + threadIdx.x;
    uint4 v[kChunkPerThread];
#pragma unroll
    for (uint32_t u = 0; u < kChunkPerThread; u++) {
        const uint64_t c = base + (uint64_t) u * kThreads;
        v[u] = c < nchunks ? src[c] : make_uint4(0, 0, 0, 0);
    }
#pragma unroll
    for (uint32_t u = 0; u < kChunkPerThread; u++) {
        const uint64_t c = base + (uint64_t) u * kThreads;
        if (c < nchunks)
            dst[c] = v[u];
    }
}

__global__ void __launch_bounds__(kThreads) k_calib_copy_lds(const uint8_t* __restrict__ src, uint4* __restrict__ dst,
    uint64_t nchunks)
{
    __shared__ __attribute__((aligned(16))) uint4 stage[kThreads * kChunkPerThread];
    const uint32_t wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const uint64_t c0 = (uint64_t) blockIdx.x * kThreads * kChunkPerThread;
    // the block's chunks as one range-checked view: reads past the array return zeros
    const uint64_t left = nchunks > c0 ? nchunks - c0 : 0;
    const uint64_t nb = (left < kThreads * kChunkPerThread ? left : kThreads * kChunkPerThread) * 16;
    const __amdgpu_buffer_rsrc_t v = __builtin_amdgcn_make_buffer_rsrc((void*) (src + c0 * 16), (short) 0, (int) nb,
        0x00020000);
#pragma unroll
    for (uint32_t u = 0; u < kChunkPerThread; u++) {
        // wave wv, pass u: chunks [(u * 4 + wv) * 64, +64) of the block
        const uint32_t run = u * (kThreads / 64) + wv;
        __builtin_amdgcn_raw_ptr_buffer_load_lds(v, (__attribute__((address_space(3))) void*) (stage + run * 64), 16,
            (run * 64 + lane) * 16, 0, 0, 0);
    }
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
#pragma unroll
    for (uint32_t u = 0; u < kChunkPerThread; u++) {
        const uint32_t i = u * kThreads + threadIdx.x;
        const uint64_t c = c0 + i;
        if (c < nchunks)
            dst[c] = stage[i];
    }
}

}  // namespace

// Copy `bytes` (a multiple of 16; src and dst 16-byte aligned, not
// overlapping) with the given load path.  No counterpart in the reference:
// a measurement aid for the HBM traffic figures.
extern "C" int zccl_calib_copy(const void* src, void* dst, size_t bytes, int mode, void* stream)
{
    if (!src || !dst || (bytes & 15) || (((uintptr_t) src | (uintptr_t) dst) & 15) || mode < 0 || mode > 1)
        return -1;
    const uint64_t nchunks = bytes / 16;
    const uint64_t per = (uint64_t) kThreads * kChunkPerThread;
    const uint32_t grid = (uint32_t) ((nchunks + per - 1) / per);
    if (!grid)
        return 0;
    if (mode == 0)
        hipLaunchKernelGGL(k_calib_copy_global, dim3(grid), dim3(kThreads), 0, (hipStream_t) stream,
            (const uint4*) src, (uint4*) dst, nchunks);
    else
        hipLaunchKernelGGL(k_calib_copy_lds, dim3(grid), dim3(kThreads), 0, (hipStream_t) stream,
            (const uint8_t*) src, (uint4*) dst, nchunks);
    return hipGetLastError() == hipSuccess ? 0 : -2;
}
