// zccl_kernels.h -- kernel argument blocks and launchers (internal).
#pragma once

#include <hip/hip_runtime.h>

#include "zccl_internal.h"

namespace zk {

struct CompressArgs {
    Geom g;
    const float* x;
    double inv;                    // (double)(1.0f / eb), FP32 division done on the host
    uint8_t* out;
    uint64_t out_cap;
    uint8_t* index;                // output index (Geom::ix_bytes)
    WsHeader* ws;
    unsigned long long* size_out;  // optional device copy of the stream size
};

struct Operand {
    const uint8_t* bytes;
    const uint8_t* index;
};

struct HomsumArgs {
    Geom g;
    int k;
    Operand in[kMaxOperands];
    uint8_t* out;
    uint64_t out_cap;
    uint8_t* index;
    WsHeader* ws;
    unsigned long long* size_out;
};

struct DecompressArgs {
    Geom g;
    const uint8_t* bytes;
    const uint8_t* index;
    float eb;
    float* out;
    WsHeader* ws;
};

struct ParseArgs {
    Geom g;
    const uint8_t* bytes;
    uint64_t in_cap;               // bytes of `bytes` that may be read
    uint8_t* index;
    WsHeader* ws;
    unsigned long long* size_out;
};

hipError_t launch_compress(const CompressArgs& a, hipStream_t s);
hipError_t launch_homsum(const HomsumArgs& a, hipStream_t s);
hipError_t launch_decompress(const DecompressArgs& a, hipStream_t s);
hipError_t launch_parse(const ParseArgs& a, hipStream_t s);
hipError_t launch_case_count(const uint8_t* w1, const uint8_t* w2, uint64_t nblocks,
    unsigned long long* counts, hipStream_t s);
hipError_t launch_ws_init(WsHeader* ws, uint64_t capacity, hipStream_t s);

}  // namespace zk
