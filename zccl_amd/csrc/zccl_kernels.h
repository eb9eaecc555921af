// zccl_kernels.h -- kernel argument blocks and launchers (internal).
#pragma once

#include <hip/hip_runtime.h>

#include "zccl_internal.h"
#include "zccl_sz3d.h"

namespace zk {

// Workgroup size bound of every tile kernel: a tile holds at most kTileUnits
// 4-value units and a thread owns four of them.
// Workgroup size targets (threads): compress and decompress tiles fill 8
// waves, the homomorphic sum (more state per thread, k operands' bytes in LDS)
// fills 4 so three of its workgroups share a CU.
constexpr uint32_t kCompressThreads = 512;
constexpr uint32_t kDecompressThreads = 512;
constexpr uint32_t kHomsumThreads = 256;

// LDS stage of the homomorphic sum (bytes): two halves, one being decoded
// while the next round of operand bytes lands in the other.  With the tables
// and static state a workgroup needs just under 26 KB, so six 256-thread
// workgroups share a CU's 160 KB at the allocation granularity (measured:
// 17.5 KB beats 18 KB, which drops to five workgroups, 24 KB with five and
// 32 KB with four).
constexpr uint32_t kStageMin = 17 * 1024 + 512;

struct Operand {
    const uint8_t* bytes;
    const uint8_t* index;
};

// Pointers of one operation; geometry travels separately so a launch whose
// jobs share one geometry keeps it in kernel arguments.
struct CompressIO {
    const float* x;
    uint8_t* out;
    uint64_t out_cap;
    uint8_t* index;                // output index (Geom::ix_bytes)
    unsigned long long* size_out;  // optional device copy of the stream size
};

struct HomsumIO {
    Operand in[kMaxOperands];
    uint8_t* out;
    uint64_t out_cap;
    uint8_t* index;
    unsigned long long* size_out;
};

struct DecompressIO {
    Operand in[1];
    float* out;
};

// Scalar parameters of an operation.
struct OpParams {
    double inv;           // compress: (double)(1.0f / eb), FP32 division done on the host
    float eb;             // decompress: error bound
    int k;                // homomorphic sum: operands
    uint32_t stage_bytes; // sum / decompress: LDS stage area
};

struct CompressJob {
    Geom g;
    Tiling tl;
    OpParams p;
    CompressIO io;
};

struct HomsumJob {
    Geom g;
    Tiling tl;
    OpParams p;
    HomsumIO io;
};

struct DecompressJob {
    Geom g;
    Tiling tl;
    OpParams p;
    DecompressIO io;
};

struct CompressArgs {
    CompressJob job;
    WsHeader* ws;
};

struct HomsumArgs {
    HomsumJob job;
    WsHeader* ws;
};

struct DecompressArgs {
    DecompressJob job;
    WsHeader* ws;
};

// Many jobs of one kind in one launch.  Job j owns tiles [job_tile0[j],
// job_tile0[j+1]) of the launch.  With `uniform` every job shares g / tl / p
// (kernel arguments) and `jobs` holds only their IO blocks; otherwise `jobs`
// holds full CompressJob / HomsumJob / DecompressJob descriptors.  With
// `interleave` (uniform jobs) workgroup b serves tile b / njobs of job
// b % njobs, running the jobs' lookback chains side by side.
enum class BatchKind { Compress, Homsum, Decompress };

struct BatchArgs {
    const void* jobs;
    const unsigned long long* job_tile0;  // njobs + 1 entries
    uint32_t njobs;
    uint32_t uniform;
    uint32_t interleave;
    uint32_t total_chains;
    uint32_t max_segments;  // the most segments of any job
    uint32_t pad0;
    uint64_t total_tiles;
    Geom g;
    Tiling tl;
    OpParams p;
    WsHeader* ws;
};

struct ParseArgs {
    Geom g;
    const uint8_t* bytes;
    uint64_t in_cap;               // bytes of `bytes` that may be read
    uint8_t* index;
    WsHeader* ws;
    uint64_t ws_bytes;             // the workspace's size (the parallel parse keeps its tables there)
    unsigned long long* size_out;
};

// Parallel parse (zccl_kernels.hip): a chunk's block region (the bytes after
// its seed) in tiles of kParseTile bytes; per tile a table of candidate first
// block offsets (0 .. block_bytes(32, L) - 1), then per tile its convergence
// word, entry and block count / base.  The tables live in the workspace
// after its header and are zeroed again when the parse ends.
constexpr uint32_t kParseTile = 2048;
constexpr uint32_t kParseMaxL = 64;         // at most 265 candidates per tile
constexpr uint32_t kParseMinBlocks = 256;   // chunks of fewer blocks take the serial walk
constexpr uint32_t kParseDead = 0xFFFFFFFFu;   // no entry / no common exit
constexpr uint32_t kParseNoExit = 0xFFFFu;     // table: the walk met a width byte above 32

struct ParseLayout {
    uint32_t maxb;     // candidates per tile: the largest block's bytes
    uint32_t tpc;      // tiles per chunk: its worst-case block region / kParseTile
    uint64_t ntiles;   // T * tpc
    uint64_t flag_off;  // a word set when any pass finds the stream corrupt
    uint64_t table_off, conv_off, entry_off, base_off, bytes;  // byte offsets from the workspace start
};

ZK_HD ParseLayout parse_layout(const Geom& g)
{
    ParseLayout p;
    p.maxb = block_bytes(32, g.L);
    p.tpc = (uint32_t) ceil_div((uint64_t) g.nbT * p.maxb, kParseTile);
    p.ntiles = (uint64_t) g.T * p.tpc;
    uint64_t at = sizeof(WsHeader);
    p.flag_off = at;
    at += 16;
    p.table_off = at;
    at += p.ntiles * p.maxb * 4;
    p.conv_off = at;
    at += p.ntiles * 4;
    p.entry_off = at;
    at += p.ntiles * 4;
    p.base_off = at;
    at += p.ntiles * 4;
    p.bytes = align_up(at, 16);
    return p;
}

ZK_HD bool parse_parallel_ok(const Geom& g) { return g.L <= kParseMaxL && g.nbT >= kParseMinBlocks; }

// One call of the 3-D codec (zccl_sz3d.hip).  The workspace holds the status
// words (status[0] stream size, status[1] error bits), the per-block size
// words, their scan, the scan's group sums and the sub-volume table.
struct Sz3Args {
    Geom3 g;
    const Part3* parts;
    double inv;                      // compress: 0.5 / eb
    double tw;                       // decompress: 2 * eb
    const float* x;                  // compress input
    uint8_t* out;                    // compress output
    uint64_t out_cap;
    const uint8_t* in;               // decompress input
    uint64_t in_size;
    float* y;                        // decompress output (int32 residuals until the last pass)
    uint32_t* words;                 // per block: payload bytes | width << 26
    unsigned long long* offs;        // per block: exclusive scan of payload bytes
    unsigned long long* part_sums;   // per scan group (+ the total)
    unsigned long long* status;
    unsigned long long* size_out;    // optional device copy of the stream size
};

size_t compress_lds_bytes(const Tiling& tl);
uint32_t stage_bytes_for(const Tiling& tl, int k);
// Smallest operand stage of the slot-layout sum (bytes, a multiple of 32; 0:
// the default); larger stages take more operands per round and fewer
// workgroups per CU.
bool set_stage_min_bytes(uint32_t bytes);
uint32_t get_stage_min_bytes();
size_t operand_lds_bytes(const Tiling& tl, int k, uint32_t stage);

hipError_t launch_compress(const CompressArgs& a, hipStream_t s);
hipError_t launch_homsum(const HomsumArgs& a, hipStream_t s);
hipError_t launch_decompress(const DecompressArgs& a, hipStream_t s);
hipError_t launch_batch(BatchKind kind, const BatchArgs& b, uint32_t threads, size_t lds, hipStream_t s);
hipError_t launch_parse(const ParseArgs& a, hipStream_t s);
// Compute units of the current device (queried once per device).
uint32_t device_cu_count();
// Lane layout (zccl_lane.hip): tilings from make_lane_tiling, 36-value blocks.
size_t lane_lds_bytes();
hipError_t launch_homsum_lane(const HomsumArgs& a, hipStream_t s);
hipError_t launch_homsum_lane_batch(const BatchArgs& b, hipStream_t s);
hipError_t launch_homsum_lane_wg(const HomsumArgs& a, hipStream_t s);
hipError_t launch_homsum_lane_wg_batch(const BatchArgs& b, hipStream_t s);
hipError_t launch_compress_lane(const CompressArgs& a, hipStream_t s);
hipError_t launch_compress_lane_batch(const BatchArgs& b, hipStream_t s);
hipError_t launch_decompress_lane(const DecompressArgs& a, hipStream_t s);
hipError_t launch_decompress_lane_batch(const BatchArgs& b, hipStream_t s);
// Workgroups per CU of the persistent lane decompress grid (1 ..
// kDecGroupsMax; 0 restores the default, 4).
constexpr int kDecGroupsMax = 5;
bool set_decompress_groups_per_cu(int groups);
int get_decompress_groups_per_cu();

// Two-pass lane sum.  Pass one sums and encodes every tile into a slot of a
// stage area with no wait on other tiles; pass two scans the tile sizes (one
// workgroup per sum) and places the bytes (one wave per tile, with the index
// entries that depend on stream offsets).
// A stage area for R tiles holds R stream offsets (u64), R records, then R
// slots.
struct StageRec {
    uint32_t total;          // encoded bytes of the tile
    uint16_t e16, e32, e48;  // offsets of the tile's blocks 16, 32 and 48 inside the tile
    uint16_t pad;
    uint32_t seed;           // summed seed (a segment's first tile)
};
static_assert(sizeof(StageRec) == 16, "stage records are 16 bytes");
constexpr uint64_t kStageSlot = 9600;  // 64 blocks of at most 150 bytes

ZK_HD uint64_t stage_area_bytes(uint64_t tiles) { return tiles * (8 + sizeof(StageRec) + kStageSlot); }

// Single sum (a.job) and uniform batch.
hipError_t launch_homsum_lane_staged(const HomsumArgs& a, uint8_t* stage, hipStream_t s);
hipError_t launch_homsum_lane_batch_staged(const BatchArgs& b, uint8_t* stage, hipStream_t s);
hipError_t launch_case_count(const uint8_t* w1, const uint8_t* w2, uint64_t nblocks,
    unsigned long long* counts, hipStream_t s);
hipError_t launch_ws_init(WsHeader* ws, uint64_t capacity, hipStream_t s);
hipError_t launch_sz3_table(const Geom3& g, Part3* parts, unsigned long long* status, hipStream_t s);
hipError_t launch_sz3_compress(const Sz3Args& a, hipStream_t s);
hipError_t launch_sz3_decompress(const Sz3Args& a, hipStream_t s);

}  // namespace zk
