// zccl_capi.hip -- C ABI: the reference's drop-in entry points and the device
// API declared in include/zccl_amd.h.  Host code only; all stream work runs in
// the kernels of zccl_kernels.hip.
#include <dlfcn.h>
#include <hip/hip_runtime.h>
#include <sched.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <atomic>
#include <mutex>
#include <vector>

#include "zccl_amd.h"
#include "zccl_internal.h"
#include "zccl_kernels.h"

using namespace zk;

// ------------------------------------------------------------ status

static thread_local int g_last = ZCCL_OK;

static int set_last(int s)
{
    g_last = s;
    return s;
}

extern "C" int zccl_last_error(void) { return g_last; }

extern "C" const char* zccl_strerror(int status)
{
    switch (status) {
    case ZCCL_OK: return "ok";
    case ZCCL_EINVAL: return "invalid argument";
    case ZCCL_ENOSPC: return "output capacity too small";
    case ZCCL_ECORRUPT: return "malformed stream";
    case ZCCL_EDEVICE: return "HIP runtime error";
    case ZCCL_ENOMEM: return "device allocation failed";
    case ZCCL_EWORKSPACE: return "workspace too small or not initialised";
    default: return "unknown status";
    }
}

static int map_errors(unsigned long long e)
{
    if (e & ZK_ERR_CAPACITY)
        return ZCCL_EWORKSPACE;
    if (e & ZK_ERR_NOSPACE)
        return ZCCL_ENOSPC;
    if (e & (ZK_ERR_WIDTH | ZK_ERR_MISMATCH | ZK_ERR_OVERRUN))
        return ZCCL_ECORRUPT;
    if (e & ZK_ERR_SPIN)
        return ZCCL_EDEVICE;
    return ZCCL_OK;
}

#define HIP_TRY(expr)                                                                              \
    do {                                                                                           \
        if ((expr) != hipSuccess) {                                                                \
            (void) hipGetLastError();                                                              \
            return ZCCL_EDEVICE;                                                                   \
        }                                                                                          \
    } while (0)

// ------------------------------------------------------------ chunk count

static std::atomic<int> g_chunks_override{0};

// The reference's chunk count is the size of its OpenMP team, i.e.
// omp_get_max_threads() of libgomp in the calling process
// (ZCCL/src/ZCCL_float.c:533).  Ask libgomp itself so both libraries agree.
static int omp_team_size()
{
    using fn_t = int (*)(void);
    static fn_t fn = []() -> fn_t {
        void* h = dlopen("libgomp.so.1", RTLD_LAZY | RTLD_NOLOAD);
        if (!h)
            h = dlopen("libgomp.so.1", RTLD_LAZY);
        return h ? (fn_t) dlsym(h, "omp_get_max_threads") : nullptr;
    }();
    if (fn)
        return fn();
    const char* env = getenv("OMP_NUM_THREADS");
    if (env && atoi(env) > 0)
        return atoi(env);
    cpu_set_t set;
    if (sched_getaffinity(0, sizeof(set), &set) == 0)
        return CPU_COUNT(&set);
    return 1;
}

extern "C" void zccl_set_num_chunks(int nchunks) { g_chunks_override = nchunks > 0 ? nchunks : 0; }

extern "C" int zccl_get_num_chunks(void)
{
    int o = g_chunks_override.load();
    return o > 0 ? o : omp_team_size();
}

// ------------------------------------------------------------ tunables

// Workgroup size targets per operation (0 compress, 1 homomorphic sum,
// 2 decompress); each is capped by the kernel's compiled launch bound.
static std::atomic<int> g_threads[3] = {{(int) kCompressThreads}, {(int) kHomsumThreads}, {(int) kDecompressThreads}};
static const int kThreadsMax[3] = {(int) kCompressThreads, (int) kHomsumThreads, (int) kDecompressThreads};

static uint32_t op_threads(int op) { return (uint32_t) g_threads[op].load(); }

extern "C" int zccl_tune_threads(int op, int threads)
{
    if (op < 0 || op > 2)
        return ZCCL_EINVAL;
    if (threads == 0)
        threads = kThreadsMax[op];
    if (threads < 64 || threads > kThreadsMax[op] || threads % 64)
        return ZCCL_EINVAL;
    g_threads[op].store(threads);
    return ZCCL_OK;
}

extern "C" int zccl_tune_get_threads(int op) { return op < 0 || op > 2 ? 0 : g_threads[op].load(); }

// Layout of compress (op 0), the homomorphic sum (op 1) and decompress (op
// 2) for 36-value blocks: 0 one lane per block (zccl_lane.hip, the default),
// 1 workgroup tiles (slot layout); for the sum, 2 one lane per block in
// single-pass workgroup tiles of four waves.
static std::atomic<int> g_layout[3] = {{0}, {0}, {0}};

extern "C" int zccl_tune_layout(int op, int layout)
{
    if (op < 0 || op > 2 || layout < 0 || layout > (op == 1 ? 2 : 1))
        return ZCCL_EINVAL;
    g_layout[op].store(layout);
    return ZCCL_OK;
}

extern "C" int zccl_tune_get_layout(int op) { return op < 0 || op > 2 ? -1 : g_layout[op].load(); }

extern "C" int zccl_tune_stage_bytes(unsigned int bytes) { return set_stage_min_bytes(bytes) ? ZCCL_OK : ZCCL_EINVAL; }

extern "C" unsigned int zccl_tune_get_stage_bytes(void) { return get_stage_min_bytes(); }

// Tiling of operation op for geometry g.
static Tiling op_tiling(int op, const Geom& g)
{
    if (((op == 0 && g_layout[0].load() == 0) || (op == 1 && g_layout[1].load() == 2)) && lane_layout_ok(g))
        return make_lane_wg_tiling(g);
    if ((op == 1 || op == 2) && g_layout[op].load() == 0 && lane_layout_ok(g))
        return make_lane_tiling(g);
    return make_tiling(g, op_threads(op));
}

static bool is_lane_tiling(const Geom& g, const Tiling& tl) { return lane_layout_ok(g) && tl.threads == 64 && tl.G == kLaneTileBlocks; }

// The lane-layout compress tiling (a workgroup per kLaneWaves x 64 blocks).
static bool is_lane_wg_tiling(const Geom& g, const Tiling& tl)
{
    return lane_layout_ok(g) && tl.threads == 64 * kLaneWaves && tl.G == kLaneWaves * kLaneTileBlocks;
}

// ------------------------------------------------------------ sizes

static bool geom_of(size_t n, int L, int T, Geom* g)
{
    return L > 0 && T > 0 && make_geom((uint64_t) n, (uint32_t) L, (uint32_t) T, g);
}

extern "C" size_t zccl_worst_case_bytes(size_t n, int L, int T)
{
    Geom g;
    return geom_of(n, L, T, &g) ? (size_t) worst_case_bytes(g) : 0;
}

extern "C" size_t zccl_index_bytes(size_t n, int L, int T)
{
    Geom g;
    return geom_of(n, L, T, &g) ? (size_t) g.ix_bytes : 0;
}

// Tiles of the finest tiling any kernel uses for (n, L, T): a workspace of
// this many tiles serves every operation on such a stream.
extern "C" size_t zccl_num_tiles(size_t n, int L, int T)
{
    Geom g;
    if (!geom_of(n, L, T, &g))
        return 0;
    uint64_t a = op_tiling(0, g).ntiles, b = op_tiling(1, g).ntiles, c = op_tiling(2, g).ntiles;
    uint64_t m = std::max(a, std::max(b, c));
    if (lane_layout_ok(g))
        m = std::max<uint64_t>(m, make_lane_tiling(g).ntiles);  // either layout of the sum
    return (size_t) m;
}

extern "C" size_t zccl_num_blocks(size_t n, int L, int T)
{
    Geom g;
    return geom_of(n, L, T, &g) ? (size_t) g.nblocks : 0;
}

extern "C" size_t zccl_index_size_offset(size_t n, int L, int T)
{
    Geom g;
    return geom_of(n, L, T, &g) ? (size_t) g.ix_meta : 0;
}

extern "C" size_t zccl_workspace_bytes(size_t max_tiles) { return (size_t) workspace_bytes(max_tiles); }

// ------------------------------------------------------------ device API

static int check_ws(void* ws, size_t ws_bytes, uint64_t tiles)
{
    if (!ws)
        return ZCCL_EWORKSPACE;
    if (ws_bytes < workspace_bytes(tiles))
        return ZCCL_EWORKSPACE;
    return ZCCL_OK;
}

static double inverse_bound(float eb)
{
    volatile float one = 1.0f;
    float invf = one / eb;  // FP32 division, as the reference's int / float
    return (double) invf;
}

static OpParams op_params(const Tiling& tl, double inv, float eb, int k)
{
    OpParams p;
    p.inv = inv;
    p.eb = eb;
    p.k = k;
    p.stage_bytes = stage_bytes_for(tl, k);
    return p;
}

extern "C" int zccl_workspace_init(void* ws, size_t ws_bytes, void* stream)
{
    if (!ws || ws_bytes < sizeof(WsHeader))
        return ZCCL_EINVAL;
    uint64_t cap = workspace_capacity(ws_bytes);
    HIP_TRY(launch_ws_init((WsHeader*) ws, cap, (hipStream_t) stream));
    return ZCCL_OK;
}

extern "C" int zccl_workspace_errors(void* ws, unsigned long long* errors, int clear, void* stream)
{
    if (!ws || !errors)
        return ZCCL_EINVAL;
    HIP_TRY(hipStreamSynchronize((hipStream_t) stream));
    HIP_TRY(hipMemcpy(errors, &((WsHeader*) ws)->errors, 8, hipMemcpyDeviceToHost));
    if (clear && *errors)
        HIP_TRY(hipMemset(&((WsHeader*) ws)->errors, 0, 8));
    return ZCCL_OK;
}

extern "C" int zccl_compress_dev(const float* x, size_t n, float eb, int L, int T, unsigned char* out,
    size_t out_cap, void* index, unsigned long long* d_size, void* ws, size_t ws_bytes, void* stream)
{
    Geom g;
    if (!x || !out || !index || !geom_of(n, L, T, &g))
        return ZCCL_EINVAL;
    CompressArgs a;
    a.job.g = g;
    a.job.tl = op_tiling(0, g);
    int rc = check_ws(ws, ws_bytes, a.job.tl.ntiles);
    if (rc)
        return rc;
    a.job.p = op_params(a.job.tl, inverse_bound(eb), eb, 1);
    a.job.io.x = x;
    a.job.io.out = out;
    a.job.io.out_cap = out_cap;
    a.job.io.index = (uint8_t*) index;
    a.job.io.size_out = d_size;
    a.ws = (WsHeader*) ws;
    if (is_lane_wg_tiling(g, a.job.tl))
        HIP_TRY(launch_compress_lane(a, (hipStream_t) stream));
    else
        HIP_TRY(launch_compress(a, (hipStream_t) stream));
    return ZCCL_OK;
}

extern "C" int zccl_decompress_dev(const unsigned char* in, const void* index, size_t n, float eb,
    int L, int T, float* out, void* ws, size_t ws_bytes, void* stream)
{
    Geom g;
    if (!in || !index || !out || !geom_of(n, L, T, &g))
        return ZCCL_EINVAL;
    DecompressArgs a;
    a.job.g = g;
    a.job.tl = op_tiling(2, g);
    int rc = check_ws(ws, ws_bytes, a.job.tl.ntiles);
    if (rc)
        return rc;
    a.job.p = op_params(a.job.tl, 0.0, eb, 1);
    a.job.io.in[0].bytes = in;
    a.job.io.in[0].index = (const uint8_t*) index;
    a.job.io.out = out;
    a.ws = (WsHeader*) ws;
    if (is_lane_tiling(g, a.job.tl))
        HIP_TRY(launch_decompress_lane(a, (hipStream_t) stream));
    else
        HIP_TRY(launch_decompress(a, (hipStream_t) stream));
    return ZCCL_OK;
}

static int homsum_args(const unsigned char* const* ins, const void* const* indices, int k, size_t n, int L,
    int T, unsigned char* out, size_t out_cap, void* index, unsigned long long* d_size, void* ws, size_t ws_bytes,
    HomsumArgs* a)
{
    Geom g;
    if (!ins || !indices || k < 1 || k > kMaxOperands || !out || !index || !geom_of(n, L, T, &g))
        return ZCCL_EINVAL;
    memset(a, 0, sizeof(*a));
    a->job.g = g;
    a->job.tl = op_tiling(1, g);
    int rc = check_ws(ws, ws_bytes, a->job.tl.ntiles);
    if (rc)
        return rc;
    a->job.p = op_params(a->job.tl, 0.0, 0.0f, k);
    for (int r = 0; r < k; r++) {
        if (!ins[r] || !indices[r] || ins[r] == out)
            return ZCCL_EINVAL;
        a->job.io.in[r].bytes = ins[r];
        a->job.io.in[r].index = (const uint8_t*) indices[r];
    }
    a->job.io.out = out;
    a->job.io.out_cap = out_cap;
    a->job.io.index = (uint8_t*) index;
    a->job.io.size_out = d_size;
    a->ws = (WsHeader*) ws;
    return ZCCL_OK;
}

extern "C" int zccl_homsum_dev(const unsigned char* const* ins, const void* const* indices, int k,
    size_t n, int L, int T, unsigned char* out, size_t out_cap, void* index,
    unsigned long long* d_size, void* ws, size_t ws_bytes, void* stream)
{
    HomsumArgs a;
    int rc = homsum_args(ins, indices, k, n, L, T, out, out_cap, index, d_size, ws, ws_bytes, &a);
    if (rc)
        return rc;
    if (is_lane_tiling(a.job.g, a.job.tl))
        HIP_TRY(launch_homsum_lane(a, (hipStream_t) stream));
    else if (is_lane_wg_tiling(a.job.g, a.job.tl))
        HIP_TRY(launch_homsum_lane_wg(a, (hipStream_t) stream));
    else
        HIP_TRY(launch_homsum(a, (hipStream_t) stream));
    return ZCCL_OK;
}

extern "C" size_t zccl_sum_stage_bytes(size_t max_tiles) { return (size_t) stage_area_bytes(max_tiles); }

// The two-pass form runs for sums of at least this many tiles; smaller ones
// take the single-pass form (one launch instead of three).  Tunable.
constexpr uint64_t kStagedMinTiles = 1024;
static std::atomic<unsigned long long> g_staged_min_tiles{kStagedMinTiles};

extern "C" int zccl_tune_two_pass_min_tiles(unsigned long long min_tiles)
{
    g_staged_min_tiles.store(min_tiles ? min_tiles : kStagedMinTiles);
    return ZCCL_OK;
}

extern "C" unsigned long long zccl_tune_get_two_pass_min_tiles(void) { return g_staged_min_tiles.load(); }

extern "C" int zccl_tune_decompress_groups(int groups_per_cu)
{
    return set_decompress_groups_per_cu(groups_per_cu) ? ZCCL_OK : ZCCL_EINVAL;
}

extern "C" int zccl_tune_get_decompress_groups(void) { return get_decompress_groups_per_cu(); }

// A stage area serves the two-pass form when the sum is large enough and the
// area holds every tile of it and is 16-byte aligned.
static bool stage_ok(const void* stage, size_t stage_bytes, uint64_t tiles)
{
    return tiles >= g_staged_min_tiles.load() && stage && ((uintptr_t) stage & 15) == 0 &&
           stage_bytes >= stage_area_bytes(tiles);
}

extern "C" int zccl_homsum_dev_staged(const unsigned char* const* ins, const void* const* indices, int k,
    size_t n, int L, int T, unsigned char* out, size_t out_cap, void* index,
    unsigned long long* d_size, void* ws, size_t ws_bytes, void* stage, size_t stage_bytes, void* stream)
{
    HomsumArgs a;
    int rc = homsum_args(ins, indices, k, n, L, T, out, out_cap, index, d_size, ws, ws_bytes, &a);
    if (rc)
        return rc;
    if (is_lane_tiling(a.job.g, a.job.tl) && stage_ok(stage, stage_bytes, a.job.tl.ntiles))
        HIP_TRY(launch_homsum_lane_staged(a, (uint8_t*) stage, (hipStream_t) stream));
    else if (is_lane_tiling(a.job.g, a.job.tl))
        HIP_TRY(launch_homsum_lane(a, (hipStream_t) stream));
    else if (is_lane_wg_tiling(a.job.g, a.job.tl))
        HIP_TRY(launch_homsum_lane_wg(a, (hipStream_t) stream));
    else
        HIP_TRY(launch_homsum(a, (hipStream_t) stream));
    return ZCCL_OK;
}

// ------------------------------------------------------------ allocation-free sum
//
// zccl_homsum_dev without a caller workspace, for callers that cannot
// allocate (graph capture, serving loops).  The lookback workspace is this
// file-scope device array: each GPU holds its own copy until the process
// ends.  Every call on a GPU shares that copy, so calls that overlap on one
// GPU race; each call re-initialises the part it uses with kernels (no
// hipMemsetAsync), so a call never depends on what an earlier one left.

static constexpr size_t kScratchBytes = 1u << 20;  // serves launches of up to 65 520 tiles
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

static std::mutex g_scratch_mu;
static std::vector<void*> g_scratch_ptr;  // device address of `scratch`, per device (null: not yet looked up)

// Device address of this GPU's `scratch` (looked up once per device, with that
// device current); null when the runtime cannot provide it.
static void* scratch_ptr()
{
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) {
        (void) hipGetLastError();
        return nullptr;
    }
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if ((int) g_scratch_ptr.size() <= dev)
        g_scratch_ptr.resize(dev + 1, nullptr);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) {
            (void) hipGetLastError();
            return nullptr;
        }
        g_scratch_ptr[dev] = p;
    }
    return g_scratch_ptr[dev];
}

extern "C" int zccl_homsum_dev_noalloc(const unsigned char* const* ins, const void* const* indices, int k,
    size_t n, int L, int T, unsigned char* out, size_t out_cap, void* index, unsigned long long* d_size,
    void* stream)
{
    Geom g;
    if (!geom_of(n, L, T, &g))
        return ZCCL_EINVAL;
    const uint64_t tiles = op_tiling(1, g).ntiles;
    const uint64_t need = workspace_bytes(tiles);
    if (need > kScratchBytes)
        return ZCCL_EWORKSPACE;  // beyond the fixed scratch: use zccl_homsum_dev with a workspace
    void* ws = scratch_ptr();
    if (!ws)
        return ZCCL_EDEVICE;
    HIP_TRY(launch_ws_init((WsHeader*) ws, tiles, (hipStream_t) stream));
    return zccl_homsum_dev(ins, indices, k, n, L, T, out, out_cap, index, d_size, ws, (size_t) need, stream);
}

// ------------------------------------------------------------ batch plans
//
// A plan is built on the host: jobs are validated and their device
// descriptors packed into plan-owned host memory ([IO or Job] x njobs, then
// njobs+1 u64 tile starts).  The descriptors reach device memory either by the
// owning create() calls (hipMalloc + copy) or by the caller (any async copy)
// followed by zccl_batch_bind(); the latter makes no allocating or
// synchronising HIP call, which a serving loop with work in flight needs.

struct zccl_batch {
    BatchKind kind;
    BatchArgs args;
    uint32_t threads;
    size_t lds;
    std::vector<unsigned char> host;  // packed descriptors
    size_t tile0_off;                 // byte offset of the tile starts in `host`
    void* owned;                      // device copy owned by the plan (create()), else null
    bool bound;
    bool lane;                        // compress, homomorphic sum or decompress in the lane layout
    bool wg;                          // homomorphic sum in single-pass workgroup tiles (layout 2)
};

namespace {

template <typename IO, typename Job>
zccl_batch* batch_plan(BatchKind kind, const std::vector<Job>& jobs, uint32_t chains, int* status)
{
    const size_t nj = jobs.size();
    bool uniform = true;
    for (size_t j = 1; j < nj; j++) {
        const Job& a = jobs[0];
        const Job& b = jobs[j];
        if (memcmp(&a.g, &b.g, sizeof(Geom)) || memcmp(&a.tl, &b.tl, sizeof(Tiling)) ||
            memcmp(&a.p, &b.p, sizeof(OpParams)))
            uniform = false;
    }
    zccl_batch* b = new zccl_batch();
    b->kind = kind;
    b->owned = nullptr;
    b->bound = false;
    b->wg = kind == BatchKind::Homsum && nj > 0 && is_lane_wg_tiling(jobs[0].g, jobs[0].tl);
    b->lane = nj > 0 && (kind == BatchKind::Compress ? is_lane_wg_tiling(jobs[0].g, jobs[0].tl)
                         : kind == BatchKind::Homsum
                             ? is_lane_tiling(jobs[0].g, jobs[0].tl) || is_lane_wg_tiling(jobs[0].g, jobs[0].tl)
                             : is_lane_tiling(jobs[0].g, jobs[0].tl));
    uint32_t threads = 64;
    size_t lds = 0;
    std::vector<unsigned long long> tile0(nj + 1, 0);
    for (size_t j = 0; j < nj; j++) {
        tile0[j + 1] = tile0[j] + jobs[j].tl.ntiles;
        threads = std::max(threads, jobs[j].tl.threads);
        size_t need = kind == BatchKind::Compress ? compress_lds_bytes(jobs[j].tl)
                      : kind == BatchKind::Homsum ? operand_lds_bytes(jobs[j].tl, jobs[j].p.k, jobs[j].p.stage_bytes)
                                                  : operand_lds_bytes(jobs[j].tl, 1, jobs[j].p.stage_bytes);
        lds = std::max(lds, need);
    }
    b->threads = threads;
    b->lds = lds;
    const size_t desc = uniform ? nj * sizeof(IO) : nj * sizeof(Job);
    b->tile0_off = (size_t) align_up(desc, 16);
    b->host.assign(b->tile0_off + 8 * (nj + 1), 0);
    if (uniform) {
        for (size_t j = 0; j < nj; j++)
            memcpy(b->host.data() + j * sizeof(IO), &jobs[j].io, sizeof(IO));
    } else {
        memcpy(b->host.data(), jobs.data(), desc);
    }
    memcpy(b->host.data() + b->tile0_off, tile0.data(), 8 * (nj + 1));
    BatchArgs& a = b->args;
    memset(&a, 0, sizeof(a));
    a.njobs = (uint32_t) nj;
    a.uniform = uniform;
    a.interleave = uniform && nj > 1;
    a.total_chains = chains;
    a.total_tiles = tile0[nj];
    for (size_t j = 0; j < nj; j++)
        a.max_segments = std::max(a.max_segments, jobs[j].g.S);
    if (uniform) {
        a.g = jobs[0].g;
        a.tl = jobs[0].tl;
        a.p = jobs[0].p;
    }
    *status = ZCCL_OK;
    return b;
}

// Owning create(): device copy of the descriptors (allocates, synchronises).
zccl_batch* batch_own(zccl_batch* b, int* status)
{
    if (!b)
        return nullptr;
    if (hipMalloc(&b->owned, b->host.size()) != hipSuccess) {
        (void) hipGetLastError();
        b->owned = nullptr;
        zccl_batch_destroy(b);
        *status = ZCCL_ENOMEM;
        return nullptr;
    }
    if (hipMemcpy(b->owned, b->host.data(), b->host.size(), hipMemcpyHostToDevice) != hipSuccess) {
        (void) hipGetLastError();
        zccl_batch_destroy(b);
        *status = ZCCL_EDEVICE;
        return nullptr;
    }
    zccl_batch_bind(b, b->owned);
    return b;
}

int* status_or(int* status, int* dummy) { return status ? status : dummy; }

}  // namespace

extern "C" zccl_batch* zccl_batch_homsum_plan(int njobs, const int* ks, const size_t* nbEle, int L, int T,
    const unsigned char* const* ins, const void* const* indices, unsigned char* const* outs,
    const size_t* out_caps, void* const* out_indices, int* status)
{
    int dummy;
    int* st = status_or(status, &dummy);
    if (njobs < 1 || !ks || !nbEle || !ins || !indices || !outs || !out_caps || !out_indices) {
        *st = ZCCL_EINVAL;
        return nullptr;
    }
    std::vector<HomsumJob> jobs(njobs);
    size_t pos = 0;
    for (int j = 0; j < njobs; j++) {
        HomsumJob& h = jobs[j];
        memset(&h, 0, sizeof(h));
        const int k = ks[j];
        if (k < 1 || k > kMaxOperands || !geom_of(nbEle[j], L, T, &h.g) || !outs[j] || !out_indices[j]) {
            *st = ZCCL_EINVAL;
            return nullptr;
        }
        h.tl = op_tiling(1, h.g);
        h.p = op_params(h.tl, 0.0, 0.0f, k);
        for (int r = 0; r < k; r++, pos++) {
            if (!ins[pos] || !indices[pos] || ins[pos] == outs[j]) {
                *st = ZCCL_EINVAL;
                return nullptr;
            }
            h.io.in[r].bytes = ins[pos];
            h.io.in[r].index = (const uint8_t*) indices[pos];
        }
        h.io.out = outs[j];
        h.io.out_cap = out_caps[j];
        h.io.index = (uint8_t*) out_indices[j];
        h.io.size_out = nullptr;
    }
    return batch_plan<HomsumIO>(BatchKind::Homsum, jobs, (uint32_t) njobs, st);
}

extern "C" zccl_batch* zccl_batch_homsum_create(int njobs, int k, const size_t* nbEle, int L, int T,
    const unsigned char* const* ins, const void* const* indices, unsigned char* const* outs,
    const size_t* out_caps, void* const* out_indices, int* status)
{
    int dummy;
    int* st = status_or(status, &dummy);
    if (njobs < 1) {
        *st = ZCCL_EINVAL;
        return nullptr;
    }
    std::vector<int> ks((size_t) njobs, k);
    return batch_own(zccl_batch_homsum_plan(njobs, ks.data(), nbEle, L, T, ins, indices, outs, out_caps,
                         out_indices, st), st);
}

extern "C" zccl_batch* zccl_batch_compress_plan(int njobs, const size_t* nbEle, const float* ebs, int L, int T,
    const float* const* xs, unsigned char* const* outs, const size_t* out_caps, void* const* out_indices,
    int* status)
{
    int dummy;
    int* st = status_or(status, &dummy);
    if (njobs < 1 || !nbEle || !ebs || !xs || !outs || !out_caps || !out_indices) {
        *st = ZCCL_EINVAL;
        return nullptr;
    }
    std::vector<CompressJob> jobs(njobs);
    for (int j = 0; j < njobs; j++) {
        CompressJob& c = jobs[j];
        memset(&c, 0, sizeof(c));
        if (!geom_of(nbEle[j], L, T, &c.g) || !xs[j] || !outs[j] || !out_indices[j]) {
            *st = ZCCL_EINVAL;
            return nullptr;
        }
        c.tl = op_tiling(0, c.g);
        c.p = op_params(c.tl, inverse_bound(ebs[j]), ebs[j], 1);
        c.io.x = xs[j];
        c.io.out = outs[j];
        c.io.out_cap = out_caps[j];
        c.io.index = (uint8_t*) out_indices[j];
        c.io.size_out = nullptr;
    }
    return batch_plan<CompressIO>(BatchKind::Compress, jobs, (uint32_t) njobs, st);
}

extern "C" zccl_batch* zccl_batch_compress_create(int njobs, const size_t* nbEle, const float* ebs, int L,
    int T, const float* const* xs, unsigned char* const* outs, const size_t* out_caps,
    void* const* out_indices, int* status)
{
    int dummy;
    int* st = status_or(status, &dummy);
    return batch_own(zccl_batch_compress_plan(njobs, nbEle, ebs, L, T, xs, outs, out_caps, out_indices, st), st);
}

extern "C" zccl_batch* zccl_batch_decompress_plan(int njobs, const size_t* nbEle, const float* ebs, int L,
    int T, const unsigned char* const* ins, const void* const* indices, float* const* outs, int* status)
{
    int dummy;
    int* st = status_or(status, &dummy);
    if (njobs < 1 || !nbEle || !ebs || !ins || !indices || !outs) {
        *st = ZCCL_EINVAL;
        return nullptr;
    }
    std::vector<DecompressJob> jobs(njobs);
    uint32_t chains = 0;
    for (int j = 0; j < njobs; j++) {
        DecompressJob& d = jobs[j];
        memset(&d, 0, sizeof(d));
        if (!geom_of(nbEle[j], L, T, &d.g) || !ins[j] || !indices[j] || !outs[j]) {
            *st = ZCCL_EINVAL;
            return nullptr;
        }
        d.tl = op_tiling(2, d.g);
        d.p = op_params(d.tl, 0.0, ebs[j], 1);
        d.io.in[0].bytes = ins[j];
        d.io.in[0].index = (const uint8_t*) indices[j];
        d.io.out = outs[j];
        chains += d.g.S;
    }
    return batch_plan<DecompressIO>(BatchKind::Decompress, jobs, chains, st);
}

extern "C" zccl_batch* zccl_batch_decompress_create(int njobs, const size_t* nbEle, const float* ebs,
    int L, int T, const unsigned char* const* ins, const void* const* indices, float* const* outs,
    int* status)
{
    int dummy;
    int* st = status_or(status, &dummy);
    return batch_own(zccl_batch_decompress_plan(njobs, nbEle, ebs, L, T, ins, indices, outs, st), st);
}

extern "C" size_t zccl_batch_desc_bytes(const zccl_batch* b) { return b ? b->host.size() : 0; }

extern "C" const void* zccl_batch_desc_host(const zccl_batch* b) { return b ? b->host.data() : nullptr; }

extern "C" int zccl_batch_bind(zccl_batch* b, void* dev_desc)
{
    if (!b || !dev_desc || ((uintptr_t) dev_desc & 15))
        return ZCCL_EINVAL;
    b->args.jobs = dev_desc;
    b->args.job_tile0 = (const unsigned long long*) ((unsigned char*) dev_desc + b->tile0_off);
    b->bound = true;
    return ZCCL_OK;
}

extern "C" size_t zccl_batch_tiles(const zccl_batch* b) { return b ? (size_t) b->args.total_tiles : 0; }

extern "C" int zccl_batch_launch(const zccl_batch* b, void* ws, size_t ws_bytes, void* stream)
{
    if (!b || !b->bound)
        return ZCCL_EINVAL;
    int rc = check_ws(ws, ws_bytes, b->args.total_tiles);
    if (rc)
        return rc;
    BatchArgs a = b->args;
    a.ws = (WsHeader*) ws;
    if (b->wg)
        HIP_TRY(launch_homsum_lane_wg_batch(a, (hipStream_t) stream));
    else if (b->lane && b->kind == BatchKind::Homsum)
        HIP_TRY(launch_homsum_lane_batch(a, (hipStream_t) stream));
    else if (b->lane && b->kind == BatchKind::Compress)
        HIP_TRY(launch_compress_lane_batch(a, (hipStream_t) stream));
    else if (b->lane)
        HIP_TRY(launch_decompress_lane_batch(a, (hipStream_t) stream));
    else
        HIP_TRY(launch_batch(b->kind, a, b->threads, b->lds, (hipStream_t) stream));
    return ZCCL_OK;
}

extern "C" int zccl_batch_launch_staged(const zccl_batch* b, void* ws, size_t ws_bytes, void* stage,
    size_t stage_bytes, void* stream)
{
    if (!b || !b->bound)
        return ZCCL_EINVAL;
    if (!(b->lane && !b->wg && b->kind == BatchKind::Homsum && b->args.uniform &&
          stage_ok(stage, stage_bytes, b->args.total_tiles)))
        return zccl_batch_launch(b, ws, ws_bytes, stream);
    int rc = check_ws(ws, ws_bytes, b->args.total_tiles);
    if (rc)
        return rc;
    BatchArgs a = b->args;
    a.ws = (WsHeader*) ws;
    HIP_TRY(launch_homsum_lane_batch_staged(a, (uint8_t*) stage, (hipStream_t) stream));
    return ZCCL_OK;
}

extern "C" void zccl_batch_destroy(zccl_batch* b)
{
    if (!b)
        return;
    if (b->owned)
        (void) hipFree(b->owned);
    delete b;
}

extern "C" int zccl_parse_dev(const unsigned char* in, size_t in_cap, size_t n, int L, int T,
    void* index, unsigned long long* d_size, void* ws, size_t ws_bytes, void* stream)
{
    Geom g;
    if (!in || !index || !geom_of(n, L, T, &g))
        return ZCCL_EINVAL;
    if (!ws || ws_bytes < sizeof(WsHeader))
        return ZCCL_EWORKSPACE;
    ParseArgs a;
    a.g = g;
    a.bytes = in;
    a.in_cap = in_cap;
    a.index = (uint8_t*) index;
    a.ws = (WsHeader*) ws;
    a.ws_bytes = ws_bytes;
    a.size_out = d_size;
    HIP_TRY(launch_parse(a, (hipStream_t) stream));
    return ZCCL_OK;
}

extern "C" size_t zccl_parse_workspace_bytes(size_t n, int L, int T)
{
    Geom g;
    if (!geom_of(n, L, T, &g))
        return 0;
    const size_t base = (size_t) workspace_bytes(1024);
    return parse_parallel_ok(g) ? std::max<size_t>(base, (size_t) parse_layout(g).bytes) : base;
}

extern "C" int zccl_case_count_dev(const void* index1, const void* index2, size_t n, int L, int T,
    unsigned long long* d_counts4, void* stream)
{
    Geom g;
    if (!index1 || !index2 || !d_counts4 || !geom_of(n, L, T, &g))
        return ZCCL_EINVAL;
    HIP_TRY(launch_case_count((const uint8_t*) index1, (const uint8_t*) index2, g.nblocks, d_counts4,
        (hipStream_t) stream));
    return ZCCL_OK;
}

// ------------------------------------------------------------ drop-in context

namespace {

struct DevBuf {
    void* p = nullptr;
    size_t cap = 0;
};

struct Ctx {
    std::mutex mu;
    hipStream_t stream = nullptr;
    DevBuf x, out, in[2], ix[3], ws, small;
    uint64_t ws_tiles = 0;
};

int ensure(DevBuf& b, size_t bytes)
{
    if (b.cap >= bytes && b.p)
        return ZCCL_OK;
    if (b.p) {
        (void) hipFree(b.p);
        b.p = nullptr;
        b.cap = 0;
    }
    size_t want = std::max<size_t>(bytes, 256);
    if (hipMalloc(&b.p, want) != hipSuccess) {
        (void) hipGetLastError();
        b.p = nullptr;
        return ZCCL_ENOMEM;
    }
    b.cap = want;
    return ZCCL_OK;
}

Ctx* ctx_current(int* rc)
{
    static std::mutex mu;
    static std::vector<Ctx*> ctxs;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) {
        (void) hipGetLastError();
        *rc = ZCCL_EDEVICE;
        return nullptr;
    }
    std::lock_guard<std::mutex> lk(mu);
    if ((int) ctxs.size() <= dev)
        ctxs.resize(dev + 1, nullptr);
    if (!ctxs[dev]) {
        Ctx* c = new Ctx();
        if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
            (void) hipGetLastError();
            delete c;
            *rc = ZCCL_EDEVICE;
            return nullptr;
        }
        ctxs[dev] = c;
    }
    *rc = ZCCL_OK;
    return ctxs[dev];
}

// A workspace of at least `tiles` tiles that also holds the parallel parse's
// tables for the geometry (n, L, T) when that parse applies.
int ensure_ws(Ctx& c, uint64_t tiles, size_t n = 0, int L = 0, int T = 0)
{
    Geom g;
    if (n && geom_of(n, L, T, &g) && parse_parallel_ok(g))
        tiles = std::max<uint64_t>(tiles, workspace_capacity(parse_layout(g).bytes) + 1);
    if (c.ws.p && c.ws_tiles >= tiles)
        return ZCCL_OK;
    uint64_t want = std::max<uint64_t>(tiles, 1024);
    int rc = ensure(c.ws, workspace_bytes(want));
    if (rc)
        return rc;
    HIP_TRY(launch_ws_init((WsHeader*) c.ws.p, want, c.stream));
    c.ws_tiles = want;
    return ZCCL_OK;
}

// Device pointer?  If so, *avail receives the bytes from p to the end of its
// allocation (so kernels never read or write past it).
bool device_ptr(const void* p, size_t* avail)
{
    hipPointerAttribute_t a;
    if (hipPointerGetAttributes(&a, p) != hipSuccess) {
        (void) hipGetLastError();
        return false;
    }
    if (a.type != hipMemoryTypeDevice)
        return false;
    hipDeviceptr_t base = nullptr;
    size_t size = 0;
    if (hipMemGetAddressRange(&base, &size, (hipDeviceptr_t) p) != hipSuccess) {
        (void) hipGetLastError();
        *avail = 0;
        return true;
    }
    *avail = (size_t) ((char*) base + size - (char*) p);
    return true;
}

// Synchronise, then turn the workspace's sticky error bits into a status
// (and clear them for the next call).
int finish(Ctx& c)
{
    HIP_TRY(hipStreamSynchronize(c.stream));
    unsigned long long e = 0;
    HIP_TRY(hipMemcpy(&e, &((WsHeader*) c.ws.p)->errors, 8, hipMemcpyDeviceToHost));
    if (e)
        HIP_TRY(hipMemset(&((WsHeader*) c.ws.p)->errors, 0, 8));
    return map_errors(e);
}

unsigned long long* small_slot(Ctx& c, int i) { return (unsigned long long*) c.small.p + i; }

// Make stream bytes available on the device and build their index.  A host
// stream is copied exactly: cap bytes when its size is stated (cap_exact),
// else the size its own header and width bytes prove.  For a device stream
// the allocation bounds the read.
int stage_stream(Ctx& c, const Geom& g, const unsigned char* bytes, size_t cap, bool cap_exact,
    int slot, const unsigned char** dptr)
{
    size_t avail = 0;
    const bool dev = device_ptr(bytes, &avail);
    if (dev) {
        if (8ull * g.T > avail)
            return ZCCL_ECORRUPT;
        cap = cap_exact ? std::min(cap, avail) : avail;
    } else if (cap_exact) {
        if (cap < 8ull * g.T)
            return ZCCL_ECORRUPT;
    } else {
        cap = (size_t) host_stream_bytes(bytes, g);
        if (!cap)
            return ZCCL_ECORRUPT;
    }
    const unsigned char* d = bytes;
    if (!dev) {
        int rc = ensure(c.in[slot], cap + 16);
        if (rc)
            return rc;
        HIP_TRY(hipMemcpyAsync(c.in[slot].p, bytes, cap, hipMemcpyHostToDevice, c.stream));
        d = (const unsigned char*) c.in[slot].p;
    }
    int rc = ensure(c.ix[slot], g.ix_bytes);
    if (rc)
        return rc;
    ParseArgs a;
    a.g = g;
    a.bytes = d;
    a.in_cap = cap;
    a.index = (uint8_t*) c.ix[slot].p;
    a.ws = (WsHeader*) c.ws.p;
    a.ws_bytes = c.ws.cap;
    a.size_out = nullptr;
    HIP_TRY(launch_parse(a, c.stream));
    *dptr = d;
    return ZCCL_OK;
}

int dropin_compress(unsigned char* outputBytes, const float* x, size_t* outSize, float eb, size_t n,
    int L, int T)
{
    if (!outputBytes || !x || !outSize)
        return ZCCL_EINVAL;
    *outSize = 0;
    Geom g;
    if (!geom_of(n, L, T, &g))
        return ZCCL_EINVAL;
    int rc;
    Ctx* c = ctx_current(&rc);
    if (!c)
        return rc;
    std::lock_guard<std::mutex> lk(c->mu);
    if ((rc = ensure_ws(*c, zccl_num_tiles(n, L, T), n, L, T)) || (rc = ensure(c->small, 64)) || (rc = ensure(c->ix[0], g.ix_bytes)))
        return rc;
    size_t avail_x = 0, avail_o = 0;
    bool xdev = device_ptr(x, &avail_x), odev = device_ptr(outputBytes, &avail_o);
    const float* dx = x;
    if (!xdev) {
        if ((rc = ensure(c->x, 4 * n)))
            return rc;
        HIP_TRY(hipMemcpyAsync(c->x.p, x, 4 * n, hipMemcpyHostToDevice, c->stream));
        dx = (const float*) c->x.p;
    } else if (avail_x < 4 * n) {
        return ZCCL_EINVAL;
    }
    unsigned char* dout = outputBytes;
    size_t dcap = avail_o;
    if (!odev) {
        dcap = worst_case_bytes(g);
        if ((rc = ensure(c->out, dcap)))
            return rc;
        dout = (unsigned char*) c->out.p;
    }
    rc = zccl_compress_dev(dx, n, eb, L, T, dout, dcap, c->ix[0].p, small_slot(*c, 0), c->ws.p,
        c->ws.cap, c->stream);
    if (rc)
        return rc;
    if ((rc = finish(*c)))
        return rc;
    unsigned long long size = 0;
    HIP_TRY(hipMemcpy(&size, small_slot(*c, 0), 8, hipMemcpyDeviceToHost));
    if (!odev)
        HIP_TRY(hipMemcpy(outputBytes, dout, size, hipMemcpyDeviceToHost));
    *outSize = size;
    return ZCCL_OK;
}

int dropin_decompress(float* newData, size_t n, float eb, int L, int T, const unsigned char* cmp,
    size_t cmpSize, bool sized)
{
    if (!newData || !cmp)
        return ZCCL_EINVAL;
    Geom g;
    if (!geom_of(n, L, T, &g))
        return ZCCL_EINVAL;
    int rc;
    Ctx* c = ctx_current(&rc);
    if (!c)
        return rc;
    std::lock_guard<std::mutex> lk(c->mu);
    if ((rc = ensure_ws(*c, zccl_num_tiles(n, L, T), n, L, T)))
        return rc;
    const unsigned char* din = nullptr;
    if ((rc = stage_stream(*c, g, cmp, cmpSize, sized, 0, &din)))
        return rc;
    size_t avail = 0;
    bool odev = device_ptr(newData, &avail);
    float* dout = newData;
    if (odev && avail < 4 * n)
        return ZCCL_EINVAL;
    if (!odev) {
        if ((rc = ensure(c->x, 4 * n)))
            return rc;
        dout = (float*) c->x.p;
    }
    rc = zccl_decompress_dev(din, c->ix[0].p, n, eb, L, T, dout, c->ws.p, c->ws.cap, c->stream);
    if (rc)
        return rc;
    if ((rc = finish(*c)))
        return rc;
    if (!odev)
        HIP_TRY(hipMemcpy(newData, dout, 4 * n, hipMemcpyDeviceToHost));
    return ZCCL_OK;
}

int dropin_add(unsigned char* out, size_t out_cap, bool out_cap_known, size_t* out_size, bool accumulate,
    size_t n, int L, int T, const unsigned char* a, size_t asz, const unsigned char* b, size_t bsz,
    bool sized, int* counters)
{
    if (!out || !out_size || !a || !b)
        return ZCCL_EINVAL;
    Geom g;
    if (!geom_of(n, L, T, &g))
        return ZCCL_EINVAL;
    int rc;
    Ctx* c = ctx_current(&rc);
    if (!c)
        return rc;
    std::lock_guard<std::mutex> lk(c->mu);
    if ((rc = ensure_ws(*c, zccl_num_tiles(n, L, T), n, L, T)) || (rc = ensure(c->small, 64)) || (rc = ensure(c->ix[2], g.ix_bytes)))
        return rc;
    const unsigned char* da = nullptr;
    const unsigned char* db = nullptr;
    if ((rc = stage_stream(*c, g, a, asz, sized, 0, &da)) || (rc = stage_stream(*c, g, b, bsz, sized, 1, &db)))
        return rc;
    uint64_t cap = worst_case_bytes(g);
    if ((rc = ensure(c->out, cap)))
        return rc;
    const unsigned char* ins[2] = {da, db};
    const void* ixs[2] = {c->ix[0].p, c->ix[1].p};
    rc = zccl_homsum_dev(ins, ixs, 2, n, L, T, (unsigned char*) c->out.p, cap, c->ix[2].p,
        small_slot(*c, 0), c->ws.p, c->ws.cap, c->stream);
    if (rc)
        return rc;
    if (counters)
        HIP_TRY(launch_case_count((const uint8_t*) c->ix[0].p, (const uint8_t*) c->ix[1].p, g.nblocks,
            small_slot(*c, 2), c->stream));
    if ((rc = finish(*c)))
        return rc;
    unsigned long long size = 0;
    HIP_TRY(hipMemcpy(&size, small_slot(*c, 0), 8, hipMemcpyDeviceToHost));
    size_t avail = 0;
    bool odev = device_ptr(out, &avail);
    if ((odev && size > avail) || (out_cap_known && size > out_cap))
        return ZCCL_ENOSPC;
    HIP_TRY(hipMemcpy(out, c->out.p, size, odev ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost));
    if (counters) {
        unsigned long long cnt[4];
        HIP_TRY(hipMemcpy(cnt, small_slot(*c, 2), 32, hipMemcpyDeviceToHost));
        for (int i = 0; i < 4; i++)
            counters[i] += (int) cnt[i];
    }
    *out_size = accumulate ? *out_size + size : size;
    return ZCCL_OK;
}

}  // namespace

// ------------------------------------------------------------ drop-in ABI

extern "C" void ZCCL_float_openmp_threadblock_arg(unsigned char* outputBytes, float* oriData,
    size_t* outSize, float absErrBound, size_t nbEle, int blockSize)
{
    set_last(dropin_compress(outputBytes, oriData, outSize, absErrBound, nbEle, blockSize,
        zccl_get_num_chunks()));
}

extern "C" void ZCCL_float_single_thread_arg(unsigned char* outputBytes, float* oriData,
    size_t* outSize, float absErrBound, size_t nbEle, int blockSize)
{
    set_last(dropin_compress(outputBytes, oriData, outSize, absErrBound, nbEle, blockSize, 1));
}

extern "C" unsigned char* ZCCL_float_openmp_threadblock(float* oriData, size_t* outSize,
    float absErrBound, size_t nbEle, int blockSize)
{
    int T = zccl_get_num_chunks();
    size_t cap = zccl_worst_case_bytes(nbEle, blockSize, T);
    if (!cap || !outSize) {
        set_last(ZCCL_EINVAL);
        return nullptr;
    }
    unsigned char* out = (unsigned char*) malloc(cap);
    if (!out) {
        set_last(ZCCL_ENOMEM);
        return nullptr;
    }
    int rc = dropin_compress(out, oriData, outSize, absErrBound, nbEle, blockSize, T);
    set_last(rc);
    if (rc) {
        free(out);
        return nullptr;
    }
    return out;
}

extern "C" void ZCCL_float_decompress_openmp_threadblock_arg(float* newData, size_t nbEle,
    float absErrBound, int blockSize, unsigned char* cmpBytes)
{
    set_last(dropin_decompress(newData, nbEle, absErrBound, blockSize, zccl_get_num_chunks(), cmpBytes, 0,
        false));
}

extern "C" void ZCCL_float_decompress_single_thread_arg(float* newData, size_t nbEle,
    float absErrBound, int blockSize, unsigned char* cmpBytes)
{
    set_last(dropin_decompress(newData, nbEle, absErrBound, blockSize, 1, cmpBytes, 0, false));
}

extern "C" void ZCCL_float_decompress_openmp_threadblock(float** newData, size_t nbEle,
    float absErrBound, int blockSize, unsigned char* cmpBytes)
{
    if (!newData) {
        set_last(ZCCL_EINVAL);
        return;
    }
    *newData = (float*) malloc(sizeof(float) * (nbEle ? nbEle : 1));
    if (!*newData) {
        set_last(ZCCL_ENOMEM);
        return;
    }
    set_last(dropin_decompress(*newData, nbEle, absErrBound, blockSize, zccl_get_num_chunks(), cmpBytes, 0,
        false));
}

extern "C" void ZCCL_float_homomophic_add_openmp_threadblock(unsigned char* final_cmpBytes,
    size_t* final_cmpSize, size_t nbEle, float absErrBound, int blockSize, unsigned char* cmpBytes,
    unsigned char* cmpBytes2)
{
    (void) absErrBound;  // unused by the reference too
    set_last(dropin_add(final_cmpBytes, 0, false, final_cmpSize, true, nbEle, blockSize,
        zccl_get_num_chunks(), cmpBytes, 0, cmpBytes2, 0, false, nullptr));
}

extern "C" void ZCCL_float_homomophic_add_single_thread(unsigned char* final_cmpBytes,
    size_t* final_cmpSize, size_t nbEle, float absErrBound, int blockSize, unsigned char* cmpBytes,
    unsigned char* cmpBytes2)
{
    (void) absErrBound;
    set_last(dropin_add(final_cmpBytes, 0, false, final_cmpSize, true, nbEle, blockSize, 1, cmpBytes, 0,
        cmpBytes2, 0, false, nullptr));
}

extern "C" void ZCCL_float_homomophic_add_openmp_threadblock_record(unsigned char* final_cmpBytes,
    size_t* final_cmpSize, size_t nbEle, float absErrBound, int blockSize, unsigned char* cmpBytes,
    unsigned char* cmpBytes2, int* total_counter_1, int* total_counter_2, int* total_counter_3,
    int* total_counter_4)
{
    (void) absErrBound;
    int cnt[4] = {0, 0, 0, 0};
    int rc = dropin_add(final_cmpBytes, 0, false, final_cmpSize, true, nbEle, blockSize,
        zccl_get_num_chunks(), cmpBytes, 0, cmpBytes2, 0, false, cnt);
    set_last(rc);
    if (rc == ZCCL_OK) {
        if (total_counter_1) *total_counter_1 += cnt[0];
        if (total_counter_2) *total_counter_2 += cnt[1];
        if (total_counter_3) *total_counter_3 += cnt[2];
        if (total_counter_4) *total_counter_4 += cnt[3];
    }
}

extern "C" int zccl_decompress_sized(float* newData, size_t nbEle, float absErrBound, int blockSize,
    int nchunks, const unsigned char* cmpBytes, size_t cmpSize)
{
    return set_last(dropin_decompress(newData, nbEle, absErrBound, blockSize, nchunks, cmpBytes, cmpSize,
        true));
}

extern "C" int zccl_homomorphic_add_sized(unsigned char* final_cmpBytes, size_t final_cap,
    size_t* final_size, size_t nbEle, int blockSize, int nchunks, const unsigned char* cmpBytes,
    size_t cmpSize, const unsigned char* cmpBytes2, size_t cmpSize2)
{
    return set_last(dropin_add(final_cmpBytes, final_cap, true, final_size, false, nbEle, blockSize, nchunks,
        cmpBytes, cmpSize, cmpBytes2, cmpSize2, true, nullptr));
}

// ------------------------------------------------------------ 3-D SZp codec

static bool geom3_of(size_t d1, size_t d2, size_t d3, int Bs, int T, Geom3* g, std::vector<Part3>* parts)
{
    if (T < 1 || T > kMaxParts3)
        return false;
    parts->resize((size_t) T);
    return make_geom3(d1, d2, d3, Bs, T, g, parts->data());
}

extern "C" int zccl_sz3d_sizes(size_t d1, size_t d2, size_t d3, int Bs, int T, size_t* worst_bytes,
    size_t* ws_bytes)
{
    Geom3 g;
    std::vector<Part3> parts;
    if (!geom3_of(d1, d2, d3, Bs, T, &g, &parts))
        return set_last(ZCCL_EINVAL);
    if (worst_bytes)
        *worst_bytes = (size_t) worst_case_bytes3(g, parts.data());
    if (ws_bytes)
        *ws_bytes = (size_t) ws3_layout(g.nblocks, g.T).bytes;
    return set_last(ZCCL_OK);
}

static int sz3_args(size_t d1, size_t d2, size_t d3, int Bs, int T, void* ws, size_t ws_bytes, Sz3Args* a,
    hipStream_t s)
{
    std::vector<Part3> parts;
    if (!ws || !geom3_of(d1, d2, d3, Bs, T, &a->g, &parts))
        return ZCCL_EINVAL;
    const Ws3 L = ws3_layout(a->g.nblocks, a->g.T);
    if (ws_bytes < L.bytes)
        return ZCCL_EWORKSPACE;
    char* w = (char*) ws;
    a->status = (unsigned long long*) w;
    a->words = (uint32_t*) (w + L.words);
    a->offs = (unsigned long long*) (w + L.offs);
    a->part_sums = (unsigned long long*) (w + L.sums);
    a->parts = (const Part3*) (w + L.parts);
    HIP_TRY(launch_sz3_table(a->g, (Part3*) a->parts, a->status, s));
    return ZCCL_OK;
}

extern "C" int zccl_sz3d_compress_dev(const float* x, size_t d1, size_t d2, size_t d3, int Bs, double eb, int T,
    unsigned char* out, size_t out_cap, void* ws, size_t ws_bytes, unsigned long long* d_size, void* stream)
{
    if (!x || !out || !(eb > 0))
        return set_last(ZCCL_EINVAL);
    Sz3Args a = {};
    int rc = sz3_args(d1, d2, d3, Bs, T, ws, ws_bytes, &a, (hipStream_t) stream);
    if (rc)
        return set_last(rc);
    a.x = x;
    a.out = out;
    a.out_cap = out_cap;
    a.inv = 0.5 / eb;  // SZp_compress3D_fast_openmp's inver_eb (ZCCL/src/ZCCL_float.c:1869)
    a.size_out = d_size;
    if (launch_sz3_compress(a, (hipStream_t) stream) != hipSuccess) {
        (void) hipGetLastError();
        return set_last(ZCCL_EDEVICE);
    }
    return set_last(ZCCL_OK);
}

extern "C" int zccl_sz3d_decompress_dev(const unsigned char* in, size_t in_bytes, size_t d1, size_t d2, size_t d3,
    int Bs, double eb, int T, float* out, void* ws, size_t ws_bytes, void* stream)
{
    if (!in || !out || !(eb > 0) || in_bytes < 8ull * (T > 0 ? T : 1))
        return set_last(ZCCL_EINVAL);
    Sz3Args a = {};
    int rc = sz3_args(d1, d2, d3, Bs, T, ws, ws_bytes, &a, (hipStream_t) stream);
    if (rc)
        return set_last(rc);
    a.in = in;
    a.in_size = in_bytes;
    a.y = out;
    a.tw = 2 * eb;  // twice_eb (ZCCL/src/ZCCLd_float.c:2885)
    if (launch_sz3_decompress(a, (hipStream_t) stream) != hipSuccess) {
        (void) hipGetLastError();
        return set_last(ZCCL_EDEVICE);
    }
    return set_last(ZCCL_OK);
}

extern "C" int zccl_sz3d_status(const void* ws, size_t* stream_bytes)
{
    if (!ws)
        return set_last(ZCCL_EINVAL);
    unsigned long long st[2] = {0, 0};
    if (hipMemcpy(st, ws, sizeof(st), hipMemcpyDeviceToHost) != hipSuccess) {
        (void) hipGetLastError();
        return set_last(ZCCL_EDEVICE);
    }
    if (stream_bytes)
        *stream_bytes = (size_t) st[0];
    return set_last(map_errors(st[1]));
}

namespace {

struct Ctx3 {
    std::mutex mu;
    hipStream_t stream = nullptr;
    DevBuf x, out, ws;
};

Ctx3* ctx3_current(int* rc)
{
    static std::mutex mu;
    static std::vector<Ctx3*> ctxs;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) {
        (void) hipGetLastError();
        *rc = ZCCL_EDEVICE;
        return nullptr;
    }
    std::lock_guard<std::mutex> lk(mu);
    if ((int) ctxs.size() <= dev)
        ctxs.resize(dev + 1, nullptr);
    if (!ctxs[dev]) {
        Ctx3* c = new Ctx3();
        if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
            (void) hipGetLastError();
            delete c;
            *rc = ZCCL_EDEVICE;
            return nullptr;
        }
        ctxs[dev] = c;
    }
    *rc = ZCCL_OK;
    return ctxs[dev];
}

int dropin_compress3(const float* x, unsigned char* out, size_t d1, size_t d2, size_t d3, int Bs, double eb,
    size_t* size)
{
    if (!x || !out || !size)
        return ZCCL_EINVAL;
    *size = 0;  // ZCCL/src/ZCCL_float.c:1881
    const int T = zccl_get_num_chunks();
    Geom3 g;
    std::vector<Part3> parts;
    if (!geom3_of(d1, d2, d3, Bs, T, &g, &parts))
        return ZCCL_EINVAL;
    const uint64_t n = (uint64_t) d1 * d2 * d3, worst = worst_case_bytes3(g, parts.data());
    int rc;
    Ctx3* c = ctx3_current(&rc);
    if (!c)
        return rc;
    std::lock_guard<std::mutex> lk(c->mu);
    if ((rc = ensure(c->ws, ws3_layout(g.nblocks, g.T).bytes)))
        return rc;
    size_t avail_x = 0, avail_o = 0;
    const bool xdev = device_ptr(x, &avail_x), odev = device_ptr(out, &avail_o);
    const float* dx = x;
    if (!xdev) {
        if ((rc = ensure(c->x, 4 * n)))
            return rc;
        HIP_TRY(hipMemcpyAsync(c->x.p, x, 4 * n, hipMemcpyHostToDevice, c->stream));
        dx = (const float*) c->x.p;
    } else if (avail_x < 4 * n) {
        return ZCCL_EINVAL;
    }
    unsigned char* dout = out;
    size_t cap = avail_o;
    if (!odev) {
        if ((rc = ensure(c->out, worst)))
            return rc;
        dout = (unsigned char*) c->out.p;
        cap = worst;
    }
    if ((rc = zccl_sz3d_compress_dev(dx, d1, d2, d3, Bs, eb, T, dout, cap, c->ws.p, c->ws.cap, nullptr, c->stream)))
        return rc;
    HIP_TRY(hipStreamSynchronize(c->stream));
    size_t bytes = 0;
    if ((rc = zccl_sz3d_status(c->ws.p, &bytes)))
        return rc;
    if (!odev)
        HIP_TRY(hipMemcpy(out, dout, bytes, hipMemcpyDeviceToHost));
    *size = bytes;
    return ZCCL_OK;
}

// Bytes of a host stream: the last sub-volume's offset from the header, then
// its width bytes give its payload size (ZCCL/src/ZCCLd_float.c:2946-2951).
uint64_t host_stream_bytes3(const unsigned char* in, const Geom3& g, const std::vector<Part3>& parts,
    uint64_t worst)
{
    const Part3& p = parts[g.T - 1];
    uint64_t off;
    memcpy(&off, in + 8ull * (g.T - 1), 8);
    const uint64_t nb = (uint64_t) p.b1 * p.b2 * p.b3;
    if (off > worst || 8ull * g.T + off + nb > worst)
        return 0;  // malformed header: the walk would leave the worst-case stream
    const uint64_t w0 = 8ull * g.T + off;
    uint64_t tot = w0 + nb, lb = 0;
    for (uint32_t bx = 0; bx < p.b1; bx++)
        for (uint32_t by = 0; by < p.b2; by++)
            for (uint32_t bz = 0; bz < p.b3; bz++, lb++) {
                const uint32_t ex = std::min(g.Bs, p.sx - bx * g.Bs), ey = std::min(g.Bs, p.sy - by * g.Bs),
                               ez = std::min(g.Bs, p.sz - bz * g.Bs);
                const uint32_t w = in[w0 + lb];
                tot += payload3(w > 32 ? 32 : w, ex * ey * ez);
            }
    return tot <= worst ? tot : 0;
}

int dropin_decompress3(float* y, const unsigned char* in, size_t d1, size_t d2, size_t d3, int Bs, double eb)
{
    if (!y || !in)
        return ZCCL_EINVAL;
    const int T = zccl_get_num_chunks();
    Geom3 g;
    std::vector<Part3> parts;
    if (!geom3_of(d1, d2, d3, Bs, T, &g, &parts))
        return ZCCL_EINVAL;
    const uint64_t n = (uint64_t) d1 * d2 * d3;
    int rc;
    Ctx3* c = ctx3_current(&rc);
    if (!c)
        return rc;
    std::lock_guard<std::mutex> lk(c->mu);
    if ((rc = ensure(c->ws, ws3_layout(g.nblocks, g.T).bytes)))
        return rc;
    size_t avail_i = 0, avail_y = 0;
    const bool idev = device_ptr(in, &avail_i), ydev = device_ptr(y, &avail_y);
    const unsigned char* din = in;
    uint64_t in_bytes = avail_i;
    if (!idev) {
        size_t worst = 0, wsb = 0;
        if ((rc = zccl_sz3d_sizes(d1, d2, d3, Bs, T, &worst, &wsb)))
            return rc;
        in_bytes = host_stream_bytes3(in, g, parts, worst);
        if (!in_bytes)
            return ZCCL_ECORRUPT;
        if ((rc = ensure(c->out, in_bytes)))
            return rc;
        HIP_TRY(hipMemcpyAsync(c->out.p, in, in_bytes, hipMemcpyHostToDevice, c->stream));
        din = (const unsigned char*) c->out.p;
    }
    float* dy = y;
    if (!ydev) {
        if ((rc = ensure(c->x, 4 * n)))
            return rc;
        dy = (float*) c->x.p;
    } else if (avail_y < 4 * n) {
        return ZCCL_EINVAL;
    }
    if ((rc = zccl_sz3d_decompress_dev(din, in_bytes, d1, d2, d3, Bs, eb, T, dy, c->ws.p, c->ws.cap, c->stream)))
        return rc;
    HIP_TRY(hipStreamSynchronize(c->stream));
    if ((rc = zccl_sz3d_status(c->ws.p, nullptr)))
        return rc;
    if (!ydev)
        HIP_TRY(hipMemcpy(y, dy, 4 * n, hipMemcpyDeviceToHost));
    return ZCCL_OK;
}

}  // namespace

extern "C" void SZp_compress3D_fast_openmp(const float* oriData, unsigned char* cmpData, size_t dim1, size_t dim2,
    size_t dim3, int blockSideLength, double errorBound, size_t* cmpSize)
{
    set_last(dropin_compress3(oriData, cmpData, dim1, dim2, dim3, blockSideLength, errorBound, cmpSize));
}

extern "C" void SZp_decompress3D_fast_openmp(float* decData, unsigned char* cmpData, size_t dim1, size_t dim2,
    size_t dim3, int blockSideLength, double errorBound)
{
    set_last(dropin_decompress3(decData, cmpData, dim1, dim2, dim3, blockSideLength, errorBound));
}
