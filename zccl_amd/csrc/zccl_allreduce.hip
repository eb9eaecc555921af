// zccl_allreduce.hip -- the hZCCL allreduce data flow of P simulated ranks
// on one device, as C host code over the library's own C ABI (no kernels
// here).
//
// The reference (MPI_Allreduce_ZCCL_RI2_mt_oa_ho_record,
// ZCCL/src/ZCCL_ring_ho.c:177-317) runs on each of P ranks: split `count`
// values into P segments of ceil(count/P) (the last short, :218-233),
// compress its P segments (:247-256), P-1 ring steps that add the received
// compressed partial sum to its own segment homomorphically (:258-290), a
// compressed ring allgather of the owned segments (:292, :12-175) and the
// decompression of all P segments (:163-169).
//
// Here the P ranks are P device send buffers and the run is three batch
// launches on the caller's stream:
//   1. one compress launch of all P x P (rank, segment) pairs;
//   2. one fused-sum launch: for every segment, the P ranks' streams summed
//      in one k = P sum (byte-equal to the ring's chain of pairwise adds: the
//      encoding of a sum is canonical), in the two-pass form when large;
//   3. one decompress launch of the P reduced segments into recvbuf.
// On one device every rank would receive the same reduced segments, so the
// allgather moves no bytes: recvbuf is what each rank would hold.
//
// Device memory comes from one caller workspace, laid out at create time;
// run() allocates nothing (it can be captured in a HIP graph once its
// buffers are fixed: plans are rebuilt only when send or receive pointers
// change, with a synchronous copy of their descriptors).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <cstring>
#include <vector>

#include "zccl_amd.h"

namespace {

constexpr size_t kAlign = 256;

size_t align(size_t x) { return (x + kAlign - 1) / kAlign * kAlign; }

// Segment sizes and offsets, ZCCL/src/ZCCL_ring_ho.c:218-233.
void segments(int P, size_t count, std::vector<size_t>& cnts, std::vector<size_t>& displs)
{
    cnts.assign(P, 0);
    displs.assign(P, 0);
    size_t total = 0;
    for (int i = 0; i < P; i++) {
        size_t c = (count + P - 1) / P;
        if (total + c > count) {
            cnts[i] = count - total;
            break;
        }
        cnts[i] = c;
        total += c;
    }
    for (int i = 1; i < P; i++)
        displs[i] = displs[i - 1] + cnts[i - 1];
}

// Where everything lives inside the workspace, for (P, count, L, T).
struct Layout {
    std::vector<size_t> cnts, displs;
    std::vector<size_t> cap;       // stream capacity of segment j
    std::vector<size_t> ixb;       // index bytes of segment j
    std::vector<size_t> stream_off;  // [r * P + j]: rank r's stream of segment j
    std::vector<size_t> index_off;
    std::vector<size_t> sum_off, sum_index_off;  // [j]
    size_t lookback_off = 0, lookback_bytes = 0;
    size_t stage_off = 0, stage_bytes = 0;
    size_t desc_off[3] = {0, 0, 0}, desc_bytes[3] = {0, 0, 0};
    size_t total = 0;
    size_t max_tiles = 0;
};

// Descriptor bytes of the three plans for this geometry (their sizes depend
// only on job counts and geometry, not on pointers).
bool plan_desc_bytes(int P, const Layout& l, int L, int T, size_t out[3]);

bool make_layout(int P, size_t count, int L, int T, Layout* l)
{
    segments(P, count, l->cnts, l->displs);
    size_t at = 0;
    l->cap.resize(P);
    l->ixb.resize(P);
    for (int j = 0; j < P; j++) {
        if (l->cnts[j] < (size_t) T)
            return false;  // every segment must hold at least T values
        l->cap[j] = zccl_worst_case_bytes(l->cnts[j], L, T);
        l->ixb[j] = zccl_index_bytes(l->cnts[j], L, T);
        if (!l->cap[j] || !l->ixb[j])
            return false;
        l->max_tiles = std::max(l->max_tiles, zccl_num_tiles(l->cnts[j], L, T));
    }
    l->stream_off.resize((size_t) P * P);
    l->index_off.resize((size_t) P * P);
    for (int r = 0; r < P; r++)
        for (int j = 0; j < P; j++) {
            l->stream_off[(size_t) r * P + j] = at;
            at += align(l->cap[j]);
            l->index_off[(size_t) r * P + j] = at;
            at += align(l->ixb[j]);
        }
    l->sum_off.resize(P);
    l->sum_index_off.resize(P);
    for (int j = 0; j < P; j++) {
        l->sum_off[j] = at;
        at += align(l->cap[j]);
        l->sum_index_off[j] = at;
        at += align(l->ixb[j]);
    }
    // every launch has at most P * P * max_tiles tiles (the compress batch)
    const size_t launch_tiles = (size_t) P * P * l->max_tiles;
    l->lookback_off = at;
    l->lookback_bytes = zccl_workspace_bytes(launch_tiles);
    at += align(l->lookback_bytes);
    l->stage_off = at;
    l->stage_bytes = zccl_sum_stage_bytes((size_t) P * l->max_tiles);
    at += align(l->stage_bytes);
    if (!plan_desc_bytes(P, *l, L, T, l->desc_bytes))
        return false;
    for (int i = 0; i < 3; i++) {
        l->desc_off[i] = at;
        at += align(l->desc_bytes[i]);
    }
    l->total = at;
    return true;
}

}  // namespace

struct zccl_allreduce {
    int P = 0, L = 0, T = 0;
    size_t count = 0;
    float eb = 0.0f;
    Layout lay;
    unsigned char* ws = nullptr;
    zccl_batch* plan[3] = {nullptr, nullptr, nullptr};  // compress, sum, decompress
    std::vector<const float*> last_send;
    float* last_recv = nullptr;
};

namespace {

unsigned char* at(const zccl_allreduce* a, size_t off) { return a->ws + off; }

void free_plan(zccl_allreduce* a, int i)
{
    if (a->plan[i]) {
        zccl_batch_destroy(a->plan[i]);
        a->plan[i] = nullptr;
    }
}

// Build plan i (0 compress, 1 sum, 2 decompress) against the workspace (or,
// with ws == nullptr, against placeholder pointers, to size descriptors).
zccl_batch* build_plan(int i, int P, const Layout& l, int L, int T, float eb, unsigned char* ws,
    const float* const* send, float* recv, int* st)
{
    // placeholders only need to be distinct, non-null and 16-byte aligned
    auto ptr = [&](size_t off) -> unsigned char* { return ws ? ws + off : (unsigned char*) (uintptr_t) (kAlign + off); };
    if (i == 0) {
        std::vector<size_t> n;
        std::vector<float> ebs;
        std::vector<const float*> xs;
        std::vector<unsigned char*> outs;
        std::vector<size_t> caps;
        std::vector<void*> ixs;
        for (int r = 0; r < P; r++)
            for (int j = 0; j < P; j++) {
                n.push_back(l.cnts[j]);
                ebs.push_back(eb);
                xs.push_back(send ? send[r] + l.displs[j] : (const float*) ptr(16 * (r + 1)));
                outs.push_back(ptr(l.stream_off[(size_t) r * P + j]));
                caps.push_back(l.cap[j]);
                ixs.push_back(ptr(l.index_off[(size_t) r * P + j]));
            }
        return zccl_batch_compress_plan(P * P, n.data(), ebs.data(), L, T, xs.data(), outs.data(), caps.data(),
            ixs.data(), st);
    }
    if (i == 1) {
        std::vector<int> ks(P, P);
        std::vector<size_t> n(l.cnts.begin(), l.cnts.end());
        std::vector<const unsigned char*> ins;
        std::vector<const void*> ixs;
        std::vector<unsigned char*> outs;
        std::vector<size_t> caps;
        std::vector<void*> oix;
        for (int j = 0; j < P; j++) {
            // the order of the operands does not change the bytes; the ring's
            // order (ZCCL_ring_ho.c:259-260) adds rank j+1's segment first
            for (int q = 0; q < P; q++) {
                const int r = (j + 1 + q) % P;
                ins.push_back(ptr(l.stream_off[(size_t) r * P + j]));
                ixs.push_back(ptr(l.index_off[(size_t) r * P + j]));
            }
            outs.push_back(ptr(l.sum_off[j]));
            caps.push_back(l.cap[j]);
            oix.push_back(ptr(l.sum_index_off[j]));
        }
        return zccl_batch_homsum_plan(P, ks.data(), n.data(), L, T, ins.data(), ixs.data(), outs.data(), caps.data(),
            oix.data(), st);
    }
    std::vector<size_t> n(l.cnts.begin(), l.cnts.end());
    std::vector<float> ebs(P, eb);
    std::vector<const unsigned char*> ins;
    std::vector<const void*> ixs;
    std::vector<float*> outs;
    for (int j = 0; j < P; j++) {
        ins.push_back(ptr(l.sum_off[j]));
        ixs.push_back(ptr(l.sum_index_off[j]));
        outs.push_back(recv ? recv + l.displs[j] : (float*) ptr(32 * (j + 1)));
    }
    return zccl_batch_decompress_plan(P, n.data(), ebs.data(), L, T, ins.data(), ixs.data(), outs.data(), st);
}

bool plan_desc_bytes(int P, const Layout& l, int L, int T, size_t out[3])
{
    for (int i = 0; i < 3; i++) {
        int st = 0;
        zccl_batch* b = build_plan(i, P, l, L, T, 1.0f, nullptr, nullptr, nullptr, &st);
        if (!b)
            return false;
        out[i] = zccl_batch_desc_bytes(b);
        zccl_batch_destroy(b);
    }
    return true;
}

// (Re)build plan i against the workspace and bind it to its descriptor area.
int bind_plan(zccl_allreduce* a, int i, const float* const* send, float* recv)
{
    free_plan(a, i);
    int st = 0;
    zccl_batch* b = build_plan(i, a->P, a->lay, a->L, a->T, a->eb, a->ws, send, recv, &st);
    if (!b)
        return st ? st : ZCCL_EINVAL;
    const size_t nb = zccl_batch_desc_bytes(b);
    if (nb > a->lay.desc_bytes[i]) {
        zccl_batch_destroy(b);
        return ZCCL_EWORKSPACE;
    }
    unsigned char* d = at(a, a->lay.desc_off[i]);
    if (hipMemcpy(d, zccl_batch_desc_host(b), nb, hipMemcpyHostToDevice) != hipSuccess) {
        (void) hipGetLastError();
        zccl_batch_destroy(b);
        return ZCCL_EDEVICE;
    }
    int rc = zccl_batch_bind(b, d);
    if (rc) {
        zccl_batch_destroy(b);
        return rc;
    }
    a->plan[i] = b;
    return ZCCL_OK;
}

}  // namespace

extern "C" size_t zccl_allreduce_workspace_bytes(int nranks, size_t count, int blockSize, int nchunks)
{
    if (nranks < 1 || count < 1 || nchunks < 1)
        return 0;
    Layout l;
    return make_layout(nranks, count, blockSize, nchunks, &l) ? l.total : 0;
}

extern "C" zccl_allreduce* zccl_allreduce_create(int nranks, size_t count, float absErrBound, int blockSize,
    int nchunks, void* workspace, size_t workspace_bytes, int* status)
{
    int dummy;
    int* st = status ? status : &dummy;
    *st = ZCCL_OK;
    if (nranks < 1 || nranks > 32 || count < 1 || nchunks < 1 || !(absErrBound > 0.0f) || !workspace ||
        ((uintptr_t) workspace % kAlign)) {
        *st = ZCCL_EINVAL;
        return nullptr;
    }
    zccl_allreduce* a = new zccl_allreduce();
    a->P = nranks;
    a->count = count;
    a->eb = absErrBound;
    a->L = blockSize;
    a->T = nchunks;
    a->ws = (unsigned char*) workspace;
    if (!make_layout(nranks, count, blockSize, nchunks, &a->lay)) {
        delete a;
        *st = ZCCL_EINVAL;
        return nullptr;
    }
    if (workspace_bytes < a->lay.total) {
        delete a;
        *st = ZCCL_EWORKSPACE;
        return nullptr;
    }
    int rc = zccl_workspace_init(at(a, a->lay.lookback_off), a->lay.lookback_bytes, nullptr);
    if (!rc && hipDeviceSynchronize() != hipSuccess) {
        (void) hipGetLastError();
        rc = ZCCL_EDEVICE;
    }
    if (!rc)
        rc = bind_plan(a, 1, nullptr, nullptr);  // the sum plan depends on the workspace only
    if (rc) {
        zccl_allreduce_destroy(a);
        *st = rc;
        return nullptr;
    }
    return a;
}

extern "C" int zccl_allreduce_run(zccl_allreduce* a, const float* const* sendbufs, float* recvbuf, void* stream)
{
    if (!a || !sendbufs || !recvbuf)
        return ZCCL_EINVAL;
    for (int r = 0; r < a->P; r++)
        if (!sendbufs[r])
            return ZCCL_EINVAL;
    const bool same_send = a->plan[0] && a->last_send.size() == (size_t) a->P &&
                           std::equal(a->last_send.begin(), a->last_send.end(), sendbufs);
    if (!same_send) {
        int rc = bind_plan(a, 0, sendbufs, nullptr);
        if (rc)
            return rc;
        a->last_send.assign(sendbufs, sendbufs + a->P);
    }
    if (!a->plan[2] || a->last_recv != recvbuf) {
        int rc = bind_plan(a, 2, nullptr, recvbuf);
        if (rc)
            return rc;
        a->last_recv = recvbuf;
    }
    void* lb = at(a, a->lay.lookback_off);
    const size_t lbb = a->lay.lookback_bytes;
    int rc = zccl_batch_launch(a->plan[0], lb, lbb, stream);
    if (!rc)
        rc = zccl_batch_launch_staged(a->plan[1], lb, lbb, at(a, a->lay.stage_off), a->lay.stage_bytes, stream);
    if (!rc)
        rc = zccl_batch_launch(a->plan[2], lb, lbb, stream);
    return rc;
}

extern "C" int zccl_allreduce_segment(const zccl_allreduce* a, int j, const unsigned char** bytes, const void** index,
    size_t* nbEle, size_t* offset)
{
    if (!a || j < 0 || j >= a->P)
        return ZCCL_EINVAL;
    if (bytes)
        *bytes = a->ws + a->lay.sum_off[j];
    if (index)
        *index = a->ws + a->lay.sum_index_off[j];
    if (nbEle)
        *nbEle = a->lay.cnts[j];
    if (offset)
        *offset = a->lay.displs[j];
    return ZCCL_OK;
}

extern "C" int zccl_allreduce_errors(const zccl_allreduce* a, unsigned long long* errors, int clear, void* stream)
{
    if (!a)
        return ZCCL_EINVAL;
    return zccl_workspace_errors(a->ws + a->lay.lookback_off, errors, clear, stream);
}

extern "C" void zccl_allreduce_destroy(zccl_allreduce* a)
{
    if (!a)
        return;
    for (int i = 0; i < 3; i++)
        free_plan(a, i);
    delete a;
}
