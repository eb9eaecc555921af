/*
 * zccl_amd.h -- C ABI of the MI355X (gfx950) SZp fastmode1 codec and hZCCL
 * homomorphic sum.  Library: zccl_amd/libzccl_amd.so.
 *
 * Two layers:
 *
 *  1. Drop-in entry points with the reference's own names and signatures.
 *     A program that links the reference's libZCCL for these calls can link
 *     this library instead.  Pointers may be host memory (copied through the
 *     GPU) or device memory (used in place).  Calls are synchronous.  Like the
 *     reference they return void; the status of the last call on the calling
 *     thread is available from zccl_last_error().  The chunk count T that
 *     shapes the stream (the reference's OpenMP team size,
 *     ZCCL/src/ZCCL_float.c:533) is omp_get_max_threads() of the process's
 *     libgomp, or the value set with zccl_set_num_chunks().
 *
 *  2. A device API (zccl_*_dev): device pointers only, explicit chunk count,
 *     asynchronous on a caller-supplied hipStream_t, caller-owned workspace.
 *     Every device stream travels with an index (block widths, tile and seed
 *     offsets, stream size) that the kernels use instead of walking the bytes.
 *
 * Streams are byte-identical to the reference's for the same input, error
 * bound, block length and chunk count.
 */
#ifndef ZCCL_AMD_H
#define ZCCL_AMD_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------ status */

#define ZCCL_OK 0
#define ZCCL_EINVAL (-1)     /* bad argument: nbEle < chunks, blockSize out of range, null pointer */
#define ZCCL_ENOSPC (-2)     /* output capacity too small */
#define ZCCL_ECORRUPT (-3)   /* stream malformed: width > 32, index/stream mismatch, overrun */
#define ZCCL_EDEVICE (-4)    /* HIP runtime error */
#define ZCCL_ENOMEM (-5)     /* device allocation failed */
#define ZCCL_EWORKSPACE (-6) /* workspace too small or not initialised */

int zccl_last_error(void);
const char *zccl_strerror(int status);

/* Chunk count used by the drop-in entry points; 0 restores the default
 * (omp_get_max_threads() of libgomp, i.e. the reference's team size). */
void zccl_set_num_chunks(int nchunks);
int zccl_get_num_chunks(void);

/* ------------------------------------------------- drop-in (reference ABI) */

/* ZCCL/include/ZCCL_float.h:31-32 (ZCCL/src/ZCCL_float.c:501-807) */
void ZCCL_float_openmp_threadblock_arg(unsigned char *outputBytes, float *oriData, size_t *outSize,
    float absErrBound, size_t nbEle, int blockSize);

/* ZCCL/include/ZCCL_float.h:34-35 (ZCCL/src/ZCCL_float.c:809-988): chunk count 1 */
void ZCCL_float_single_thread_arg(unsigned char *outputBytes, float *oriData, size_t *outSize,
    float absErrBound, size_t nbEle, int blockSize);

/* ZCCL/include/ZCCL_float.h:27-29: allocating form; result from malloc(),
 * caller frees.  Produces the same bytes as the _arg form (the reference's
 * allocating form mis-sizes the last chunk's remainder block,
 * ZCCL/src/ZCCL_float.c:452,456; that defect is not reproduced). */
unsigned char *ZCCL_float_openmp_threadblock(float *oriData, size_t *outSize, float absErrBound,
    size_t nbEle, int blockSize);

/* ZCCL/include/ZCCLd_float.h:21 (ZCCL/src/ZCCLd_float.c:538-754) */
void ZCCL_float_decompress_openmp_threadblock_arg(float *newData, size_t nbEle, float absErrBound,
    int blockSize, unsigned char *cmpBytes);

/* ZCCL/include/ZCCLd_float.h:23 (ZCCL/src/ZCCLd_float.c:247-374): chunk count 1 */
void ZCCL_float_decompress_single_thread_arg(float *newData, size_t nbEle, float absErrBound,
    int blockSize, unsigned char *cmpBytes);

/* ZCCL/include/ZCCLd_float.h:19: allocating form, *newData from malloc(). */
void ZCCL_float_decompress_openmp_threadblock(float **newData, size_t nbEle, float absErrBound,
    int blockSize, unsigned char *cmpBytes);

/* ZCCL/include/ZCCLd_float.h:29 (ZCCL/src/ZCCLd_float.c:988-1679).
 * *final_cmpSize is incremented by the stream size, as in the reference
 * (:1027, :1655).  final_cmpBytes may alias cmpBytes2 (in-place ring use,
 * ZCCL/src/ZCCL_ring_ho.c:281-287). */
void ZCCL_float_homomophic_add_openmp_threadblock(unsigned char *final_cmpBytes,
    size_t *final_cmpSize, size_t nbEle, float absErrBound, int blockSize, unsigned char *cmpBytes,
    unsigned char *cmpBytes2);

/* ZCCL/include/ZCCLd_float.h:33 (ZCCL/src/ZCCLd_float.c:1681-2066): chunk count 1 */
void ZCCL_float_homomophic_add_single_thread(unsigned char *final_cmpBytes, size_t *final_cmpSize,
    size_t nbEle, float absErrBound, int blockSize, unsigned char *cmpBytes,
    unsigned char *cmpBytes2);

/* ZCCL/include/ZCCLd_float.h:31 (ZCCL/src/ZCCLd_float.c:2067-2800): as the
 * add above, and the counters are incremented by the number of blocks where
 * both widths are zero / only the first is zero / only the second is zero /
 * neither is zero. */
void ZCCL_float_homomophic_add_openmp_threadblock_record(unsigned char *final_cmpBytes,
    size_t *final_cmpSize, size_t nbEle, float absErrBound, int blockSize, unsigned char *cmpBytes,
    unsigned char *cmpBytes2, int *total_counter_1, int *total_counter_2, int *total_counter_3,
    int *total_counter_4);

/* Drop-in variants with the stream size stated (no upper bound inferred from
 * the header): reads exactly cmpSize bytes of each stream.  No reference
 * counterpart; for callers that know their sizes. */
int zccl_decompress_sized(float *newData, size_t nbEle, float absErrBound, int blockSize,
    int nchunks, const unsigned char *cmpBytes, size_t cmpSize);
int zccl_homomorphic_add_sized(unsigned char *final_cmpBytes, size_t final_cap, size_t *final_size,
    size_t nbEle, int blockSize, int nchunks, const unsigned char *cmpBytes, size_t cmpSize,
    const unsigned char *cmpBytes2, size_t cmpSize2);

/* ------------------------------------------------------------ sizes */

size_t zccl_worst_case_bytes(size_t nbEle, int blockSize, int nchunks);
size_t zccl_index_bytes(size_t nbEle, int blockSize, int nchunks);
size_t zccl_num_tiles(size_t nbEle, int blockSize, int nchunks);
size_t zccl_num_blocks(size_t nbEle, int blockSize, int nchunks);
/* Byte offset, inside an index, of the u64 stream size. */
size_t zccl_index_size_offset(size_t nbEle, int blockSize, int nchunks);
/* Workspace able to serve launches of up to max_tiles tiles. */
size_t zccl_workspace_bytes(size_t max_tiles);

/* ------------------------------------------------------------ device API */

/* Zero a workspace (device memory of zccl_workspace_bytes(max_tiles) bytes)
 * before its first use.  A workspace serves one launch at a time: launches
 * that share one must be ordered on one stream. */
int zccl_workspace_init(void *workspace, size_t workspace_bytes, void *stream);
/* Synchronise `stream`, read (and optionally clear) the workspace's sticky
 * error bits: 1 lookback wait bound hit, 2 output capacity, 4 width > 32,
 * 8 index/stream mismatch, 16 stream overrun, 32 workspace capacity. */
int zccl_workspace_errors(void *workspace, unsigned long long *errors, int clear, void *stream);

/* Compress nbEle floats at x into out (capacity out_cap).  Writes the index
 * and, if d_size is not null, the stream size to device memory. */
int zccl_compress_dev(const float *x, size_t nbEle, float absErrBound, int blockSize, int nchunks,
    unsigned char *out, size_t out_cap, void *index, unsigned long long *d_size, void *workspace,
    size_t workspace_bytes, void *stream);

/* Decompress a device stream described by `index` into out[nbEle]. */
int zccl_decompress_dev(const unsigned char *in, const void *index, size_t nbEle, float absErrBound,
    int blockSize, int nchunks, float *out, void *workspace, size_t workspace_bytes, void *stream);

/* Fused homomorphic sum of k streams (1 <= k <= 32) sharing (nbEle,
 * blockSize, nchunks): one pass over all operands, output canonical (equal to
 * the reference's chain of k-1 pairwise adds in any order).  `ins` and
 * `indices` are host arrays of k device pointers.  `out` must not alias an
 * input. */
int zccl_homsum_dev(const unsigned char *const *ins, const void *const *indices, int k,
    size_t nbEle, int blockSize, int nchunks, unsigned char *out, size_t out_cap, void *index,
    unsigned long long *d_size, void *workspace, size_t workspace_bytes, void *stream);

/* No counterpart in the reference: the two-pass form of zccl_homsum_dev.
 * Pass one sums and encodes every tile into a caller stage area of
 * zccl_sum_stage_bytes(zccl_num_tiles(...)) bytes (16-byte aligned) without
 * waiting on other tiles; pass two scans the tile sizes and moves the bytes
 * to their stream offsets.  Same output as zccl_homsum_dev.  Used when
 * blockSize is 36, the sum has at least 1024 tiles (zccl_num_tiles) and the
 * stage area is large enough; otherwise this is zccl_homsum_dev. */
size_t zccl_sum_stage_bytes(size_t max_tiles);
int zccl_homsum_dev_staged(const unsigned char *const *ins, const void *const *indices, int k,
    size_t nbEle, int blockSize, int nchunks, unsigned char *out, size_t out_cap, void *index,
    unsigned long long *d_size, void *workspace, size_t workspace_bytes, void *stage, size_t stage_bytes,
    void *stream);

/* ------------------------------------------------------------ staged allreduce
 *
 * The hZCCL allreduce data flow (MPI_Allreduce_ZCCL_RI2_mt_oa_ho_record,
 * ZCCL/src/ZCCL_ring_ho.c:177-317) for nranks simulated ranks on one device:
 * segments of ceil(count/nranks) values (the last short), every rank's
 * segments compressed in one launch, each segment's nranks streams summed
 * homomorphically in one launch (the ring's pairwise chain gives the same
 * bytes), the reduced segments decompressed into recvbuf in one launch.  On
 * one device the compressed allgather moves nothing: recvbuf is what every
 * rank holds after it.  absErrBound is the reference's absolute bound (its
 * `compressionRatio` argument, :195); blockSize and nchunks as in the codec.
 * All device memory is one caller workspace of
 * zccl_allreduce_workspace_bytes() bytes (256-byte aligned); run() launches
 * on `stream`, allocates nothing and returns before the work completes
 * (kernel-side errors: zccl_allreduce_errors).  Every segment must hold at
 * least nchunks values; at most 32 ranks. */
typedef struct zccl_allreduce zccl_allreduce;
size_t zccl_allreduce_workspace_bytes(int nranks, size_t count, int blockSize, int nchunks);
zccl_allreduce *zccl_allreduce_create(int nranks, size_t count, float absErrBound, int blockSize, int nchunks,
    void *workspace, size_t workspace_bytes, int *status);
/* sendbufs: host array of nranks device pointers to count floats each;
 * recvbuf: device, count floats. */
int zccl_allreduce_run(zccl_allreduce *allreduce, const float *const *sendbufs, float *recvbuf, void *stream);
/* The reduced stream of segment j after a run (device bytes and index), its
 * element count and its offset in recvbuf. */
int zccl_allreduce_segment(const zccl_allreduce *allreduce, int j, const unsigned char **bytes,
    const void **index, size_t *nbEle, size_t *offset);
int zccl_allreduce_errors(const zccl_allreduce *allreduce, unsigned long long *errors, int clear, void *stream);
void zccl_allreduce_destroy(zccl_allreduce *allreduce);

/* No counterpart in the reference: a measurement aid.  Copies `bytes` (a
 * multiple of 16, 16-byte aligned buffers) from src to dst with 16-byte
 * global loads (mode 0) or 16-byte buffer loads into LDS (mode 1, the load
 * path of the sum kernels), so rocprofv3's byte counters can be scaled
 * against a launch of known traffic.  0 on success. */
int zccl_calib_copy(const void *src, void *dst, size_t bytes, int mode, void *stream);

/* No counterpart in the reference: zccl_homsum_dev for callers that cannot
 * allocate (code under HIP graph capture, tight serving loops).  It takes no
 * workspace and allocates nothing: the lookback state lives in a fixed 1 MiB
 * device array inside the library, one per GPU, so geometries needing more
 * than 65 520 tiles (about 1.5e8 values at blockSize 36) return
 * ZCCL_EWORKSPACE.  Calls that overlap on one GPU race on that array.
 * Kernel-side errors (workspace_errors bits) are not reported; a sum that did
 * not fit out_cap leaves size 0 in the index and *d_size. */
int zccl_homsum_dev_noalloc(const unsigned char *const *ins, const void *const *indices, int k,
    size_t nbEle, int blockSize, int nchunks, unsigned char *out, size_t out_cap, void *index,
    unsigned long long *d_size, void *stream);

/* Rebuild the index of a device stream that has none (e.g. produced by the
 * CPU library); reads at most in_cap bytes of `in`.  The reference walks
 * each chunk sequentially (ZCCL/src/ZCCLd_float.c:597-745); with a workspace
 * of zccl_parse_workspace_bytes() bytes and chunks of at least 256 blocks
 * (blockSize <= 64) the walk runs in parallel over 4 KB tiles of every
 * chunk, otherwise one wave walks each chunk. */
int zccl_parse_dev(const unsigned char *in, size_t in_cap, size_t nbEle, int blockSize,
    int nchunks, void *index, unsigned long long *d_size, void *workspace, size_t workspace_bytes,
    void *stream);
size_t zccl_parse_workspace_bytes(size_t nbEle, int blockSize, int nchunks);

/* ------------------------------------------------------------ batches
 *
 * A batch plan puts many independent operations of one kind into a single
 * launch: every job keeps its own stream geometry and lookback chain, a
 * block-to-job table maps workgroups to jobs.  create() validates the jobs
 * and copies their descriptors to device memory once (it synchronises);
 * launch() is one kernel launch on `stream` that may be repeated or captured
 * in a hipGraph while the buffers the plan names stay alive.  The workspace
 * must hold zccl_batch_tiles() tiles.  No reference counterpart: the
 * reference runs each codec call separately. */
typedef struct zccl_batch zccl_batch;

/* njobs fused sums of k operands each; ins / indices are job-major
 * (job j operand r at [j*k + r]). */
zccl_batch *zccl_batch_homsum_create(int njobs, int k, const size_t *nbEle, int blockSize,
    int nchunks, const unsigned char *const *ins, const void *const *indices,
    unsigned char *const *outs, const size_t *out_caps, void *const *out_indices, int *status);
zccl_batch *zccl_batch_compress_create(int njobs, const size_t *nbEle, const float *absErrBounds,
    int blockSize, int nchunks, const float *const *xs, unsigned char *const *outs,
    const size_t *out_caps, void *const *out_indices, int *status);
zccl_batch *zccl_batch_decompress_create(int njobs, const size_t *nbEle, const float *absErrBounds,
    int blockSize, int nchunks, const unsigned char *const *ins, const void *const *indices,
    float *const *outs, int *status);
size_t zccl_batch_tiles(const zccl_batch *batch);
int zccl_batch_launch(const zccl_batch *batch, void *workspace, size_t workspace_bytes, void *stream);
/* The two-pass form (see zccl_homsum_dev_staged) of a homomorphic-sum batch
 * whose jobs share one geometry with blockSize 36 and that has at least 1024
 * tiles, given a stage area of zccl_sum_stage_bytes(zccl_batch_tiles())
 * bytes; any other batch launches as zccl_batch_launch. */
int zccl_batch_launch_staged(const zccl_batch *batch, void *workspace, size_t workspace_bytes, void *stage,
    size_t stage_bytes, void *stream);
void zccl_batch_destroy(zccl_batch *batch);

/* Host-only plans: validate and pack the descriptors without any HIP call.
 * The caller copies zccl_batch_desc_bytes() bytes from zccl_batch_desc_host()
 * to 16-byte aligned device memory (any stream-ordered copy) and binds that
 * address before launching; destroy() then frees host memory only.  The
 * homomorphic-sum plan takes a per-job operand count ks[j]; ins / indices
 * hold job 0's ks[0] operands, then job 1's, and so on. */
zccl_batch *zccl_batch_homsum_plan(int njobs, const int *ks, const size_t *nbEle, int blockSize,
    int nchunks, const unsigned char *const *ins, const void *const *indices,
    unsigned char *const *outs, const size_t *out_caps, void *const *out_indices, int *status);
zccl_batch *zccl_batch_compress_plan(int njobs, const size_t *nbEle, const float *absErrBounds,
    int blockSize, int nchunks, const float *const *xs, unsigned char *const *outs,
    const size_t *out_caps, void *const *out_indices, int *status);
zccl_batch *zccl_batch_decompress_plan(int njobs, const size_t *nbEle, const float *absErrBounds,
    int blockSize, int nchunks, const unsigned char *const *ins, const void *const *indices,
    float *const *outs, int *status);
size_t zccl_batch_desc_bytes(const zccl_batch *batch);
const void *zccl_batch_desc_host(const zccl_batch *batch);
int zccl_batch_bind(zccl_batch *batch, void *device_desc);

/* ------------------------------------------------------------ tuning
 *
 * Workgroup size target (threads, a multiple of 64) of operation op: 0
 * compress (at most 512), 1 homomorphic sum (at most 256), 2 decompress (at
 * most 512); 0 restores the default (the maximum).  Affects launches and
 * plans created afterwards, and zccl_num_tiles(); never the stream bytes. */
int zccl_tune_threads(int op, int threads);
int zccl_tune_get_threads(int op);

/* Layout of compress (op 0), the homomorphic sum (op 1) and decompress (op 2)
 * for 36-value blocks: 0 one lane per block (the default; a wave per 64-block
 * tile, compress a workgroup of four waves per 256-block tile); 1 workgroup
 * tiles of the zccl_tune_threads size; 2 (sum only) one lane per block in
 * single-pass workgroup tiles of four waves (no two-pass form).  Affects launches and
 * plans created afterwards, and zccl_num_tiles() never shrinks below either
 * layout's need; never the stream bytes.  No reference counterpart. */
int zccl_tune_layout(int op, int layout);
int zccl_tune_get_layout(int op);

/* Smallest operand stage, in bytes, of the slot-layout homomorphic sum
 * (layout 1): a multiple of 32 up to 65536, 0 for the default.  Larger
 * stages take more operands per staging round and fewer workgroups per CU.
 * Affects launches and plans created afterwards.  No reference counterpart. */
int zccl_tune_stage_bytes(unsigned int bytes);
unsigned int zccl_tune_get_stage_bytes(void);

/* Sums of at least this many tiles take the two-pass form when a stage area
 * is given (zccl_homsum_dev_staged, zccl_batch_launch_staged); smaller ones
 * the single-pass form.  0 restores the default (1024).  No reference
 * counterpart. */
int zccl_tune_two_pass_min_tiles(unsigned long long min_tiles);
unsigned long long zccl_tune_get_two_pass_min_tiles(void);

/* Workgroups per CU of the persistent lane-layout decompress grid (1..5; 0
 * restores the default, 4).  No reference counterpart. */
int zccl_tune_decompress_groups(int groups_per_cu);
int zccl_tune_get_decompress_groups(void);

/* Block-case histogram of two indexed streams (see the _record add). */
int zccl_case_count_dev(const void *index1, const void *index2, size_t nbEle, int blockSize,
    int nchunks, unsigned long long *d_counts4, void *stream);

/* ------------------------------------------------------------ 3-D SZp codec
 *
 * Volume dim1 x dim2 x dim3 (dim3 fastest), cubic blocks of side
 * blockSideLength (1..64), absolute bound errorBound, split over the OpenMP
 * team size T as the reference splits it over its threads (the drop-ins take
 * T from omp_get_max_threads(), or zccl_set_num_chunks()). */

/* ZCCL/include/ZCCL_float.h:98-101 (ZCCL/src/ZCCL_float.c:1862-2048).
 * *cmpSize is set to the stream size; oriData and cmpData may be host or
 * device pointers. */
void SZp_compress3D_fast_openmp(const float *oriData, unsigned char *cmpData, size_t dim1, size_t dim2,
    size_t dim3, int blockSideLength, double errorBound, size_t *cmpSize);

/* ZCCL/include/ZCCLd_float.h:45-48 (ZCCL/src/ZCCLd_float.c:2873-3040). */
void SZp_decompress3D_fast_openmp(float *decData, unsigned char *cmpData, size_t dim1, size_t dim2,
    size_t dim3, int blockSideLength, double errorBound);

/* Worst-case stream bytes and device workspace bytes of one 3-D call. */
int zccl_sz3d_sizes(size_t dim1, size_t dim2, size_t dim3, int blockSideLength, int nchunks,
    size_t *worst_bytes, size_t *ws_bytes);
/* Device API: device pointers, asynchronous on `stream`.  The stream size is
 * written to *d_size (device, optional) and to the workspace's status, which
 * zccl_sz3d_status() reads (blocking) together with the sticky error bits. */
int zccl_sz3d_compress_dev(const float *x, size_t dim1, size_t dim2, size_t dim3, int blockSideLength,
    double errorBound, int nchunks, unsigned char *out, size_t out_cap, void *ws, size_t ws_bytes,
    unsigned long long *d_size, void *stream);
int zccl_sz3d_decompress_dev(const unsigned char *in, size_t in_bytes, size_t dim1, size_t dim2,
    size_t dim3, int blockSideLength, double errorBound, int nchunks, float *out, void *ws,
    size_t ws_bytes, void *stream);
int zccl_sz3d_status(const void *ws, size_t *stream_bytes);

#ifdef __cplusplus
}
#endif

#endif /* ZCCL_AMD_H */
