"""Benchmark: fused homomorphic sum (hZCCL reduce-scatter) on one MI355X.

Workload (one step): the reduce-scatter of the hZCCL allreduce data flow
(ZCCL/src/ZCCL_ring_ho.c:258-290) for 8 simulated ranks, each holding a 512^3
FP32 Gaussian random field with P(k) ~ k^-3 (shared seed) plus an independent
N(0, 0.05 std) perturbation per rank; relative error bound 1e-3 of the largest
value range; block length L = 36; chunk count T = 16.  Every rank's field is
compressed into 8 segments of 2^24 values beforehand (inputs resident in HBM as
compressed streams); a step is one batched launch of the 8 fused
homomorphic sums (one per segment, each adding the 8 ranks' streams).
``value`` = GB/s of input represented (8 ranks x 512 MiB per step) / step time.

Also reported: the same sums as 8 separate fused launches, the plain per-call
path (the reference's ring order: 7 pairwise adds per segment, one launch
each), compress and decompress rates, the roofline of the fused
kernel (algorithmic bytes per launch / average launch time from HIP events,
against 8 TB/s), HBM traffic per launch from rocprofv3 counters when
profiles/homsum_traffic.json exists, and the compiled reference's OpenMP
homomorphic add (the same 8-operand chain on segment 0) timed on the host
cores, checked byte for byte against the GPU output.

    python bench.py [--gpus 1] [--steps 10] [--warmup 3]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBPS = 8000.0
METRIC = "GB/s input (fused homomorphic sum, 1 GPU) vs plain; serve p50/p99 ms"


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def parse_args():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--side", type=int, default=512, help="grid side of each rank's field")
    ap.add_argument("--ranks", type=int, default=8)
    ap.add_argument("--rel-eb", type=float, default=1e-3)
    ap.add_argument("--L", type=int, default=36)
    ap.add_argument("--T", type=int, default=16)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=12.0, help="bound on the CPU baseline sample")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a 256 MiB elementwise copy kernel (rocprof counter calibration)")
    ap.add_argument("--traffic", default=os.path.join(ROOT, "profiles", "homsum_traffic.json"))
    ap.add_argument("--no-batched", action="store_true", help="skip the uneven-size batched workload")
    ap.add_argument("--no-serve", action="store_true", help="skip the serving workload")
    ap.add_argument("--batch-buffers", type=int, default=4096)
    ap.add_argument("--batch-ranks", type=int, default=16)
    ap.add_argument("--serve-requests", type=int, default=400)
    ap.add_argument("--no-3d", action="store_true", help="skip the 3-D codec measurement")
    ap.add_argument("--side3d", type=int, default=512, help="grid side of the 3-D codec's field")
    ap.add_argument("--bs3d", type=int, default=4, help="block side of the 3-D codec")
    ap.add_argument("--dropin-only", action="store_true",
                    help="only the reference-ABI drop-in calls (device pointers) and the parse they run, on "
                         "the first segment of the headline fields; one JSON line")
    ap.add_argument("--codec-only", action="store_true",
                    help="only the batched compress (every rank's P segments) and decompress (rank 0's P "
                         "segments) of the headline fields, timed with HIP events; one JSON line")
    ap.add_argument("--tune-out", default=None,
                    help="write the autotuned configurations (per op, blockSize, operand range, size class) "
                         "to this JSON table")
    ap.add_argument("--no-tune", action="store_true",
                    help="skip the autotuner sweeps; knobs from the committed table where it has the workload")
    ap.add_argument("--layout", choices=("lane", "tiles"), default="lane",
                    help="sum layout for 36-value blocks: one lane per block, or workgroup tiles")
    return ap.parse_args()


def cpu_baseline(streams0, n, L, T, gpu_sum, seconds):
    """Time the compiled reference's pairwise homomorphic add (OpenMP team of
    T) over the ring chain of segment 0's 8 streams; bounded to ~`seconds`."""
    import ctypes

    import numpy as np

    from oracle import oracle as O

    kind = "reference" if O.ref_available() else "port"
    hs = [s.to_host() for s in streams0]
    if kind == "reference":
        R = O.ref()
        O._set_T(T)
        pad = [np.zeros(h.size + 256, dtype=np.uint8) for h in hs]
        for p, h in zip(pad, hs):
            p[: h.size] = h
        cap = max(O.worst_case_bytes(n, L, T), 4 * n) + 64
        acc_a = np.zeros(cap, dtype=np.uint8)
        acc_b = np.zeros(cap, dtype=np.uint8)
        if not O.ref_buffer_fits(gpu_sum, n, T):
            return None

        def chain():
            size = ctypes.c_size_t(0)
            cur = pad[0]
            bufs = (acc_a, acc_b)
            for i, p in enumerate(pad[1:]):
                out = bufs[i % 2]
                size.value = 0
                R.ZCCL_float_homomophic_add_openmp_threadblock(
                    out.ctypes.data, ctypes.byref(size), n, ctypes.c_float(0.0), L, cur.ctypes.data, p.ctypes.data)
                cur = out
            return cur[: size.value]
    else:
        def chain():
            return O.homsum(hs, n, L, T)

    out = chain()  # warm-up, and the parity check
    bitexact = bool(np.array_equal(out, gpu_sum))
    times = []
    t_end = time.perf_counter() + seconds
    while len(times) < 3 or (time.perf_counter() < t_end and len(times) < 50):
        t0 = time.perf_counter()
        chain()
        times.append(time.perf_counter() - t0)
    med = float(np.median(times))
    gb = len(hs) * 4.0 * n / 1e9
    return {
        "value": round(gb / med, 3),
        "unit": "GB/s",
        "cores": T,
        "kind": kind,
        "sample": (f"segment 0 only: ring chain of {len(hs) - 1} pairwise homomorphic adds over "
                   f"{len(hs)} streams of {n} values (L={L}, T={T}), median of {len(times)} runs"),
        "bitexact_vs_gpu": bitexact,
    }


def bench_sz3d(side, Bs, T, rel_eb, steps, warmup, cpu_seconds, with_cpu):
    """3-D SZp codec (SZp_compress3D_fast_openmp / SZp_decompress3D_fast_openmp)
    on one side^3 FP32 GRF (k^-3): GPU compress and decompress with inputs in
    HBM, HIP events on the launch stream; the round trip is checked against
    its closed form (float)(floor(x * 0.5/eb + 0.5) * 2 eb).  The compiled
    reference (OpenMP team of T on the host cores) is timed on a bounded
    (side/2)^3 sample and its stream compared byte for byte with the GPU's."""
    import numpy as np
    import torch

    from oracle import oracle as O
    from zccl_amd import synth, sz3d

    x = synth.grf3d(side, seed=20261203).reshape(side, side, side).contiguous()
    eb = float((x.max() - x.min()).item()) * rel_eb
    worst, wsb = sz3d.sizes(x.shape, Bs, T)
    data = torch.empty(worst, dtype=torch.uint8, device="cuda")
    ws = torch.empty(wsb, dtype=torch.uint8, device="cuda")
    y = torch.empty_like(x)
    for _ in range(warmup):
        s = sz3d.compress(x, Bs, eb, T, out=data, ws=ws)
        sz3d.decompress(s, out=y, ws=ws)
    stream = torch.cuda.current_stream()

    def timed_events(fn):
        ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(steps)]
        for a, b in ev:
            a.record(stream)
            fn()
            b.record(stream)
        torch.cuda.synchronize()
        return float(np.mean([a.elapsed_time(b) for a, b in ev])) / 1e3

    t_c = timed_events(lambda: sz3d.compress(x, Bs, eb, T, out=data, ws=ws))
    s = sz3d.Stream3D(data, ws, tuple(x.shape), Bs, eb, T)
    size = s.size
    t_d = timed_events(lambda: sz3d.decompress(s, out=y, ws=ws, nbytes=size))
    sz3d.check(ws)
    q = torch.floor(x.double() * (0.5 / eb) + 0.5)
    closed = bool(torch.equal(y.view(torch.int32), (q * (2 * eb)).float().view(torch.int32)))
    del q
    n = side ** 3
    res = {
        "dims": [side, side, side], "Bs": Bs, "T": T, "abs_eb": eb, "rel_eb": rel_eb,
        "compression_ratio": round(4 * n / size, 3),
        "compress_GBps_input": round(4 * n / 1e9 / t_c, 1),
        "decompress_GBps_output": round(4 * n / 1e9 / t_d, 1),
        "compress_ms": round(t_c * 1e3, 3), "decompress_ms": round(t_d * 1e3, 3),
        "alg_GBps_compress": round((4 * n + size) / 1e9 / t_c, 1),
        "alg_GBps_decompress": round((4 * n + size) / 1e9 / t_d, 1),
        "round_trip_equals_closed_form": closed,
        "cpu_reference": None,
    }
    if with_cpu and O.ref_available():
        h = side // 2
        xs = x[:h, :h, :h].contiguous()
        xh = xs.cpu().numpy()
        gs = sz3d.compress(xs, Bs, eb, T)
        ghost = gs.to_host()
        ref = O.ref_compress3d(xh, (h, h, h), Bs, eb, T)  # checks the reference's buffer fits, then runs it
        # time the reference's own calls on preallocated buffers
        import ctypes

        R = O.ref()
        O._set_T(T)
        cbuf = np.zeros(max(ref.size, 4 * h ** 3) + 4096, dtype=np.uint8)
        dbuf = np.zeros(ref.size + 256, dtype=np.uint8)
        dbuf[: ref.size] = ref
        yh = np.zeros(h ** 3, dtype=np.float32)
        csize = ctypes.c_size_t(0)
        tc, td = [], []
        t_end = time.perf_counter() + cpu_seconds
        while len(tc) < 3 or (time.perf_counter() < t_end and len(tc) < 20):
            t0 = time.perf_counter()
            R.SZp_compress3D_fast_openmp(xh.ctypes.data, cbuf.ctypes.data, h, h, h, Bs, eb, ctypes.byref(csize))
            t1 = time.perf_counter()
            R.SZp_decompress3D_fast_openmp(yh.ctypes.data, dbuf.ctypes.data, h, h, h, Bs, eb)
            tc.append(t1 - t0)
            td.append(time.perf_counter() - t1)
        ref = cbuf[: csize.value]
        res["cpu_reference"] = {
            "compress_GBps_input": round(4 * h ** 3 / 1e9 / float(np.median(tc)), 3),
            "decompress_GBps_output": round(4 * h ** 3 / 1e9 / float(np.median(td)), 3),
            "cores": T, "kind": "reference",
            "sample": f"{h}^3 corner of the field, median of {len(tc)} runs",
            "bitexact_vs_gpu": bool(np.array_equal(ref, ghost)),
        }
    del x, y, data, ws
    torch.cuda.empty_cache()
    return res


def testfloat_leg(a, b, rel_eb, L, T, timed):
    """The reference driver's two-operand comparison
    (examples/testfloat_homocompress_fastmode1_arg.c:181-282) on two FP32
    fields of one size, on the GPU: eb = larger value range x rel (FP32);
    DOC = decompress both + FP32 add + compress the sum; homomorphic = one add
    of the two streams.  Throughputs in the driver's convention (4 N bytes /
    time), and the quality of both results against the FP32 sum."""
    import torch

    from zccl_amd import codec, metrics, synth

    n = a.numel()
    eb = synth.rel_to_abs_eb([a, b], rel_eb)
    ca = codec.compress(a, eb, L, T)
    cb = codec.compress(b, eb, L, T)
    da = torch.empty(n, dtype=torch.float32, device="cuda")
    db = torch.empty_like(da)
    dsum = torch.empty_like(da)
    cdoc = codec.empty_stream(n, L, T)
    chomo = codec.empty_stream(n, L, T)

    def doc():
        codec.decompress(ca, eb, out=da)
        codec.decompress(cb, eb, out=db)
        torch.add(da, db, out=dsum)
        codec.compress(dsum, eb, L, T, out=cdoc)

    def homo():
        codec.homsum([ca, cb], out=chomo)

    doc()
    homo()
    torch.cuda.synchronize()
    codec.check_errors()
    t_doc = timed(doc, 10)
    t_homo = timed(homo, 10)
    ori = a + b
    q_doc = metrics.error_evaluation(ori, codec.decompress(cdoc, eb), cdoc.size, eb)
    q_homo = metrics.error_evaluation(ori, codec.decompress(chomo, eb), chomo.size, eb)
    return {
        "elements": n, "abs_eb": eb, "rel_eb": rel_eb,
        "doc_ms": round(t_doc * 1e3, 4), "homomorphic_ms": round(t_homo * 1e3, 4),
        "doc_GBps": round(n * 4 / t_doc / 1e9, 1), "homomorphic_GBps": round(n * 4 / t_homo / 1e9, 1),
        "speedup": round(t_doc / t_homo, 2),
        "doc_quality": q_doc, "homomorphic_quality": q_homo,
    }


def dropin_leg(a, b, eb, L, T, timed):
    """The reference-ABI calls with device pointers on two FP32 fields of one
    size (the drop-in decompress, ZCCL/include/ZCCLd_float.h:21, and
    homomorphic add, :29), timed per call, beside the index rebuild (parse)
    every such call runs on a stream that carries no index, timed alone."""
    import ctypes

    import torch

    from zccl_amd import codec
    from zccl_amd._lib import lib

    n = a.numel()
    lib().zccl_set_num_chunks(T)
    try:
        ca = codec.compress(a, eb, L, T)
        cb = codec.compress(b, eb, L, T)
        torch.cuda.synchronize()
        y = torch.empty(n, dtype=torch.float32, device="cuda")
        out = torch.zeros(codec.worst_case_bytes(n, L, T), dtype=torch.uint8, device="cuda")
        size = ctypes.c_size_t(0)
        f_eb = ctypes.c_float(eb)

        def dec():
            lib().ZCCL_float_decompress_openmp_threadblock_arg(ctypes.c_void_p(y.data_ptr()), n, f_eb, L,
                                                               ctypes.c_void_p(ca.data.data_ptr()))

        def add():
            size.value = 0
            lib().ZCCL_float_homomophic_add_openmp_threadblock(ctypes.c_void_p(out.data_ptr()), ctypes.byref(size), n,
                                                               f_eb, L, ctypes.c_void_p(ca.data.data_ptr()),
                                                               ctypes.c_void_p(cb.data.data_ptr()))

        dec()
        add()
        assert lib().zccl_last_error() == 0
        ok = bool(torch.equal(y, codec.decompress(ca, eb))) and \
            bool(torch.equal(out[: size.value], codec.homsum([ca, cb]).data[: size.value]))
        t_dec = timed(dec, 10)
        t_add = timed(add, 10)
        t_parse = timed(lambda: codec.from_bytes(ca.data, n, L, T), 10)
        t_kdec = timed(lambda: codec.decompress(ca, eb, out=y), 10)
    finally:
        lib().zccl_set_num_chunks(0)
    return {
        "elements": n, "T": T,
        "decompress_ms": round(t_dec * 1e3, 4), "homomorphic_add_ms": round(t_add * 1e3, 4),
        "parse_ms": round(t_parse * 1e3, 4), "device_api_decompress_ms": round(t_kdec * 1e3, 4),
        "parse_share_of_decompress": round(t_parse / t_dec, 3),
        "equal_device_api": ok,
    }


def codec_leg(ar, xs, steps, warmup):
    """Batched compress of every rank's P segments (P*P jobs) and batched
    decompress of rank 0's P segments (P jobs), each one launch, timed with
    HIP events on the launch stream; fractions of HBM peak on the algorithmic
    bytes (floats in + stream bytes out, and the reverse)."""
    import numpy as np
    import torch

    from zccl_amd import codec

    P = ar.P
    stream = torch.cuda.current_stream()
    y = torch.empty(ar.count, dtype=torch.float32, device="cuda")

    def events(fn):
        for _ in range(warmup):
            fn()
        ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(steps)]
        torch.cuda.synchronize()
        for a, b in ev:
            a.record(stream)
            fn()
            b.record(stream)
        torch.cuda.synchronize()
        return float(np.mean([a.elapsed_time(b) for a, b in ev])) / 1e3

    t_c = events(lambda: ar.compress_all(xs))
    codec.check_errors()
    cbytes = sum(int(s.size) for row in ar.streams for s in row)
    segs = [ar.streams[0][j] for j in range(P)]
    t_d = events(lambda: ar.decompress_all(segs, out=y))
    codec.check_errors()
    alg_c = P * ar.count * 4 + cbytes
    dbytes = sum(int(s.size) for s in segs)
    alg_d = ar.count * 4 + dbytes
    return {"workload": f"compress {P * P} x {ar.cnts[0]} values (one batched launch), decompress {P} x "
                        f"{ar.cnts[0]} values (one batched launch), L={ar.L}, T={ar.T}",
            "compress_us": round(t_c * 1e6, 1), "compress_alg_bytes": alg_c,
            "compress_GBps": round(alg_c / t_c / 1e9, 1), "compress_frac": round(alg_c / t_c / 1e9 / HBM_PEAK_GBPS, 4),
            "decompress_us": round(t_d * 1e6, 1), "decompress_alg_bytes": alg_d,
            "decompress_GBps": round(alg_d / t_d / 1e9, 1),
            "decompress_frac": round(alg_d / t_d / 1e9 / HBM_PEAK_GBPS, 4),
            "steps": steps, "warmup": warmup}


def main():
    args = parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "one GPU only: --gpus above 1 is not supported; the multi-GPU "
                                   "segment exchange is designed but not built, and the value "
                                   "at N=1 is one GPU's share of it"}))
        sys.exit(2)

    import numpy as np
    import torch

    from zccl_amd import batch, codec, synth
    from zccl_amd.allreduce import StagedAllreduce

    torch.cuda.set_device(0)
    from zccl_amd._lib import lib as _zl
    _zl().zccl_tune_layout(1, 0 if args.layout == "lane" else 1)
    if args.copy_calibration:
        # copies reading 256 MiB and writing 256 MiB per launch: torch's
        # elementwise kernel, and the library's calibration copies with 16-byte
        # global loads (mode 0) and 16-byte buffer loads into LDS (mode 1, the
        # sum kernels' staging path)
        import ctypes

        a = torch.ones(64 << 20, dtype=torch.float32, device="cuda")
        b = torch.empty_like(a)
        for _ in range(3):
            torch.mul(a, 2.0, out=b)
        s = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
        for mode in (0, 1):
            for _ in range(3):
                rc = _zl().zccl_calib_copy(ctypes.c_void_p(a.data_ptr()), ctypes.c_void_p(b.data_ptr()),
                                           a.numel() * 4, mode, s)
                if rc:
                    raise RuntimeError(f"calibration copy failed ({rc})")
        torch.cuda.synchronize()
        assert torch.equal(a, b)
        print(json.dumps({"copy_read_bytes_per_launch": a.numel() * 4,
                          "copy_write_bytes_per_launch": b.numel() * 4}))
        return

    P, side, L, T = args.ranks, args.side, args.L, args.T
    count = side ** 3
    log(f"generating {P} fields of {side}^3")
    base = synth.grf3d(side, seed=20261202)
    xs = [synth.perturb(base, 0.05, seed=1000 + r) for r in range(P)]
    del base
    eb = synth.rel_to_abs_eb(xs, args.rel_eb)
    ar = StagedAllreduce(P=P, count=count, eb=eb, L=L, T=T)
    seg_n = ar.cnts[0]
    log(f"eb={eb:.6g} segments={ar.cnts[0]}x{P}")
    if args.codec_only:
        print(json.dumps(codec_leg(ar, xs, args.steps, args.warmup)), flush=True)
        return
    if args.dropin_only:
        def timed_only(fn, reps):
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            for _ in range(reps):
                fn()
            torch.cuda.synchronize()
            return (time.perf_counter() - t0) / reps
        print(json.dumps(dropin_leg(xs[0][: seg_n], xs[1][: seg_n], eb, L, T, timed_only)), flush=True)
        return

    def timed(fn, reps):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            fn()
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) / reps

    # compress every rank's segments (inputs of the timed step); the
    # workgroup size of the compress launch is autotuned on this workload
    from zccl_amd import autotune

    def compress_fresh():
        ar._cplan_key = None  # plans capture the tiling: rebuild per candidate
        ar.compress_all(xs)

    table = autotune.load()  # the committed table of winners (zccl_amd/tuned.json)
    table_out = {}
    device_name = torch.cuda.get_device_name(0)
    if args.no_tune:
        tune_c = {"op": "compress", "config": autotune.apply(table, "compress", L, 1, seg_n) or
                  autotune.get_knobs("compress"), "source": "table"}
    else:
        tune_c = autotune.tune("compress", autotune.timed(compress_fresh))
        autotune.record(table_out, "compress", L, 1, seg_n, tune_c, device_name)
    ar._cplan_key = None
    ar.compress_all(xs)  # builds the batch plan
    t_comp = timed(lambda: ar.compress_all(xs), 3)
    streams = ar.streams
    codec.check_errors()
    in_sizes = [[s.size for s in row] for row in streams]
    total_c = sum(sum(r) for r in in_sizes)
    cr = P * count * 4 / total_c
    log(f"compressed: total {total_c / 1e6:.1f} MB, CR {cr:.2f}, compress {t_comp * 1e3:.2f} ms")

    outs = [codec.empty_stream(seg_n, L, T) for _ in range(P)]
    sum_ops = [[streams[r][j] for r in range(P)] for j in range(P)]

    # the sum's knobs (layout, two-pass threshold, slot workgroup size and
    # stage) autotuned on this workload: each candidate builds its plan
    # untimed and times one launch
    def homsum_run():
        p = batch.homsum_plan(sum_ops, outs)
        w = codec.Workspace(p.tiles + 1024)
        p.launch(w)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        p.launch(w)
        torch.cuda.synchronize()
        return time.perf_counter() - t0

    if args.layout == "tiles":
        tune_h = {"op": "homsum", "config": autotune.get_knobs("homsum"), "source": "--layout tiles"}
    elif args.no_tune:
        tune_h = {"op": "homsum", "config": autotune.apply(table, "homsum", L, P, seg_n) or
                  autotune.get_knobs("homsum"), "source": "table"}
    else:
        tune_h = autotune.tune("homsum", homsum_run)
        autotune.record(table_out, "homsum", L, P, seg_n, tune_h, device_name)
        codec.check_errors()
    # one launch sums all P segments (P operands each), lookback chains interleaved
    plan = batch.homsum_plan(sum_ops, outs)
    ws = codec.Workspace(plan.tiles + 1024)

    for _ in range(args.warmup):
        plan.launch(ws)
    torch.cuda.synchronize()
    codec.check_errors(ws)

    # timed region: K steps of one launch each, HIP events on the launch stream
    stream = torch.cuda.current_stream()
    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(args.steps)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(args.steps):
        ev[k][0].record(stream)
        plan.launch(ws)
        ev[k][1].record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    codec.check_errors(ws)
    ms_per_step = wall / args.steps * 1e3
    launch_ms = [a.elapsed_time(b) for a, b in ev]
    avg_launch_s = float(np.mean(launch_ms)) / 1e3
    out_sizes = [o.size for o in outs]
    alg_per_launch = float(sum(sum(in_sizes[r][j] for r in range(P)) + out_sizes[j] for j in range(P)))
    achieved = alg_per_launch / avg_launch_s / 1e9
    input_gb = P * count * 4 / 1e9
    value = input_gb / (ms_per_step / 1e3)
    log(f"fused (1 launch): {ms_per_step:.3f} ms/step, launch {avg_launch_s * 1e6:.1f} us, "
        f"{achieved:.0f} GB/s algorithmic")

    # the same sums as P separate fused launches
    outs_pc = [codec.empty_stream(seg_n, L, T) for _ in range(P)]
    t_pc = timed(lambda: ar.reduce_scatter_fused(outs_pc), max(2, args.steps // 2))
    per_call_value = input_gb / t_pc

    # plain per-call path: reference ring order, pairwise launches
    scratch = ar.ring_scratch()
    ring_res = ar.reduce_scatter_ring(scratch)
    torch.cuda.synchronize()
    codec.check_errors()
    same = all(torch.equal(a.data[: a.size], b.data[: b.size]) for a, b in zip(ring_res, outs))
    same = same and all(torch.equal(a.data[: a.size], b.data[: b.size]) for a, b in zip(outs_pc, outs))
    t_ring = timed(lambda: ar.reduce_scatter_ring(scratch), max(2, args.steps // 2))
    plain_value = input_gb / t_ring

    # decompress + quality of the reduced field
    y = torch.empty(count, dtype=torch.float32, device="cuda")
    def decompress_fresh():
        ar._dplan_key = None
        ar.decompress_all(outs, out=y)

    if args.no_tune:
        tune_d = {"op": "decompress", "config": autotune.apply(table, "decompress", L, 1, seg_n) or
                  autotune.get_knobs("decompress"), "source": "table"}
    else:
        tune_d = autotune.tune("decompress", autotune.timed(decompress_fresh))
        autotune.record(table_out, "decompress", L, 1, seg_n, tune_d, device_name)
    ar._dplan_key = None
    ar.decompress_all(outs, out=y)  # builds the batch plan
    t_dec = timed(lambda: ar.decompress_all(outs, out=y), 3)
    codec.check_errors()
    # the batched compress and decompress alone, HIP events, fractions of peak
    codec_fracs = codec_leg(ar, xs, 5, 2)
    log(f"codec: compress {codec_fracs['compress_us']:.1f} us ({codec_fracs['compress_frac']:.3f} of peak), "
        f"decompress {codec_fracs['decompress_us']:.1f} us ({codec_fracs['decompress_frac']:.3f})")
    truth = torch.zeros(count, dtype=torch.float64, device="cuda")
    for x in xs:
        truth += x.double()
    err = (y.double() - truth).abs()
    max_err = float(err.max().item())
    n_over = int((err > P * eb).sum().item())
    del truth, err
    # the reference driver's quality metrics of the reduced field against
    # the FP32 sum of the ranks' fields in rank order (add_vectors)
    from zccl_amd import metrics

    ori32 = xs[0].clone()
    for x in xs[1:]:
        ori32 += x
    quality = metrics.error_evaluation(ori32, y, sum(o.size for o in outs), P * eb)
    del ori32
    testfloat = testfloat_leg(xs[0][: seg_n], xs[1][: seg_n], args.rel_eb, L, T, timed)
    dropin = dropin_leg(xs[0][: seg_n], xs[1][: seg_n], eb, L, T, timed)
    log(f"drop-in (device pointers, 2^24): {dropin}")

    # full staged allreduce (compress all + fused reduce-scatter + decompress all)
    t_all = timed(lambda: (ar.compress_all(xs), plan.launch(ws), ar.decompress_all(outs, out=y)), 3)
    # the same data flow run by the library's C host code (three launches)
    from zccl_amd.allreduce import DeviceAllreduce

    car = DeviceAllreduce(P, count, eb, L, T)
    yc = car.run(xs)
    torch.cuda.synchronize()
    c_equal = bool(torch.equal(yc, y)) and car.errors() == 0
    t_all_c = timed(lambda: car.run(xs, out=yc), 5)
    car.close()
    del car, yc
    # the reference's decompress-operate-compress ring on the same data
    t_doc = timed(lambda: ar.reduce_scatter_doc(xs), 2)
    t_homo_rs = t_comp + ms_per_step / 1e3

    # the uniform batch of 36-value blocks runs, as the knobs chose, the
    # two-pass lane-layout sum (three kernels: sum into the stage area, scan
    # of tile sizes, placement), the single-pass lane sum, or the slot-layout
    # sum at 256 threads
    hk = autotune.get_knobs("homsum")
    two_pass = L == 36 and hk["layout"] == 0 and plan.tiles >= hk["two_pass"]
    if L == 36 and hk["layout"] == 0:
        homsum_kernel = ("zk::k_homsum_lane_batch_staged + zk::k_scan_lane_batch + zk::k_place_lane_batch"
                         if two_pass else "zk::k_homsum_lane_batch")
    elif L == 36 and hk["threads"] == 256:
        homsum_kernel = "zk::k_homsum_batch_slots"
    else:
        homsum_kernel = "zk::k_homsum_batch"
    roof = {
        "bound": "hbm",
        "achieved": round(achieved, 1),
        "peak": HBM_PEAK_GBPS,
        "unit": "GB/s",
        "frac": round(achieved / HBM_PEAK_GBPS, 4),
        "traffic": None,
        "kernel": homsum_kernel,
        "alg_bytes_per_launch": int(alg_per_launch),
        "avg_launch_us": round(avg_launch_s * 1e6, 2),
    }
    if os.path.exists(args.traffic) and two_pass:
        # (the committed counters are those of the two-pass lane kernels)
        with open(args.traffic) as f:
            tr = json.load(f)
        roof["traffic"] = tr.get("hbm_bytes_per_launch")
        roof["traffic_source"] = os.path.relpath(args.traffic, ROOT)

    # uneven-size batched form and the serving loop (extra workloads, reported
    # beside the headline line)
    from zccl_amd import workloads

    batched = None
    if not args.no_batched:
        log("batched uneven-size workload")
        batched = workloads.uneven_batch(nbuf=args.batch_buffers, ranks=args.batch_ranks, L=L, T=T)
    served = None
    if not args.no_serve:
        log("serving workload")
        served = workloads.serving(nreq=args.serve_requests, L=L, T=T)

    sz3 = None
    if not args.no_3d:
        log("3-D codec")
        sz3 = bench_sz3d(args.side3d, args.bs3d, T, args.rel_eb, args.steps, args.warmup,
                         min(args.cpu_seconds, 10.0), not args.no_cpu_baseline)

    cpu = None
    if not args.no_cpu_baseline:
        log("cpu baseline")
        cpu = cpu_baseline([streams[r][0] for r in range(P)], seg_n, L, T,
                           outs[0].to_host(), args.cpu_seconds)

    res = {
        "metric": METRIC,
        "value": round(value, 2),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int32",
        "data": "synthetic",
        "config": {
            "workload": (f"hZCCL fused homomorphic reduce-scatter: {P} simulated ranks x {side}^3 fp32 "
                         f"GRF k^-3 (+N(0,0.05 std) per rank), rel eb {args.rel_eb:g}, L={L}, T={T}"),
            "ranks": P, "elements_per_rank": count, "segment_elements": seg_n,
            "L": L, "T": T, "abs_eb": eb, "rel_eb": args.rel_eb,
        },
        "roofline": roof,
        "cpu_baseline": cpu,
        "fused_per_call_value": round(per_call_value, 2),
        "plain_value": round(plain_value, 2),
        "plain_ms_per_step": round(t_ring * 1e3, 3),
        "fused_vs_plain": round(value / plain_value, 2),
        "ring_bytes_equal_fused": bool(same),
        "compression_ratio": round(cr, 3),
        "compress_GBps_input": round(P * count * 4 / 1e9 / t_comp, 1),
        "decompress_GBps_output": round(count * 4 / 1e9 / t_dec, 1),
        "codec": codec_fracs,
        "allreduce_ms": round(t_all * 1e3, 3),
        "allreduce_c_ms": round(t_all_c * 1e3, 3),
        "allreduce_c_equal": c_equal,
        "autotune": {"chosen": {t["op"]: t["config"] for t in (tune_c, tune_h, tune_d)},
                     "keys": {op: autotune.key(op, L, k, seg_n) for op, k in
                              (("compress", 1), ("homsum", P), ("decompress", 1))},
                     "sweeps": [t for t in (tune_c, tune_h, tune_d) if "median_s" in t],
                     "table": os.path.relpath(autotune.DEFAULT_TABLE, ROOT)},
        "homomorphic_reduce_scatter_ms": round(t_homo_rs * 1e3, 3),
        "doc_reduce_scatter_ms": round(t_doc * 1e3, 3),
        "homomorphic_vs_doc": round(t_doc / t_homo_rs, 2),
        "quality": quality,
        "testfloat_two_operand": testfloat,
        "dropin_device_pointers": dropin,
        "max_abs_err": max_err,
        "max_err_over_P_eb": round(max_err / (P * eb), 4),
        "n_err_over_P_eb": n_over,
        "batched_uneven": batched,
        "serve": served,
        "sz3d": sz3,
    }
    if served:
        res["serve_p50_ms"] = served["loads"][-1]["p50_ms"]
        res["serve_p99_ms"] = served["loads"][-1]["p99_ms"]
    if args.tune_out and table_out:
        res["autotune"]["written"] = autotune.save(table_out, args.tune_out)
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()
