"""Synthetic workloads of the batched and served forms, shared by bench.py and
the GPU tests.

* ``uneven_batch``: many FP32 buffers of log-uniform sizes (gradient-bucket
  like values: N(0, 1e-2) with 1 % outliers at 100x), replicated over R
  simulated ranks.  One GPU's share of a compressed reduce-scatter is
  measured: compress every buffer of the rank, then sum, for the buffers this
  rank owns (every R-th), the R ranks' compressed copies.  Batched form: one
  compress launch plus one fused-sum launch (autotuned geometry), captured in
  a HIP graph and replayed.  Plain form: one compress call per buffer and the
  reference's ring of R-1 pairwise adds per owned buffer, one launch each.
* ``serving``: Poisson arrivals of mixed compress / decompress / k-operand
  sum requests over log-uniform sizes, served by ``serve.Server`` under an
  HBM budget; p50/p99 latency and throughput at several offered loads, and
  the same requests handled one at a time by per-call launches.
"""
from __future__ import annotations

import math
import time
from typing import Dict, List, Optional

import numpy as np
import torch

from . import autotune, batch, codec, serve, synth


def _log(*a):
    import sys

    print("[workload]", *a, file=sys.stderr, flush=True)


def log_uniform_sizes(n: int, lo_bytes: int, hi_bytes: int, seed: int) -> List[int]:
    rng = np.random.default_rng(seed)
    b = np.exp(rng.uniform(math.log(lo_bytes), math.log(hi_bytes), size=n))
    return [max(int(v) // 4, 64) for v in b]


def gradient_buffers(sizes: List[int], seed: int):
    """One flat device tensor holding all buffers back to back (N(0,1e-2), 1 %
    outliers at 100x) and the per-buffer views."""
    total = sum(sizes)
    g = torch.Generator(device="cuda")
    g.manual_seed(seed)
    flat = torch.randn(total, generator=g, device="cuda", dtype=torch.float32).mul_(1e-2)
    chunk = 1 << 28
    for c in range(0, total, chunk):
        part = flat[c: c + chunk]
        part[torch.rand(part.numel(), generator=g, device="cuda") < 0.01] *= 100.0
    views, off = [], 0
    for n in sizes:
        views.append(flat[off: off + n])
        off += n
    return flat, views


def per_buffer_eb(views, rel: float) -> List[float]:
    mins = torch.stack([v.min() for v in views]).cpu().numpy()
    maxs = torch.stack([v.max() for v in views]).cpu().numpy()
    return [float(np.float32((np.float32(hi) - np.float32(lo)) * np.float32(rel))) or rel
            for lo, hi in zip(mins, maxs)]


def uneven_batch(nbuf: int = 4096, ranks: int = 16, rel: float = 1e-3, L: int = 36, T: int = 16,
                 reps: int = 3, seed: int = 4096, lo: int = 4096, hi: int = 64 << 20) -> Dict[str, object]:
    sizes = log_uniform_sizes(nbuf, lo, hi, seed)
    sizes = [max(n, T) for n in sizes]
    flat, xs = gradient_buffers(sizes, seed)
    ebs = per_buffer_eb(xs, rel)
    owned = list(range(0, nbuf, ranks))
    _log(f"uneven batch: {nbuf} buffers, {sum(sizes) * 4 / 1e9:.2f} GB, owned {len(owned)}, ranks {ranks}")

    outs = [codec.empty_stream(n, L, T) for n in sizes]

    # autotune the workgroup size of both kernels on this workload
    def compress_once():
        p = batch.compress_plan(xs, ebs, outs)
        ws = codec.Workspace(p.tiles)
        p.launch(ws)

    tune_c = autotune.tune("compress", autotune.timed(compress_once))
    compress_once()
    torch.cuda.synchronize()
    codec.check_errors()
    # the other ranks' copies of the owned buffers' streams (replicated data,
    # distinct memory)
    copies = [[codec.DeviceStream(o.data[: o.size].clone(), o.index.clone(), o.n, L, T)
               for o in (outs[b] for b in owned)] for _ in range(ranks)]
    sums = [codec.empty_stream(sizes[b], L, T) for b in owned]
    jobs = [[copies[r][i] for r in range(ranks)] for i in range(len(owned))]

    def homsum_once():
        p = batch.homsum_plan(jobs, sums)
        ws = codec.Workspace(p.tiles)
        p.launch(ws)

    tune_h = autotune.tune("homsum", autotune.timed(homsum_once))

    # batched form: both launches captured in one graph
    cp = batch.compress_plan(xs, ebs, outs)
    hp = batch.homsum_plan(jobs, sums)
    ws = codec.Workspace(max(cp.tiles, hp.tiles))
    cp.launch(ws)
    hp.launch(ws)
    torch.cuda.synchronize()
    codec.check_errors(ws)
    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(side):
        cp.launch(ws)
        hp.launch(ws)
    torch.cuda.current_stream().wait_stream(side)
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        cp.launch(ws)
        hp.launch(ws)
    g.replay()
    torch.cuda.synchronize()
    t = []
    for _ in range(reps):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        g.replay()
        torch.cuda.synchronize()
        t.append(time.perf_counter() - t0)
    codec.check_errors(ws)
    t_batched = sorted(t)[len(t) // 2]
    batched_sums = [s.data[: s.size].clone() for s in sums]

    # plain form: one call per buffer, ring order of pairwise adds per owned buffer
    pws = codec.Workspace(max(codec.num_tiles(n, L, T) for n in sizes))

    def plain():
        res = []
        for x, e in zip(xs, ebs):
            codec.compress(x, e, L, T, ws=pws)
        for i in range(len(owned)):
            acc = codec.homsum([copies[1 % ranks][i], copies[0][i]], ws=pws)
            for r in range(2, ranks):
                acc = codec.homsum([acc, copies[r][i]], ws=pws)
            res.append(acc)
        return res

    plain()
    torch.cuda.synchronize()
    t = []
    for _ in range(reps):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        res = plain()
        torch.cuda.synchronize()
        t.append(time.perf_counter() - t0)
    codec.check_errors(pws)
    t_plain = sorted(t)[len(t) // 2]
    same = all(torch.equal(a.data[: a.size], b) for a, b in zip(res, batched_sums))
    in_bytes = 4 * sum(sizes) + 4 * ranks * sum(sizes[b] for b in owned)
    out = {
        "buffers": nbuf, "ranks": ranks, "owned_buffers": len(owned),
        "input_GB": round(in_bytes / 1e9, 3),
        "batched_ms": round(t_batched * 1e3, 3), "plain_ms": round(t_plain * 1e3, 3),
        "batched_GBps_input": round(in_bytes / t_batched / 1e9, 1),
        "plain_GBps_input": round(in_bytes / t_plain / 1e9, 1),
        "batched_vs_plain": round(t_plain / t_batched, 2),
        "bytes_equal": bool(same),
        "autotune": [tune_c, tune_h],
        "launches_batched": 2, "launches_plain": nbuf + len(owned) * (ranks - 1),
        "graph": True,
    }
    for op in ("compress", "homsum"):
        autotune.set_threads(op, 0)
    del flat, xs, outs, copies, sums, jobs, cp, hp, g, res, batched_sums
    torch.cuda.synchronize()
    return out


def serving_requests(nreq: int, seed: int, L: int, T: int, field: torch.Tensor,
                     lo: int = 64 << 10, hi: int = 512 << 20) -> List[serve.Request]:
    """Mix of 60 % compress / 20 % decompress / 20 % k-operand sums (k in
    [2, 32]); sizes log-uniform over [lo, hi] bytes; relative eb drawn from
    {1e-2, 1e-3, 1e-4}.  Values are windows of a k^-3 GRF; decompress sources
    and sum operands are compressed beforehand (device-resident state)."""
    rng = np.random.default_rng(seed)
    nmax = field.numel()
    reqs = []
    for i in range(nreq):
        kind = str(rng.choice(["compress", "decompress", "homsum"], p=[0.6, 0.2, 0.2]))
        n = min(max(int(math.exp(rng.uniform(math.log(lo), math.log(hi)))) // 4, T), nmax)
        rel = float(rng.choice([1e-2, 1e-3, 1e-4]))

        def window(shift):
            off = int(rng.integers(0, nmax - n + 1))
            v = field[off: off + n]
            return v if shift == 0 else (v + 0.01 * shift)

        x = window(0)
        eb = float(np.float32(float((x.max() - x.min()).item())) * np.float32(rel)) or rel
        if kind == "compress":
            reqs.append(serve.Request(i, kind, n, eb, x=x))
        elif kind == "decompress":
            reqs.append(serve.Request(i, kind, n, eb, src=[resident(codec.compress(x, eb, L, T))]))
        else:
            k = int(rng.integers(2, 33))
            ops = [resident(codec.compress(window(j).contiguous(), eb, L, T)) for j in range(k)]
            reqs.append(serve.Request(i, kind, n, eb, src=ops))
    torch.cuda.synchronize()
    return reqs


def resident(s: codec.DeviceStream) -> codec.DeviceStream:
    """Keep only the stream's bytes (kernels read operands up to their size)."""
    return codec.DeviceStream(s.data[: s.size].clone(), s.index, s.n, s.L, s.T)


def _reset(reqs):
    for r in reqs:
        r.out = None
        r.t_admit = r.t_done = float("nan")


def serving(nreq: int = 400, L: int = 36, T: int = 16, slots: Optional[int] = None,
            max_jobs: Optional[int] = None, budget: float = 200e9, loads=(0.5, 0.8, 0.95), seed: int = 5,
            side: int = 512, bypass: int = 4, slot_choices=(2, 4, 8),
            job_choices=(16, 64, 256)) -> Dict[str, object]:
    """The serving mix at the given budget.  Slots and jobs per launch not
    given are chosen by the closed-loop makespan over slot_choices x
    job_choices (the same requests, everything arriving at once); the loads
    then run at the chosen pair."""
    field = synth.grf3d(side, seed=seed)
    reqs = serving_requests(nreq, seed, L, T, field)
    resident = sum(sum(s.data.numel() + s.index.numel() for s in (r.src or [])) for r in reqs)
    resident += field.numel() * 4
    _log(f"serving: {nreq} requests, resident {resident / 1e9:.1f} GB")
    s_opts = (slots,) if slots else tuple(slot_choices)
    j_opts = (max_jobs,) if max_jobs else tuple(job_choices)
    ex = serve.Executor(L, T, slots=max(s_opts))

    # closed loop: everything arrives at t = 0 (capacity)
    for r in reqs:
        r.arrival = 0.0
    sweep = {}
    best = None
    for s_ in s_opts:
        for j_ in j_opts:
            srv_ = serve.Server(ex, hbm_budget=budget, slots=s_, max_jobs=j_, resident_bytes=resident,
                                bypass=bypass)
            srv_.run(reqs)  # warm-up: allocations, plan paths
            _reset(reqs)
            torch.cuda.synchronize()
            st_ = srv_.run(reqs)
            _reset(reqs)
            torch.cuda.synchronize()
            sweep[f"slots={s_},max_jobs={j_}"] = round(st_.throughput_GBps, 1)
            if best is None or st_.makespan_s < best[2].makespan_s:
                best = (s_, j_, st_, srv_)
    slots, max_jobs, cap, srv = best
    out: Dict[str, object] = {"requests": nreq, "slots": slots, "max_jobs_per_launch": max_jobs, "bypass": bypass,
                              "hbm_budget_GB": budget / 1e9, "resident_GB": round(resident / 1e9, 2)}
    if len(sweep) > 1:
        out["tuned"] = {"closed_loop_GBps": sweep, "chosen": {"slots": slots, "max_jobs": max_jobs}}
    out["closed_loop"] = cap.as_dict()
    rate = nreq / cap.makespan_s
    # plain per-call: one request at a time, each its own launch, completed
    # before the next is taken
    _reset(reqs)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for r in reqs:
        if r.kind == "compress":
            r.out = codec.compress(r.x, r.eb, L, T)
        elif r.kind == "decompress":
            r.out = codec.decompress(r.src[0], r.eb)
        else:
            r.out = codec.homsum(r.src)
        torch.cuda.current_stream().synchronize()
    t_plain = time.perf_counter() - t0
    tot = sum(r.input_bytes for r in reqs)
    out["plain_per_call"] = {"makespan_s": round(t_plain, 4), "throughput_GBps": round(tot / t_plain / 1e9, 1)}
    out["served_vs_plain"] = round(t_plain / cap.makespan_s, 2)
    loads_out = []
    for i, load in enumerate(loads):
        arr = serve.poisson_arrivals(nreq, rate * load, seed + 100 + i)
        for r, a in zip(reqs, arr):
            r.arrival = float(a)
        _reset(reqs)
        torch.cuda.synchronize()
        st = srv.run(reqs)
        d = st.as_dict()
        d["offered_load"] = load
        d["offered_rate_per_s"] = round(rate * load, 1)
        loads_out.append(d)
    out["loads"] = loads_out
    codec.check_errors()
    _reset(reqs)
    del reqs, field
    torch.cuda.synchronize()
    return out
