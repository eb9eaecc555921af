"""Continuous-batching request loop for compress / decompress / homomorphic-sum
requests against device-resident compressed state.

No reference counterpart: the reference runs one codec call at a time on the
CPU.  The loop here:

* works on device-resident state: compress inputs, the operands of sums and
  every result stay in HBM (requests hold DeviceStreams / tensors);
* admits arrived requests in arrival order while their output memory fits an
  HBM budget and an execution slot (a HIP stream with its own lookback
  workspace) is free; a request that does not fit waits at the head of the
  queue until completions release memory, and with ``bypass`` > 0 later
  requests that fit are admitted past it, at most ``bypass`` groups in a
  row, after which the head is served before anything else (no starvation);
* launches each admitted group as at most three batch launches (one per
  operation kind), so many small requests share one kernel;
* keeps up to ``slots`` groups in flight: requests that arrive while others
  run are admitted into the next free slot and their work overlaps on the
  device;
* records arrival, admission and completion times; latency is completion
  minus arrival.

The loop is independent of the GPU: ``Executor`` does the device work, and
the admission logic can be driven by any object with the same four methods
(the CPU tests use a simulated one).
"""
from __future__ import annotations

import collections
import time
from dataclasses import dataclass, field
from typing import Any, Callable, Dict, List, Optional, Sequence

import numpy as np

KINDS = ("compress", "decompress", "homsum")


@dataclass
class Request:
    rid: int
    kind: str
    n: int
    eb: float                        # absolute error bound
    arrival: float = 0.0             # seconds after the run starts
    x: Any = None                    # compress: float32 device tensor of n values
    src: Optional[list] = None       # decompress: [stream]; homsum: k operand streams
    out: Any = None                  # result: DeviceStream or float32 tensor
    t_admit: float = float("nan")
    t_done: float = float("nan")
    out_bytes: int = 0               # device memory reserved for the result

    @property
    def input_bytes(self) -> int:
        """Uncompressed bytes the request represents (k*4n for a k-operand sum)."""
        k = len(self.src) if self.kind == "homsum" else 1
        return 4 * self.n * k

    @property
    def latency(self) -> float:
        return self.t_done - self.arrival


@dataclass
class Step:
    slot: int
    reqs: List[Request]
    handle: Any
    reserved: int


@dataclass
class RunStats:
    requests: int
    makespan_s: float
    input_bytes: int
    throughput_GBps: float
    p50_ms: float
    p99_ms: float
    mean_ms: float
    steps: int
    max_inflight: int
    peak_reserved_bytes: int
    per_kind: Dict[str, dict] = field(default_factory=dict)

    def as_dict(self) -> dict:
        return {k: (round(v, 4) if isinstance(v, float) else v) for k, v in self.__dict__.items()}


def percentile_ms(lat_s: Sequence[float], q: float) -> float:
    return float(np.percentile(np.asarray(lat_s) * 1e3, q)) if len(lat_s) else float("nan")


class Server:
    """Admission and batching loop.

    ``executor.reserve(req)`` returns the bytes a request's result needs;
    ``executor.launch(slot, reqs)`` starts the device work of a group and
    returns a handle; ``executor.done(handle)`` polls it without blocking;
    ``executor.finish(handle)`` is called once after completion (errors are
    raised there)."""

    def __init__(self, executor, hbm_budget: float, slots: int = 4, max_jobs: int = 256,
                 resident_bytes: int = 0, clock: Callable[[], float] = time.perf_counter,
                 idle: Optional[Callable[[float], None]] = None, bypass: int = 0):
        if slots < 1 or max_jobs < 1 or bypass < 0:
            raise ValueError("need at least one slot and one job per step, and bypass >= 0")
        self.bypass = bypass
        self.ex = executor
        # resident state (inputs, operands already in HBM) counts against the budget
        self.budget = float(hbm_budget) - float(resident_bytes)
        self.slots = slots
        self.max_jobs = max_jobs
        self.clock = clock
        self.idle = idle or _spin

    def run(self, requests: Sequence[Request]) -> RunStats:
        reqs = sorted(requests, key=lambda r: (r.arrival, r.rid))
        for r in reqs:
            r.out_bytes = int(self.ex.reserve(r))
            if r.out_bytes > self.budget:
                raise ValueError(f"request {r.rid} needs {r.out_bytes} B, above the HBM budget")
        waiting: collections.deque = collections.deque()
        inflight: List[Step] = []
        free = list(range(self.slots - 1, -1, -1))
        used = 0
        peak = 0
        nsteps = 0
        max_inflight = 0
        passed = 0  # groups admitted in a row past a head that did not fit
        i = 0
        t0 = self.clock()
        while i < len(reqs) or waiting or inflight:
            now = self.clock() - t0
            while i < len(reqs) and reqs[i].arrival <= now:
                waiting.append(reqs[i])
                i += 1
            progressed = False
            for st in list(inflight):
                if self.ex.done(st.handle):
                    t = self.clock() - t0
                    self.ex.finish(st.handle)
                    for r in st.reqs:
                        r.t_done = t
                    used -= st.reserved
                    free.append(st.slot)
                    inflight.remove(st)
                    progressed = True
            if free and waiting:
                group: List[Request] = []
                reserved = 0
                while waiting and len(group) < self.max_jobs:
                    r = waiting[0]
                    if used + reserved + r.out_bytes > self.budget:
                        break  # head of line waits for memory
                    waiting.popleft()
                    reserved += r.out_bytes
                    group.append(r)
                if waiting and len(group) < self.max_jobs and passed < self.bypass:
                    # the head does not fit: later requests that do go ahead of it
                    head = waiting.popleft()
                    keep: collections.deque = collections.deque([head])
                    while waiting and len(group) < self.max_jobs:
                        r = waiting.popleft()
                        if used + reserved + r.out_bytes <= self.budget:
                            reserved += r.out_bytes
                            group.append(r)
                        else:
                            keep.append(r)
                    keep.extend(waiting)
                    waiting = keep
                    if group:
                        passed += 1
                elif group:
                    passed = 0
                if group:
                    slot = free.pop()
                    t = self.clock() - t0
                    for r in group:
                        r.t_admit = t
                    used += reserved
                    peak = max(peak, used)
                    inflight.append(Step(slot, group, self.ex.launch(slot, group), reserved))
                    nsteps += 1
                    max_inflight = max(max_inflight, len(inflight))
                    progressed = True
            if not progressed:
                nxt = reqs[i].arrival - (self.clock() - t0) if i < len(reqs) else 0.0
                self.idle(nxt if not inflight and not (free and waiting) else 0.0)
        makespan = max((r.t_done for r in reqs), default=0.0) - min((r.arrival for r in reqs), default=0.0)
        tot = sum(r.input_bytes for r in reqs)
        lat = [r.latency for r in reqs]
        per_kind = {}
        for k in KINDS:
            lk = [r.latency for r in reqs if r.kind == k]
            if lk:
                per_kind[k] = {"requests": len(lk), "p50_ms": round(percentile_ms(lk, 50), 3),
                               "p99_ms": round(percentile_ms(lk, 99), 3)}
        return RunStats(requests=len(reqs), makespan_s=makespan, input_bytes=tot,
                        throughput_GBps=tot / makespan / 1e9 if makespan > 0 else float("nan"),
                        p50_ms=percentile_ms(lat, 50), p99_ms=percentile_ms(lat, 99),
                        mean_ms=float(np.mean(lat) * 1e3) if lat else float("nan"),
                        steps=nsteps, max_inflight=max_inflight, peak_reserved_bytes=int(peak),
                        per_kind=per_kind)


def _spin(dt: float) -> None:
    """Wait up to dt (at most 50 us) without yielding: completions are polled."""
    end = time.perf_counter() + min(max(dt, 0.0), 5e-5)
    while time.perf_counter() < end:
        pass


def poisson_arrivals(n: int, rate_per_s: float, seed: int) -> np.ndarray:
    rng = np.random.default_rng(seed)
    return np.cumsum(rng.exponential(1.0 / rate_per_s, size=n))


class Executor:
    """Device work of the loop: each slot is a HIP stream with its own
    workspace; a group becomes one batch launch per operation kind."""

    def __init__(self, L: int, T: int, slots: int, max_tiles: int = 1 << 22):
        import torch

        from . import codec

        self.torch = torch
        self.codec = codec
        self.L, self.T = L, T
        self.streams = [torch.cuda.Stream() for _ in range(slots)]
        self.ws = []
        for s in self.streams:
            with torch.cuda.stream(s):
                self.ws.append(codec.Workspace(max_tiles))
        torch.cuda.synchronize()

    def reserve(self, r: Request) -> int:
        c = self.codec
        if r.kind == "decompress":
            return 4 * r.n
        return c.worst_case_bytes(r.n, self.L, self.T) + c.index_bytes(r.n, self.L, self.T)

    def launch(self, slot: int, reqs: List[Request]):
        torch = self.torch
        s = self.streams[slot]
        plans = []
        with torch.cuda.stream(s):
            comp = [r for r in reqs if r.kind == "compress"]
            dec = [r for r in reqs if r.kind == "decompress"]
            hs = [r for r in reqs if r.kind == "homsum"]
            for group, make in ((comp, self._compress), (hs, self._homsum), (dec, self._decompress)):
                for part in _split_tiles(group, self.ws[slot].max_tiles, self):
                    p = make(part)
                    p.launch(self.ws[slot])
                    plans.append(p)
            ev = torch.cuda.Event()
            ev.record(s)
        return (slot, ev, plans, reqs)

    def _compress(self, rs):
        from . import batch

        for r in rs:
            r.out = self.codec.empty_stream(r.n, self.L, self.T)
        return batch.compress_plan([r.x for r in rs], [r.eb for r in rs], [r.out for r in rs])

    def _homsum(self, rs):
        from . import batch

        for r in rs:
            r.out = self.codec.empty_stream(r.n, self.L, self.T)
        return batch.homsum_plan([r.src for r in rs], [r.out for r in rs])

    def _decompress(self, rs):
        from . import batch

        for r in rs:
            r.out = self.torch.empty(r.n, dtype=self.torch.float32, device="cuda")
        return batch.decompress_plan([r.src[0] for r in rs], [r.eb for r in rs], [r.out for r in rs])

    def tiles(self, r: Request) -> int:
        return self.codec.num_tiles(r.n, self.L, self.T)

    def done(self, h) -> bool:
        return h[1].query()

    def finish(self, h) -> None:
        torch = self.torch
        slot = h[0]
        with torch.cuda.stream(self.streams[slot]):
            e = self.ws[slot].errors()
        if e:
            raise RuntimeError(f"device error bits {e:#x} in serving slot {slot}")
        # results were allocated on the slot's stream; they are read on others
        cur = torch.cuda.current_stream()
        for r in h[3]:
            if isinstance(r.out, torch.Tensor):
                r.out.record_stream(cur)
            else:
                r.out.data.record_stream(cur)
                r.out.index.record_stream(cur)


def _split_tiles(group: List[Request], max_tiles: int, ex) -> List[List[Request]]:
    """Cut a kind's group into launches whose tiles fit the slot workspace."""
    parts, cur, tiles = [], [], 0
    for r in group:
        t = ex.tiles(r)
        if cur and tiles + t > max_tiles:
            parts.append(cur)
            cur, tiles = [], 0
        cur.append(r)
        tiles += t
    if cur:
        parts.append(cur)
    return parts
