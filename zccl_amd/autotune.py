"""Measurement-driven choice of launch geometry.

The kernels' workgroup size (and with it the blocks per tile and the length
of each lookback chain) is a tunable of the library
(``zccl_tune_threads``); it never changes the bytes.  ``tune`` times a
caller-supplied run at each candidate and keeps the fastest.  ``run`` must
build whatever depends on the tiling (batch plans, workspaces) itself, since
plans capture the tiling when they are created.
"""
from __future__ import annotations

import time
from typing import Callable, Dict, Sequence

from ._lib import check, lib

OPS = {"compress": 0, "homsum": 1, "decompress": 2}
CANDIDATES = {"compress": (256, 512), "homsum": (128, 256), "decompress": (256, 512)}


def set_threads(op: str, threads: int) -> None:
    check(lib().zccl_tune_threads(OPS[op], int(threads)), f"tune {op} {threads}")


def get_threads(op: str) -> int:
    return int(lib().zccl_tune_get_threads(OPS[op]))


def tune(op: str, run: Callable[[], float], candidates: Sequence[int] = (), reps: int = 3) -> Dict[str, object]:
    """Time ``run`` (returns seconds of one timed execution, device-synchronised)
    ``reps`` times per candidate; keep the candidate with the lowest median.
    The library is left set to the winner."""
    cands = list(candidates) or list(CANDIDATES[op])
    times = {}
    for c in cands:
        set_threads(op, c)
        run()  # warm-up (plan build, caches)
        ts = sorted(run() for _ in range(reps))
        times[c] = ts[len(ts) // 2]
    best = min(times, key=times.get)
    set_threads(op, best)
    return {"op": op, "threads": best, "median_s": {str(k): round(v, 6) for k, v in times.items()}}


def timed(fn: Callable[[], None]) -> Callable[[], float]:
    """Wrap fn as a device-synchronised timer."""
    import torch

    def run() -> float:
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        fn()
        torch.cuda.synchronize()
        return time.perf_counter() - t0

    return run
