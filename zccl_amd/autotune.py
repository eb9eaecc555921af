"""Measurement-driven choice of the library's launch knobs, with a persisted
table of winners.

Knobs (C ABI ``zccl_tune_*``; none changes the stream bytes):

* ``threads``  -- workgroup size of the workgroup-tiled (slot) kernels, and
                  with it the blocks per tile and the length of each
                  lookback chain (ops compress, homsum, decompress);
* ``layout``   -- 0 one lane per block (a wave per 64-block tile; compress a
                  workgroup of four waves per 256 blocks); 1 the
                  workgroup-tiled kernels (every op, blockSize 36); 2 (sum
                  only) one lane per block, single pass, a lookback tile per
                  workgroup of four waves;
* ``stage``    -- smallest operand stage of the slot-layout sum, in bytes;
* ``two_pass`` -- sums of at least this many tiles take the two-pass form;
* ``groups``   -- workgroups per CU of the persistent lane decompress.

``tune`` times a caller-supplied run at each configuration of an op's space
and keeps the fastest.  ``run`` must build whatever depends on the tiling
(batch plans, workspaces) itself, since plans capture the tiling when they
are created.  Winners are stored per (op, blockSize, operand-count range,
size class) in a JSON table (``save`` / ``load``); ``apply`` sets the knobs
from a table entry before a run.
"""
from __future__ import annotations

import json
import math
import os
import time
from typing import Callable, Dict, List, Optional, Sequence

from ._lib import check, lib

OPS = {"compress": 0, "homsum": 1, "decompress": 2}
# the slot-layout candidates of the workgroup size (kept for callers that tune it alone)
CANDIDATES = {"compress": (256, 512), "homsum": (128, 256), "decompress": (256, 512)}

# configuration spaces: every knob an op reads; the layout-1 knobs only vary
# where layout 1 runs
SPACES: Dict[str, List[Dict[str, int]]] = {
    "compress": [{"layout": 0}] + [{"layout": 1, "threads": t} for t in (256, 512)],
    "homsum": ([{"layout": 0, "two_pass": tp} for tp in (1024, 1 << 40)] + [{"layout": 2}] +
               [{"layout": 1, "threads": t, "stage": st} for t in (128, 256) for st in (0, 32768)]),
    "decompress": ([{"layout": 0, "groups": g} for g in (3, 4, 5)] +
                   [{"layout": 1, "threads": t} for t in (256, 512)]),
}

DEFAULT_TABLE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tuned.json")


def set_threads(op: str, threads: int) -> None:
    check(lib().zccl_tune_threads(OPS[op], int(threads)), f"tune {op} threads {threads}")


def get_threads(op: str) -> int:
    return int(lib().zccl_tune_get_threads(OPS[op]))


def set_knobs(op: str, cfg: Dict[str, int]) -> None:
    """Set the knobs of op named in cfg (others keep their values)."""
    L = lib()
    if "threads" in cfg:
        set_threads(op, cfg["threads"])
    if "layout" in cfg:
        check(L.zccl_tune_layout(OPS[op], int(cfg["layout"])), f"tune {op} layout")
    if "stage" in cfg:
        check(L.zccl_tune_stage_bytes(int(cfg["stage"])), "tune stage")
    if "two_pass" in cfg:
        check(L.zccl_tune_two_pass_min_tiles(int(cfg["two_pass"])), "tune two_pass")
    if "groups" in cfg:
        check(L.zccl_tune_decompress_groups(int(cfg["groups"])), "tune groups")


def get_knobs(op: str) -> Dict[str, int]:
    """The library's current values of every knob op reads."""
    L = lib()
    k = {"threads": get_threads(op), "layout": int(L.zccl_tune_get_layout(OPS[op]))}
    if op == "homsum":
        k["stage"] = int(L.zccl_tune_get_stage_bytes())
        k["two_pass"] = int(L.zccl_tune_get_two_pass_min_tiles())
    if op == "decompress":
        k["groups"] = int(L.zccl_tune_get_decompress_groups())
    return k


def reset(op: str) -> None:
    """Library defaults for op's knobs."""
    cfg = {"threads": 0, "layout": 0}
    if op == "homsum":
        cfg.update(stage=0, two_pass=0)
    if op == "decompress":
        cfg["groups"] = 0
    set_knobs(op, cfg)


def label(cfg: Dict[str, int]) -> str:
    return ",".join(f"{k}={v}" for k, v in sorted(cfg.items()))


def tune(op: str, run: Callable[[], float], candidates: Sequence = (), reps: int = 3) -> Dict[str, object]:
    """Time ``run`` (returns seconds of one timed execution, device-synchronised)
    ``reps`` times per configuration; keep the one with the lowest median.
    ``candidates``: configuration dicts, or workgroup sizes (ints), default the
    op's whole space.  The library is left set to the winner."""
    cands: List[Dict[str, int]] = [({"threads": int(c)} if not isinstance(c, dict) else dict(c)) for c in candidates]
    if not cands:
        cands = [dict(c) for c in SPACES[op]]
    times = {}
    for c in cands:
        set_knobs(op, c)
        run()  # warm-up (plan build, caches)
        ts = sorted(run() for _ in range(reps))
        times[label(c)] = (ts[len(ts) // 2], c)
    best_label = min(times, key=lambda k: times[k][0])
    best = times[best_label][1]
    set_knobs(op, best)
    out = {"op": op, "config": best, "median_s": {k: round(v[0], 6) for k, v in times.items()}}
    if "threads" in best:
        out["threads"] = best["threads"]
    return out


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


# ------------------------------------------------------------ persisted table

def k_range(k: int) -> str:
    """Operand-count class: 1, 2, 3-4, 5-8, 9-16, 17-32, 33+."""
    if k <= 2:
        return str(max(k, 1))
    hi = 1 << math.ceil(math.log2(k))
    return f"{hi // 2 + 1}-{hi}" if hi <= 32 else "33+"


def size_class(n: int) -> str:
    """Values per job, to the nearest power of two."""
    return f"2^{int(round(math.log2(max(n, 1))))}"


def key(op: str, L: int, k: int, n: int) -> str:
    return f"{op}/L{L}/k{k_range(k)}/n{size_class(n)}"


def load(path: Optional[str] = None) -> Dict[str, Dict[str, object]]:
    p = path or os.environ.get("ZCCL_AMD_TUNE_TABLE", DEFAULT_TABLE)
    if not os.path.exists(p):
        return {}
    with open(p) as f:
        return json.load(f)


def save(table: Dict[str, Dict[str, object]], path: Optional[str] = None) -> str:
    p = path or os.environ.get("ZCCL_AMD_TUNE_TABLE", DEFAULT_TABLE)
    with open(p, "w") as f:
        json.dump(table, f, indent=1, sort_keys=True)
        f.write("\n")
    return p


def record(table: Dict[str, Dict[str, object]], op: str, L: int, k: int, n: int, result: Dict[str, object],
           device: str = "") -> None:
    """Store a tune() result under its key."""
    table[key(op, L, k, n)] = {"config": result["config"], "median_s": result["median_s"], "device": device}


def apply(table: Dict[str, Dict[str, object]], op: str, L: int, k: int, n: int) -> Optional[Dict[str, int]]:
    """Set op's knobs from the table entry for (L, k, n); returns the
    configuration, or None (knobs unchanged) when the table has no entry."""
    e = table.get(key(op, L, k, n))
    if not e:
        return None
    cfg = {str(a): int(b) for a, b in dict(e["config"]).items()}
    set_knobs(op, cfg)
    return cfg
