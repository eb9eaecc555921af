"""Batch plans: many codec operations in one launch (zccl_batch_* in
include/zccl_amd.h).

A plan is packed on the host by the library (no HIP call), its descriptors are
copied to a device tensor with a stream-ordered copy from pinned memory, and
the plan is bound to that tensor.  Building a plan therefore never allocates
with hipMalloc nor synchronises, so plans can be built while other work runs
(the serving loop does).  Each launch is a single kernel on the current
stream, which can be replayed inside a captured HIP graph.  The plan keeps
references to the tensors it names so they outlive it.
"""
from __future__ import annotations

import ctypes
from typing import List, Optional, Sequence

import torch

from . import _lib
from ._lib import check, lib
from .codec import STAGED_L, DeviceStream, Workspace, _stream_handle, default_stage, default_workspace, staged_min_tiles


class BatchPlan:
    def __init__(self, handle: int, keep: list, outs: Optional[List[DeviceStream]] = None,
                 staged: bool = False):
        if not handle:
            raise _lib.ZcclError(-1, "batch plan")
        self.handle = ctypes.c_void_p(handle)
        L = lib()
        self.tiles = int(L.zccl_batch_tiles(self.handle))
        nbytes = int(L.zccl_batch_desc_bytes(self.handle))
        host = torch.empty(nbytes, dtype=torch.uint8, pin_memory=True)
        ctypes.memmove(host.data_ptr(), L.zccl_batch_desc_host(self.handle), nbytes)
        self.desc = torch.empty(nbytes, dtype=torch.uint8, device="cuda")
        self.desc.copy_(host, non_blocking=True)
        check(L.zccl_batch_bind(self.handle, ctypes.c_void_p(self.desc.data_ptr())), "batch bind")
        self._keep = keep
        self.outs = outs or []
        # a homomorphic-sum batch of one geometry and operand count at
        # blockSize 36: the two-pass form applies
        self.staged = staged

    def launch(self, ws: Optional[Workspace] = None, stage=None) -> None:
        """stage: a StageArea for the two-pass sum (None: the default one,
        False: the single-pass form); ignored by other batches."""
        w = ws if ws is not None else default_workspace(self.tiles)
        if w.max_tiles < self.tiles:
            raise _lib.ZcclError(-6, f"workspace holds {w.max_tiles} tiles, batch needs {self.tiles}")
        for o in self.outs:
            o.invalidate()
        if self.staged and stage is not False and self.tiles >= staged_min_tiles():
            st = stage if stage is not None else default_stage(self.tiles)
            if st.max_tiles < self.tiles:
                raise _lib.ZcclError(-6, f"stage area holds {st.max_tiles} tiles, batch needs {self.tiles}")
            check(lib().zccl_batch_launch_staged(self.handle, w.ptr, w.nbytes, st.ptr, st.nbytes, _stream_handle()),
                  "batch launch")
        else:
            check(lib().zccl_batch_launch(self.handle, w.ptr, w.nbytes, _stream_handle()), "batch launch")

    def close(self) -> None:
        if self.handle is not None:
            lib().zccl_batch_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _ptrs(ts) -> ctypes.Array:
    return (ctypes.c_void_p * len(ts))(*[t.data_ptr() for t in ts])


def homsum_plan(jobs: Sequence[Sequence[DeviceStream]], outs: Sequence[DeviceStream]) -> BatchPlan:
    """jobs[j] is the list of operand streams (1 to 32) summed into outs[j];
    jobs may have different operand counts and sizes."""
    nj = len(jobs)
    if len(outs) != nj or nj == 0:
        raise ValueError("one output per job")
    L, T = outs[0].L, outs[0].T
    for j, ops in enumerate(jobs):
        if not 1 <= len(ops) <= 32:
            raise ValueError("1 to 32 operands per job")
        for s in ops:
            if (s.n, s.L, s.T) != (outs[j].n, L, T):
                raise ValueError("operands of a job must share (n, L, T) with its output")
    ks = (ctypes.c_int * nj)(*[len(ops) for ops in jobs])
    n = (ctypes.c_size_t * nj)(*[o.n for o in outs])
    ins = _ptrs([s.data for ops in jobs for s in ops])
    ixs = _ptrs([s.index for ops in jobs for s in ops])
    od = _ptrs([o.data for o in outs])
    ocap = (ctypes.c_size_t * nj)(*[o.data.numel() for o in outs])
    oix = _ptrs([o.index for o in outs])
    st = ctypes.c_int(0)
    h = lib().zccl_batch_homsum_plan(nj, ks, n, L, T, ins, ixs, od, ocap, oix, ctypes.byref(st))
    check(st.value, "batch homsum plan")
    staged = L == STAGED_L and len({(o.n, len(ops)) for o, ops in zip(outs, jobs)}) == 1
    return BatchPlan(h, [list(jobs), list(outs)], list(outs), staged=staged)


def compress_plan(xs: Sequence[torch.Tensor], ebs: Sequence[float], outs: Sequence[DeviceStream]) -> BatchPlan:
    nj = len(xs)
    L, T = outs[0].L, outs[0].T
    for x, o in zip(xs, outs):
        if x.numel() != o.n or x.dtype != torch.float32 or not x.is_contiguous() or (o.L, o.T) != (L, T):
            raise ValueError("inputs must be contiguous float32 matching their output streams")
    n = (ctypes.c_size_t * nj)(*[o.n for o in outs])
    eb = (ctypes.c_float * nj)(*ebs)
    st = ctypes.c_int(0)
    h = lib().zccl_batch_compress_plan(nj, n, eb, L, T, _ptrs(xs), _ptrs([o.data for o in outs]),
                                        (ctypes.c_size_t * nj)(*[o.data.numel() for o in outs]),
                                        _ptrs([o.index for o in outs]), ctypes.byref(st))
    check(st.value, "batch compress plan")
    return BatchPlan(h, [list(xs), list(outs)], list(outs))


def decompress_plan(streams: Sequence[DeviceStream], ebs: Sequence[float], outs: Sequence[torch.Tensor]) -> BatchPlan:
    nj = len(streams)
    L, T = streams[0].L, streams[0].T
    for s, y in zip(streams, outs):
        if y.numel() != s.n or y.dtype != torch.float32 or not y.is_contiguous() or (s.L, s.T) != (L, T):
            raise ValueError("outputs must be contiguous float32 matching their streams")
    n = (ctypes.c_size_t * nj)(*[s.n for s in streams])
    eb = (ctypes.c_float * nj)(*ebs)
    st = ctypes.c_int(0)
    h = lib().zccl_batch_decompress_plan(nj, n, eb, L, T, _ptrs([s.data for s in streams]),
                                          _ptrs([s.index for s in streams]), _ptrs(outs), ctypes.byref(st))
    check(st.value, "batch decompress plan")
    return BatchPlan(h, [list(streams), list(outs)])
