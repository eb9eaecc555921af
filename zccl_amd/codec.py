"""Device-resident codec: compress / decompress / fused homomorphic sum on HBM.

Thin Python host mirror of the device C API (``zccl_*_dev`` in
include/zccl_amd.h).  PyTorch supplies device memory and the current HIP
stream only; every byte of the stream is produced and consumed by the HIP
kernels in libzccl_amd.so.

Reference operations mirrored (ZCCL/include/ZCCL_float.h:31-32,
ZCCL/include/ZCCLd_float.h:21,29):

* ``compress``   -- ZCCL_float_openmp_threadblock_arg with an explicit chunk count T
* ``decompress`` -- ZCCL_float_decompress_openmp_threadblock_arg
* ``homsum``     -- ZCCL_float_homomophic_add_openmp_threadblock generalised to k
                    operands in one pass (bytes equal the reference's chain of
                    pairwise adds in any order)
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass, field
from typing import Dict, Optional, Sequence, Tuple

import numpy as np
import torch

from . import _lib
from ._lib import check, lib

MAX_OPERANDS = 32


def _stream_handle() -> ctypes.c_void_p:
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def _ptr(t: torch.Tensor) -> ctypes.c_void_p:
    return ctypes.c_void_p(t.data_ptr())


def worst_case_bytes(n: int, L: int, T: int) -> int:
    v = lib().zccl_worst_case_bytes(n, L, T)
    if v == 0:
        raise _lib.ZcclError(-1, f"geometry n={n} L={L} T={T}")
    return int(v)


def index_bytes(n: int, L: int, T: int) -> int:
    return int(lib().zccl_index_bytes(n, L, T))


def num_tiles(n: int, L: int, T: int) -> int:
    return int(lib().zccl_num_tiles(n, L, T))


def num_blocks(n: int, L: int, T: int) -> int:
    return int(lib().zccl_num_blocks(n, L, T))


class Workspace:
    """Decoupled-lookback state for launches of up to ``max_tiles`` tiles.

    One launch at a time: launches sharing a workspace must be ordered on one
    stream (the default workspaces are per device and stream)."""

    def __init__(self, max_tiles: int, device: Optional[torch.device] = None):
        self.max_tiles = max(int(max_tiles), 1)
        self.nbytes = int(lib().zccl_workspace_bytes(self.max_tiles))
        self.buf = torch.empty(self.nbytes, dtype=torch.uint8, device=device or "cuda")
        check(lib().zccl_workspace_init(_ptr(self.buf), self.nbytes, _stream_handle()), "workspace_init")

    @property
    def ptr(self) -> ctypes.c_void_p:
        return _ptr(self.buf)

    def errors(self, clear: bool = True) -> int:
        e = ctypes.c_ulonglong(0)
        check(lib().zccl_workspace_errors(self.ptr, ctypes.byref(e), int(clear), _stream_handle()),
              "workspace_errors")
        return int(e.value)


_DEFAULT_WS: Dict[Tuple[int, int], Workspace] = {}


def default_workspace(tiles: int) -> Workspace:
    dev = torch.cuda.current_device()
    key = (dev, torch.cuda.current_stream().cuda_stream)
    ws = _DEFAULT_WS.get(key)
    if ws is None or ws.max_tiles < tiles:
        ws = Workspace(max(tiles, 1 << 16))
        _DEFAULT_WS[key] = ws
    return ws


class StageArea:
    """Stage area of the two-pass homomorphic sum (blockSize 36): a record
    and a 9600-byte slot per tile, for sums of up to ``max_tiles`` tiles.
    Like a workspace, it serves one launch at a time."""

    def __init__(self, max_tiles: int, device: Optional[torch.device] = None):
        self.max_tiles = max(int(max_tiles), 1)
        self.nbytes = int(lib().zccl_sum_stage_bytes(self.max_tiles))
        self.buf = torch.empty(self.nbytes, dtype=torch.uint8, device=device or "cuda")

    @property
    def ptr(self) -> ctypes.c_void_p:
        return _ptr(self.buf)


_DEFAULT_STAGE: Dict[Tuple[int, int], StageArea] = {}


def default_stage(tiles: int) -> StageArea:
    dev = torch.cuda.current_device()
    key = (dev, torch.cuda.current_stream().cuda_stream)
    st = _DEFAULT_STAGE.get(key)
    if st is None or st.max_tiles < tiles:
        st = StageArea(tiles)
        _DEFAULT_STAGE[key] = st
    return st


# the two-pass sum: blockSize 36, sums of at least staged_min_tiles() tiles
# (a library tunable, 1024 by default)
STAGED_L = 36
STAGED_MIN_TILES = 1024


def staged_min_tiles() -> int:
    return int(lib().zccl_tune_get_two_pass_min_tiles())


@dataclass
class DeviceStream:
    """A compressed stream resident in HBM together with its index.

    ``data[:size]`` are the stream bytes in the reference wire format;
    ``index`` holds the per-block widths, tile and seed offsets and the size.
    """

    data: torch.Tensor
    index: torch.Tensor
    n: int
    L: int
    T: int
    _size: Optional[int] = field(default=None, repr=False)

    @property
    def size_offset(self) -> int:
        return int(lib().zccl_index_size_offset(self.n, self.L, self.T))

    def size_tensor(self) -> torch.Tensor:
        """Device u64 view of the stream size (no synchronisation)."""
        off = self.size_offset
        return self.index[off: off + 8].view(torch.int64)

    @property
    def size(self) -> int:
        if self._size is None:
            self._size = int(self.size_tensor().item())
        return self._size

    def widths(self) -> torch.Tensor:
        return self.index[: num_blocks(self.n, self.L, self.T)]

    def to_host(self) -> np.ndarray:
        return self.data[: self.size].cpu().numpy()

    def invalidate(self) -> None:
        self._size = None


def empty_stream(n: int, L: int, T: int, capacity: Optional[int] = None,
                 device: Optional[torch.device] = None) -> DeviceStream:
    cap = worst_case_bytes(n, L, T) if capacity is None else int(capacity)
    data = torch.empty(cap, dtype=torch.uint8, device=device or "cuda")
    index = torch.empty(index_bytes(n, L, T), dtype=torch.uint8, device=device or "cuda")
    return DeviceStream(data, index, n, L, T)


def _ws_for(n: int, L: int, T: int, ws: Optional[Workspace]) -> Workspace:
    tiles = num_tiles(n, L, T)
    if ws is None:
        return default_workspace(tiles)
    if ws.max_tiles < tiles:
        raise _lib.ZcclError(-6, f"workspace holds {ws.max_tiles} tiles, launch needs {tiles}")
    return ws


def compress(x: torch.Tensor, eb: float, L: int, T: int, out: Optional[DeviceStream] = None,
             ws: Optional[Workspace] = None) -> DeviceStream:
    """Compress a contiguous float32 CUDA tensor (asynchronous on the current stream)."""
    if x.dtype != torch.float32 or not x.is_cuda or not x.is_contiguous():
        raise ValueError("x must be a contiguous float32 CUDA tensor")
    n = x.numel()
    s = out if out is not None else empty_stream(n, L, T, device=x.device)
    if (s.n, s.L, s.T) != (n, L, T):
        raise ValueError("output stream geometry differs")
    w = _ws_for(n, L, T, ws)
    s.invalidate()
    check(lib().zccl_compress_dev(_ptr(x), n, ctypes.c_float(eb), L, T, _ptr(s.data), s.data.numel(),
                                  _ptr(s.index), None, w.ptr, w.nbytes, _stream_handle()), "compress")
    return s


def decompress(s: DeviceStream, eb: float, out: Optional[torch.Tensor] = None,
               ws: Optional[Workspace] = None) -> torch.Tensor:
    y = out if out is not None else torch.empty(s.n, dtype=torch.float32, device=s.data.device)
    if y.numel() != s.n or y.dtype != torch.float32 or not y.is_contiguous():
        raise ValueError("out must be a contiguous float32 tensor of n elements")
    w = _ws_for(s.n, s.L, s.T, ws)
    check(lib().zccl_decompress_dev(_ptr(s.data), _ptr(s.index), s.n, ctypes.c_float(eb), s.L, s.T,
                                    _ptr(y), w.ptr, w.nbytes, _stream_handle()), "decompress")
    return y


def homsum(streams: Sequence[DeviceStream], out: Optional[DeviceStream] = None,
           ws: Optional[Workspace] = None, stage=None) -> DeviceStream:
    """Fused homomorphic sum of k streams.  Up to 32 operands go through one
    launch; more are summed in groups (the encoding is canonical, so grouping
    does not change the bytes).  For blockSize 36 the two-pass form runs with
    ``stage`` (a StageArea; None: the per-stream default one); stage=False
    runs the single-pass lookback form."""
    if not streams:
        raise ValueError("need at least one stream")
    n, L, T = streams[0].n, streams[0].L, streams[0].T
    for s in streams:
        if (s.n, s.L, s.T) != (n, L, T):
            raise ValueError("operands differ in (n, L, T)")
    w = _ws_for(n, L, T, ws)
    st = None
    tiles = num_tiles(n, L, T)
    if L == STAGED_L and stage is not False and tiles >= staged_min_tiles():
        st = stage if stage is not None else default_stage(tiles)
        if st.max_tiles < tiles:
            raise _lib.ZcclError(-6, f"stage area holds {st.max_tiles} tiles, sum needs {tiles}")
    ops = list(streams)
    while len(ops) > MAX_OPERANDS:
        part = _homsum_launch(ops[:MAX_OPERANDS], None, w, st)
        ops = [part] + ops[MAX_OPERANDS:]
    return _homsum_launch(ops, out, w, st)


def _homsum_launch(ops: Sequence[DeviceStream], out: Optional[DeviceStream], w: Workspace,
                   st: Optional[StageArea] = None) -> DeviceStream:
    n, L, T = ops[0].n, ops[0].L, ops[0].T
    o = out if out is not None else empty_stream(n, L, T, device=ops[0].data.device)
    k = len(ops)
    ins = (ctypes.c_void_p * k)(*[s.data.data_ptr() for s in ops])
    ixs = (ctypes.c_void_p * k)(*[s.index.data_ptr() for s in ops])
    o.invalidate()
    if st is not None:
        check(lib().zccl_homsum_dev_staged(ins, ixs, k, n, L, T, _ptr(o.data), o.data.numel(), _ptr(o.index),
                                           None, w.ptr, w.nbytes, st.ptr, st.nbytes, _stream_handle()),
              "homsum")
    else:
        check(lib().zccl_homsum_dev(ins, ixs, k, n, L, T, _ptr(o.data), o.data.numel(), _ptr(o.index),
                                    None, w.ptr, w.nbytes, _stream_handle()), "homsum")
    return o


def from_bytes(data: torch.Tensor, n: int, L: int, T: int, ws: Optional[Workspace] = None) -> DeviceStream:
    """Adopt stream bytes already on the device (e.g. written by the CPU
    library) and rebuild their index on the GPU."""
    if data.dtype != torch.uint8 or not data.is_cuda:
        raise ValueError("data must be a uint8 CUDA tensor")
    index = torch.empty(index_bytes(n, L, T), dtype=torch.uint8, device=data.device)
    if ws is None:
        # room for the parallel parse's tables
        need = int(lib().zccl_parse_workspace_bytes(n, L, T))
        w = default_workspace(max(num_tiles(n, L, T), (need - 256) // 16 + 1))
    else:
        w = _ws_for(n, L, T, ws)
    check(lib().zccl_parse_dev(_ptr(data), data.numel(), n, L, T, _ptr(index), None, w.ptr, w.nbytes,
                               _stream_handle()), "parse")
    return DeviceStream(data, index, n, L, T)


def upload(stream: np.ndarray, n: int, L: int, T: int) -> DeviceStream:
    """Copy a host stream to HBM and index it."""
    t = torch.from_numpy(np.ascontiguousarray(stream, dtype=np.uint8)).to("cuda")
    return from_bytes(t, n, L, T)


def check_errors(ws: Optional[Workspace] = None) -> None:
    """Synchronise and raise if any kernel on the workspace flagged an error."""
    wss = [ws] if ws is not None else list(_DEFAULT_WS.values())
    for w in wss:
        e = w.errors(clear=True)
        if e:
            status = -6 if e & 32 else -2 if e & 2 else -3 if e & (4 | 8 | 16) else -4
            raise _lib.ZcclError(status, f"device error bits 0x{e:x}")
