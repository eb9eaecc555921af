"""3-D SZp codec on the GPU (the reference's SZp_compress3D_fast_openmp /
SZp_decompress3D_fast_openmp, ZCCL/src/ZCCL_float.c:1862-2048 and
ZCCL/src/ZCCLd_float.c:2873-3040), through the C ABI of libzccl_amd.so.

A volume is a contiguous float32 CUDA tensor of shape (dim1, dim2, dim3),
dim3 fastest.  ``T`` is the reference's OpenMP team size: the volume is split
into T sub-volumes exactly as the reference splits it over its threads, so a
stream made here with T decompresses on the CPU with an OpenMP team of T and
vice versa.  Calls are asynchronous on the current torch stream; ``size`` and
``check`` synchronise.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass
from typing import Optional, Sequence, Tuple

import torch

from . import _lib
from ._lib import lib


def _stream() -> ctypes.c_void_p:
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def _dims(dims: Sequence[int]) -> Tuple[int, int, int]:
    if len(dims) != 3:
        raise ValueError("dims must be (dim1, dim2, dim3)")
    return int(dims[0]), int(dims[1]), int(dims[2])


def sizes(dims: Sequence[int], Bs: int, T: int) -> Tuple[int, int]:
    """(worst-case stream bytes, workspace bytes) of one call."""
    d1, d2, d3 = _dims(dims)
    worst, ws = ctypes.c_size_t(0), ctypes.c_size_t(0)
    rc = lib().zccl_sz3d_sizes(d1, d2, d3, Bs, T, ctypes.byref(worst), ctypes.byref(ws))
    if rc:
        raise _lib.ZcclError(rc, f"3-D geometry dims={dims} Bs={Bs} T={T}")
    return int(worst.value), int(ws.value)


@dataclass
class Stream3D:
    """A compressed volume resident in HBM: ``data[:size]`` are the stream
    bytes in the reference's format; ``ws`` is the workspace of the call that
    made it (its status words hold the size and the error bits)."""

    data: torch.Tensor
    ws: torch.Tensor
    dims: Tuple[int, int, int]
    Bs: int
    eb: float
    T: int
    _size: Optional[int] = None

    @property
    def size(self) -> int:
        if self._size is None:
            self._size = check(self.ws)
        return self._size

    def to_host(self):
        return self.data[: self.size].cpu().numpy()


def check(ws: torch.Tensor) -> int:
    """Block until the call that used ``ws`` finished; raise on its error
    bits, else return the stream size it recorded."""
    torch.cuda.current_stream().synchronize()
    n = ctypes.c_size_t(0)
    rc = lib().zccl_sz3d_status(ctypes.c_void_p(ws.data_ptr()), ctypes.byref(n))
    if rc:
        raise _lib.ZcclError(rc, "3-D codec")
    return int(n.value)


def compress(x: torch.Tensor, Bs: int, eb: float, T: int, out: Optional[torch.Tensor] = None,
             ws: Optional[torch.Tensor] = None) -> Stream3D:
    """Compress a (dim1, dim2, dim3) float32 CUDA volume."""
    if x.dtype != torch.float32 or not x.is_cuda or not x.is_contiguous() or x.dim() != 3:
        raise ValueError("x must be a contiguous 3-D float32 CUDA tensor")
    dims = _dims(x.shape)
    worst, wsb = sizes(dims, Bs, T)
    data = out if out is not None else torch.empty(worst, dtype=torch.uint8, device=x.device)
    w = ws if ws is not None else torch.empty(wsb, dtype=torch.uint8, device=x.device)
    rc = lib().zccl_sz3d_compress_dev(ctypes.c_void_p(x.data_ptr()), dims[0], dims[1], dims[2], Bs, float(eb), T,
                                      ctypes.c_void_p(data.data_ptr()), data.numel(),
                                      ctypes.c_void_p(w.data_ptr()), w.numel(), None, _stream())
    if rc:
        raise _lib.ZcclError(rc, "zccl_sz3d_compress_dev")
    return Stream3D(data, w, dims, Bs, float(eb), T)


def decompress(s: Stream3D, out: Optional[torch.Tensor] = None, ws: Optional[torch.Tensor] = None,
               nbytes: Optional[int] = None) -> torch.Tensor:
    """Decompress into a (dim1, dim2, dim3) float32 CUDA tensor.  ``nbytes``
    bounds the stream's readable bytes (default: its buffer)."""
    d1, d2, d3 = s.dims
    _, wsb = sizes(s.dims, s.Bs, s.T)
    y = out if out is not None else torch.empty((d1, d2, d3), dtype=torch.float32, device=s.data.device)
    if y.numel() != d1 * d2 * d3 or y.dtype != torch.float32 or not y.is_contiguous():
        raise ValueError("out must be a contiguous float32 tensor of the volume's size")
    w = ws if ws is not None else torch.empty(wsb, dtype=torch.uint8, device=s.data.device)
    nb = s.data.numel() if nbytes is None else int(nbytes)
    rc = lib().zccl_sz3d_decompress_dev(ctypes.c_void_p(s.data.data_ptr()), nb, d1, d2, d3, s.Bs, s.eb, s.T,
                                        ctypes.c_void_p(y.data_ptr()), ctypes.c_void_p(w.data_ptr()), w.numel(),
                                        _stream())
    if rc:
        raise _lib.ZcclError(rc, "zccl_sz3d_decompress_dev")
    return y


def upload(stream, dims: Sequence[int], Bs: int, eb: float, T: int) -> Stream3D:
    """A host stream (numpy uint8) as a device Stream3D."""
    data = torch.from_numpy(stream.copy()).cuda()
    _, wsb = sizes(dims, Bs, T)
    w = torch.zeros(wsb, dtype=torch.uint8, device=data.device)
    return Stream3D(data, w, _dims(dims), Bs, float(eb), T, _size=int(stream.size))
