"""hZCCL allreduce data flow staged on one GPU.

The reference's MPI_Allreduce_ZCCL_RI2_mt_oa_ho_record
(ZCCL/src/ZCCL_ring_ho.c:177-317) runs, on each of P ranks:

1. split ``count`` values into P segments of ceil(count/P) (the last short,
   :218-233) and compress every segment of its own data (:247-256);
2. P-1 ring steps: receive the partial compressed sum of segment
   (rank-2-i) mod P from the left neighbour, add the own compressed segment
   homomorphically in place (:258-290) -- rank r ends owning segment r;
3. compressed ring allgather of the owned segments (:292, :12-175) and
   decompression of all P segments (:163-169).

Here the P ranks are P sets of device-resident buffers on one GPU and no
bytes travel between GPUs.  Two forms of step 2:

* ``reduce_scatter_ring``  -- the reference's order: P-1 pairwise adds per
  segment, one launch each (the plain per-call path);
* ``reduce_scatter_fused`` -- one k=P homomorphic-sum launch per segment.
  The encoding is canonical, so both give the same bytes.

Step 3's allgather moves the P owned streams to every rank; on one GPU every
"rank" reads the same resident streams, so the allgather is the identity and
``decompress_all`` decompresses each segment once into the full output.
"""
from __future__ import annotations

from dataclasses import dataclass
from typing import List, Optional, Sequence

import torch

from . import codec


def segments(count: int, P: int):
    """(cnts, displs) as ZCCL/src/ZCCL_ring_ho.c:218-233 computes them."""
    cnts = [0] * P
    total = 0
    for i in range(P):
        c = (count + P - 1) // P
        if total + c > count:
            cnts[i] = count - total
            break
        cnts[i] = c
        total += c
    displs = [0] * P
    for i in range(1, P):
        displs[i] = displs[i - 1] + cnts[i - 1]
    return cnts, displs


def ring_order(P: int, j: int) -> List[int]:
    """Ranks in the order the ring adds their copy of segment j: rank j+1
    sends its raw segment at step 0, ranks j+2, ..., j+P (= j) each add theirs
    (recv_rank = (r-2-i) mod P at step i, ZCCL/src/ZCCL_ring_ho.c:259-260)."""
    return [(j + 1 + i) % P for i in range(P)]


@dataclass
class StagedAllreduce:
    P: int
    count: int
    eb: float
    L: int = 36
    T: int = 16

    def __post_init__(self):
        self.cnts, self.displs = segments(self.count, self.P)
        # a segment of 0 elements carries no stream (count < P); the
        # reference would also produce none for it
        self.active = [j for j in range(self.P) if self.cnts[j] >= self.T]
        if len(self.active) != self.P:
            raise ValueError("every segment must hold at least T elements")
        self.streams: Optional[List[List[codec.DeviceStream]]] = None

    # step 1
    def compress_all(self, xs: Sequence[torch.Tensor], batched: bool = True) -> List[List[codec.DeviceStream]]:
        """Compress every rank's P segments (ZCCL_ring_ho.c:247-256): one
        batched launch for all P*P segments, or one call per segment."""
        if len(xs) != self.P:
            raise ValueError("one input per rank")
        if self.streams is None:
            self.streams = [[codec.empty_stream(self.cnts[j], self.L, self.T) for j in range(self.P)]
                            for _ in range(self.P)]
        segs = [x[self.displs[j]: self.displs[j] + self.cnts[j]] for x in xs for j in range(self.P)]
        if batched:
            from . import batch

            key = tuple(s.data_ptr() for s in segs)
            if getattr(self, "_cplan_key", None) != key:
                outs = [self.streams[r][j] for r in range(self.P) for j in range(self.P)]
                self._cplan = batch.compress_plan(segs, [self.eb] * len(segs), outs)
                self._cplan_key = key
            self._cplan.launch()
        else:
            for i, seg in enumerate(segs):
                codec.compress(seg, self.eb, self.L, self.T, out=self.streams[i // self.P][i % self.P])
        return self.streams

    # step 2, fused
    def reduce_scatter_fused(self, outs: Optional[List[codec.DeviceStream]] = None) -> List[codec.DeviceStream]:
        res = []
        for j in range(self.P):
            ops = [self.streams[r][j] for r in range(self.P)]
            res.append(codec.homsum(ops, out=outs[j] if outs else None))
        return res

    # step 2, reference order
    def reduce_scatter_ring(self, scratch: Optional[List[List[codec.DeviceStream]]] = None) -> List[codec.DeviceStream]:
        res = []
        for j in range(self.P):
            order = ring_order(self.P, j)
            acc = self.streams[order[0]][j]
            for i, r in enumerate(order[1:]):
                out = scratch[j][i % 2] if scratch else None
                acc = codec.homsum([acc, self.streams[r][j]], out=out)
            res.append(acc)
        return res

    def ring_scratch(self) -> List[List[codec.DeviceStream]]:
        return [[codec.empty_stream(self.cnts[j], self.L, self.T) for _ in range(2)] for j in range(self.P)]

    # step 3
    def decompress_all(self, sums: Sequence[codec.DeviceStream], out: Optional[torch.Tensor] = None,
                       batched: bool = True) -> torch.Tensor:
        """Decompress the P reduced segments into the full output
        (ZCCL_ring_ho.c:163-169); one batched launch or one call each."""
        y = out if out is not None else torch.empty(self.count, dtype=torch.float32, device="cuda")
        views = [y[self.displs[j]: self.displs[j] + self.cnts[j]] for j in range(self.P)]
        if batched:
            from . import batch

            # (the index pointers too: a stream re-indexed over the same bytes
            # needs a new plan)
            key = (tuple(v.data_ptr() for v in views) + tuple(s.data.data_ptr() for s in sums) +
                   tuple(s.index.data_ptr() for s in sums))
            if getattr(self, "_dplan_key", None) != key:
                self._dplan = batch.decompress_plan(list(sums), [self.eb] * self.P, views)
                self._dplan_key = key
            self._dplan.launch()
        else:
            for j in range(self.P):
                codec.decompress(sums[j], self.eb, out=views[j])
        return y

    # the reference's decompress-operate-compress ring (DOC), for comparison
    def reduce_scatter_doc(self, xs: Sequence[torch.Tensor]) -> List[torch.Tensor]:
        """MPI_Allreduce_ZCCL_RI2_mt_oa_record's reduce-scatter
        (ZCCL/src/ZCCL_ring.c:979-1108) with all P ranks on one device: at
        each of the P-1 steps a rank compresses its partial FP32 sum of the
        segment it sends (:1049-1055), the receiver decompresses it (:1066-1067)
        and adds its own FP32 segment (MPI_Reduce_local, :1068-1069).  Returns
        the FP32 segment each rank ends with (rank j owns segment j)."""
        P = self.P
        part = [[xs[r][self.displs[j]: self.displs[j] + self.cnts[j]].clone() for j in range(P)] for r in range(P)]
        tmp = [codec.empty_stream(self.cnts[j], self.L, self.T) for j in range(P)]
        for i in range(P - 1):
            # every rank r sends segment (r-1-i) mod P to rank r+1
            for r in range(P):
                send = (P + r - 1 - i) % P
                dst = (r + 1) % P
                codec.compress(part[r][send], self.eb, self.L, self.T, out=tmp[send])
                y = codec.decompress(tmp[send], self.eb)
                part[dst][send] += y
        return [part[(j) % P][j] for j in range(P)]


class DeviceAllreduce:
    """The same data flow as StagedAllreduce, run by the library's C host
    code (zccl_allreduce_*: one compress launch, one fused-sum launch and one
    decompress launch over a caller workspace)."""

    def __init__(self, P: int, count: int, eb: float, L: int = 36, T: int = 16):
        import ctypes

        from ._lib import check, lib

        self.P, self.count, self.eb, self.L, self.T = P, count, eb, L, T
        nbytes = int(lib().zccl_allreduce_workspace_bytes(P, count, L, T))
        if not nbytes:
            raise ValueError("invalid allreduce geometry (every segment needs at least T values)")
        self.ws = torch.empty(nbytes + 256, dtype=torch.uint8, device="cuda")
        base = self.ws.data_ptr()
        self._ws_ptr = (base + 255) // 256 * 256
        st = ctypes.c_int(0)
        h = lib().zccl_allreduce_create(P, count, ctypes.c_float(eb), L, T, ctypes.c_void_p(self._ws_ptr), nbytes,
                                        ctypes.byref(st))
        check(st.value, "allreduce create")
        self.handle = ctypes.c_void_p(h)
        self.cnts, self.displs = segments(count, P)

    def run(self, xs: Sequence[torch.Tensor], out: Optional[torch.Tensor] = None) -> torch.Tensor:
        import ctypes

        from ._lib import check, lib

        if len(xs) != self.P or any(x.numel() != self.count or x.dtype != torch.float32 or not x.is_contiguous()
                                    for x in xs):
            raise ValueError("one contiguous float32 buffer of count values per rank")
        y = out if out is not None else torch.empty(self.count, dtype=torch.float32, device="cuda")
        send = (ctypes.c_void_p * self.P)(*[x.data_ptr() for x in xs])
        check(lib().zccl_allreduce_run(self.handle, send, ctypes.c_void_p(y.data_ptr()),
                                       ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)), "allreduce run")
        return y

    def segment_stream(self, j: int) -> codec.DeviceStream:
        """The reduced stream of segment j after a run: views of its bytes
        and index inside the workspace (valid until the next run)."""
        import ctypes

        from ._lib import check, lib

        b, ix, n, off = ctypes.c_void_p(), ctypes.c_void_p(), ctypes.c_size_t(), ctypes.c_size_t()
        check(lib().zccl_allreduce_segment(self.handle, j, ctypes.byref(b), ctypes.byref(ix), ctypes.byref(n),
                                           ctypes.byref(off)), "allreduce segment")
        base = self.ws.data_ptr()
        cap = codec.worst_case_bytes(n.value, self.L, self.T)
        ixb = codec.index_bytes(n.value, self.L, self.T)
        data = self.ws[b.value - base: b.value - base + cap]
        index = self.ws[ix.value - base: ix.value - base + ixb]
        return codec.DeviceStream(data, index, n.value, self.L, self.T)

    def errors(self) -> int:
        import ctypes

        from ._lib import check, lib

        e = ctypes.c_ulonglong(0)
        check(lib().zccl_allreduce_errors(self.handle, ctypes.byref(e), 1,
                                          ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)), "allreduce errors")
        return int(e.value)

    def close(self) -> None:
        from ._lib import lib

        if getattr(self, "handle", None) is not None:
            lib().zccl_allreduce_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
