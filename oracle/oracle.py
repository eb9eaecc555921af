"""ctypes bindings to the CPU checker libraries.

TEST INFRASTRUCTURE ONLY.  Imported by tests/, ``__graft_entry__.smoke()`` and
the ``cpu_baseline`` leg of ``bench.py``; the product package ``zccl_amd``
never imports this module.

Two libraries:

* ``oracle/liboracle.so`` -- the clean-room C restatement (``zccl_oracle.c``),
  with the chunk count ``T`` as an explicit argument.
* ``oracle/_ref/libzccl_ref.so`` -- the reference codec compiled from its own
  sources by ``oracle/Makefile``'s ``ref`` target.  Its chunk count is the
  OpenMP team size (ZCCL/src/ZCCL_float.c:533), so calls set
  ``omp_set_num_threads(T)`` first.  Optional: ``ref_available()`` says
  whether it was built.
"""
from __future__ import annotations

import ctypes
import os
from typing import Sequence

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None
_REF = None
_GOMP = None

_u8p = ctypes.POINTER(ctypes.c_uint8)
_u64p = ctypes.POINTER(ctypes.c_uint64)


class Geom(ctypes.Structure):
    _fields_ = [
        ("n", ctypes.c_uint64),
        ("L", ctypes.c_int32),
        ("T", ctypes.c_int32),
        ("tb", ctypes.c_uint64),
        ("R", ctypes.c_uint64),
        ("S", ctypes.c_int32),
        ("nbT", ctypes.c_uint64),
        ("nbR", ctypes.c_uint64),
        ("nblocks", ctypes.c_uint64),
    ]


class OracleError(RuntimeError):
    pass


def _ptr(a: np.ndarray):
    return ctypes.c_void_p(a.ctypes.data)


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            raise OracleError(f"{path} not built (run `make -C oracle`)")
        L = ctypes.CDLL(path)
        vp, u64, i32, f32 = ctypes.c_void_p, ctypes.c_uint64, ctypes.c_int, ctypes.c_float
        L.zo_geometry.argtypes = [u64, i32, i32, ctypes.POINTER(Geom)]
        L.zo_geometry.restype = i32
        L.zo_width.argtypes = [ctypes.c_uint32]
        L.zo_width.restype = i32
        L.zo_width_array.argtypes = [vp, u64, vp]
        L.zo_width_array.restype = None
        L.zo_block_bytes.argtypes = [i32, ctypes.c_uint32]
        L.zo_block_bytes.restype = u64
        L.zo_worst_case_bytes.argtypes = [u64, i32, i32]
        L.zo_worst_case_bytes.restype = u64
        L.zo_quantize.argtypes = [vp, u64, f32, vp]
        L.zo_quantize.restype = None
        L.zo_dequantize.argtypes = [vp, u64, f32, vp]
        L.zo_dequantize.restype = None
        L.zo_encode_block.argtypes = [vp, ctypes.c_uint32, vp]
        L.zo_encode_block.restype = u64
        L.zo_encode_block_bitwise.argtypes = [vp, ctypes.c_uint32, vp]
        L.zo_encode_block_bitwise.restype = u64
        L.zo_decode_block.argtypes = [vp, u64, ctypes.c_uint32, vp]
        L.zo_decode_block.restype = ctypes.c_int64
        L.zo_encode_q.argtypes = [vp, u64, i32, i32, vp, u64]
        L.zo_encode_q.restype = ctypes.c_int64
        L.zo_compress.argtypes = [vp, u64, f32, i32, i32, vp, u64]
        L.zo_compress.restype = ctypes.c_int64
        L.zo_decode_q.argtypes = [vp, u64, u64, i32, i32, vp]
        L.zo_decode_q.restype = i32
        L.zo_decompress.argtypes = [vp, u64, u64, f32, i32, i32, vp]
        L.zo_decompress.restype = i32
        L.zo_homsum.argtypes = [vp, vp, i32, u64, i32, i32, vp, u64]
        L.zo_homsum.restype = ctypes.c_int64
        L.zo_parse.argtypes = [vp, u64, u64, i32, i32, vp, vp, vp, vp]
        L.zo_parse.restype = i32
        L.zo_doc_chain.argtypes = [vp, i32, u64, f32, i32, i32, vp, vp]
        L.zo_doc_chain.restype = i32
        f64 = ctypes.c_double
        L.zo_dims_create.argtypes = [i32, vp]
        L.zo_dims_create.restype = None
        L.zo_partition3d.argtypes = [u64, u64, u64, i32, vp]
        L.zo_partition3d.restype = i32
        L.zo_quantize3d.argtypes = [vp, u64, f64, vp]
        L.zo_quantize3d.restype = None
        L.zo_dequantize3d.argtypes = [vp, u64, f64, vp]
        L.zo_dequantize3d.restype = None
        L.zo_worst_case_bytes3d.argtypes = [u64, u64, u64, i32, i32]
        L.zo_worst_case_bytes3d.restype = u64
        L.zo_compress3d.argtypes = [vp, u64, u64, u64, i32, f64, i32, vp, u64]
        L.zo_compress3d.restype = ctypes.c_int64
        L.zo_decompress3d.argtypes = [vp, u64, u64, u64, u64, i32, f64, i32, vp]
        L.zo_decompress3d.restype = i32
        _LIB = L
    return _LIB


def _check(rc, what):
    if rc < 0:
        raise OracleError(f"{what} failed with status {rc}")
    return rc


def geometry(n: int, L: int, T: int) -> Geom:
    g = Geom()
    _check(lib().zo_geometry(n, L, T, ctypes.byref(g)), "geometry")
    return g


def width(m) -> np.ndarray:
    m = np.ascontiguousarray(m, dtype=np.uint32)
    out = np.empty(m.shape, dtype=np.uint8)
    lib().zo_width_array(_ptr(m), m.size, _ptr(out))
    return out


def block_bytes(b: int, n: int) -> int:
    return int(lib().zo_block_bytes(b, n))


def worst_case_bytes(n: int, L: int, T: int) -> int:
    return int(lib().zo_worst_case_bytes(n, L, T))


def quantize(x: np.ndarray, eb: float) -> np.ndarray:
    x = np.ascontiguousarray(x, dtype=np.float32)
    q = np.empty(x.shape, dtype=np.int32)
    lib().zo_quantize(_ptr(x), x.size, eb, _ptr(q))
    return q


def dequantize(q: np.ndarray, eb: float) -> np.ndarray:
    q = np.ascontiguousarray(q, dtype=np.int32)
    x = np.empty(q.shape, dtype=np.float32)
    lib().zo_dequantize(_ptr(q), q.size, eb, _ptr(x))
    return x


def encode_block(d: np.ndarray, bitwise: bool = False) -> bytes:
    d = np.ascontiguousarray(d, dtype=np.int32)
    out = np.zeros(1 + (d.size + 7) // 8 + 4 * d.size + 8, dtype=np.uint8)
    fn = lib().zo_encode_block_bitwise if bitwise else lib().zo_encode_block
    n = fn(_ptr(d), d.size, _ptr(out))
    return out[:n].tobytes()


def decode_block(buf: bytes, n: int) -> np.ndarray:
    a = np.frombuffer(buf, dtype=np.uint8).copy()
    d = np.empty(n, dtype=np.int32)
    _check(lib().zo_decode_block(_ptr(a), a.size, n, _ptr(d)), "decode_block")
    return d


def encode_q(q: np.ndarray, L: int, T: int) -> np.ndarray:
    q = np.ascontiguousarray(q, dtype=np.int32)
    cap = worst_case_bytes(q.size, L, T)
    if cap == 0:
        raise OracleError("invalid geometry")
    out = np.empty(cap, dtype=np.uint8)
    n = _check(lib().zo_encode_q(_ptr(q), q.size, L, T, _ptr(out), cap), "encode_q")
    return out[:n].copy()


def compress(x: np.ndarray, eb: float, L: int, T: int) -> np.ndarray:
    x = np.ascontiguousarray(x, dtype=np.float32)
    cap = worst_case_bytes(x.size, L, T)
    if cap == 0:
        raise OracleError("invalid geometry")
    out = np.empty(cap, dtype=np.uint8)
    n = _check(lib().zo_compress(_ptr(x), x.size, eb, L, T, _ptr(out), cap), "compress")
    return out[:n].copy()


def decode_q(stream: np.ndarray, n: int, L: int, T: int) -> np.ndarray:
    s = np.ascontiguousarray(stream, dtype=np.uint8)
    q = np.empty(n, dtype=np.int32)
    _check(lib().zo_decode_q(_ptr(s), s.size, n, L, T, _ptr(q)), "decode_q")
    return q


def decompress(stream: np.ndarray, n: int, eb: float, L: int, T: int) -> np.ndarray:
    s = np.ascontiguousarray(stream, dtype=np.uint8)
    x = np.empty(n, dtype=np.float32)
    _check(lib().zo_decompress(_ptr(s), s.size, n, eb, L, T, _ptr(x)), "decompress")
    return x


def homsum(streams: Sequence[np.ndarray], n: int, L: int, T: int) -> np.ndarray:
    ss = [np.ascontiguousarray(s, dtype=np.uint8) for s in streams]
    k = len(ss)
    ptrs = (ctypes.c_void_p * k)(*[s.ctypes.data for s in ss])
    sizes = (ctypes.c_uint64 * k)(*[s.size for s in ss])
    cap = worst_case_bytes(n, L, T)
    out = np.empty(cap, dtype=np.uint8)
    m = _check(lib().zo_homsum(ptrs, sizes, k, n, L, T, _ptr(out), cap), "homsum")
    return out[:m].copy()


def doc_chain(xs: Sequence[np.ndarray], eb: float, L: int, T: int):
    """The reference's DOC ring reduction of one segment (ranks' copies in ring
    order): returns (last rank's FP32 partial sum, value after the compressed
    allgather)."""
    xs = [np.ascontiguousarray(x, dtype=np.float32) for x in xs]
    n = xs[0].size
    ptrs = (ctypes.c_void_p * len(xs))(*[x.ctypes.data for x in xs])
    acc = np.empty(n, dtype=np.float32)
    fin = np.empty(n, dtype=np.float32)
    _check(lib().zo_doc_chain(ptrs, len(xs), n, ctypes.c_float(eb), L, T, _ptr(acc), _ptr(fin)), "doc_chain")
    return acc, fin


def parse(stream: np.ndarray, n: int, L: int, T: int):
    """Return (widths[nblocks] u8, block_off[nblocks] u64, seed_off[S] u64, total)."""
    g = geometry(n, L, T)
    s = np.ascontiguousarray(stream, dtype=np.uint8)
    w = np.empty(g.nblocks, dtype=np.uint8)
    bo = np.empty(g.nblocks, dtype=np.uint64)
    so = np.empty(g.S, dtype=np.uint64)
    tot = ctypes.c_uint64(0)
    _check(lib().zo_parse(_ptr(s), s.size, n, L, T, _ptr(w), _ptr(bo), _ptr(so), ctypes.byref(tot)),
           "parse")
    return w, bo, so, int(tot.value)


# ------------------------------------------------------------------- 3-D codec

def dims_create(T: int):
    d = np.zeros(3, dtype=np.int32)
    lib().zo_dims_create(T, _ptr(d))
    return tuple(int(v) for v in d)


def partition3d(dims, T: int) -> np.ndarray:
    """Sub-volumes (ox, oy, oz, sx, sy, sz) of the T-way split, rows in thread order."""
    p = np.zeros((T, 6), dtype=np.uint64)
    _check(lib().zo_partition3d(dims[0], dims[1], dims[2], T, _ptr(p)), "partition3d")
    return p.astype(np.int64)


def quantize3d(x: np.ndarray, eb: float) -> np.ndarray:
    x = np.ascontiguousarray(x, dtype=np.float32)
    q = np.empty(x.shape, dtype=np.int32)
    lib().zo_quantize3d(_ptr(x), x.size, eb, _ptr(q))
    return q


def dequantize3d(q: np.ndarray, eb: float) -> np.ndarray:
    q = np.ascontiguousarray(q, dtype=np.int32)
    x = np.empty(q.shape, dtype=np.float32)
    lib().zo_dequantize3d(_ptr(q), q.size, eb, _ptr(x))
    return x


def worst_case_bytes3d(dims, Bs: int, T: int) -> int:
    return int(lib().zo_worst_case_bytes3d(dims[0], dims[1], dims[2], Bs, T))


def compress3d(x: np.ndarray, dims, Bs: int, eb: float, T: int) -> np.ndarray:
    x = np.ascontiguousarray(x, dtype=np.float32).reshape(-1)
    cap = worst_case_bytes3d(dims, Bs, T)
    if cap == 0:
        raise OracleError("invalid 3-D geometry")
    out = np.empty(cap, dtype=np.uint8)
    n = _check(lib().zo_compress3d(_ptr(x), dims[0], dims[1], dims[2], Bs, eb, T, _ptr(out), cap), "compress3d")
    return out[:n].copy()


def decompress3d(stream: np.ndarray, dims, Bs: int, eb: float, T: int) -> np.ndarray:
    s = np.ascontiguousarray(stream, dtype=np.uint8)
    x = np.empty(int(dims[0]) * int(dims[1]) * int(dims[2]), dtype=np.float32)
    _check(lib().zo_decompress3d(_ptr(s), s.size, dims[0], dims[1], dims[2], Bs, eb, T, _ptr(x)), "decompress3d")
    return x


# ----------------------------------------------------------------- reference

def ref_path() -> str:
    return os.path.join(_HERE, "_ref", "libzccl_ref.so")


def ref_available() -> bool:
    return os.path.exists(ref_path())


def ref():
    global _REF, _GOMP
    if _REF is None:
        if not ref_available():
            raise OracleError("oracle/_ref/libzccl_ref.so not built (make -C oracle ref)")
        R = ctypes.CDLL(ref_path())
        vp, sz, f32, i32 = ctypes.c_void_p, ctypes.c_size_t, ctypes.c_float, ctypes.c_int
        R.ZCCL_float_openmp_threadblock_arg.argtypes = [vp, vp, vp, f32, sz, i32]
        R.ZCCL_float_openmp_threadblock_arg.restype = None
        R.ZCCL_float_single_thread_arg.argtypes = [vp, vp, vp, f32, sz, i32]
        R.ZCCL_float_single_thread_arg.restype = None
        R.ZCCL_float_decompress_openmp_threadblock_arg.argtypes = [vp, sz, f32, i32, vp]
        R.ZCCL_float_decompress_openmp_threadblock_arg.restype = None
        R.ZCCL_float_decompress_single_thread_arg.argtypes = [vp, sz, f32, i32, vp]
        R.ZCCL_float_decompress_single_thread_arg.restype = None
        R.ZCCL_float_homomophic_add_openmp_threadblock.argtypes = [vp, vp, sz, f32, i32, vp, vp]
        R.ZCCL_float_homomophic_add_openmp_threadblock.restype = None
        R.ZCCL_float_homomophic_add_single_thread.argtypes = [vp, vp, sz, f32, i32, vp, vp]
        R.ZCCL_float_homomophic_add_single_thread.restype = None
        R.SZp_compress3D_fast_openmp.argtypes = [vp, vp, sz, sz, sz, i32, ctypes.c_double, vp]
        R.SZp_compress3D_fast_openmp.restype = None
        R.SZp_decompress3D_fast_openmp.argtypes = [vp, vp, sz, sz, sz, i32, ctypes.c_double]
        R.SZp_decompress3D_fast_openmp.restype = None
        _GOMP = ctypes.CDLL("libgomp.so.1")
        _GOMP.omp_set_num_threads.argtypes = [ctypes.c_int]
        _GOMP.omp_get_max_threads.restype = ctypes.c_int
        _REF = R
    return _REF


def _set_T(T: int):
    ref()
    _GOMP.omp_set_num_threads(int(T))


def chunk_payload_sizes(stream: np.ndarray, T: int) -> np.ndarray:
    """Bytes of each chunk payload (header offsets; the last runs to the end)."""
    h = np.frombuffer(np.ascontiguousarray(stream[: 8 * T]).tobytes(), dtype=np.uint64).astype(np.int64)
    ends = np.append(h[1:], stream.size - 8 * T)
    return ends - h


def ref_buffer_fits(stream: np.ndarray, n: int, T: int) -> bool:
    """Whether the reference's per-thread scratch of ceil(4n/T) bytes
    (ZCCL/src/ZCCL_float.c:538,549; ZCCL/src/ZCCLd_float.c:1032,1044) holds
    every chunk payload of ``stream``, counting the slack glibc gives the
    allocation.  The reference writes past that buffer (heap overflow) when a
    chunk compresses to more than 4 bytes per element -- and for n = T+1,
    where chunk T-1 carries two seeds -- so reference runs are only made
    where this holds."""
    need = int(chunk_payload_sizes(stream, T).max())
    return need <= _usable_bytes((4 * n + T - 1) // T)


_LIBC = None


def _usable_bytes(request: int) -> int:
    """Bytes glibc actually provides for malloc(request) (malloc_usable_size):
    writes up to this size stay inside the reference's own allocation."""
    global _LIBC
    if _LIBC is None:
        _LIBC = ctypes.CDLL("libc.so.6")
        _LIBC.malloc.argtypes = [ctypes.c_size_t]
        _LIBC.malloc.restype = ctypes.c_void_p
        _LIBC.malloc_usable_size.argtypes = [ctypes.c_void_p]
        _LIBC.malloc_usable_size.restype = ctypes.c_size_t
        _LIBC.free.argtypes = [ctypes.c_void_p]
    p = _LIBC.malloc(max(1, request))
    try:
        return int(_LIBC.malloc_usable_size(p))
    finally:
        _LIBC.free(p)


def ref_compress(x: np.ndarray, eb: float, L: int, T: int, single_thread: bool = False) -> np.ndarray:
    """Reference ZCCL_float_openmp_threadblock_arg with an OpenMP team of T.

    Raises OracleError instead of calling the reference when its per-thread
    buffer would overflow (see ``ref_buffer_fits``)."""
    x = np.ascontiguousarray(x, dtype=np.float32)
    if not ref_buffer_fits(compress(x, eb, L, T), x.size, 1 if single_thread else T):
        raise OracleError("reference per-thread buffer would overflow for this input")
    R = ref()
    _set_T(T)
    # the reference's per-thread scratch is 4*ceil(n/T) bytes and its callers
    # size outputs at 4n; give it the worst case so nothing overruns
    cap = max(worst_case_bytes(x.size, L, T), 4 * x.size) + 64
    out = np.zeros(cap, dtype=np.uint8)
    size = ctypes.c_size_t(0)
    fn = R.ZCCL_float_single_thread_arg if single_thread else R.ZCCL_float_openmp_threadblock_arg
    fn(_ptr(out), _ptr(x), ctypes.byref(size), eb, x.size, L)
    return out[: size.value].copy()


def ref_decompress(stream: np.ndarray, n: int, eb: float, L: int, T: int,
                   single_thread: bool = False) -> np.ndarray:
    s = np.ascontiguousarray(stream, dtype=np.uint8)
    buf = np.zeros(s.size + 256, dtype=np.uint8)  # the 3-bit unpacker reads one byte past a block
    buf[: s.size] = s
    R = ref()
    _set_T(T)
    x = np.zeros(n, dtype=np.float32)
    fn = (R.ZCCL_float_decompress_single_thread_arg if single_thread
          else R.ZCCL_float_decompress_openmp_threadblock_arg)
    fn(_ptr(x), n, eb, L, _ptr(buf))
    return x


def ref_add2(a: np.ndarray, b: np.ndarray, n: int, L: int, T: int, single_thread: bool = False,
             initial_size: int = 0):
    """Reference pairwise homomorphic add; returns (bytes, reported size).

    The reference *adds* to ``*final_cmpSize`` (ZCCL/src/ZCCLd_float.c:1027,1655),
    so the reported size is ``initial_size`` + the stream size.
    """
    if not ref_buffer_fits(homsum([a, b], n, L, T), n, T):
        raise OracleError("reference per-thread buffer would overflow for this input")
    R = ref()
    _set_T(T)
    pa = np.zeros(a.size + 256, dtype=np.uint8)
    pa[: a.size] = a
    pb = np.zeros(b.size + 256, dtype=np.uint8)
    pb[: b.size] = b
    cap = max(worst_case_bytes(n, L, T), 4 * n) + 64
    out = np.zeros(cap, dtype=np.uint8)
    size = ctypes.c_size_t(initial_size)
    fn = (R.ZCCL_float_homomophic_add_single_thread if single_thread
          else R.ZCCL_float_homomophic_add_openmp_threadblock)
    fn(_ptr(out), ctypes.byref(size), n, ctypes.c_float(0.0), L, _ptr(pa), _ptr(pb))
    m = size.value - initial_size
    return out[:m].copy(), size.value


def ref_compress3d(x: np.ndarray, dims, Bs: int, eb: float, T: int) -> np.ndarray:
    """Reference SZp_compress3D_fast_openmp with an OpenMP team of T.

    Its per-thread scratch holds ceil(4n/T) bytes (ZCCL/src/ZCCL_float.c:1896,
    1948); callers here keep every sub-volume's payload below that (the oracle's
    stream is checked first) so the reference never writes past it."""
    x = np.ascontiguousarray(x, dtype=np.float32).reshape(-1)
    mine = compress3d(x, dims, Bs, eb, T)
    offs = np.frombuffer(mine[: 8 * T].tobytes(), dtype=np.uint64).astype(np.int64)
    sizes = np.append(offs[1:], mine.size - 8 * T) - offs
    if int(sizes.max()) > _usable_bytes((4 * x.size + T - 1) // T):
        raise OracleError("reference per-thread buffer would overflow for this input")
    R = ref()
    _set_T(T)
    out = np.zeros(max(worst_case_bytes3d(dims, Bs, T), 4 * x.size) + 64, dtype=np.uint8)
    size = ctypes.c_size_t(0)
    R.SZp_compress3D_fast_openmp(_ptr(x), _ptr(out), dims[0], dims[1], dims[2], Bs, eb, ctypes.byref(size))
    return out[: size.value].copy()


def ref_decompress3d(stream: np.ndarray, dims, Bs: int, eb: float, T: int) -> np.ndarray:
    s = np.ascontiguousarray(stream, dtype=np.uint8)
    buf = np.zeros(s.size + 256, dtype=np.uint8)  # the unpackers read a byte past a block
    buf[: s.size] = s
    R = ref()
    _set_T(T)
    x = np.zeros(int(dims[0]) * int(dims[1]) * int(dims[2]), dtype=np.float32)
    R.SZp_decompress3D_fast_openmp(_ptr(x), _ptr(buf), dims[0], dims[1], dims[2], Bs, eb)
    return x
