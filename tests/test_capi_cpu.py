"""CPU-side checks of the C ABI library: it loads without a GPU, exports every
symbol include/zccl_amd.h declares, and its size queries agree with the oracle's
geometry.  No compute entry point is called here."""
import ctypes
import os
import re

import pytest

from oracle import oracle as O
from zccl_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "zccl_amd.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    text = re.sub(r"^#.*$", "", text, flags=re.M)
    return sorted(set(re.findall(r"\b((?:zccl|ZCCL|SZp)_\w+)\s*\(", text)))


def test_library_loads_and_exports_every_declared_symbol():
    L = _lib.lib()
    decl = declared_symbols()
    assert len(decl) >= 25
    for name in decl:
        assert hasattr(L, name), name
    assert set(decl) == set(_lib.EXPORTS)


@pytest.mark.parametrize("n,L,T", [(1 << 20, 36, 16), (1000, 7, 3), (8, 36, 8), (9, 36, 8), (5000, 128, 1),
                                   (100000, 4096, 4), (33, 1, 2)])
def test_sizes_agree_with_oracle(n, L, T):
    lib = _lib.lib()
    assert lib.zccl_worst_case_bytes(n, L, T) == O.worst_case_bytes(n, L, T)
    assert lib.zccl_num_blocks(n, L, T) == O.geometry(n, L, T).nblocks
    assert lib.zccl_index_bytes(n, L, T) > lib.zccl_num_blocks(n, L, T)


def test_invalid_geometry_reports_zero():
    lib = _lib.lib()
    assert lib.zccl_worst_case_bytes(3, 36, 8) == 0      # n < T
    assert lib.zccl_worst_case_bytes(100, 0, 1) == 0     # L < 1
    assert lib.zccl_worst_case_bytes(100, 4097, 1) == 0  # L above the tile kernels' limit


def test_chunk_count_defaults_to_openmp_team_and_can_be_overridden():
    lib = _lib.lib()
    gomp = ctypes.CDLL("libgomp.so.1")
    assert lib.zccl_get_num_chunks() == gomp.omp_get_max_threads()
    lib.zccl_set_num_chunks(5)
    assert lib.zccl_get_num_chunks() == 5
    lib.zccl_set_num_chunks(0)
    assert lib.zccl_get_num_chunks() == gomp.omp_get_max_threads()


@pytest.mark.parametrize("dims,Bs,T", [((16, 16, 16), 4, 8), ((13, 11, 17), 5, 3), ((512, 512, 512), 4, 16),
                                       ((40, 1, 1), 4, 2), ((9, 8, 10), 3, 12)])
def test_3d_sizes_agree_with_oracle(dims, Bs, T):
    lib = _lib.lib()
    worst, ws = ctypes.c_size_t(0), ctypes.c_size_t(0)
    assert lib.zccl_sz3d_sizes(*dims, Bs, T, ctypes.byref(worst), ctypes.byref(ws)) == 0
    assert worst.value == O.worst_case_bytes3d(dims, Bs, T)
    assert ws.value > 0


def test_3d_invalid_geometry_rejected():
    lib = _lib.lib()
    worst = ctypes.c_size_t(0)
    assert lib.zccl_sz3d_sizes(2, 2, 2, 4, 16, ctypes.byref(worst), None) != 0   # empty sub-volume
    assert lib.zccl_sz3d_sizes(8, 8, 8, 0, 1, ctypes.byref(worst), None) != 0    # block side < 1
    assert lib.zccl_sz3d_sizes(8, 8, 8, 65, 1, ctypes.byref(worst), None) != 0   # block side > 64


def test_reference_headers_agree_with_ours(tmp_path):
    """One C translation unit includes the reference's own prototypes
    (ZCCL/include/ZCCL_float.h, ZCCLd_float.h) and include/zccl_amd.h: any
    drop-in declaration whose types differ from the reference's fails to
    compile as conflicting types.  Skipped where the reference tree is
    absent."""
    import subprocess

    ref_inc = "/root/reference/ZCCL/include"
    if not os.path.isdir(ref_inc):
        pytest.skip("reference headers not present")
    src = tmp_path / "both.c"
    src.write_text('#include "ZCCL_float.h"\n#include "ZCCLd_float.h"\n#include "zccl_amd.h"\n'
                   "void *keep[] = {\n"
                   "  (void *) ZCCL_float_openmp_threadblock_arg, (void *) ZCCL_float_single_thread_arg,\n"
                   "  (void *) ZCCL_float_openmp_threadblock,\n"
                   "  (void *) ZCCL_float_decompress_openmp_threadblock_arg,\n"
                   "  (void *) ZCCL_float_decompress_single_thread_arg,\n"
                   "  (void *) ZCCL_float_decompress_openmp_threadblock,\n"
                   "  (void *) ZCCL_float_homomophic_add_openmp_threadblock,\n"
                   "  (void *) ZCCL_float_homomophic_add_single_thread,\n"
                   "  (void *) ZCCL_float_homomophic_add_openmp_threadblock_record,\n"
                   "};\n")
    out = subprocess.run(["gcc", "-std=gnu99", "-fsyntax-only", "-I", ref_inc, "-I",
                          os.path.join(ROOT, "include"), "-I", "/opt/conda/include", str(src)],
                         capture_output=True, text=True)
    assert out.returncode == 0, out.stderr
