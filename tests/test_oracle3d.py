"""Pins the oracle's restatement of the 3-D SZp codec (SZp_compress3D_fast_openmp,
ZCCL/src/ZCCL_float.c:1862-2048; SZp_decompress3D_fast_openmp,
ZCCL/src/ZCCLd_float.c:2873-3040) against the committed reference fixtures,
the compiled reference where it is built, and closed forms that need neither:
the quantizer against numpy, decompress(compress(x)) against
dequantize(quantize(x)) for every split, and the block widths against a numpy
3-D difference of the quantized field."""
import ctypes
import json
import os

import numpy as np
import pytest

from oracle import oracle as O
from tests.golden import gen

HERE = os.path.dirname(os.path.abspath(__file__))
FIX = os.path.join(HERE, "golden")


def fixture_cases():
    with open(os.path.join(FIX, "fixtures_3d.json")) as f:
        meta = json.load(f)["cases"]
    arrs = np.load(os.path.join(FIX, "fixtures_3d.npz"))
    return [(m, arrs[m["name"] + "/x"], arrs[m["name"] + "/c"], arrs[m["name"] + "/d"]) for m in meta]


@pytest.mark.parametrize("case", fixture_cases(), ids=lambda c: c[0]["name"])
def test_oracle3d_matches_reference_fixtures(case):
    m, x, c, d = case
    dims, Bs, eb, T = tuple(m["dims"]), m["Bs"], m["eb"], m["T"]
    got = O.compress3d(x, dims, Bs, eb, T)
    assert got.size == c.size and np.array_equal(got, c)
    y = O.decompress3d(c, dims, Bs, eb, T)
    assert np.array_equal(y.view(np.uint32), d.view(np.uint32))


def test_dims_create_known_values():
    want = {1: (1, 1, 1), 2: (2, 1, 1), 6: (3, 2, 1), 8: (2, 2, 2), 12: (3, 2, 2), 16: (4, 2, 2),
            30: (5, 3, 2), 64: (4, 4, 4), 7: (7, 1, 1), 36: (4, 3, 3)}
    for T, d in want.items():
        assert O.dims_create(T) == d, T


@pytest.mark.skipif(not O.ref_available(), reason="compiled reference not built")
def test_dims_create_equals_reference():
    R = O.ref()
    R.My_Dims_create.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_void_p]
    R.My_Dims_create.restype = None
    for T in range(1, 65):
        d = np.zeros(3, dtype=np.int32)
        R.My_Dims_create(T, 3, d.ctypes.data)
        assert O.dims_create(T) == tuple(int(v) for v in d), T


@pytest.mark.skipif(not O.ref_available(), reason="compiled reference not built")
@pytest.mark.parametrize("T", [1, 2, 3, 5, 8, 12])
@pytest.mark.parametrize("Bs", [2, 4, 6])
def test_oracle3d_equals_reference_random(T, Bs):
    rng = np.random.default_rng(100 * T + Bs)
    dims = tuple(int(v) for v in rng.integers(max(3, T), 21, 3))
    x = gen.smooth3d(dims, 7 * T + Bs).reshape(-1) * np.float32(rng.uniform(0.5, 20))
    for eb in (1e-3, 1e-1):
        c = O.compress3d(x, dims, Bs, eb, T)
        assert np.array_equal(c, O.ref_compress3d(x, dims, Bs, eb, T))
        y = O.decompress3d(c, dims, Bs, eb, T)
        assert np.array_equal(y.view(np.uint32), O.ref_decompress3d(c, dims, Bs, eb, T).view(np.uint32))


def numpy_quantize3d(x, eb):
    p = np.floor(x.astype(np.float64) * (0.5 / eb) + 0.5)
    ok = (p >= -2.0 ** 31) & (p < 2.0 ** 31)
    return np.where(ok, np.where(ok, p, 0).astype(np.int64), -2 ** 31).astype(np.int32)


def test_quantize3d_closed_form():
    rng = np.random.default_rng(5)
    x = np.concatenate([rng.normal(0, 10, 5000), [0.0, -0.0, 0.25, -0.25, 0.75, 1e30, -1e30, np.nan, np.inf,
                                                  -np.inf, 2.0 ** 31, -(2.0 ** 31)]]).astype(np.float32)
    for eb in (1e-3, 0.5, 3.0):
        assert np.array_equal(O.quantize3d(x, eb), numpy_quantize3d(x, eb))
    q = numpy_quantize3d(x[:5000], 1e-3)
    assert np.array_equal(O.dequantize3d(q, 1e-3), (q.astype(np.float64) * 2e-3).astype(np.float32))


@pytest.mark.parametrize("T", [1, 4, 7, 16])
@pytest.mark.parametrize("Bs", [1, 3, 4, 8])
def test_round_trip_is_dequantized_quantizer_for_every_split(T, Bs):
    """The residuals are an exact 3-D difference and the decoder its exact
    inverse, so the output depends only on x and eb: (float)(q * 2 eb)."""
    dims = (17, 12, 19)
    x = gen.smooth3d(dims, 11).reshape(-1) * np.float32(5.0)
    eb = 1e-3
    y = O.decompress3d(O.compress3d(x, dims, Bs, eb, T), dims, Bs, eb, T)
    want = O.dequantize3d(numpy_quantize3d(x, eb), eb)
    assert np.array_equal(y.view(np.uint32), want.view(np.uint32))
    assert float(np.abs(y.astype(np.float64) - x).max()) <= eb * (1 + 1e-6) + 2.0 ** -23 * float(np.abs(x).max())


def test_block_widths_are_clz_of_numpy_3d_difference():
    """T = 1: the width header holds 32 - clz(max |dx dy dz q|) per block,
    blocks in x, y, z order; the difference is taken with zero padding."""
    dims, Bs, eb = (9, 10, 11), 4, 1e-3
    x = gen.smooth3d(dims, 3).reshape(-1) * np.float32(3.0)
    q = numpy_quantize3d(x, eb).reshape(dims).astype(np.int64)
    r = q.copy()
    for ax in range(3):
        r = np.diff(r, axis=ax, prepend=0)
    r = (r + 2 ** 31) % 2 ** 32 - 2 ** 31  # int32 wrap
    c = O.compress3d(x, dims, Bs, eb, 1)
    nb = [(d - 1) // Bs + 1 for d in dims]
    widths = c[8: 8 + nb[0] * nb[1] * nb[2]]
    i = 0
    for bx in range(nb[0]):
        for by in range(nb[1]):
            for bz in range(nb[2]):
                blk = np.abs(r[bx * Bs:(bx + 1) * Bs, by * Bs:(by + 1) * Bs, bz * Bs:(bz + 1) * Bs]).max()
                assert widths[i] == (int(blk).bit_length() if blk else 0)
                i += 1
