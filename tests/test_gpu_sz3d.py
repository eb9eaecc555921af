"""GPU parity of the 3-D SZp codec (SZp_compress3D_fast_openmp /
SZp_decompress3D_fast_openmp) against the committed reference fixtures and
the oracle: stream bytes and decompressed bits equal, every value within the
bound, the drop-in entry points with host and device pointers, and the
size-independent properties at 512^3 (the round trip equals
(float)(floor(x * 0.5/eb + 0.5) * 2 eb) for any split)."""
import ctypes
import json
import os

import numpy as np
import pytest

torch = pytest.importorskip("torch")

from oracle import oracle as O  # noqa: E402
from tests.golden import gen  # noqa: E402

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))


def say(*a):
    print("[gpu3d]", *a, flush=True)


@pytest.fixture(scope="module")
def sz():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    torch.cuda.set_device(0)
    from zccl_amd import sz3d

    return sz3d


def bits(a):
    return np.ascontiguousarray(a).view(np.uint32)


def fixture_cases():
    with open(os.path.join(HERE, "golden", "fixtures_3d.json")) as f:
        return json.load(f)["cases"]


@pytest.mark.parametrize("m", fixture_cases(), ids=lambda m: m["name"])
def test_fixtures_bit_exact(sz, m):
    arrs = np.load(os.path.join(HERE, "golden", "fixtures_3d.npz"))
    x, c, d = arrs[m["name"] + "/x"], arrs[m["name"] + "/c"], arrs[m["name"] + "/d"]
    dims, Bs, eb, T = tuple(m["dims"]), m["Bs"], m["eb"], m["T"]
    say(m["name"])
    s = sz.compress(torch.from_numpy(x).reshape(dims).cuda(), Bs, eb, T)
    got = s.to_host()
    assert got.size == c.size and np.array_equal(got, c)
    y = sz.decompress(s).cpu().numpy().reshape(-1)
    assert np.array_equal(bits(y), bits(d))
    # the reference's stream, uploaded, decodes to the same bits
    y2 = sz.decompress(sz.upload(c, dims, Bs, eb, T)).cpu().numpy().reshape(-1)
    assert np.array_equal(bits(y2), bits(d))


@pytest.mark.parametrize("T", [1, 3, 8, 16])
@pytest.mark.parametrize("Bs", [2, 4, 5, 8])
def test_random_shapes_equal_oracle(sz, T, Bs):
    rng = np.random.default_rng(31 * T + Bs)
    dims = tuple(int(v) for v in rng.integers(max(4, T), 40, 3))
    x = gen.smooth3d(dims, T + 100 * Bs) * np.float32(rng.uniform(1, 50))
    for eb in (1e-3, 1e-1):
        say("random", dims, Bs, T, eb)
        s = sz.compress(torch.from_numpy(x).cuda(), Bs, eb, T)
        want = O.compress3d(x, dims, Bs, eb, T)
        assert np.array_equal(s.to_host(), want)
        y = sz.decompress(s).cpu().numpy().reshape(-1)
        assert np.array_equal(bits(y), bits(O.decompress3d(want, dims, Bs, eb, T)))
        assert float(np.abs(y.astype(np.float64) - x.reshape(-1)).max()) <= eb * (1 + 1e-6) + \
            2.0 ** -23 * float(np.abs(x).max())


@pytest.mark.parametrize("dims,Bs,T", [((9, 10, 300), 4, 2), ((5, 7, 517), 4, 1), ((12, 9, 150), 8, 3),
                                       ((17, 16, 64), 8, 1), ((8, 8, 65), 8, 1)])
def test_chunked_columns_equal_oracle(sz, dims, Bs, T):
    """Columns of blocks longer than one LDS chunk along z (256 values for
    block side 4, 64 for 8), with partial last chunks and blocks: stream
    bytes and decoded bits equal the oracle's."""
    x = gen.smooth3d(dims, 7 * Bs + T) * np.float32(20)
    say("chunked", dims, Bs, T)
    s = sz.compress(torch.from_numpy(x).cuda(), Bs, 1e-3, T)
    want = O.compress3d(x, dims, Bs, 1e-3, T)
    assert np.array_equal(s.to_host(), want)
    y = sz.decompress(s).cpu().numpy().reshape(-1)
    assert np.array_equal(bits(y), bits(O.decompress3d(want, dims, Bs, 1e-3, T)))


def test_special_values_and_wide_residuals(sz):
    """Zeros, huge values (widths up to 32), NaN and infinities (INT_MIN
    quantization) give the oracle's bytes."""
    dims, Bs, T = (12, 9, 14), 4, 4
    rng = np.random.default_rng(9)
    x = rng.normal(0, 1, dims).astype(np.float32)
    x[0] = 0
    x[1, :, :3] = 3e9
    x[2, 4, 5] = np.nan
    x[3, 1, 1] = np.inf
    x[4, 2, 2] = -np.inf
    x[5] = rng.normal(0, 1e6, (9, 14)).astype(np.float32)
    s = sz.compress(torch.from_numpy(x).cuda(), Bs, 1e-3, T)
    want = O.compress3d(x, dims, Bs, 1e-3, T)
    assert np.array_equal(s.to_host(), want)
    y = sz.decompress(s).cpu().numpy().reshape(-1)
    assert np.array_equal(bits(y), bits(O.decompress3d(want, dims, Bs, 1e-3, T)))


def test_dropin_host_and_device_pointers(sz):
    from zccl_amd._lib import lib

    L = lib()
    dims, Bs, eb, T = (20, 24, 28), 4, 1e-3, 6
    x = gen.smooth3d(dims, 5).reshape(-1)
    L.zccl_set_num_chunks(T)
    try:
        want = O.compress3d(x, dims, Bs, eb, T)
        out = np.zeros(want.size + 4096, dtype=np.uint8)
        size = ctypes.c_size_t(123)
        L.SZp_compress3D_fast_openmp(x.ctypes.data, out.ctypes.data, *dims, Bs, eb, ctypes.byref(size))
        assert L.zccl_last_error() == 0
        assert size.value == want.size and np.array_equal(out[: size.value], want)
        y = np.zeros(x.size, dtype=np.float32)
        L.SZp_decompress3D_fast_openmp(y.ctypes.data, out.ctypes.data, *dims, Bs, eb)
        assert L.zccl_last_error() == 0
        assert np.array_equal(bits(y), bits(O.decompress3d(want, dims, Bs, eb, T)))
        # device pointers
        dx = torch.from_numpy(x).cuda()
        dout = torch.zeros(want.size + 4096, dtype=torch.uint8, device="cuda")
        L.SZp_compress3D_fast_openmp(dx.data_ptr(), dout.data_ptr(), *dims, Bs, eb, ctypes.byref(size))
        assert L.zccl_last_error() == 0 and size.value == want.size
        assert np.array_equal(dout[: size.value].cpu().numpy(), want)
        dy = torch.zeros(x.size, dtype=torch.float32, device="cuda")
        L.SZp_decompress3D_fast_openmp(dy.data_ptr(), dout.data_ptr(), *dims, Bs, eb)
        assert L.zccl_last_error() == 0
        assert np.array_equal(bits(dy.cpu().numpy()), bits(y))
    finally:
        L.zccl_set_num_chunks(0)


def test_errors_reported(sz):
    from zccl_amd._lib import ZcclError

    with pytest.raises(ZcclError):
        sz.sizes((2, 2, 2), 4, 16)  # a sub-volume would be empty
    with pytest.raises(ZcclError):
        sz.sizes((8, 8, 8), 65, 1)  # block side out of range
    x = torch.zeros((8, 8, 8), dtype=torch.float32, device="cuda")
    x.normal_()
    s = sz.compress(x, 4, 1e-6, 2, out=torch.empty(100, dtype=torch.uint8, device="cuda"))
    with pytest.raises(ZcclError):
        s.size  # output capacity too small


@pytest.mark.parametrize("T", [1, 16])
def test_512cube_round_trip_closed_form(sz, T):
    """Full-size property: decompress(compress(x)) is (float)(q * 2 eb) with
    q = floor(x * 0.5/eb + 0.5), whatever the split, checked on the GPU
    against the same closed form computed by torch in float64."""
    from zccl_amd import synth

    side, Bs, eb = 512, 4, 1e-3
    x = synth.grf3d(side, seed=77).reshape(side, side, side).contiguous()
    say("512^3", T)
    s = sz.compress(x, Bs, eb, T)
    y = sz.decompress(s)
    q = torch.floor(x.double() * (0.5 / eb) + 0.5)
    want = (q * (2 * eb)).float()
    assert torch.equal(y.view(torch.int32), want.view(torch.int32))
    say("bytes", s.size, "CR", 4 * side ** 3 / s.size)
