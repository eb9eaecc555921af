"""The allocation-free sum (zccl_homsum_dev_noalloc) against the workspace
entry point it varies (zccl_homsum_dev): same stream bytes, same index, same
size, on calls in a row whose second input is smaller than the first (the
device array keeps the last call's contents), and against the oracle."""
import ctypes

import numpy as np
import pytest

torch = pytest.importorskip("torch")

from oracle import oracle as O  # noqa: E402

pytestmark = pytest.mark.gpu


def say(*a):
    print("[gpu]", *a, flush=True)


@pytest.fixture(scope="module")
def codec():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from zccl_amd import codec as c

    torch.cuda.set_device(0)
    return c


def field(n, seed, r):
    rng = np.random.default_rng(seed)
    return (np.cumsum(rng.normal(0, 1, n)) * 0.05 + 0.01 * r).astype(np.float32)


def zeroed_stream(codec, n, L, T):
    # (index padding is never written: zero it so whole indexes compare)
    s = codec.empty_stream(n, L, T)
    s.index.zero_()
    return s


def noalloc_sum(codec, streams):
    from zccl_amd._lib import lib

    s0 = streams[0]
    k = len(streams)
    out = zeroed_stream(codec, s0.n, s0.L, s0.T)
    ins = (ctypes.c_void_p * k)(*[s.data.data_ptr() for s in streams])
    ixs = (ctypes.c_void_p * k)(*[s.index.data_ptr() for s in streams])
    d_size = torch.zeros(1, dtype=torch.int64, device="cuda")
    rc = lib().zccl_homsum_dev_noalloc(ins, ixs, k, s0.n, s0.L, s0.T, ctypes.c_void_p(out.data.data_ptr()),
                                       out.data.numel(), ctypes.c_void_p(out.index.data_ptr()),
                                       ctypes.c_void_p(d_size.data_ptr()),
                                       ctypes.c_void_p(torch.cuda.current_stream().cuda_stream))
    assert rc == 0, rc
    torch.cuda.synchronize()
    return out, int(d_size.item())


@pytest.mark.parametrize("L", [36, 32])
def test_noalloc_sum_equals_workspace_entry_point(codec, L):
    T, eb = 8, 1e-3
    # the second call sums fewer, shorter streams than the first
    for n, k, seed in [(1_000_003, 5, 11), (120_001, 3, 12)]:
        say("noalloc", "L", L, "n", n, "k", k)
        xs = [field(n, seed, r) for r in range(k)]
        streams = [codec.compress(torch.from_numpy(x).cuda(), eb, L, T) for x in xs]
        ref = codec.homsum(streams, out=zeroed_stream(codec, n, L, T))
        torch.cuda.synchronize()
        codec.check_errors()
        out, size = noalloc_sum(codec, streams)
        assert size == ref.size
        assert out.size == ref.size
        assert torch.equal(out.data[:size], ref.data[: ref.size])
        assert torch.equal(out.index, ref.index)
        if n < 200_000:
            hosts = [s.to_host() for s in streams]
            assert np.array_equal(out.to_host(), O.homsum(hosts, n, L, T))


def test_noalloc_sum_rejects_geometry_beyond_its_array(codec):
    from zccl_amd._lib import lib

    n = (1 << 31) - 1  # needs far more lookback tiles than the fixed array holds
    rc = lib().zccl_homsum_dev_noalloc(None, None, 2, n, 36, 1, None, 0, None, None, None)
    assert rc != 0
