"""Parity at the workload shapes: the single-rank 2^24 sinusoid case
(bit-exact against the oracle at full size), cosmology-like mixed fields
(bit-exact at 128^3, properties at 512^3 x 4 ranks).  Every byte comes from
the HIP library; the oracle is the checker."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

from oracle import oracle as O  # noqa: E402
from tests.golden import gen  # noqa: E402

pytestmark = pytest.mark.gpu


def say(*a):
    print("[gpu]", *a, flush=True)


@pytest.fixture(scope="module")
def codec():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    torch.cuda.set_device(0)
    from zccl_amd import codec as c

    return c


def bits(a):
    return np.ascontiguousarray(a).view(np.uint32)


@pytest.mark.parametrize("L", [36, 32, 128])
def test_sinusoids_2p24_bit_exact(codec, L):
    """2^24 values: 16 random-phase sinusoids (periods 64..65536) + N(0,1e-3),
    absolute eb 1e-4: stream bytes, decompressed bits and the count of
    elements over eb equal the oracle's."""
    n, eb, T = 1 << 24, 1e-4, 16
    say("sines", n, L)
    x = gen.make("cfg1", n, 20261202)
    s = codec.compress(torch.from_numpy(x).cuda(), eb, L, T)
    want = O.compress(x, eb, L, T)
    got = s.to_host()
    assert got.size == want.size and np.array_equal(got, want)
    y = codec.decompress(s, eb).cpu().numpy()
    yo = O.decompress(want, n, eb, L, T)
    assert np.array_equal(bits(y), bits(yo))
    err = np.abs(x.astype(np.float64) - y.astype(np.float64))
    assert int((err > eb).sum()) == int((np.abs(x.astype(np.float64) - yo) > eb).sum())
    assert bool((err <= eb + 2.0 ** -21 * np.abs(x)).all())


def fields(side, seed, device="cuda"):
    """Lognormal density exp(1.5 g), a velocity component 300 g and a field
    with 70 % exact zeros in contiguous slabs along the slowest axis."""
    from zccl_amd import synth

    g = synth.grf3d(side, seed=seed, device=device)
    dens = torch.exp(1.5 * g)
    vel = 300.0 * synth.grf3d(side, seed=seed + 1, device=device)
    slab = synth.grf3d(side, seed=seed + 2, device=device).reshape(side, side * side).clone()
    zero = torch.rand(side, generator=torch.Generator(device=device).manual_seed(seed), device=device) < 0.7
    slab[zero] = 0.0
    return {"density": dens, "velocity": vel, "slabs": slab.reshape(-1).contiguous()}


@pytest.mark.parametrize("eb", [1e-2, 1e-4])
def test_mixed_fields_128cube_bit_exact(codec, eb):
    """4 ranks x 128^3 of each field type: compress, the fused 4-operand sum
    and decompress of the sum equal the oracle byte for byte."""
    L, T, P = 36, 16, 4
    per_rank = [fields(128, 40 + 10 * r) for r in range(P)]
    for name in per_rank[0]:
        say("field", name, eb)
        xs = [f[name] for f in per_rank]
        hs = [x.cpu().numpy() for x in xs]
        ss = [codec.compress(x, eb, L, T) for x in xs]
        want = [O.compress(h, eb, L, T) for h in hs]
        for s, w in zip(ss, want):
            assert np.array_equal(s.to_host(), w), name
        n = hs[0].size
        h = codec.homsum(ss)
        hw = O.homsum(want, n, L, T)
        assert np.array_equal(h.to_host(), hw), name
        y = codec.decompress(h, eb).cpu().numpy()
        assert np.array_equal(bits(y), bits(O.decompress(hw, n, eb, L, T))), name
        widths = h.widths().cpu().numpy()
        say(name, "max width", int(widths.max()), "zero blocks", float((widths == 0).mean()))
    codec.check_errors()


@pytest.mark.parametrize("eb", [1e-2, 1e-4])
def test_mixed_fields_512cube_properties(codec, eb):
    """4 ranks x 512^3 (full size): the fused sum equals the ring chain of
    pairwise adds byte for byte, and every decompressed value is within the
    stated bound (single field: eb + 2^-21|x|; sum of 4: 4 eb + 2^-21 sum|x|,
    against the FP64 sum)."""
    from zccl_amd.allreduce import StagedAllreduce

    L, T, P, side = 36, 16, 4, 512
    for name in ("density", "velocity", "slabs"):
        say("field 512^3", name, eb)
        xs = [fields(side, 90 + 10 * r)[name] for r in range(P)]
        ar = StagedAllreduce(P=P, count=side ** 3, eb=eb, L=L, T=T)
        ar.compress_all(xs)
        fused = ar.reduce_scatter_fused()
        ring = ar.reduce_scatter_ring(ar.ring_scratch())
        for a, b in zip(fused, ring):
            assert torch.equal(a.data[: a.size], b.data[: b.size]), name
        y = ar.decompress_all(fused)
        truth = torch.zeros(side ** 3, dtype=torch.float64, device="cuda")
        absum = torch.zeros_like(truth)
        for x in xs:
            truth += x.double()
            absum += x.double().abs()
        ok = (y.double() - truth).abs() <= P * eb + 2.0 ** -21 * absum
        assert bool(ok.all()), name
        del truth, absum
        s0 = ar.streams[0][0]
        x0 = xs[0][: ar.cnts[0]]
        y0 = codec.decompress(s0, eb)
        assert bool(((y0.double() - x0.double()).abs() <= eb + 2.0 ** -21 * x0.double().abs()).all()), name
        codec.check_errors()
        del xs, ar, fused, ring, y
        torch.cuda.empty_cache()
