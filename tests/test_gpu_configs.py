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


def test_eight_512cube_fields_full_size_bit_exact(codec):
    """The headline workload at full size: 8 ranks x 512^3 FP32 GRF (k^-3,
    shared seed) + N(0, 0.05 std) per rank, relative eb 1e-3, L = 36, T = 16.
    All 64 segment streams (batched compress), the 8 fused sums of the
    bench's one-launch batch (two-pass form) and the decompressed field
    equal the oracle's bytes and bits; the per-call fused sums and the
    decompressed field's error bound agree."""
    from zccl_amd import batch, synth
    from zccl_amd.allreduce import StagedAllreduce

    P, side, L, T = 8, 512, 36, 16
    count = side ** 3
    say("8 x 512^3 fields")
    base = synth.grf3d(side, seed=20261202)
    xs = [synth.perturb(base, 0.05, seed=1000 + r) for r in range(P)]
    del base
    eb = synth.rel_to_abs_eb(xs, 1e-3)
    ar = StagedAllreduce(P=P, count=count, eb=eb, L=L, T=T)
    ar.compress_all(xs)
    outs = [codec.empty_stream(ar.cnts[j], L, T) for j in range(P)]
    plan = batch.homsum_plan([[ar.streams[r][j] for r in range(P)] for j in range(P)], outs)
    assert plan.staged and plan.tiles >= codec.STAGED_MIN_TILES
    ws = codec.Workspace(plan.tiles)
    plan.launch(ws)
    torch.cuda.synchronize()
    codec.check_errors(ws)
    per_call = ar.reduce_scatter_fused()
    y = ar.decompress_all(outs)
    torch.cuda.synchronize()
    codec.check_errors()
    for j in range(P):
        assert torch.equal(per_call[j].data[: per_call[j].size], outs[j].data[: outs[j].size]), j
    say("8 x 512^3 oracle", eb)
    yh = y.cpu().numpy()
    for j in range(P):
        a, n = ar.displs[j], ar.cnts[j]
        want = []
        for r in range(P):
            c = O.compress(xs[r][a: a + n].cpu().numpy(), eb, L, T)
            assert np.array_equal(ar.streams[r][j].to_host(), c), (r, j)
            want.append(c)
        hw = O.homsum(want, n, L, T)
        assert np.array_equal(outs[j].to_host(), hw), j
        assert np.array_equal(bits(yh[a: a + n]), bits(O.decompress(hw, n, eb, L, T))), j
        say("segment", j, "ok", hw.size)
    truth = torch.zeros(count, dtype=torch.float64, device="cuda")
    absum = torch.zeros_like(truth)
    for x in xs:
        truth += x.double()
        absum += x.double().abs()
    assert bool(((y.double() - truth).abs() <= P * eb + 2.0 ** -21 * absum).all())


def test_uneven_4096_buffers_graph_vs_oracle(codec):
    """The batched form: 16 simulated ranks x 32 buffers of log-uniform sizes
    (4 KB..8 MB of gradient-like values: N(0, 1e-2), 1 % outliers at 100x,
    distinct per rank), relative eb 1e-3 per buffer.  One compress launch of
    all 512 buffers and one fused-sum launch of the 32 16-operand sums (a
    mixed batch: single-pass form), captured in a HIP graph and replayed;
    then a uniform batch of 2 x 16 operands of 2 400 017 values (two-pass
    form) in a graph.  Every replay equals the eager result and the oracle."""
    from zccl_amd import batch, workloads

    L, T, ranks, nslot = 36, 16, 16, 32
    sizes = [max(n, T) for n in workloads.log_uniform_sizes(nslot, 4096, 8 << 20, 77)]
    say("uneven batch sizes", min(sizes), max(sizes), sum(sizes) * ranks * 4 / 1e9, "GB")
    flat, views = workloads.gradient_buffers([n for _ in range(ranks) for n in sizes], 4097)
    xs = [[views[r * nslot + i] for i in range(nslot)] for r in range(ranks)]
    ebs = workloads.per_buffer_eb([torch.cat([xs[r][i] for r in range(ranks)]) for i in range(nslot)], 1e-3)

    def run_graph(cp, hp, ws, outs, sums):
        cp.launch(ws)
        hp.launch(ws)
        torch.cuda.synchronize()
        codec.check_errors(ws)
        eager = [s.data[: s.size].clone() for s in sums]
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g):
            cp.launch(ws)
            hp.launch(ws)
        for rep in range(2):
            for o in outs + sums:
                o.data.zero_()
            g.replay()
            torch.cuda.synchronize()
            for o in outs + sums:
                o.invalidate()
            codec.check_errors(ws)
            for e, s in zip(eager, sums):
                assert torch.equal(s.data[: s.size], e), rep
            _ = flat[:1024].sum().item()  # eager work between replays

    outs = [[codec.empty_stream(n, L, T) for n in sizes] for _ in range(ranks)]
    sums = [codec.empty_stream(n, L, T) for n in sizes]
    cp = batch.compress_plan([x for r in range(ranks) for x in xs[r]], ebs * ranks,
                             [o for r in range(ranks) for o in outs[r]])
    hp = batch.homsum_plan([[outs[r][i] for r in range(ranks)] for i in range(nslot)], sums)
    assert not hp.staged
    ws = codec.Workspace(max(cp.tiles, hp.tiles))
    run_graph(cp, hp, ws, [o for r in range(ranks) for o in outs[r]], sums)
    say("uneven batch oracle (mixed)")
    for i, n in enumerate(sizes):
        want = []
        for r in range(ranks):
            c = O.compress(xs[r][i].cpu().numpy(), ebs[i], L, T)
            assert np.array_equal(outs[r][i].to_host(), c), (r, i)
            want.append(c)
        assert np.array_equal(sums[i].to_host(), O.homsum(want, n, L, T)), i

    # uniform jobs above the two-pass threshold
    n2, njob = 2_400_017, 2
    say("uneven batch, uniform two-pass batch")
    flat2, v2 = workloads.gradient_buffers([n2] * (ranks * njob), 4098)
    eb2 = workloads.per_buffer_eb([flat2], 1e-3)[0]
    outs2 = [codec.empty_stream(n2, L, T) for _ in range(ranks * njob)]
    sums2 = [codec.empty_stream(n2, L, T) for _ in range(njob)]
    cp2 = batch.compress_plan(v2, [eb2] * len(v2), outs2)
    hp2 = batch.homsum_plan([[outs2[r * njob + j] for r in range(ranks)] for j in range(njob)], sums2)
    assert hp2.staged and hp2.tiles >= codec.STAGED_MIN_TILES
    ws2 = codec.Workspace(max(cp2.tiles, hp2.tiles))
    run_graph(cp2, hp2, ws2, outs2, sums2)
    for j in range(njob):
        want = [O.compress(v2[r * njob + j].cpu().numpy(), eb2, L, T) for r in range(ranks)]
        for r in range(ranks):
            assert np.array_equal(outs2[r * njob + j].to_host(), want[r]), (r, j)
        assert np.array_equal(sums2[j].to_host(), O.homsum(want, n2, L, T)), j
