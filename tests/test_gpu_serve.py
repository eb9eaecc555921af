"""The serving loop on the GPU: every result of a mixed, overlapping,
batched run equals the result of the same request run alone (bit-exact), and
a sample equals the CPU oracle."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

from oracle import oracle as O  # noqa: E402

pytestmark = pytest.mark.gpu


def say(*a):
    print("[gpu]", *a, flush=True)


@pytest.fixture(scope="module")
def mods():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    torch.cuda.set_device(0)
    from zccl_amd import codec, serve, synth

    return codec, serve, synth


def build_requests(codec, serve, synth, nreq, seed, L, T):
    rng = np.random.default_rng(seed)
    field = synth.grf3d(128, seed=seed).reshape(-1)  # 2 097 152 values
    arrivals = serve.poisson_arrivals(nreq, 4000.0, seed)
    reqs = []
    for i in range(nreq):
        kind = str(rng.choice(["compress", "decompress", "homsum"], p=[0.6, 0.2, 0.2]))
        n = int(np.exp(rng.uniform(np.log(64), np.log(400_000))))
        n = max(n, T)
        rel = float(rng.choice([1e-2, 1e-3, 1e-4]))
        def slice_(j):
            off = int(rng.integers(0, field.numel() - n))
            return (field[off: off + n] + 0.01 * j).contiguous()
        x = slice_(0)
        eb = rel * float(x.max() - x.min()) or rel
        if kind == "compress":
            reqs.append(serve.Request(i, kind, n, eb, float(arrivals[i]), x=x))
        elif kind == "decompress":
            reqs.append(serve.Request(i, kind, n, eb, float(arrivals[i]), src=[codec.compress(x, eb, L, T)]))
        else:
            k = int(rng.integers(2, 33))
            ops = [codec.compress(slice_(j), eb, L, T) for j in range(k)]
            reqs.append(serve.Request(i, kind, n, eb, float(arrivals[i]), src=ops))
    torch.cuda.synchronize()
    return reqs


def test_results_same_alone_and_in_mix(mods):
    codec, serve, synth = mods
    L, T = 36, 16
    reqs = build_requests(codec, serve, synth, 64, 2027, L, T)
    say("requests", len(reqs), {k: sum(r.kind == k for r in reqs) for k in serve.KINDS})
    ex = serve.Executor(L, T, slots=3, max_tiles=1 << 18)
    srv = serve.Server(ex, hbm_budget=2e9, slots=3, max_jobs=8)
    st = srv.run(reqs)
    torch.cuda.synchronize()
    say("stats", st.as_dict())
    assert st.max_inflight >= 2, "requests overlapped on the device"
    assert st.steps < len(reqs), "requests were batched"
    for r in reqs:
        if r.kind == "compress":
            alone = codec.compress(r.x, r.eb, L, T)
            assert torch.equal(r.out.data[: r.out.size], alone.data[: alone.size]), r.rid
        elif r.kind == "decompress":
            alone = codec.decompress(r.src[0], r.eb)
            assert torch.equal(r.out.view(torch.int32), alone.view(torch.int32)), r.rid
        else:
            alone = codec.homsum(r.src)
            assert torch.equal(r.out.data[: r.out.size], alone.data[: alone.size]), r.rid
    codec.check_errors()
    # a sample against the CPU oracle
    for r in [q for q in reqs if q.kind == "compress"][:4]:
        want = O.compress(r.x.cpu().numpy(), r.eb, L, T)
        assert np.array_equal(r.out.to_host(), want), r.rid
    for r in [q for q in reqs if q.kind == "homsum"][:2]:
        want = O.homsum([s.to_host() for s in r.src], r.n, L, T)
        assert np.array_equal(r.out.to_host(), want), r.rid


def _check_alone(codec, reqs, L, T):
    for r in reqs:
        if r.kind == "compress":
            alone = codec.compress(r.x, r.eb, L, T)
            assert torch.equal(r.out.data[: r.out.size], alone.data[: alone.size]), r.rid
        elif r.kind == "decompress":
            alone = codec.decompress(r.src[0], r.eb)
            assert torch.equal(r.out.view(torch.int32), alone.view(torch.int32)), r.rid
        else:
            alone = codec.homsum(r.src)
            assert torch.equal(r.out.data[: r.out.size], alone.data[: alone.size]), r.rid
        del alone


def test_full_size_range_and_budget(mods):
    """The serving mix at its full request sizes: 64 KiB to 512 MiB per
    request (plus three requests of exactly 512 MiB), groups of at most 8
    jobs on 4 slots, the 200 GB budget with everything arriving at once, then
    Poisson arrivals, then a budget that holds only the two largest results
    at a time, first in arrival order, then with bypass of a blocked head.
    Every result equals the same request run alone, and the reserved bytes
    never pass the budget."""
    codec, serve, synth = mods
    from zccl_amd import workloads

    L, T = 36, 16
    field = synth.grf3d(512, seed=11)  # 134 217 728 values = 512 MiB
    reqs = workloads.serving_requests(20, 11, L, T, field)
    reqs += workloads.serving_requests(3, 12, L, T, field, lo=512 << 20, hi=512 << 20)
    for i, r in enumerate(reqs):
        r.rid = i
    sizes = sorted(r.n * 4 for r in reqs)
    say("requests", len(reqs), {k: sum(r.kind == k for r in reqs) for k in serve.KINDS},
        "bytes min", sizes[0], "max", sizes[-1])
    assert sizes[0] >= 64 << 10 and sizes[-1] == 512 << 20
    resident = sum(sum(s.data.numel() + s.index.numel() for s in (r.src or [])) for r in reqs)
    resident += field.numel() * 4
    ex = serve.Executor(L, T, slots=4)

    def run(budget, arrivals, tag, bypass=0):
        for r, a in zip(reqs, arrivals):
            r.arrival, r.out = float(a), None
        srv = serve.Server(ex, hbm_budget=budget, slots=4, max_jobs=8, resident_bytes=resident, bypass=bypass)
        st = srv.run(reqs)
        torch.cuda.synchronize()
        say(tag, st.as_dict())
        assert st.peak_reserved_bytes <= budget - resident
        _check_alone(codec, reqs, L, T)
        codec.check_errors()
        return st

    st = run(200e9, [0.0] * len(reqs), "closed loop, 200 GB")
    assert 3 <= st.steps < len(reqs)  # groups of at most 8 jobs
    rate = len(reqs) / st.makespan_s
    run(200e9, serve.poisson_arrivals(len(reqs), 0.8 * rate, 13), "poisson 0.8")
    need = sorted(ex.reserve(r) for r in reqs)
    tight = resident + need[-1] + need[-2]
    st = run(tight, [0.0] * len(reqs), f"budget {tight / 1e9:.2f} GB")
    assert st.peak_reserved_bytes <= need[-1] + need[-2]
    assert st.steps >= 3
    # the same budget with smaller requests admitted past a blocked head
    st = run(tight, [0.0] * len(reqs), f"budget {tight / 1e9:.2f} GB, bypass 4", bypass=4)
    assert st.peak_reserved_bytes <= need[-1] + need[-2]
