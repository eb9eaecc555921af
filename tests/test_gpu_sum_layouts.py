"""The lane sum's three forms for 36-value blocks (layout 0 two-pass and
single-pass per wave tile, layout 2 single-pass per workgroup tile) and the
slot layout give the oracle's bytes: single sums over edge shapes and every
width, and uniform and mixed batch plans."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

from oracle import oracle as O  # noqa: E402
from tests.golden import gen  # noqa: E402
from tests.test_gpu_codec import check_full_index  # noqa: E402

pytestmark = pytest.mark.gpu

L = 36
FORMS = [{"layout": 0, "two_pass": 1}, {"layout": 0, "two_pass": 1 << 40}, {"layout": 2}, {"layout": 1}]
FORM_IDS = ["two-pass", "wave", "workgroup", "slot"]
CASES = [("grf", 1_000_003, 1e-3, 16, 8), ("walk", 100_003, 1e-2, 7, 3), ("ladder", 300_000, 1.0, 3, 2),
         ("wide", 200_001, 1.0, 1, 5), ("zeros", 50_000, 1e-2, 4, 2), ("outliers", 333_333, 1e-3, 16, 17),
         ("sines", 37, 1e-4, 1, 2), ("grf", 16, 1e-3, 16, 4), ("slabs", 2_000_000, 1e-2, 8, 4)]
IDS = [f"{c[0]}-{c[1]}-T{c[3]}-k{c[4]}" for c in CASES]


def say(*a):
    print("[gpu]", *a, flush=True)


@pytest.fixture(scope="module")
def codec():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from zccl_amd import codec as c

    torch.cuda.set_device(0)
    return c


@pytest.fixture
def knobs():
    from zccl_amd import autotune

    yield autotune
    autotune.reset("homsum")


@pytest.mark.parametrize("form", FORMS, ids=FORM_IDS)
@pytest.mark.parametrize("kind,n,eb,T,k", CASES, ids=IDS)
def test_sum_forms_bit_exact(codec, knobs, form, kind, n, eb, T, k):
    knobs.set_knobs("homsum", form)
    say("sum form", form, kind, n, T, k)
    x = gen.make(kind, n, 40)
    xs = [x] + [gen.perturb(x, 41 + r) for r in range(k - 1)]
    refs = [O.compress(v, eb, L, T) for v in xs]
    streams = [codec.upload(r, n, L, T) for r in refs]
    s = codec.homsum(streams)
    torch.cuda.synchronize()
    codec.check_errors()
    want = O.homsum(refs, n, L, T)
    assert s.size == want.size
    assert np.array_equal(s.to_host(), want)
    check_full_index(s, want)


@pytest.mark.parametrize("form", FORMS, ids=FORM_IDS)
@pytest.mark.parametrize("mixed", [False, True], ids=["uniform", "mixed"])
def test_sum_forms_batch_bit_exact(codec, knobs, form, mixed):
    """Batched sums at each form, launched twice (the workspace's lookback
    state carries over between launches)."""
    from zccl_amd import batch

    knobs.set_knobs("homsum", form)
    T = 16
    ns = [300_001, 40_000, 1_000_000, 17] if mixed else [250_000] * 4
    ks = [3, 8, 2, 5] if mixed else [6] * 4
    say("sum batch form", form, "mixed" if mixed else "uniform")
    jobs, wants = [], []
    for i, (n, k) in enumerate(zip(ns, ks)):
        x = gen.make("grf", n, 50 + i)
        refs = [O.compress(v, 1e-3, L, T) for v in [x] + [gen.perturb(x, 60 + r) for r in range(k - 1)]]
        jobs.append([codec.upload(r, n, L, T) for r in refs])
        wants.append(O.homsum(refs, n, L, T))
    outs = [codec.empty_stream(n, L, T) for n in ns]
    plan = batch.homsum_plan(jobs, outs)
    ws = codec.Workspace(plan.tiles + 64)
    for _ in range(2):
        plan.launch(ws)
        torch.cuda.synchronize()
        codec.check_errors(ws)
        for o, w in zip(outs, wants):
            assert np.array_equal(o.to_host(), w)
