"""Both launch layouts of compress and decompress for 36-value blocks (0:
one lane per block, the lane kernels; 1: workgroup tiles, the slot kernels)
give the oracle's bytes and values: single streams over edge shapes and
every width, and uniform and mixed batch plans."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

from oracle import oracle as O  # noqa: E402
from tests.golden import gen  # noqa: E402
from tests.test_gpu_codec import check_full_index  # noqa: E402

pytestmark = pytest.mark.gpu

L = 36
# kind, n, eb, T: chunks of one tile and of many, a remainder segment
# (n mod T), a segment's short last block, n = T, every width to 32
CASES = [("grf", 1_000_003, 1e-3, 16), ("walk", 100_003, 1e-2, 7), ("ladder", 300_000, 1.0, 3),
         ("wide", 200_001, 1.0, 1), ("zeros", 50_000, 1e-2, 4), ("special", 20_000, 1e-3, 5),
         ("outliers", 333_333, 1e-3, 16), ("sines", 37, 1e-4, 1), ("grf", 16, 1e-3, 16),
         ("slabs", 2_000_000, 1e-2, 8)]
IDS = [f"{k}-{n}-T{t}" for k, n, _, t in CASES]


def say(*a):
    print("[gpu]", *a, flush=True)


def bits(a):
    return np.ascontiguousarray(a, dtype=np.float32).view(np.uint32)


@pytest.fixture(scope="module")
def codec():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from zccl_amd import codec as c

    torch.cuda.set_device(0)
    return c


@pytest.fixture
def layouts():
    from zccl_amd import autotune

    yield autotune
    for op in autotune.OPS:
        autotune.reset(op)


@pytest.mark.parametrize("layout", [0, 1])
@pytest.mark.parametrize("kind,n,eb,T", CASES, ids=IDS)
def test_compress_layouts_bit_exact(codec, layouts, layout, kind, n, eb, T):
    layouts.set_knobs("compress", {"layout": layout})
    say("compress layout", layout, kind, n, T)
    x = gen.make(kind, n, 21)
    ref = O.compress(x, eb, L, T)
    s = codec.compress(torch.from_numpy(x).cuda(), eb, L, T)
    torch.cuda.synchronize()
    codec.check_errors()
    assert s.size == ref.size
    assert np.array_equal(s.to_host(), ref)
    check_full_index(s, ref)


@pytest.mark.parametrize("layout", [0, 1])
@pytest.mark.parametrize("kind,n,eb,T", CASES, ids=IDS)
def test_decompress_layouts_bit_exact(codec, layouts, layout, kind, n, eb, T):
    layouts.set_knobs("decompress", {"layout": layout})
    say("decompress layout", layout, kind, n, T)
    x = gen.make(kind, n, 22)
    ref = O.compress(x, eb, L, T)
    s = codec.upload(ref, n, L, T)
    y = codec.decompress(s, eb)
    torch.cuda.synchronize()
    codec.check_errors()
    assert np.array_equal(bits(y.cpu().numpy()), bits(O.decompress(ref, n, eb, L, T)))


@pytest.mark.parametrize("layout", [0, 1])
@pytest.mark.parametrize("mixed", [False, True], ids=["uniform", "mixed"])
def test_batch_layouts_bit_exact(codec, layouts, layout, mixed):
    """Batched compress and decompress plans at each layout: every job's
    bytes and values equal the oracle's (a mixed plan varies sizes and error
    bounds)."""
    from zccl_amd import batch

    layouts.set_knobs("compress", {"layout": layout})
    layouts.set_knobs("decompress", {"layout": layout})
    T = 16
    ns = [300_001, 40_000, 1_000_000, 17] if mixed else [250_000] * 5
    ebs = [1e-3, 1e-2, 1e-4, 1e-3] if mixed else [1e-3] * 5
    say("batch layout", layout, "mixed" if mixed else "uniform", ns)
    xs = [gen.make("grf", n, 30 + i) for i, n in enumerate(ns)]
    outs = [codec.empty_stream(n, L, T) for n in ns]
    plan = batch.compress_plan([torch.from_numpy(x).cuda() for x in xs], ebs, outs)
    ws = codec.Workspace(plan.tiles + 64)
    plan.launch(ws)
    torch.cuda.synchronize()
    codec.check_errors(ws)
    refs = [O.compress(x, eb, L, T) for x, eb in zip(xs, ebs)]
    for s, r in zip(outs, refs):
        assert np.array_equal(s.to_host(), r)
    ys = [torch.empty(n, dtype=torch.float32, device="cuda") for n in ns]
    dplan = batch.decompress_plan(outs, ebs, ys)
    ws2 = codec.Workspace(dplan.tiles + 64)
    for _ in range(2):  # twice: the tile sums' workspace area is reused
        dplan.launch(ws2)
        torch.cuda.synchronize()
        codec.check_errors(ws2)
        for y, r, n, eb in zip(ys, refs, ns, ebs):
            assert np.array_equal(bits(y.cpu().numpy()), bits(O.decompress(r, n, eb, L, T)))
