"""The staged allreduce run by the library's C host code
(zccl_amd/csrc/zccl_allreduce.hip, zccl_allreduce_*): its reduced segment
streams and the decompressed result equal the oracle's (compress every rank's
segments, k-operand sum per segment, decompress) and the Python
orchestration's; re-runs with new send buffers rebuild its plans, and a HIP
graph of run() replays to the same bytes."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

from oracle import oracle as O  # noqa: E402
from tests.golden import gen  # noqa: E402

pytestmark = pytest.mark.gpu


def say(*a):
    print("[gpu]", *a, flush=True)


@pytest.fixture(scope="module")
def mods():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    torch.cuda.set_device(0)
    from zccl_amd import allreduce, codec

    return allreduce, codec


def bits(a):
    return np.ascontiguousarray(a).view(np.uint32)


def ranks_data(P, count, seed):
    base = gen.make("grf", count, seed)
    return [gen.perturb(base, seed + 100 + r) for r in range(P)]


def check_against_oracle(ar, hs, y, eb, L, T, segs=None):
    P = ar.P
    yh = y.cpu().numpy()
    for j in (range(P) if segs is None else segs):
        a, n = ar.displs[j], ar.cnts[j]
        want = O.homsum([O.compress(h[a: a + n], eb, L, T) for h in hs], n, L, T)
        assert np.array_equal(ar.segment_stream(j).to_host(), want), j
        assert np.array_equal(bits(yh[a: a + n]), bits(O.decompress(want, n, eb, L, T))), j


@pytest.mark.parametrize("P,count,T", [(4, 400_003, 16), (3, 100_000, 8), (8, 8 * 2_400_017, 16)])
def test_c_allreduce_matches_oracle_and_python(mods, P, count, T):
    allreduce, codec = mods
    L, eb = 36, 1e-3
    say("C allreduce", P, count, T)
    hs = ranks_data(P, count, 31 + P)
    xs = [torch.from_numpy(h).cuda() for h in hs]
    ar = allreduce.DeviceAllreduce(P, count, eb, L, T)
    y = ar.run(xs)
    torch.cuda.synchronize()
    assert ar.errors() == 0
    # the Python orchestration gives the same reduced bytes and values
    py = allreduce.StagedAllreduce(P=P, count=count, eb=eb, L=L, T=T)
    py.compress_all(xs)
    sums = py.reduce_scatter_fused()
    y2 = py.decompress_all(sums)
    torch.cuda.synchronize()
    for j in range(P):
        s = ar.segment_stream(j)
        assert torch.equal(s.data[: s.size], sums[j].data[: sums[j].size]), j
    assert torch.equal(y, y2)
    check_against_oracle(ar, hs, y, eb, L, T, segs=None if count < 10_000_000 else [0, P - 1])


def test_c_allreduce_new_buffers_and_graph_replay(mods):
    allreduce, codec = mods
    P, count, L, T, eb = 4, 300_001, 36, 8, 2e-3
    ar = allreduce.DeviceAllreduce(P, count, eb, L, T)
    hs1 = ranks_data(P, count, 5)
    hs2 = ranks_data(P, count, 6)
    x1 = [torch.from_numpy(h).cuda() for h in hs1]
    x2 = [torch.from_numpy(h).cuda() for h in hs2]
    y = torch.empty(count, dtype=torch.float32, device="cuda")
    ar.run(x1, out=y)
    torch.cuda.synchronize()
    check_against_oracle(ar, hs1, y, eb, L, T)
    ar.run(x2, out=y)  # other send buffers: plans rebuilt
    torch.cuda.synchronize()
    check_against_oracle(ar, hs2, y, eb, L, T)
    want = y.clone()
    # captured run() on the same buffers: every replay equals the eager result
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(s):
        ar.run(x2, out=y)
    torch.cuda.current_stream().wait_stream(s)
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        ar.run(x2, out=y)
    for rep in range(3):
        y.zero_()
        g.replay()
        torch.cuda.synchronize()
        assert torch.equal(y, want), rep
        _ = x1[0].sum().item()
    assert ar.errors() == 0
