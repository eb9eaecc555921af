"""Admission, budget and overlap behaviour of the serving loop, driven by a
simulated executor on a virtual clock (no GPU)."""
import pytest

from zccl_amd.serve import Request, Server, poisson_arrivals


class SimExecutor:
    """Work of a group takes bytes / rate seconds on a virtual clock; groups in
    different slots run concurrently."""

    def __init__(self, rate=1e9, launch_cost=1e-5):
        self.t = 0.0
        self.rate = rate
        self.launch_cost = launch_cost
        self.live = {}
        self.log = []

    def clock(self):
        return self.t

    def idle(self, dt):
        self.t += max(dt, 1e-6)

    def reserve(self, r):
        return 4 * r.n if r.kind == "decompress" else 5 * r.n

    def launch(self, slot, reqs):
        assert slot not in self.live, "slot reused while busy"
        self.t += self.launch_cost
        end = self.t + sum(r.input_bytes for r in reqs) / self.rate
        self.live[slot] = end
        self.log.append((slot, self.t, end, [r.rid for r in reqs]))
        return (slot, end)

    def done(self, h):
        return self.t >= h[1]

    def finish(self, h):
        del self.live[h[0]]


def mix(n, seed=1, rate=None):
    import numpy as np

    rng = np.random.default_rng(seed)
    arr = poisson_arrivals(n, rate, seed) if rate else [0.0] * n
    out = []
    for i in range(n):
        kind = rng.choice(["compress", "decompress", "homsum"], p=[0.6, 0.2, 0.2])
        size = int(np.exp(rng.uniform(np.log(16_384), np.log(4_000_000))))
        src = [object()] * int(rng.integers(2, 33)) if kind == "homsum" else [object()]
        out.append(Request(i, str(kind), size, 1e-3, float(arr[i]), src=src))
    return out


def run(reqs, **kw):
    ex = SimExecutor()
    srv = Server(ex, clock=ex.clock, idle=ex.idle, **kw)
    return srv.run(reqs), ex


def test_every_request_completes_in_order_of_admission():
    reqs = mix(200, rate=2000.0)
    st, ex = run(reqs, hbm_budget=1e12, slots=4, max_jobs=16)
    assert st.requests == 200
    for r in reqs:
        assert r.arrival <= r.t_admit <= r.t_done
    order = sorted(reqs, key=lambda r: (r.arrival, r.rid))
    admits = [r.t_admit for r in order]
    assert admits == sorted(admits), "admission is first come, first served"
    assert st.p50_ms <= st.p99_ms
    assert st.throughput_GBps > 0


def test_budget_is_never_exceeded_and_requests_wait_for_memory():
    reqs = mix(120, seed=3)
    budget = 3 * max(5 * r.n for r in reqs)
    st, ex = run(reqs, hbm_budget=budget, slots=8, max_jobs=64)
    assert st.peak_reserved_bytes <= budget
    # at any instant the groups in flight fit the budget
    for slot, s, e, ids in ex.log:
        live = [g for g in ex.log if g[1] <= s < g[2]]
        held = sum(5 * reqs[i].n if reqs[i].kind != "decompress" else 4 * reqs[i].n
                   for g in live for i in g[3])
        assert held <= budget
    assert any(r.t_admit > r.arrival for r in reqs), "a tight budget makes requests wait"


def test_slots_bound_concurrency_and_overlap_happens():
    reqs = mix(300, seed=5, rate=20000.0)
    st, ex = run(reqs, hbm_budget=1e12, slots=3, max_jobs=4)
    assert st.max_inflight <= 3
    overlapped = any(a[1] < b[2] and b[1] < a[2] for a in ex.log for b in ex.log if a is not b)
    assert overlapped, "requests arriving during a step run beside it"


def test_groups_respect_max_jobs_and_batch_small_requests():
    reqs = mix(256, seed=7)
    st, ex = run(reqs, hbm_budget=1e12, slots=2, max_jobs=32)
    assert all(len(g[3]) <= 32 for g in ex.log)
    assert st.steps < len(reqs), "closed-loop arrivals are batched"


def test_request_above_budget_is_refused():
    reqs = [Request(0, "compress", 1_000_000, 1e-3, 0.0, src=[None])]
    with pytest.raises(ValueError):
        run(reqs, hbm_budget=1000)


def test_resident_state_counts_against_budget():
    reqs = [Request(i, "compress", 1000, 1e-3, 0.0, src=[None]) for i in range(10)]
    st, ex = run(reqs, hbm_budget=20_000, resident_bytes=10_000, slots=4, max_jobs=10)
    assert st.peak_reserved_bytes <= 10_000


def _blocked_head_case():
    """A large request at the head behind one that holds most of the budget,
    followed by small ones that fit beside the first."""
    big = Request(0, "compress", 1_000_000, 1e-3, 0.0, src=[object()])
    head = Request(1, "compress", 1_000_000, 1e-3, 0.0, src=[object()])
    small = [Request(2 + i, "compress", 10_000, 1e-3, 0.0, src=[object()]) for i in range(12)]
    return [big, head] + small, 5 * 1_000_000 + 5 * 10_000 * 4


def test_fifo_without_bypass_holds_later_requests_behind_the_head():
    reqs, budget = _blocked_head_case()
    st, ex = run(reqs, hbm_budget=budget, slots=4, max_jobs=4)
    head = reqs[1]
    assert all(r.t_admit >= head.t_admit for r in reqs[2:])


def test_bypass_admits_fitting_requests_past_a_blocked_head_and_bounds_it():
    reqs, budget = _blocked_head_case()
    st, ex = run(reqs, hbm_budget=budget, slots=4, max_jobs=4, bypass=2)
    head = reqs[1]
    early = [r for r in reqs[2:] if r.t_admit < head.t_admit]
    assert early, "small requests went ahead of the blocked head"
    # at most `bypass` groups of at most max_jobs requests pass the head
    assert len(early) <= 2 * 4
    groups_before = {tuple(g) for s, t, e, g in ex.log if t < head.t_admit and 0 not in g}
    assert len(groups_before) <= 2
    assert st.peak_reserved_bytes <= budget
    for r in reqs:
        assert r.arrival <= r.t_admit <= r.t_done


def test_bypass_rejects_negative():
    with pytest.raises(ValueError):
        Server(SimExecutor(), hbm_budget=1e9, bypass=-1)
