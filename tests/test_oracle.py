"""Pins the CPU oracle (oracle/zccl_oracle.c) before it is trusted as the checker.

Checks, in order of strength:
  * committed golden vectors written by the compiled reference
    (tests/golden/fixtures_*.{npz,json}, tests/golden/make_fixtures.py);
  * the compiled reference itself (oracle/_ref, when built) on random cases;
  * the exhaustive glibc log2f probe (tests/golden/width_thresholds.json) and
    libm's log2f called directly near every threshold;
  * closed forms: quantize = trunc((double)x * (double)(1.0f/eb)) with x86
    out-of-range -> INT_MIN, dequantize = fl32(fl32(q) * eb), and
    decompress(compress(x)) = dequantize(quantize(x)) for every T and L;
  * invariants: canonical homomorphic sum (encode(sum q) == homsum(encode q)),
    a bit-at-a-time packer, round trips, parse offsets vs block sizes.
"""
import ctypes
import hashlib
import json
import os

import numpy as np
import pytest

from oracle import oracle as O
from tests.golden import gen

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _small():
    with open(os.path.join(GOLD, "fixtures_small.json")) as f:
        meta = json.load(f)
    arrs = np.load(os.path.join(GOLD, "fixtures_small.npz"), allow_pickle=False)
    return meta, arrs


SMALL_META, SMALL = _small()
with open(os.path.join(GOLD, "fixtures_large.json")) as _f:
    LARGE_META = json.load(_f)


def bits(a):
    return np.ascontiguousarray(a, dtype=np.float32).view(np.uint32)


# ------------------------------------------------------------------ width

def test_width_table_matches_exhaustive_probe():
    with open(os.path.join(GOLD, "width_thresholds.json")) as f:
        probe = json.load(f)
    assert probe["monotone_and_w_or_w_plus_1"]
    bump = {int(k): int(v) for k, v in probe["bump_c"].items()}
    for w in range(1, 33):
        top = (1 << w) - 1
        if w in bump:
            t = (1 << w) - bump[w]
            assert O.width([t - 1])[0] == w
            assert O.width([t])[0] == w + 1
            assert O.width([min(top, (1 << 32) - 1)])[0] == w + 1
        else:
            assert O.width([top])[0] == w
        assert O.width([1 << (w - 1)])[0] == w
    assert O.width([0])[0] == 0


def test_width_vs_libm_log2f_near_thresholds():
    libm = ctypes.CDLL("libm.so.6")
    libm.log2f.argtypes = [ctypes.c_float]
    libm.log2f.restype = ctypes.c_float
    ms = []
    for w in range(1, 33):
        hi = 1 << w
        ms.extend(range(max(1, hi - 3000), min(hi + 8, 1 << 32)))
    rng = np.random.default_rng(3)
    ms.extend(rng.integers(1, 1 << 32, 20000, dtype=np.uint64).tolist())
    ms = np.unique(np.array(ms, dtype=np.uint64)).astype(np.uint32)
    expect = np.array([int(libm.log2f(float(np.float32(m)))) + 1 for m in ms.tolist()])
    assert np.array_equal(O.width(ms).astype(np.int64), expect)


# ------------------------------------------------------------- fixtures

@pytest.mark.parametrize("case", SMALL_META, ids=[m["name"] for m in SMALL_META])
def test_small_fixture(case):
    name, n, eb, L, T = case["name"], case["n"], case["eb"], case["L"], case["T"]
    x, y = SMALL[f"{name}.x"], SMALL[f"{name}.y"]
    cx = O.compress(x, eb, L, T)
    assert np.array_equal(cx, SMALL[f"{name}.cx"])
    assert np.array_equal(bits(O.decompress(cx, n, eb, L, T)), bits(SMALL[f"{name}.dx"]))
    if f"{name}.cy" in SMALL:
        assert np.array_equal(O.compress(y, eb, L, T), SMALL[f"{name}.cy"])
    hxy = O.homsum([cx, O.compress(y, eb, L, T)], n, L, T)
    assert np.array_equal(hxy, SMALL[f"{name}.hxy"])


def _sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


@pytest.mark.parametrize("case", LARGE_META, ids=[m["name"] for m in LARGE_META])
def test_large_fixture_digests(case):
    n, eb, L, T = case["n"], case["eb"], case["L"], case["T"]
    x = gen.make(case["kind"], n, case["seed"])
    assert _sha(x) == case["sha_x"], "generator drifted; fixture inputs differ"
    y = gen.perturb(x, case["seed"] + 1)
    cx = O.compress(x, eb, L, T)
    cy = O.compress(y, eb, L, T)
    assert cx.size == case["cx_size"] and _sha(cx) == case["sha_cx"]
    assert _sha(cy) == case["sha_cy"]
    assert _sha(O.decompress(cx, n, eb, L, T)) == case["sha_dx"]
    h = O.homsum([cx, cy], n, L, T)
    assert h.size == case["hxy_size"] and _sha(h) == case["sha_hxy"]


# ------------------------------------------------- compiled reference

needs_ref = pytest.mark.skipif(not O.ref_available(), reason="oracle/_ref not built")


@needs_ref
def test_random_cases_against_compiled_reference():
    rng = np.random.default_rng(11)
    done = 0
    for trial in range(160):
        T = int(rng.choice([1, 2, 3, 5, 8, 16]))
        L = int(rng.choice([1, 7, 8, 32, 36, 128]))
        n = int(rng.integers(T, 4000))
        kind = gen.KINDS[trial % len(gen.KINDS)]
        eb = 1.0 if kind in ("wide", "ladder") else (3e-39 if kind == "denormal" else
                                           float(rng.choice([1e-1, 1e-2, 1e-3, 1e-4])))
        x = gen.make(kind, n, 100 + trial)
        y = gen.perturb(x, 200 + trial)
        try:
            rx = O.ref_compress(x, eb, L, T)
            ry = O.ref_compress(y, eb, L, T)
            rh, _ = O.ref_add2(rx, ry, n, L, T)
        except O.OracleError:
            continue  # the reference would overrun its own buffer here
        cx, cy = O.compress(x, eb, L, T), O.compress(y, eb, L, T)
        assert np.array_equal(cx, rx), (kind, n, L, T, eb)
        assert np.array_equal(cy, ry)
        assert np.array_equal(bits(O.decompress(cx, n, eb, L, T)),
                              bits(O.ref_decompress(rx, n, eb, L, T)))
        assert np.array_equal(O.homsum([cx, cy], n, L, T), rh)
        done += 1
    assert done > 120


@needs_ref
def test_reference_ring_chain_equals_fused_sum():
    """hZCCL's ring applies P-1 pairwise adds per segment (ZCCL/src/ZCCL_ring_ho.c:258-290);
    the fused k-ary sum must give the same bytes."""
    n, eb, L, T, P = 5000, 1e-2, 36, 8, 5
    xs = [gen.make("walk", n, 40)] + [gen.perturb(gen.make("walk", n, 40), 41 + r) for r in range(P - 1)]
    streams = [O.compress(x, eb, L, T) for x in xs]
    acc = streams[0]
    for r in range(1, P):
        acc, _ = O.ref_add2(streams[r], acc, n, L, T)
    assert np.array_equal(acc, O.homsum(streams, n, L, T))


@needs_ref
def test_doc_chain_equals_compiled_reference_ring():
    """The DOC ring (ZCCL/src/ZCCL_ring.c:979-1108) rebuilt from the compiled
    reference's own compress / decompress calls and FP32 adds equals the
    oracle's zo_doc_chain bit for bit, before and after the allgather."""
    n, eb, L, T, P = 20_003, 1e-2, 36, 4, 5
    xs = [gen.make("walk", n, 70)] + [gen.perturb(gen.make("walk", n, 70), 71 + r) for r in range(P - 1)]
    acc = xs[0].copy()
    for m in range(1, P):
        acc = O.ref_decompress(O.ref_compress(acc, eb, L, T), n, eb, L, T) + xs[m]
    fin = O.ref_decompress(O.ref_compress(acc, eb, L, T), n, eb, L, T)
    oacc, ofin = O.doc_chain(xs, eb, L, T)
    assert np.array_equal(bits(oacc), bits(acc))
    assert np.array_equal(bits(ofin), bits(fin))


def test_doc_chain_closed_forms():
    """One rank: the partial sum is the input itself and the allgathered value
    its round trip; any P: the error against the exact sum stays within one
    round trip's bound per compression plus FP32 rounding per add."""
    n, eb, L, T = 30_001, 1e-3, 36, 8
    x = gen.make("grf", n, 80)
    acc, fin = O.doc_chain([x], eb, L, T)
    assert np.array_equal(bits(acc), bits(x))
    assert np.array_equal(bits(fin), bits(O.decompress(O.compress(x, eb, L, T), n, eb, L, T)))
    P = 6
    xs = [gen.perturb(x, 81 + r) for r in range(P)]
    acc, fin = O.doc_chain(xs, eb, L, T)
    exact = np.sum(np.stack([v.astype(np.float64) for v in xs]), axis=0)
    mag = np.sum(np.stack([np.abs(v.astype(np.float64)) for v in xs]), axis=0)
    assert np.all(np.abs(acc - exact) <= (P - 1) * eb + 2.0 ** -20 * P * mag)
    assert np.all(np.abs(fin - exact) <= P * eb + 2.0 ** -20 * (P + 1) * mag)
    # a mistake in the chain order or a skipped add shows as a larger error
    wrong = acc - xs[-1]
    assert np.max(np.abs(wrong - exact)) > 10 * eb


@needs_ref
def test_reference_add_accumulates_size():
    """ZCCL/src/ZCCLd_float.c:1027,1655 add to *final_cmpSize instead of assigning."""
    x = gen.make("walk", 3000, 5)
    a = O.compress(x, 1e-2, 36, 4)
    h, reported = O.ref_add2(a, a, 3000, 36, 4, initial_size=1234)
    assert reported == 1234 + h.size


@needs_ref
def test_single_thread_twins_fixture():
    x = SMALL["single_thread.x"]
    cx = O.compress(x, 1e-2, 36, 1)
    assert np.array_equal(cx, SMALL["single_thread.cx"])
    assert np.array_equal(bits(O.decompress(cx, x.size, 1e-2, 36, 1)), bits(SMALL["single_thread.dx"]))
    cy = O.compress(SMALL["single_thread.y"], 1e-2, 36, 1)
    assert np.array_equal(O.homsum([cx, cy], x.size, 36, 1), SMALL["single_thread.hxy"])


# ------------------------------------------------------- closed forms

def _np_quantize(x, eb):
    inv = np.float64(np.float32(1.0) / np.float32(eb))
    with np.errstate(invalid="ignore", over="ignore"):
        p = x.astype(np.float64) * inv
        ok = (p > -2147483649.0) & (p < 2147483648.0)
        q = np.full(x.shape, np.iinfo(np.int32).min, dtype=np.int64)
        q[ok] = np.trunc(p[ok]).astype(np.int64)
    return q.astype(np.int32)


@pytest.mark.parametrize("kind,eb", [("sines", 1e-4), ("special", 1e-3), ("denormal", 3e-39),
                                     ("outliers", 1e-3), ("wide", 1.0), ("grf", 1e-5)])
def test_quantize_dequantize_closed_forms(kind, eb):
    x = gen.make(kind, 20000, 9)
    q = O.quantize(x, eb)
    assert np.array_equal(q, _np_quantize(x, eb))
    d = O.dequantize(q, eb)
    assert np.array_equal(bits(d), bits(q.astype(np.float32) * np.float32(eb)))


@pytest.mark.parametrize("kind,eb", [("sines", 1e-4), ("special", 1e-3), ("denormal", 3e-39),
                                     ("zeros", 1e-3), ("wide", 1.0)])
def test_decompress_is_closed_form_for_any_T_L(kind, eb):
    n = 7777
    x = gen.make(kind, n, 21)
    expect = bits(O.dequantize(O.quantize(x, eb), eb))
    for T in (1, 4, 13):
        for L in (1, 5, 36, 200):
            s = O.compress(x, eb, L, T)
            assert np.array_equal(bits(O.decompress(s, n, eb, L, T)), expect), (T, L)


def test_denormal_outputs_preserved():
    x = gen.make("denormal", 4000, 3)
    d = O.decompress(O.compress(x, 3e-39, 36, 4), x.size, 3e-39, 36, 4)
    sub = (d != 0) & (np.abs(d) < np.finfo(np.float32).tiny)
    assert sub.sum() > 100


# -------------------------------------------------------- invariants

@pytest.mark.parametrize("k", [1, 2, 3, 5, 8])
def test_homsum_is_canonical_encoding_of_summed_q(k):
    rng = np.random.default_rng(k)
    n, L, T = 6001, 36, 8
    base = np.cumsum(rng.integers(-300, 300, n)).astype(np.int64)
    qs = []
    for r in range(k):
        q = base + rng.integers(-50, 50, n)
        q[rng.random(n) < 0.3] = 0
        if r == 1:
            q[:2000] = 2 ** 31 - 5   # drives int32 wrap in the sum
        qs.append(q.astype(np.int64).astype(np.int32))
    streams = [O.encode_q(q, L, T) for q in qs]
    total = np.zeros(n, dtype=np.int64)
    for q in qs:
        total += q
    total = (total & 0xFFFFFFFF).astype(np.uint32).view(np.int32)
    assert np.array_equal(O.homsum(streams, n, L, T), O.encode_q(total, L, T))
    assert np.array_equal(O.decode_q(O.homsum(streams, n, L, T), n, L, T), total)


def test_block_packer_matches_bitwise_writer_and_round_trips():
    rng = np.random.default_rng(7)
    for trial in range(3000):
        n = int(rng.integers(1, 140))
        w = int(rng.integers(0, 33))
        if w == 0:
            d = np.zeros(n, dtype=np.int64)
        else:
            hi = min(2 ** w, 2 ** 31)
            d = rng.integers(-hi + 1, hi, n, dtype=np.int64)
            if trial % 17 == 0:
                d[rng.integers(0, n)] = -(2 ** 31)
        d = d.astype(np.int32)
        enc = O.encode_block(d)
        assert enc == O.encode_block(d, bitwise=True)
        assert len(enc) == O.block_bytes(enc[0], n)
        assert np.array_equal(O.decode_block(enc, n), d)


def test_parse_offsets_follow_block_sizes():
    for n, L, T in [(10000, 36, 8), (10007, 7, 3), (17, 36, 16), (4097, 128, 1)]:
        x = gen.make("walk", n, n)
        s = O.compress(x, 1e-2, L, T)
        g = O.geometry(n, L, T)
        w, bo, so, tot = O.parse(s, n, L, T)
        assert tot == s.size
        assert np.array_equal(w, s[bo.astype(np.int64)])
        hdr = np.frombuffer(s[: 8 * T].tobytes(), dtype=np.uint64)
        assert np.array_equal(so[:T], hdr + 8 * T)
        # consecutive blocks inside one segment are adjacent
        for seg in range(g.S):
            first = seg * g.nbT
            cnt = g.nbT if seg < T else g.nbR
            length = g.tb if seg < T else g.R
            for j in range(cnt - 1):
                nb = min(L, length - 1 - j * L)
                assert bo[first + j + 1] - bo[first + j] == O.block_bytes(int(w[first + j]), nb)
            if cnt:
                assert bo[first] == so[seg] + 4


def test_width_inputs_cover_every_width():
    """The ladder and wide kinds, which the GPU parity tests use to reach
    every decoder branch, together give blocks of every width 0..32, and the
    oracle round-trips them exactly (integer inputs, eb = 1)."""
    seen = set()
    for kind, L, T in [("ladder", 36, 4), ("ladder", 20, 3), ("wide", 36, 4)]:
        n = 40_000
        x = gen.make(kind, n, 500)
        s = O.compress(x, 1.0, L, T)
        seen |= set(O.parse(s, n, L, T)[0].tolist())
        assert np.array_equal(O.decode_q(s, n, L, T), O.quantize(x, 1.0))
    assert seen == set(range(33))


def test_tolerance_policy_and_reference_eb_violations():
    """|x - x'| <= eb + 2^-21 |x| holds everywhere |x/eb| < 2^31; the plain
    |x - x'| <= eb does not (the reference itself violates it, 1/eb in FP32
    and fl32(q) rounding), so violations are counted, not asserted away."""
    x = gen.make("cfg1", 1 << 20, 20261202)
    eb = 1e-4
    d = O.decompress(O.compress(x, eb, 36, 8), x.size, eb, 36, 8)
    err = np.abs(d.astype(np.float64) - x.astype(np.float64))
    assert np.all(err <= eb + 2.0 ** -21 * np.abs(x))
    assert (err > eb).sum() > 0


def test_error_paths():
    x = gen.make("walk", 100, 1)
    with pytest.raises(O.OracleError):
        O.compress(x[:3], 1e-2, 36, 8)          # n < T
    s = O.compress(x, 1e-2, 36, 2)
    bad = s.copy()
    w, bo, so, tot = O.parse(s, 100, 36, 2)
    bad[int(bo[0])] = 40                          # width byte > 32
    with pytest.raises(O.OracleError):
        O.decompress(bad, 100, 1e-2, 36, 2)
    with pytest.raises(O.OracleError):
        O.decompress(s[:-3], 100, 1e-2, 36, 2)    # truncated
