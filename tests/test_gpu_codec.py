"""GPU parity of the HIP codec against the CPU oracle and the committed golden
vectors.  Bar: bit-exact streams, bit-exact decompressed FP32, bit-exact
homomorphic sums, identical indexes.  Every kernel is reached through the C ABI
(libzccl_amd.so) -- directly via ctypes for the drop-in entry points, via
zccl_amd.codec for the device API."""
import ctypes
import json
import os

import numpy as np
import pytest

torch = pytest.importorskip("torch")

from oracle import oracle as O  # noqa: E402
from tests.golden import gen  # noqa: E402

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
with open(os.path.join(GOLD, "fixtures_small.json")) as _f:
    SMALL_META = json.load(_f)
SMALL = np.load(os.path.join(GOLD, "fixtures_small.npz"), allow_pickle=False)


def say(*a):
    print("[gpu]", *a, flush=True)


@pytest.fixture(scope="module")
def codec():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from zccl_amd import codec as c

    torch.cuda.set_device(0)
    return c


def bits(a):
    return np.ascontiguousarray(a, dtype=np.float32).view(np.uint32)


def dev(x):
    return torch.from_numpy(np.ascontiguousarray(x)).cuda()


def check_index(codec, s, ref_stream):
    w, bo, so, tot = O.parse(ref_stream, s.n, s.L, s.T)
    assert s.size == tot == ref_stream.size
    assert np.array_equal(s.widths().cpu().numpy(), w)


# ------------------------------------------------------------ fixtures

@pytest.mark.parametrize("case", SMALL_META, ids=[m["name"] for m in SMALL_META])
def test_fixture_compress_decompress_add(codec, case):
    name, n, eb, L, T = case["name"], case["n"], case["eb"], case["L"], case["T"]
    say("fixture", name, n, L, T, eb)
    x, y = SMALL[f"{name}.x"], SMALL[f"{name}.y"]
    sx = codec.compress(dev(x), eb, L, T)
    sy = codec.compress(dev(y), eb, L, T)
    dx = codec.decompress(sx, eb)
    h = codec.homsum([sx, sy])
    torch.cuda.synchronize()
    codec.check_errors()
    assert np.array_equal(sx.to_host(), SMALL[f"{name}.cx"])
    check_index(codec, sx, SMALL[f"{name}.cx"])
    assert np.array_equal(bits(dx.cpu().numpy()), bits(SMALL[f"{name}.dx"]))
    assert np.array_equal(h.to_host(), SMALL[f"{name}.hxy"])
    check_index(codec, h, SMALL[f"{name}.hxy"])


# ------------------------------------------------------------ random vs oracle

CASES = [
    # kind, n, eb, L, T
    ("walk", 1 << 16, 1e-3, 36, 8),
    ("walk", 100_003, 1e-2, 36, 16),
    ("sines", 77_777, 1e-4, 32, 3),
    ("grf", 1 << 17, 1e-3, 128, 5),
    ("outliers", 50_001, 1e-3, 36, 7),
    ("special", 20_000, 1e-3, 36, 4),
    ("wide", 40_000, 1.0, 36, 4),
    ("denormal", 30_000, 3e-39, 36, 4),
    ("slabs", 60_000, 1e-3, 36, 8),
    ("zeros", 10_000, 1e-3, 36, 8),
    ("walk", 12_345, 1e-2, 1, 2),
    ("walk", 12_345, 1e-2, 5, 2),
    ("walk", 50_000, 1e-2, 100, 3),
    ("walk", 70_000, 1e-2, 4096, 3),
    ("walk", 16, 1e-2, 36, 16),      # n == T: seeds only
    ("walk", 17, 1e-2, 36, 16),      # remainder of 1
    ("walk", 40, 1e-2, 36, 16),
    ("walk", 3, 1e-2, 36, 1),
    ("ladder", 40_000, 1.0, 36, 4),  # every block width 0..24 (25..32: "wide")
    ("ladder", 40_000, 1.0, 20, 3),
]


@pytest.mark.parametrize("kind,n,eb,L,T", CASES)
def test_random_against_oracle(codec, kind, n, eb, L, T):
    say("random", kind, n, eb, L, T)
    x = gen.make(kind, n, n + L)
    y = gen.perturb(x, n + 1)
    sx = codec.compress(dev(x), eb, L, T)
    sy = codec.compress(dev(y), eb, L, T)
    dx = codec.decompress(sx, eb)
    h = codec.homsum([sx, sy])
    dh = codec.decompress(h, eb)
    torch.cuda.synchronize()
    codec.check_errors()
    cx, cy = O.compress(x, eb, L, T), O.compress(y, eb, L, T)
    assert np.array_equal(sx.to_host(), cx)
    assert np.array_equal(sy.to_host(), cy)
    check_index(codec, sx, cx)
    assert np.array_equal(bits(dx.cpu().numpy()), bits(O.decompress(cx, n, eb, L, T)))
    ch = O.homsum([cx, cy], n, L, T)
    assert np.array_equal(h.to_host(), ch)
    assert np.array_equal(bits(dh.cpu().numpy()), bits(O.decompress(ch, n, eb, L, T)))


# the two forms of the blockSize-36 sum: two-pass (stage area; sums of at
# least 1024 tiles) and single-pass (decoupled lookback)
FORMS = [None, False]
FORM_IDS = ["two_pass", "lookback"]
N_TWO_PASS = 2_400_017  # 1042 tiles at L = 36, T = 16


def form_n(codec, stage, small, L=36, T=16):
    """Element count for a test of `stage`'s form: large enough for the
    two-pass form to run (asserted), `small` for the single-pass one."""
    if stage is False:
        return small
    assert codec.num_tiles(N_TWO_PASS, L, T) >= codec.STAGED_MIN_TILES
    return N_TWO_PASS


def sum_form(codec, ss, stage, out=None):
    """codec.homsum in `stage`'s form.  The two-pass form gets a stage area
    filled with 0xFF beforehand, and the sum must have written the first
    tile's stream offset into it (so the two-pass kernels ran)."""
    n, L, T = ss[0].n, ss[0].L, ss[0].T
    if stage is False or L != codec.STAGED_L:
        return codec.homsum(ss, out=out, stage=stage)
    st = codec.StageArea(codec.num_tiles(n, L, T))
    st.buf.fill_(0xFF)
    h = codec.homsum(ss, out=out, stage=st)
    torch.cuda.synchronize()
    assert int(st.buf[:8].view(torch.int64).item()) == 8 * T + 4, "two-pass form did not run"
    return h


@pytest.mark.parametrize("stage", FORMS, ids=FORM_IDS)
@pytest.mark.parametrize("k", [1, 3, 8, 17, 32, 33, 40])
def test_fused_k_operand_sum(codec, k, stage):
    eb, L, T = 1e-3, 36, 16
    n = form_n(codec, stage, 200_001)
    say("homsum k", k, FORM_IDS[FORMS.index(stage)], n)
    base = gen.make("grf", n, 3)
    xs = [gen.perturb(base, 100 + r) for r in range(k)]
    ss = [codec.compress(dev(x), eb, L, T) for x in xs]
    h = sum_form(codec, ss, stage)
    torch.cuda.synchronize()
    codec.check_errors()
    ref = O.homsum([O.compress(x, eb, L, T) for x in xs], n, L, T)
    assert np.array_equal(h.to_host(), ref)


@pytest.mark.parametrize("stage", FORMS, ids=FORM_IDS)
@pytest.mark.parametrize("kind,L", [("ladder", 36), ("ladder", 20), ("wide", 36)])
def test_fused_sum_every_width(codec, kind, L, stage):
    """Eight operands whose blocks take every width (0..24 for the ladder,
    21..32 for the wide kind), so the decoder meets each split of a width
    into high bytes and low bits, r = 0 included; bytes equal the oracle's."""
    eb, T, k = 1.0, 4, 8
    n = form_n(codec, stage, 40_000, L, T) if L == 36 else 40_000
    say("homsum widths", kind, L, FORM_IDS[FORMS.index(stage)], n)
    xs = [gen.make(kind, n, 500 + r) for r in range(k)]
    ss = [codec.compress(dev(x), eb, L, T) for x in xs]
    h = sum_form(codec, ss, stage)
    torch.cuda.synchronize()
    codec.check_errors()
    cs = [O.compress(x, eb, L, T) for x in xs]
    w = set(O.parse(cs[0], n, L, T)[0].tolist())
    assert w >= (set(range(25)) if kind == "ladder" else set(range(21, 33)))
    assert np.array_equal(h.to_host(), O.homsum(cs, n, L, T))


@pytest.mark.parametrize("stage", FORMS, ids=FORM_IDS)
def test_fused_sum_tiles_wider_than_lds(codec, stage):
    """Integer data of magnitude up to 2^26 at eb = 1: every block is 28 to
    31 bits wide, so a 64-block tile of the sum holds more than a wave's
    8 KB LDS buffer and is encoded half a wave at a time."""
    eb, L, T, k = 1.0, 36, 4, 8
    n = form_n(codec, stage, 50_000, L, T)
    say("homsum tiles wider than LDS", FORM_IDS[FORMS.index(stage)], n)
    rng = np.random.default_rng(77)
    xs = [rng.integers(-2 ** 26, 2 ** 26, n).astype(np.float32) for _ in range(k)]
    ss = [codec.compress(dev(x), eb, L, T) for x in xs]
    h = sum_form(codec, ss, stage)
    torch.cuda.synchronize()
    codec.check_errors()
    cs = [O.compress(x, eb, L, T) for x in xs]
    want = O.homsum(cs, n, L, T)
    w = O.parse(want, n, L, T)[0].astype(np.int64)
    bb = np.where(w == 0, 1, 6 + (9 * w + 1) // 2)
    assert bb[:64].sum() > 8128  # the first tile of chunk 0 is too wide for the buffer
    assert np.array_equal(h.to_host(), want)


@pytest.mark.parametrize("stage", FORMS, ids=FORM_IDS)
def test_sum_output_too_small_is_reported(codec, stage):
    """A sum whose output does not fit out_cap raises the capacity error and
    leaves size 0, in both forms."""
    eb, L, T = 1e-3, 36, 8
    n = form_n(codec, stage, 120_000, L, T)
    say("homsum nospace", FORM_IDS[FORMS.index(stage)], n)
    x = gen.make("grf", n, 41)
    sx = codec.compress(dev(x), eb, L, T)
    torch.cuda.synchronize()
    codec.check_errors()
    small = codec.empty_stream(n, L, T, capacity=sx.size // 2)
    sum_form(codec, [sx, sx], stage, out=small)
    torch.cuda.synchronize()
    with pytest.raises(Exception):
        codec.check_errors()
    assert small.size == 0


# ------------------------------------------------------------ foreign streams

def expected_index(cx, n, L, T):
    """The index a parse must rebuild, from the oracle's walk: widths, the
    offset of every group's first block (16 blocks per group at L <= 64,
    fewer for longer blocks; an empty segment's group points past its seed),
    every segment's seed offset and the stream size, at the layout of
    zccl_internal.h make_geom."""
    w, bo, so, tot = O.parse(cx, n, L, T)
    tb, R = n // T, n % T
    nbT = -(-(tb - 1) // L)
    nbR = -(-(R - 1) // L) if R else 0
    IG = max(1, min(16, 576 // ((L + 3) // 4)))
    gT = -(-nbT // IG) if nbT else 1
    gR = (-(-nbR // IG) if nbR else 1) if R else 0
    groups = []
    for s in range(T + (1 if R else 0)):
        nb, g = (nbT, gT) if s < T else (nbR, gR)
        b0 = s * nbT
        groups += [int(bo[b0 + q * IG]) for q in range(g)] if nb else [int(so[s]) + 4]
    return w, np.array(groups, dtype=np.uint64), np.asarray(so, dtype=np.uint64), int(tot)


def check_full_index(s, cx):
    w, groups, seeds, tot = expected_index(cx, s.n, s.L, s.T)
    ix = s.index.cpu().numpy()
    nblocks = w.size
    g0 = -(-nblocks // 16) * 16
    assert np.array_equal(ix[:nblocks], w)
    assert np.array_equal(ix[g0: g0 + 8 * groups.size].view(np.uint64), groups)
    s0 = g0 + 8 * groups.size
    assert np.array_equal(ix[s0: s0 + 8 * seeds.size].view(np.uint64), seeds)
    assert s.size == tot == cx.size


PARSE_CASES = [("walk", 100_003, 1e-2, 36, 16), ("sines", 5000, 1e-4, 7, 3), ("walk", 17, 1e-2, 36, 16),
               ("zeros", 999, 1e-2, 36, 4), ("cfg1", 1 << 24, 1e-4, 36, 16), ("grf", 3_000_017, 1e-3, 36, 7),
               ("slabs", 2_000_000, 1e-2, 36, 8), ("wide", 600_000, 1.0, 36, 2), ("walk", 700_001, 1e-3, 64, 5),
               ("sines", 400_000, 1e-4, 8, 3), ("grf", 1_000_000, 1e-3, 36, 1), ("zeros", 2_000_000, 1e-2, 36, 4)]


@pytest.mark.parametrize("kind,n,eb,L,T", PARSE_CASES)
def test_parse_cpu_stream(codec, kind, n, eb, L, T):
    """A stream written by the oracle (standing in for the CPU library),
    uploaded without its index: the rebuilt index equals the oracle's walk
    in every entry, through the parallel parse (the default workspace has
    room for its tables; it serves chunks of 256 blocks and more at
    L <= 64) and through the serial walk (a workspace without that room),
    and the stream decompresses to the oracle's values."""
    say("parse", kind, n, L, T)
    x = gen.make(kind, n, 5)
    cx = O.compress(x, eb, L, T)
    s = codec.upload(cx, n, L, T)  # default workspace: parallel where it applies
    y = codec.decompress(s, eb)
    torch.cuda.synchronize()
    codec.check_errors()
    check_full_index(s, cx)
    assert np.array_equal(bits(y.cpu().numpy()), bits(O.decompress(cx, n, eb, L, T)))
    # the serial walk, forced by a workspace too small for the tables
    small = codec.Workspace(codec.num_tiles(n, L, T))
    t = torch.from_numpy(np.ascontiguousarray(cx)).cuda()
    s2 = codec.from_bytes(t, n, L, T, ws=small)
    torch.cuda.synchronize()
    codec.check_errors(small)
    check_full_index(s2, cx)


@pytest.mark.parametrize("fill", ["zeros", "width1", "random"])
def test_parse_stream_in_larger_buffer(codec, fill):
    """A foreign stream at the start of a worst-case-sized buffer (the
    reference ABI passes no stream size): whatever follows the stream (zero
    bytes, which read as one-byte blocks; bytes of 1, which read as 11-byte
    blocks; random bytes) leaves the rebuilt index and the size exactly the
    oracle's."""
    n, eb, L, T = 1_500_007, 1e-3, 36, 6
    say("parse in buffer", fill, n, L, T)
    x = gen.make("grf", n, 11)
    cx = O.compress(x, eb, L, T)
    extra = codec.worst_case_bytes(n, L, T) - cx.size
    rng = np.random.default_rng(3)
    tail = {"zeros": np.zeros(extra, np.uint8), "width1": np.ones(extra, np.uint8),
            "random": rng.integers(0, 256, extra, dtype=np.uint8)}[fill]
    s = codec.upload(np.concatenate([cx, tail]), n, L, T)
    torch.cuda.synchronize()
    codec.check_errors()
    check_full_index(s, cx)


def test_parallel_parse_rejects_corrupt_streams(codec):
    """Corrupt foreign streams through the parallel parse: a width byte above
    32 inside a chunk, and a chunk offset in the header that does not match
    where the previous chunk's blocks end, are reported; no stream size is
    left in the index."""
    n, eb, L, T = 2_000_003, 1e-3, 36, 4
    x = gen.make("grf", n, 9)
    cx = O.compress(x, eb, L, T)
    w, bo, so, tot = O.parse(cx, n, L, T)
    bad1 = cx.copy()
    bad1[int(bo[w.size // 3])] = 99
    bad2 = cx.copy()
    h = bad2[8:16].view(np.uint64)
    h[0] += 1  # chunk 1 starts one byte later than chunk 0's blocks end
    for bad in (bad1, bad2):
        t = torch.from_numpy(np.ascontiguousarray(bad)).cuda()
        s = codec.from_bytes(t, n, L, T)
        torch.cuda.synchronize()
        with pytest.raises(Exception):
            codec.check_errors()
        s.invalidate()
        assert s.size == 0


# ------------------------------------------------------------ drop-in ABI

@pytest.fixture(scope="module")
def capi(codec):
    from zccl_amd import _lib

    L = _lib.lib()
    yield L
    L.zccl_set_num_chunks(0)


def test_dropin_host_pointers(capi):
    n, eb, L, T = 300_007, 1e-3, 36, 8
    say("dropin host", n, L, T)
    capi.zccl_set_num_chunks(T)
    x = gen.make("sines", n, 9)
    y = gen.perturb(x, 10)
    out = np.zeros(4 * n, dtype=np.uint8)
    size = ctypes.c_size_t(0)
    capi.ZCCL_float_openmp_threadblock_arg(out.ctypes.data, x.ctypes.data, ctypes.byref(size), eb, n, L)
    assert capi.zccl_last_error() == 0
    cx = O.compress(x, eb, L, T)
    assert size.value == cx.size and np.array_equal(out[: size.value], cx)
    dec = np.zeros(n, dtype=np.float32)
    capi.ZCCL_float_decompress_openmp_threadblock_arg(dec.ctypes.data, n, eb, L, out.ctypes.data)
    assert capi.zccl_last_error() == 0
    assert np.array_equal(bits(dec), bits(O.decompress(cx, n, eb, L, T)))
    # in-place add (final == second operand) with the size accumulated as the reference does
    cy = O.compress(y, eb, L, T)
    buf2 = np.zeros(4 * n, dtype=np.uint8)
    buf2[: cy.size] = cy
    fsize = ctypes.c_size_t(1000)
    capi.ZCCL_float_homomophic_add_openmp_threadblock(buf2.ctypes.data, ctypes.byref(fsize), n, eb, L,
                                                      out.ctypes.data, buf2.ctypes.data)
    assert capi.zccl_last_error() == 0
    ref = O.homsum([cx, cy], n, L, T)
    assert fsize.value == 1000 + ref.size
    assert np.array_equal(buf2[: ref.size], ref)


def test_dropin_record_counters(capi):
    n, eb, L, T = 100_000, 1e-3, 36, 4
    capi.zccl_set_num_chunks(T)
    # x is zero on the first half, y on the second and fourth quarters, so
    # blocks meet all four cases: (0,0), (0,y), (x,0), (x,y)
    x = gen.make("walk", n, 2)
    y = gen.make("walk", n, 3)
    x[: n // 2] = 0.0
    y[n // 4: n // 2] = 0.0
    y[3 * n // 4:] = 0.0
    cx, cy = O.compress(x, eb, L, T), O.compress(y, eb, L, T)
    a = np.zeros(4 * n, dtype=np.uint8); a[: cx.size] = cx
    b = np.zeros(4 * n, dtype=np.uint8); b[: cy.size] = cy
    out = np.zeros(4 * n, dtype=np.uint8)
    size = ctypes.c_size_t(0)
    c = [ctypes.c_int(1) for _ in range(4)]
    capi.ZCCL_float_homomophic_add_openmp_threadblock_record(
        out.ctypes.data, ctypes.byref(size), n, eb, L, a.ctypes.data, b.ctypes.data, *[ctypes.byref(v) for v in c])
    assert capi.zccl_last_error() == 0
    wa = O.parse(cx, n, L, T)[0] != 0
    wb = O.parse(cy, n, L, T)[0] != 0
    expect = [(~wa & ~wb).sum(), (~wa & wb).sum(), (wa & ~wb).sum(), (wa & wb).sum()]
    assert [v.value - 1 for v in c] == [int(e) for e in expect]
    assert min(expect) > 0


def test_dropin_device_pointers(capi, codec):
    n, eb, L, T = 123_457, 1e-2, 36, 6
    say("dropin device", n, L, T)
    capi.zccl_set_num_chunks(T)
    x = gen.make("walk", n, 4)
    dx = dev(x)
    out = torch.zeros(4 * n, dtype=torch.uint8, device="cuda")
    size = ctypes.c_size_t(0)
    capi.ZCCL_float_openmp_threadblock_arg(out.data_ptr(), dx.data_ptr(), ctypes.byref(size), eb, n, L)
    assert capi.zccl_last_error() == 0
    cx = O.compress(x, eb, L, T)
    assert np.array_equal(out[: size.value].cpu().numpy(), cx)
    dec = torch.empty(n, dtype=torch.float32, device="cuda")
    capi.ZCCL_float_decompress_openmp_threadblock_arg(dec.data_ptr(), n, eb, L, out.data_ptr())
    assert capi.zccl_last_error() == 0
    assert np.array_equal(bits(dec.cpu().numpy()), bits(O.decompress(cx, n, eb, L, T)))


def test_single_thread_twins(capi):
    n, eb, L = 20_000, 1e-2, 36
    x = SMALL["single_thread.x"]
    n = x.size
    out = np.zeros(4 * n + 64, dtype=np.uint8)
    size = ctypes.c_size_t(0)
    capi.ZCCL_float_single_thread_arg(out.ctypes.data, x.ctypes.data, ctypes.byref(size), eb, n, L)
    assert np.array_equal(out[: size.value], SMALL["single_thread.cx"])
    dec = np.zeros(n, dtype=np.float32)
    capi.ZCCL_float_decompress_single_thread_arg(dec.ctypes.data, n, eb, L, out.ctypes.data)
    assert np.array_equal(bits(dec), bits(SMALL["single_thread.dx"]))


# ------------------------------------------------------------ errors

def test_errors_are_reported_not_faulted(codec, capi):
    n, eb, L, T = 50_000, 1e-3, 36, 4
    x = gen.make("walk", n, 1)
    # output capacity too small
    small = codec.empty_stream(n, L, T, capacity=1000)
    codec.compress(dev(x), eb, L, T, out=small)
    with pytest.raises(Exception):
        codec.check_errors()
    # corrupt width byte (> 32) in a CPU stream: parse stops and reports
    cx = O.compress(x, eb, L, T)
    w, bo, so, tot = O.parse(cx, n, L, T)
    bad = cx.copy()
    bad[int(bo[5])] = 77
    codec.upload(bad, n, L, T)
    with pytest.raises(Exception):
        codec.check_errors()
    # n < T
    with pytest.raises(Exception):
        codec.compress(dev(x[:3]), eb, L, T)
    # drop-in status: the same corrupt stream (chunk 0 broken, last chunk
    # intact, so the host-side sizing succeeds) reaches the parse and reports
    capi.zccl_set_num_chunks(T)
    try:
        dec = np.zeros(n, dtype=np.float32)
        capi.ZCCL_float_decompress_openmp_threadblock_arg(dec.ctypes.data, n, eb, L, bad.ctypes.data)
        assert capi.zccl_last_error() == -3  # ZCCL_ECORRUPT
    finally:
        capi.zccl_set_num_chunks(0)


def test_failed_parse_leaves_no_stream_size(codec):
    """A parse that fails on the last chunk must not leave an old stream size
    in the index (later kernels read up to it): the index is pre-filled with
    0xFF, the last chunk's first width byte is corrupted, and the size word and
    the optional device size must read 0 afterwards, with ECORRUPT reported."""
    from zccl_amd._lib import lib

    n, eb, L, T = 50_000, 1e-3, 36, 4
    say("failed parse")
    x = gen.make("walk", n, 3)
    cx = O.compress(x, eb, L, T)
    w, bo, so, tot = O.parse(cx, n, L, T)
    bad = cx.copy()
    nb_chunk = len(bo) // T
    bad[int(bo[(T - 1) * nb_chunk])] = 77  # first block of the last chunk
    data = dev(bad)
    index = torch.full((codec.index_bytes(n, L, T),), 0xFF, dtype=torch.uint8, device="cuda")
    d_size = torch.full((1,), -1, dtype=torch.int64, device="cuda")
    ws = codec.Workspace(1024)
    rc = lib().zccl_parse_dev(ctypes.c_void_p(data.data_ptr()), data.numel(), n, L, T,
                              ctypes.c_void_p(index.data_ptr()), ctypes.c_void_p(d_size.data_ptr()),
                              ws.ptr, ws.nbytes, codec._stream_handle())
    assert rc == 0
    torch.cuda.synchronize()
    off = int(lib().zccl_index_size_offset(n, L, T))
    size_word = int(index[off: off + 8].cpu().numpy().view(np.uint64)[0])
    assert size_word == 0
    assert int(d_size.item()) == 0
    with pytest.raises(Exception):
        codec.check_errors(ws)


@pytest.mark.parametrize("L", [36, 32])
def test_sum_reports_width_byte_that_disagrees_with_index(codec, L):
    """A stream byte changed under its index (a block's width byte set to
    another valid width) is reported by the sum (slot layout for L = 36, the
    generic kernel otherwise), not summed silently."""
    n, eb, T = 60_000, 1e-3, 4
    say("width mismatch", L)
    x = gen.make("grf", n, 5)
    y = gen.perturb(x, 6)
    sx = codec.compress(dev(x), eb, L, T)
    sy = codec.compress(dev(y), eb, L, T)
    torch.cuda.synchronize()
    codec.check_errors()
    cx = sx.to_host()
    w, bo, so, tot = O.parse(cx, n, L, T)
    j = int(np.nonzero(w)[0][7])
    sx.data[int(bo[j])] = (int(w[j]) % 32) + 1
    codec.homsum([sx, sy])
    torch.cuda.synchronize()
    with pytest.raises(Exception):
        codec.check_errors()


# ------------------------------------------------------------ batch plans

def test_batch_plans_uniform_and_mixed(codec):
    from zccl_amd import batch

    L, T, eb = 36, 8, 1e-3
    # uniform batches below and above the two-pass threshold (1024 tiles), and a mixed one
    for sizes in ([100_000] * 5, [650_000] * 4, [7, 40, 1000, 123_457, 9, 60_001]):
        say("batch sizes", sizes)
        xs = [gen.make("grf", n, 11 + i) for i, n in enumerate(sizes) if n >= T]
        ns = [x.size for x in xs]
        dxs = [dev(x) for x in xs]
        ebs = [eb * (1 + i) for i in range(len(xs))]
        outs = [codec.empty_stream(n, L, T) for n in ns]
        plan = batch.compress_plan(dxs, ebs, outs)
        plan.launch()
        plan.launch()  # replays must give the same bytes
        ys = [gen.perturb(x, 50 + i) for i, x in enumerate(xs)]
        sy = [codec.compress(dev(y), e, L, T) for y, e in zip(ys, ebs)]
        sums = [codec.empty_stream(n, L, T) for n in ns]
        hp = batch.homsum_plan([[a, b] for a, b in zip(outs, sy)], sums)
        # single-pass form first, then the two-pass form (uniform batches
        # only) over the same outputs
        hp.launch(stage=False)
        torch.cuda.synchronize()
        codec.check_errors()
        single = [s.to_host() for s in sums]
        hp.launch()
        decs = [torch.empty(n, dtype=torch.float32, device="cuda") for n in ns]
        dp = batch.decompress_plan(sums, ebs, decs)
        dp.launch()
        torch.cuda.synchronize()
        codec.check_errors()
        for i, n in enumerate(ns):
            cx = O.compress(xs[i], ebs[i], L, T)
            cy = O.compress(ys[i], ebs[i], L, T)
            assert np.array_equal(outs[i].to_host(), cx), (sizes, i)
            ch = O.homsum([cx, cy], n, L, T)
            assert np.array_equal(sums[i].to_host(), ch), (sizes, i)
            assert np.array_equal(single[i], ch), (sizes, i)
            assert np.array_equal(bits(decs[i].cpu().numpy()), bits(O.decompress(ch, n, ebs[i], L, T)))


def test_batch_homsum_ragged_operand_counts(codec):
    """One launch with jobs of 1 to 32 operands and different sizes."""
    from zccl_amd import batch

    L, T, eb = 36, 8, 1e-3
    shapes = [(50_000, 1), (123_457, 3), (8, 8), (77_777, 32), (1000, 5), (262_144, 17)]
    say("ragged", shapes)
    jobs, outs, want = [], [], []
    for j, (n, k) in enumerate(shapes):
        xs = [gen.make("grf", n, 300 + 40 * j + r) for r in range(k)]
        cs = [O.compress(x, eb, L, T) for x in xs]
        jobs.append([codec.upload(c, n, L, T) for c in cs])
        outs.append(codec.empty_stream(n, L, T))
        want.append(O.homsum(cs, n, L, T))
    plan = batch.homsum_plan(jobs, outs)
    ws = codec.Workspace(plan.tiles)
    plan.launch(ws)
    plan.launch(ws)
    torch.cuda.synchronize()
    codec.check_errors(ws)
    for j, (n, k) in enumerate(shapes):
        assert np.array_equal(outs[j].to_host(), want[j]), shapes[j]


def test_batch_owning_create_entry_points(codec):
    """The allocating create() forms give the same bytes as host-built plans."""
    from zccl_amd._lib import lib, check

    L, T, eb = 36, 8, 2e-3
    ns = [100_003, 65_536, 4_097]
    xs = [gen.make("sines", n, 500 + i) for i, n in enumerate(ns)]
    dxs = [dev(x) for x in xs]
    outs = [codec.empty_stream(n, L, T) for n in ns]
    P = lambda ts: (ctypes.c_void_p * len(ts))(*[t.data_ptr() for t in ts])  # noqa: E731
    nn = (ctypes.c_size_t * 3)(*ns)
    ebs = (ctypes.c_float * 3)(eb, eb, eb)
    st = ctypes.c_int(0)
    h = lib().zccl_batch_compress_create(3, nn, ebs, L, T, P(dxs), P([o.data for o in outs]),
                                         (ctypes.c_size_t * 3)(*[o.data.numel() for o in outs]),
                                         P([o.index for o in outs]), ctypes.byref(st))
    check(st.value, "compress create")
    sums = [codec.empty_stream(n, L, T) for n in ns]
    h2 = lib().zccl_batch_homsum_create(3, 2, nn, L, T, P([t for o in outs for t in (o.data, o.data)]),
                                        P([t for o in outs for t in (o.index, o.index)]),
                                        P([o.data for o in sums]),
                                        (ctypes.c_size_t * 3)(*[o.data.numel() for o in sums]),
                                        P([o.index for o in sums]), ctypes.byref(st))
    check(st.value, "homsum create")
    ws = codec.Workspace(max(int(lib().zccl_batch_tiles(h)), int(lib().zccl_batch_tiles(h2))))
    check(lib().zccl_batch_launch(h, ws.ptr, ws.nbytes, codec._stream_handle()), "launch")
    check(lib().zccl_batch_launch(h2, ws.ptr, ws.nbytes, codec._stream_handle()), "launch")
    torch.cuda.synchronize()
    codec.check_errors(ws)
    lib().zccl_batch_destroy(h)
    lib().zccl_batch_destroy(h2)
    for i, n in enumerate(ns):
        c = O.compress(xs[i], eb, L, T)
        assert np.array_equal(outs[i].to_host(), c)
        assert np.array_equal(sums[i].to_host(), O.homsum([c, c], n, L, T))


def test_graph_capture_replays_match_eager(codec):
    """Batch launches captured in a HIP graph: every replay, with eager work
    and synchronisation between replays, equals the eager result."""
    from zccl_amd import batch

    L, T = 36, 16
    ns = [300_000, 1_000, 65_537, 2_000_000]
    xs = [dev(gen.make("grf", n, 700 + i)) for i, n in enumerate(ns)]
    ebs = [1e-3] * len(ns)
    outs = [codec.empty_stream(n, L, T) for n in ns]
    sums = [codec.empty_stream(n, L, T) for n in ns]
    cp = batch.compress_plan(xs, ebs, outs)
    hp = batch.homsum_plan([[o, o, o] for o in outs], sums)
    ws = codec.Workspace(max(cp.tiles, hp.tiles))
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(s):
        cp.launch(ws)
        hp.launch(ws)
    torch.cuda.current_stream().wait_stream(s)
    torch.cuda.synchronize()
    codec.check_errors(ws)
    eager = [(o.data[: o.size].clone(), m.data[: m.size].clone()) for o, m in zip(outs, sums)]
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        cp.launch(ws)
        hp.launch(ws)
    for rep in range(4):
        for o, m in zip(outs, sums):
            o.data.zero_()
            m.data.zero_()
        g.replay()
        torch.cuda.synchronize()
        for o, m in zip(outs, sums):
            o.invalidate()
            m.invalidate()
        codec.check_errors(ws)
        for (e1, e2), o, m in zip(eager, outs, sums):
            assert torch.equal(o.data[: o.size], e1), rep
            assert torch.equal(m.data[: m.size], e2), rep
        _ = (xs[0] * 2).sum().item()  # eager work between replays


def test_tuned_workgroup_sizes_give_same_bytes(codec):
    from zccl_amd._lib import lib

    L, T, n, eb = 36, 8, 333_333, 1e-3
    x = gen.make("grf", n, 900)
    y = gen.make("grf", n, 901)
    want_c = O.compress(x, eb, L, T)
    want_h = O.homsum([want_c, O.compress(y, eb, L, T)], n, L, T)
    try:
        for th in ((64, 64, 64), (256, 128, 256), (512, 256, 512)):
            say("threads", th)
            for op, t in enumerate(th):
                assert lib().zccl_tune_threads(op, t) == 0
            a = codec.compress(dev(x), eb, L, T)
            b = codec.compress(dev(y), eb, L, T)
            h = codec.homsum([a, b])
            d = codec.decompress(h, eb)
            torch.cuda.synchronize()
            codec.check_errors()
            assert np.array_equal(a.to_host(), want_c)
            assert np.array_equal(h.to_host(), want_h)
            assert np.array_equal(bits(d.cpu().numpy()), bits(O.decompress(want_h, n, eb, L, T)))
    finally:
        for op in range(3):
            lib().zccl_tune_threads(op, 0)
    assert lib().zccl_tune_threads(1, 512) != 0  # above the sum kernel's launch bound


def test_staged_allreduce_forms_agree(codec):
    """hZCCL data flow on one device: batched and per-call compress agree,
    the fused reduce-scatter equals the reference's ring of pairwise adds,
    batched and per-call decompress agree, and both the homomorphic result
    and the decompress-operate-compress ring stay within their error bounds
    of the FP64 sum."""
    from zccl_amd.allreduce import StagedAllreduce

    P, count, L, T, eb = 4, 400_003, 36, 8, 1e-3
    xs = [dev(gen.make("grf", count, 1200 + r)) for r in range(P)]
    ar = StagedAllreduce(P=P, count=count, eb=eb, L=L, T=T)
    ar.compress_all(xs, batched=True)
    torch.cuda.synchronize()
    batched = [[s.data[: s.size].clone() for s in row] for row in ar.streams]
    ar.compress_all(xs, batched=False)
    torch.cuda.synchronize()
    for r in range(P):
        for j in range(P):
            assert torch.equal(ar.streams[r][j].data[: ar.streams[r][j].size], batched[r][j]), (r, j)
    fused = ar.reduce_scatter_fused()
    ring = ar.reduce_scatter_ring(ar.ring_scratch())
    for a, b in zip(fused, ring):
        assert torch.equal(a.data[: a.size], b.data[: b.size])
    y1 = ar.decompress_all(fused, batched=True)
    y2 = ar.decompress_all(fused, batched=False)
    torch.cuda.synchronize()
    codec.check_errors()
    assert torch.equal(y1.view(torch.int32), y2.view(torch.int32))
    truth = sum(x.double() for x in xs)
    bound = P * eb + 2.0 ** -21 * sum(x.double().abs() for x in xs)
    assert bool(((y1.double() - truth).abs() <= bound).all())
    doc = ar.reduce_scatter_doc(xs)
    for j, seg in enumerate(doc):
        t = truth[ar.displs[j]: ar.displs[j] + ar.cnts[j]]
        err = (seg.double() - t).abs().max().item()
        assert err <= (P - 1) * (eb + 1e-4), (j, err)


def test_doc_ring_bit_exact(codec):
    """The decompress-operate-compress ring (ZCCL/src/ZCCL_ring.c:979-1108) on
    the GPU equals the oracle's zo_doc_chain bit for bit for every segment:
    each rank's FP32 partial sum, and the value after the compressed
    allgather.  The oracle is pinned against the compiled reference
    (tests/test_oracle.py::test_doc_chain_equals_compiled_reference_ring)."""
    from zccl_amd.allreduce import StagedAllreduce, ring_order

    P, count, L, T, eb = 4, 4 * 60_001 + 3, 36, 8, 1e-3
    say("doc ring", P, count)
    hx = [gen.make("grf", count, 1300 + r) for r in range(P)]
    xs = [dev(x) for x in hx]
    ar = StagedAllreduce(P=P, count=count, eb=eb, L=L, T=T)
    doc = ar.reduce_scatter_doc(xs)
    torch.cuda.synchronize()
    codec.check_errors()
    for j, seg in enumerate(doc):
        a, b = ar.displs[j], ar.displs[j] + ar.cnts[j]
        acc, fin = O.doc_chain([hx[r][a:b] for r in ring_order(P, j)], eb, L, T)
        assert np.array_equal(bits(seg.cpu().numpy()), bits(acc)), j
        g = codec.decompress(codec.compress(seg, eb, L, T), eb)
        assert np.array_equal(bits(g.cpu().numpy()), bits(fin)), j


def test_two_operand_doc_bit_exact(codec):
    """The two-operand DOC workflow timed by examples/testfloat_homocompress_
    fastmode1_arg.c:211-243: decompress both streams, add in FP32 (add_vectors,
    :40-46), compress the sum -- equal to the oracle's composition bit for bit,
    next to the homomorphic add of the same streams."""
    n, L, T, eb = 300_007, 36, 8, 1e-3
    say("two-operand doc", n)
    x, y = gen.make("grf", n, 1400), gen.make("grf", n, 1401)
    sx, sy = codec.compress(dev(x), eb, L, T), codec.compress(dev(y), eb, L, T)
    dsum = codec.decompress(sx, eb) + codec.decompress(sy, eb)
    doc = codec.compress(dsum, eb, L, T)
    hom = codec.homsum([sx, sy])
    torch.cuda.synchronize()
    codec.check_errors()
    cx, cy = O.compress(x, eb, L, T), O.compress(y, eb, L, T)
    odsum = O.decompress(cx, n, eb, L, T) + O.decompress(cy, n, eb, L, T)
    assert np.array_equal(doc.to_host(), O.compress(odsum, eb, L, T))
    assert np.array_equal(hom.to_host(), O.homsum([cx, cy], n, L, T))
