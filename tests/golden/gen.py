"""Deterministic FP32 input generators for codec tests and fixtures (numpy only).

Every generator takes (n, seed) and returns a contiguous float32 array; the same
(kind, n, seed) gives the same bytes on the same numpy build.
"""
from __future__ import annotations

import numpy as np


def _walk(n, rng):
    return (np.cumsum(rng.normal(0.0, 1.0, n)) * 0.05).astype(np.float32)


def _sines(n, rng):
    """Sum of 16 random-phase sinusoids (periods log-uniform 64..65536 elements,
    amplitudes U[0.1,1]) plus N(0,1e-3) noise."""
    i = np.arange(n, dtype=np.float64)
    x = np.zeros(n, dtype=np.float64)
    periods = np.exp(rng.uniform(np.log(64.0), np.log(65536.0), 16))
    amps = rng.uniform(0.1, 1.0, 16)
    phases = rng.uniform(0.0, 2 * np.pi, 16)
    for p, a, ph in zip(periods, amps, phases):
        x += a * np.sin(2 * np.pi * i / p + ph)
    x += rng.normal(0.0, 1e-3, n)
    return x.astype(np.float32)


def _grf1d(n, rng, slope=-3.0):
    """1-D Gaussian random field with power spectrum P(k) ~ k^slope, unit std."""
    w = rng.normal(0.0, 1.0, n)
    f = np.fft.rfft(w)
    k = np.fft.rfftfreq(n)
    k[0] = k[1] if n > 1 else 1.0
    f *= k ** (slope / 2.0)
    f[0] = 0.0
    x = np.fft.irfft(f, n)
    s = x.std()
    return (x / (s if s > 0 else 1.0)).astype(np.float32)


def _slabs(n, rng):
    """GRF with 70% of the array replaced by contiguous zero slabs."""
    x = _grf1d(n, rng)
    nslab = 7
    edges = np.linspace(0, n, 2 * nslab + 1).astype(int)
    for s in range(nslab):
        a, b = edges[2 * s], edges[2 * s + 1]
        b = a + int((b - a) * 1.4 * 2) if s < nslab - 1 else n
        x[a:min(b, n)] = 0.0
    return x


def _wide(n, rng):
    """Integer-valued data (use eb=1: inverse bound exactly 1) whose blocks have
    max |delta| on both sides of the width-bump thresholds 2^w - c_w,
    w = 21..31, separated by flat blocks (width 0)."""
    bump = {21: 1, 22: 2, 23: 5, 24: 11, 25: 22, 26: 45, 27: 91, 28: 183, 29: 367, 30: 735}
    q = []
    targets = []
    for w, c in bump.items():
        for dc in (-1, 0, 1):
            targets.append((float(c + dc), float(2 ** w)))
    # w = 31: representable integers near 2^30 are multiples of 128; 2^31-1408
    # is above the 2^31-1471 threshold, 2^31-1536 below it
    targets.append((-float(2 ** 30 - 1408), float(2 ** 30)))
    targets.append((-float(2 ** 30 - 1536), float(2 ** 30)))
    ti = 0
    while len(q) < n:
        lo, hi = targets[ti % len(targets)]
        ti += 1
        q.extend([lo, hi] * 18)          # one block of alternating values
        q.extend([lo] * 36)              # a flat block
    x = np.array(q[:n], dtype=np.float32)
    x[0] = 0.0
    return x


def _special(n, rng):
    """Mostly a smooth walk, with NaN, +-inf, values past 2^31*eb and -0.0."""
    x = _walk(n, rng)
    idx = rng.choice(n, size=min(n, 12), replace=False)
    vals = [np.nan, np.inf, -np.inf, 3.0e6, -3.0e6, -0.0, 2.2e6, -2.2e6, np.nan, 1e30, -1e30, 0.0]
    for i, v in zip(idx, vals):
        x[i] = v
    return x.astype(np.float32)


def _denormal(n, rng):
    """Values of order 1e-37 (use eb=3e-39, itself subnormal): quantized values
    are small and most dequantized results are FP32 subnormals."""
    return (rng.normal(0.0, 1.0, n) * 1e-37).astype(np.float32)


def _outliers(n, rng):
    """Gradient-bucket-like: N(0,1e-2) with 1% outliers at 100x."""
    x = rng.normal(0.0, 1e-2, n)
    m = rng.random(n) < 0.01
    x[m] *= 100.0
    return x.astype(np.float32)


def _ladder(n, rng):
    """Width ladder for eb = 1: 36-value stretches s that alternate between 0
    and +-m with m in [2^(j-1), 2^j), j = s mod 25 (j = 0: a flat stretch),
    so deltas of magnitude m give every block width 0..24 (integers below
    2^24, exact in FP32; widths 25..32 are the "wide" kind's)."""
    x = np.zeros(n, dtype=np.float64)
    for s, i in enumerate(range(0, n, 36)):
        j = s % 25
        k = min(36, n - i)
        if j == 0:
            continue
        lo = 2 ** (j - 1)
        m = rng.integers(lo, 2 * lo, k).astype(np.float64)
        sg = np.where(rng.random(k) < 0.5, -1.0, 1.0)
        x[i:i + k] = np.where(np.arange(k) & 1, sg * m, 0.0)
    return x.astype(np.float32)


_GEN = {
    "walk": _walk,
    "sines": _sines,
    "cfg1": _sines,
    "grf": _grf1d,
    "zeros": lambda n, rng: np.zeros(n, dtype=np.float32),
    "slabs": _slabs,
    "wide": _wide,
    "special": _special,
    "denormal": _denormal,
    "outliers": _outliers,
    "ladder": _ladder,
}

KINDS = tuple(_GEN)


def make(kind: str, n: int, seed: int) -> np.ndarray:
    rng = np.random.default_rng(seed)
    return np.ascontiguousarray(_GEN[kind](n, rng), dtype=np.float32)


def perturb(x: np.ndarray, seed: int) -> np.ndarray:
    """Second operand: x plus N(0, 0.05*std) on ~60% of 256-element stretches
    (the rest left equal), so pairwise sums meet zero/non-zero block mixes."""
    rng = np.random.default_rng(seed)
    y = x.astype(np.float64).copy()
    finite = np.isfinite(y)
    s = float(np.std(y[finite])) if finite.any() else 0.0
    if s == 0.0:
        s = 1e-2
    noise = rng.normal(0.0, 0.05 * s, x.size)
    mask = np.repeat(rng.random((x.size + 255) // 256) < 0.6, 256)[: x.size]
    y[mask & finite] += noise[mask & finite]
    return y.astype(np.float32)


def smooth3d(dims, seed: int) -> np.ndarray:
    """3-D field smooth along every axis: separable random sinusoids plus
    N(0, 1e-3) noise, float32, C order (dims[0] slowest)."""
    rng = np.random.default_rng(seed)
    x = np.zeros(dims, dtype=np.float64)
    for _ in range(6):
        f = rng.uniform(0.02, 0.3, 3)
        ph = rng.uniform(0, 2 * np.pi, 3)
        a = rng.uniform(0.2, 1.0)
        gx = np.sin(2 * np.pi * f[0] * np.arange(dims[0]) + ph[0])[:, None, None]
        gy = np.sin(2 * np.pi * f[1] * np.arange(dims[1]) + ph[1])[None, :, None]
        gz = np.sin(2 * np.pi * f[2] * np.arange(dims[2]) + ph[2])[None, None, :]
        x += a * gx * gy * gz
    x += rng.normal(0.0, 1e-3, dims)
    return np.ascontiguousarray(x.astype(np.float32))
