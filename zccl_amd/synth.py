"""Synthetic FP32 fields for the benchmark workloads, generated on the GPU.

* ``grf3d``   -- Gaussian random field on an s^3 grid with power spectrum
                 P(k) ~ k^-3 (white noise filtered by |k|^-1.5 in Fourier
                 space), unit standard deviation.
* ``perturb`` -- x + N(0, sigma) from an independent generator.
* ``sines``   -- sum of 16 random-phase sinusoids plus N(0, 1e-3) noise.
* ``gradient_like`` -- N(0, 1e-2) values with 1% outliers at 100x.

Deterministic for a given seed on a given device and torch build.
"""
from __future__ import annotations

import math

import torch


def _gen(seed: int, device) -> torch.Generator:
    g = torch.Generator(device=device)
    g.manual_seed(int(seed))
    return g


def grf3d(side: int, seed: int, slope: float = -3.0, device="cuda") -> torch.Tensor:
    """Flattened (side^3,) float32 field, P(k) ~ k^slope, unit std."""
    g = _gen(seed, device)
    w = torch.randn((side, side, side), generator=g, device=device, dtype=torch.float32)
    f = torch.fft.rfftn(w)
    del w
    k1 = torch.fft.fftfreq(side, device=device)
    kz = torch.fft.rfftfreq(side, device=device)
    kk = (k1[:, None, None] ** 2 + k1[None, :, None] ** 2 + kz[None, None, :] ** 2)
    kk[0, 0, 0] = 1.0
    f *= kk.pow(slope / 4.0)          # |k|^(slope/2) = (k^2)^(slope/4)
    f[0, 0, 0] = 0
    del kk
    x = torch.fft.irfftn(f, s=(side, side, side))
    del f
    x /= x.std()
    return x.reshape(-1).contiguous()


def perturb(x: torch.Tensor, sigma: float, seed: int) -> torch.Tensor:
    g = _gen(seed, x.device)
    return (x + sigma * torch.randn(x.shape, generator=g, device=x.device, dtype=x.dtype)).contiguous()


def sines(n: int, seed: int, device="cuda") -> torch.Tensor:
    g = _gen(seed, "cpu")
    periods = torch.exp(torch.empty(16, dtype=torch.float64).uniform_(math.log(64), math.log(65536), generator=g))
    amps = torch.empty(16, dtype=torch.float64).uniform_(0.1, 1.0, generator=g)
    phases = torch.empty(16, dtype=torch.float64).uniform_(0, 2 * math.pi, generator=g)
    i = torch.arange(n, device=device, dtype=torch.float64)
    x = torch.zeros(n, device=device, dtype=torch.float64)
    for p, a, ph in zip(periods.tolist(), amps.tolist(), phases.tolist()):
        x += a * torch.sin(2 * math.pi * i / p + ph)
    gd = _gen(seed + 1, device)
    x += 1e-3 * torch.randn(n, generator=gd, device=device, dtype=torch.float64)
    return x.float().contiguous()


def gradient_like(n: int, seed: int, device="cuda") -> torch.Tensor:
    g = _gen(seed, device)
    x = 1e-2 * torch.randn(n, generator=g, device=device)
    m = torch.rand(n, generator=g, device=device) < 0.01
    x[m] *= 100.0
    return x.contiguous()


def value_range(xs) -> float:
    """max over fields of (max - min), in FP32 as the reference computes it
    (examples/testfloat_homocompress_fastmode1_arg.c:114-132, 181-191)."""
    r = max(float((x.max() - x.min()).item()) for x in xs)
    return float(torch.tensor(r, dtype=torch.float32).item())


def rel_to_abs_eb(xs, rel: float) -> float:
    """eb = (float)(largest value range) * rel, both FP32."""
    rng = torch.tensor(value_range(xs), dtype=torch.float32)
    return float((rng * torch.tensor(rel, dtype=torch.float32)).item())
