"""Quality metrics of a decompressed field, as the reference's driver prints
them (examples/testfloat_homocompress_fastmode1_arg.c:48-113,
``error_evaluation``), computed on the GPU.

The arithmetic follows the reference's types: the pointwise error
``err = fabs(dec - ori)`` is an FP32 difference, ``err * err`` an FP32
product accumulated in FP64, the means, covariance and variances FP64, and
``range = Max - Min`` an FP32 difference.  The pointwise relative error is
``err / ori`` where ``|ori| > eb`` (signed: a negative value never raises
the maximum, as in the reference) and ``err`` where ``0 < |ori| <= eb``.
Sums run in a different order than the reference's sequential loop, so the
FP64 figures may differ in their last bits.
"""
from __future__ import annotations

import math
from typing import Dict

import torch


def error_evaluation(ori: torch.Tensor, dec: torch.Tensor, cmp_size: int, eb: float) -> Dict[str, float]:
    """Metrics of ``dec`` against ``ori`` (FP32 tensors of one size) for a
    stream of ``cmp_size`` bytes: min, max, range, max absolute error, max
    relative error (max abs / range), max pointwise relative error, PSNR,
    NRMSE, acEff (correlation) and compression ratio."""
    if ori.shape != dec.shape or ori.dtype != torch.float32 or dec.dtype != torch.float32:
        raise ValueError("ori and dec must be FP32 tensors of one shape")
    n = ori.numel()
    o = ori.reshape(-1)
    d = dec.reshape(-1)
    mx = float(o.max().item())
    mn = float(o.min().item())
    rng32 = float((torch.tensor(mx, dtype=torch.float32) - torch.tensor(mn, dtype=torch.float32)).item())
    err = (d - o).abs()  # FP32
    diff_max = float(err.max().item())
    eb32 = float(torch.tensor(eb, dtype=torch.float32).item())
    o64 = o.double()
    e64 = err.double()
    rel = torch.where(o64.abs() > eb32, e64 / o64, e64)
    rel = torch.where(o64 != 0, rel, torch.zeros_like(rel))
    maxpw = max(float(rel.max().item()), 0.0)
    mean1 = float(o64.sum().item()) / n
    d64 = d.double()
    mean2 = float(d64.sum().item()) / n
    prod = float(((o64 - mean1) * (d64 - mean2)).sum().item())
    s3 = float(((o64 - mean1) ** 2).sum().item())
    s4 = float(((d64 - mean2) ** 2).sum().item())
    sq = float((err * err).double().sum().item())  # FP32 products, FP64 sum
    std1 = math.sqrt(s3 / n)
    std2 = math.sqrt(s4 / n)
    ac_eff = (prod / n) / std1 / std2 if std1 > 0 and std2 > 0 else float("nan")
    mse = sq / n
    psnr = 20 * math.log10(rng32) - 10 * math.log10(mse) if mse > 0 and rng32 > 0 else float("inf")
    nrmse = math.sqrt(mse) / rng32 if rng32 > 0 else float("nan")
    max_rel = float((torch.tensor(diff_max, dtype=torch.float32) / torch.tensor(rng32, dtype=torch.float32)).item()) \
        if rng32 > 0 else float("nan")
    return {
        "min": mn,
        "max": mx,
        "range": rng32,
        "max_abs_err": diff_max,
        "max_rel_err": max_rel,
        "max_pw_rel_err": maxpw,
        "psnr": psnr,
        "nrmse": nrmse,
        "ac_eff": ac_eff,
        "compression_ratio": n * 4 / cmp_size if cmp_size else float("nan"),
    }
