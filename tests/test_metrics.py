"""zccl_amd.metrics.error_evaluation against a line-by-line Python
restatement of the reference's loop
(examples/testfloat_homocompress_fastmode1_arg.c:48-113) on small arrays,
including its quirks: FP32 pointwise errors, a signed pointwise relative
error, the noise floor at eb."""
import math

import numpy as np
import pytest

torch = pytest.importorskip("torch")

from zccl_amd import metrics  # noqa: E402


def reference_loop(ori, dec, cmp_size, eb):
    f = np.float32
    n = ori.size
    mx = ori[0]
    mn = ori[0]
    diff_max = abs(f(dec[0] - ori[0]))
    sum1 = sum(float(v) for v in ori)
    sum2 = sum(float(v) for v in dec)
    mean1, mean2 = sum1 / n, sum2 / n
    sum3 = sum4 = s = prod = 0.0
    maxpw = 0.0
    for i in range(n):
        mx = max(mx, ori[i])
        mn = min(mn, ori[i])
        err = abs(f(dec[i] - ori[i]))
        if ori[i] != 0:
            rel = float(err) / float(ori[i]) if abs(ori[i]) > f(eb) else float(err)
            maxpw = max(maxpw, rel)
        diff_max = max(diff_max, err)
        prod += (float(ori[i]) - mean1) * (float(dec[i]) - mean2)
        sum3 += (float(ori[i]) - mean1) ** 2
        sum4 += (float(dec[i]) - mean2) ** 2
        s += float(f(err * err))
    std1, std2 = math.sqrt(sum3 / n), math.sqrt(sum4 / n)
    mse = s / n
    rng = float(f(mx - mn))
    return {
        "max_abs_err": float(diff_max), "max_rel_err": float(f(diff_max / f(mx - mn))),
        "max_pw_rel_err": maxpw, "psnr": 20 * math.log10(rng) - 10 * math.log10(mse),
        "nrmse": math.sqrt(mse) / rng, "ac_eff": (prod / n) / std1 / std2, "range": rng,
        "compression_ratio": n * 4 / cmp_size,
    }


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_error_evaluation_matches_reference_loop(seed):
    rng = np.random.default_rng(seed)
    n, eb = 3000, 1e-2
    ori = (rng.normal(0, 1, n) * (10 ** rng.uniform(-3, 1, n))).astype(np.float32)
    ori[::97] = 0.0
    dec = (ori + rng.uniform(-eb, eb, n)).astype(np.float32)
    got = metrics.error_evaluation(torch.from_numpy(ori), torch.from_numpy(dec), 4321, eb)
    want = reference_loop(ori, dec, 4321, eb)
    for k, v in want.items():
        assert got[k] == pytest.approx(v, rel=1e-9, abs=1e-12), k
    # the signed quirk: only positive values raise the pointwise maximum
    neg = -np.abs(ori) - np.float32(1.0)
    got2 = metrics.error_evaluation(torch.from_numpy(neg), torch.from_numpy(neg + np.float32(0.5)), 100, eb)
    assert got2["max_pw_rel_err"] == 0.0
