"""Write golden vectors for the 3-D SZp codec from the compiled reference.

Run where the reference sources exist (``make -C oracle ref`` builds
``oracle/_ref/libzccl_ref.so`` from them):

    python tests/golden/make_fixtures3d.py

Output (committed): ``fixtures_3d.npz`` -- for every case the FP32 input
``<case>/x``, the reference stream ``<case>/c`` (SZp_compress3D_fast_openmp with
an OpenMP team of T) and the decompressed ``<case>/d``
(SZp_decompress3D_fast_openmp); parameters in ``fixtures_3d.json``.  Inputs
keep every sub-volume's payload inside the reference's per-thread buffer of
ceil(4n/T) bytes (oracle.ref_compress3d refuses otherwise).
"""
from __future__ import annotations

import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle import oracle as O  # noqa: E402
from tests.golden import gen  # noqa: E402


def cases():
    # (name, kind, dims, Bs, eb, T)
    return [
        ("smooth_16c_B4_T1", "smooth", (16, 16, 16), 4, 1e-2, 1),
        ("smooth_16c_B4_T8", "smooth", (16, 16, 16), 4, 1e-3, 8),
        ("ragged_13x11x17_B4_T6", "smooth", (13, 11, 17), 4, 1e-2, 6),
        ("ragged_13x11x17_B5_T3", "smooth", (13, 11, 17), 5, 1e-3, 3),
        ("ragged_19x9x23_B8_T2", "smooth", (19, 9, 23), 8, 1e-2, 2),
        ("block1_6x7x5_T4", "smooth", (6, 7, 5), 1, 1e-2, 4),
        ("block3_9x8x10_T12", "smooth", (9, 8, 10), 3, 1e-2, 12),
        ("walk_10x12x14_B4_T16", "walk", (10, 12, 14), 4, 1e-2, 16),
        ("zeros_12c_B4_T8", "zeros", (12, 12, 12), 4, 1e-3, 8),
        ("slabs_8x16x16_B4_T4", "slabs", (8, 16, 16), 4, 1e-3, 4),
        ("outliers_12x10x16_B4_T2", "outliers", (12, 10, 16), 4, 1e-3, 2),
        ("wide_8x8x9_B4_T1", "wide", (8, 8, 9), 4, 1.0, 1),
        ("thin_1x1x40_B4_T1", "walk", (1, 1, 40), 4, 1e-2, 1),
        ("thin_40x1x1_B4_T2", "walk", (40, 1, 1), 4, 1e-2, 2),
    ]


def make_input(kind, dims, seed):
    n = int(np.prod(dims))
    if kind == "smooth":
        return gen.smooth3d(dims, seed).reshape(-1)
    return gen.make(kind, n, seed)


def main():
    if not O.ref_available():
        sys.exit("compiled reference not available: make -C oracle ref")
    arrays, meta = {}, []
    for i, (name, kind, dims, Bs, eb, T) in enumerate(cases()):
        x = make_input(kind, dims, 3000 + i)
        c = O.ref_compress3d(x, dims, Bs, eb, T)
        d = O.ref_decompress3d(c, dims, Bs, eb, T)
        arrays[f"{name}/x"] = x
        arrays[f"{name}/c"] = c
        arrays[f"{name}/d"] = d
        meta.append({"name": name, "kind": kind, "dims": list(dims), "Bs": Bs, "eb": eb, "T": T,
                     "seed": 3000 + i, "bytes": int(c.size)})
        print(name, c.size)
    np.savez_compressed(os.path.join(HERE, "fixtures_3d.npz"), **arrays)
    with open(os.path.join(HERE, "fixtures_3d.json"), "w") as f:
        json.dump({"source": "SZp_compress3D_fast_openmp / SZp_decompress3D_fast_openmp, compiled reference",
                   "cases": meta}, f, indent=1)


if __name__ == "__main__":
    main()
