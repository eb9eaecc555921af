"""The drop-in entry points driven by a gcc-compiled C program
(tests/dropin_driver.c, built by tests/Makefile), not through ctypes: its
compressed streams, homomorphic sum and decompressed sum equal the oracle's
on the same inputs."""
import os
import subprocess

import numpy as np
import pytest

torch = pytest.importorskip("torch")

from oracle import oracle as O  # noqa: E402

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "tests", "_bin", "dropin_driver")


@pytest.mark.parametrize("n,L,T", [(1 << 20, 36, 16), (2_400_017, 36, 16), (100_003, 32, 8)])
def test_c_program_drives_dropins(tmp_path, n, L, T):
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    assert os.path.exists(EXE), "tests/_bin/dropin_driver is missing: run __graft_entry__.build()"
    print("[gpu] C driver", n, L, T, flush=True)
    pre = str(tmp_path / "run")
    out = subprocess.run([EXE, str(n), str(L), str(T), pre], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stdout + out.stderr
    eb, sa, sb, ss, worst = open(pre + ".meta").read().split()
    eb = float(eb)
    a = np.fromfile(pre + ".a.f32", dtype=np.float32)
    b = np.fromfile(pre + ".b.f32", dtype=np.float32)
    ca = np.fromfile(pre + ".ca", dtype=np.uint8)
    cb = np.fromfile(pre + ".cb", dtype=np.uint8)
    cs = np.fromfile(pre + ".cs", dtype=np.uint8)
    ds = np.fromfile(pre + ".ds.f32", dtype=np.float32)
    ebf = float(np.float32(eb))
    wa, wb = O.compress(a, ebf, L, T), O.compress(b, ebf, L, T)
    assert ca.size == int(sa) and np.array_equal(ca, wa)
    assert cb.size == int(sb) and np.array_equal(cb, wb)
    ws = O.homsum([wa, wb], n, L, T)
    assert cs.size == int(ss) and np.array_equal(cs, ws)
    assert np.array_equal(ds.view(np.uint32), O.decompress(ws, n, ebf, L, T).view(np.uint32))
