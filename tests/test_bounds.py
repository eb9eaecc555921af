"""Host replay of the tile kernels' index arithmetic (tests/bounds_check.cpp):
every workspace, index, element and LDS offset the kernels compute stays inside
the buffer the host allocates, over a grid of stream geometries."""
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_kernel_index_arithmetic_stays_in_bounds(tmp_path):
    exe = str(tmp_path / "bounds_check")
    subprocess.run(["g++", "-O2", "-std=c++17", "-I", os.path.join(ROOT, "zccl_amd", "csrc"),
                    os.path.join(ROOT, "tests", "bounds_check.cpp"), "-o", exe], check=True)
    out = subprocess.run([exe], capture_output=True, text=True, timeout=600)
    assert out.returncode == 0, out.stdout + out.stderr
    assert out.stdout.startswith("ok ")
