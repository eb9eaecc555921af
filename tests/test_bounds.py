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


def test_slot_layout_block_encoder_matches_oracle(tmp_path):
    """The one-writer-per-byte block encoder of the slot-layout sum
    (zccl_internal.h emit_unit), run on the CPU unit by unit, equals the
    oracle's block encoder for every width and block length and writes every
    byte of the block and none outside it (tests/slot_encode_check.cpp)."""
    exe = str(tmp_path / "slot_encode_check")
    obj = str(tmp_path / "zo.o")
    subprocess.run(["gcc", "-O2", "-std=c99", "-ffp-contract=off", "-I", os.path.join(ROOT, "oracle"), "-c",
                    os.path.join(ROOT, "oracle", "zccl_oracle.c"), "-o", obj], check=True)
    subprocess.run(["g++", "-O2", "-std=c++17", "-I", os.path.join(ROOT, "zccl_amd", "csrc"),
                    "-I", os.path.join(ROOT, "oracle"), os.path.join(ROOT, "tests", "slot_encode_check.cpp"),
                    obj, "-lm", "-o", exe], check=True)
    out = subprocess.run([exe], capture_output=True, text=True, timeout=600)
    assert out.returncode == 0, out.stdout + out.stderr
    assert out.stdout.startswith("ok ")


def test_tiled_3d_work_split_covers_every_block_once(tmp_path):
    """The tiled 3-D kernels' chunks (compress) and block columns (decoder),
    replayed on the CPU through the same zccl_sz3d.h functions, cover every
    block exactly once at its stream index, with tile and row indices inside
    their LDS arrays (tests/sz3d_tiles_check.cpp)."""
    exe = str(tmp_path / "sz3d_tiles_check")
    subprocess.run(["g++", "-O2", "-std=c++17", "-I", os.path.join(ROOT, "zccl_amd", "csrc"),
                    os.path.join(ROOT, "tests", "sz3d_tiles_check.cpp"), "-o", exe], check=True)
    out = subprocess.run([exe], capture_output=True, text=True, timeout=600)
    assert out.returncode == 0, out.stdout + out.stderr
    assert out.stdout.startswith("ok ")
