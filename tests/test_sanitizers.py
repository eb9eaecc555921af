"""The oracle and the host-side C/C++ of the library built with
AddressSanitizer and UndefinedBehaviorSanitizer (gcc/g++; no GPU code is
built or run): the host stream walk of the drop-in layer on exact-size
buffers (tests/host_walk_check.cpp), the kernels' index arithmetic replayed
on the host (tests/bounds_check.cpp), the slot-layout block encoder
(tests/slot_encode_check.cpp) and the oracle's whole-stream functions
(tests/oracle_roundtrip_check.c).  Any sanitizer report fails the run."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "zccl_amd", "csrc")
ORACLE = os.path.join(ROOT, "oracle")
SAN = ["-O1", "-g", "-fno-omit-frame-pointer", "-fsanitize=address,undefined", "-fno-sanitize-recover=all"]
ENV = dict(os.environ, ASAN_OPTIONS="detect_leaks=0:abort_on_error=0", UBSAN_OPTIONS="print_stacktrace=1")


def _oracle_obj(tmp_path):
    obj = str(tmp_path / "zo_san.o")
    subprocess.run(["gcc", *SAN, "-std=c99", "-ffp-contract=off", "-I", ORACLE, "-c",
                    os.path.join(ORACLE, "zccl_oracle.c"), "-o", obj], check=True)
    return obj


def _run(exe):
    out = subprocess.run([exe], capture_output=True, text=True, timeout=900, env=ENV)
    assert out.returncode == 0, out.stdout[-2000:] + out.stderr[-4000:]
    assert out.stdout.startswith("ok "), out.stdout[-2000:]
    assert "runtime error" not in out.stderr and "AddressSanitizer" not in out.stderr, out.stderr[-4000:]


@pytest.mark.parametrize("src", ["host_walk_check.cpp", "slot_encode_check.cpp"])
def test_host_code_with_oracle_sanitized(tmp_path, src):
    exe = str(tmp_path / src.split(".")[0])
    subprocess.run(["g++", *SAN, "-std=c++17", "-I", CSRC, "-I", ORACLE, os.path.join(ROOT, "tests", src),
                    _oracle_obj(tmp_path), "-lm", "-o", exe], check=True)
    _run(exe)


def test_kernel_index_arithmetic_sanitized(tmp_path):
    exe = str(tmp_path / "bounds_check")
    subprocess.run(["g++", *SAN, "-std=c++17", "-I", CSRC, os.path.join(ROOT, "tests", "bounds_check.cpp"),
                    "-o", exe], check=True)
    _run(exe)


def test_oracle_round_trips_sanitized(tmp_path):
    exe = str(tmp_path / "oracle_roundtrip_check")
    subprocess.run(["gcc", *SAN, "-std=c99", "-I", ORACLE, os.path.join(ROOT, "tests", "oracle_roundtrip_check.c"),
                    _oracle_obj(tmp_path), "-lm", "-o", exe], check=True)
    _run(exe)
