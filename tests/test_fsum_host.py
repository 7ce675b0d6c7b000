"""The exact float-sum arithmetic of the HIP library, compiled for the CPU and checked bit for bit.

arrowmetal_amd/csrc/fsum.hpp holds the integer arithmetic of the correctly rounded grouped sum and mean
(the grid, the floor onto it, the interval test and the rounding). It is plain C++ with __host__ /
__device__ attributes, so tests/host/fsum_host_check.cpp builds it with g++ and drives the same steps
the kernels run over random groups (wide exponent ranges, subnormals, cancellations, ties, large
groups), comparing every sum and mean with the oracle's 2,304-bit exact accumulator (float64) and with
a plain integer rounding (float32 results). Planted mistakes in the rounding (tie rule, floor of
negative values, sticky bit of the division, the grid size, the exponent field) each make it fail.
"""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.mark.skipif(shutil.which("g++") is None or shutil.which("gcc") is None, reason="needs gcc and g++")
def test_fsum_arithmetic_matches_oracle(tmp_path):
    obj = str(tmp_path / "oracle.o")
    exe = str(tmp_path / "fsum_host_check")
    subprocess.run(["gcc", "-O2", "-std=gnu11", "-c", os.path.join(ROOT, "oracle", "oracle.c"), "-o", obj], check=True)
    subprocess.run(["g++", "-O2", "-std=c++17", "-D__host__=", "-D__device__=",
                    os.path.join(ROOT, "tests", "host", "fsum_host_check.cpp"), obj, "-o", exe], check=True)
    for seed in (1, 7):
        r = subprocess.run([exe, str(seed), "150000"], capture_output=True, text=True, timeout=120)
        print(r.stdout)
        assert r.returncode == 0, r.stdout + r.stderr
        assert "mismatches 0" in r.stdout
