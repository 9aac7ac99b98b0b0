"""The FFT building blocks of the PME transforms (fft_core.hpp: radix 2, 3, 4,
5, 8 butterflies and Stockham stages, run here through their host path with
the device kernels' compile-time plans) against a naive DFT, for every grid
length with a specialized kernel and both directions.  CPU only; needs hipcc."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")


@pytest.mark.skipif(not os.path.exists(HIPCC) and shutil.which("hipcc") is None, reason="hipcc not available")
def test_fft_plans_against_naive_dft(tmp_path):
    exe = tmp_path / "fft_host_check"
    inc = os.path.join(os.path.dirname(HERE), "tinker-gpu_amd", "csrc")
    cc = HIPCC if os.path.exists(HIPCC) else "hipcc"
    subprocess.run([cc, "-std=c++17", "-O2", "-I", inc, os.path.join(HERE, "fft_host_check.cpp"), "-o", str(exe)],
                   check=True, capture_output=True, timeout=300)
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
