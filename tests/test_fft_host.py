"""The register FFT templates (tc_amd/csrc/fft_reg.hpp) on the host against a direct
long-double DFT, every length the kernels instantiate (grid sides and split-pencil
sub-lengths), both signs, FP64 and FP32.  The binary is built by tc_amd/build.py."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "tests", "_bin", "fft_reg_test")


def test_register_fft_matches_direct_dft():
    if not os.path.exists(BIN):
        pytest.skip("fft_reg_test not built (run tc_amd/build.py)")
    res = subprocess.run([BIN], capture_output=True, text=True, timeout=120)
    print(res.stdout)
    assert res.returncode == 0, res.stdout + res.stderr
    assert res.stdout.count(" ok") == 17
