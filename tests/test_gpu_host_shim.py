"""The C++ host mirror calc_hamiltonian::* (tc_amd/host/) driven by a C++ program,
checked against the plane-wave closed form of the exchange operator."""
import os
import subprocess

import pytest

from conftest import ROOT

pytestmark = pytest.mark.gpu

BIN = os.path.join(ROOT, "tests", "_bin", "host_shim_test")


def test_host_shim_closed_form():
    if not os.path.exists(BIN):
        pytest.fail(f"{BIN} is not built (run __graft_entry__.build())")
    res = subprocess.run([BIN], capture_output=True, text=True, timeout=120)
    print(res.stdout, res.stderr)
    assert res.returncode == 0, res.stdout + res.stderr
    assert "OK" in res.stdout
