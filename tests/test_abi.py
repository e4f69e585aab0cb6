"""The C-ABI library: builds, loads, exports what include/tcx.h declares, and
its host-only pieces behave (no GPU calls here)."""
import ctypes
import os
import re

import numpy as np
import pytest

from conftest import ROOT
from tc_amd import tcx
from oracle import oracle as O
from tc_amd import system as S


def declared_symbols(header="tcx.h", prefix="tcx_"):
    text = open(os.path.join(ROOT, "include", header)).read()
    return sorted(set(re.findall(r"\b(" + prefix + r"[a-z0-9_]+)\s*\(", text)))


def test_header_and_binding_agree():
    assert declared_symbols() == sorted(tcx.EXPORTS)
    assert declared_symbols("tcx_host.h", "tcxh_") == sorted(tcx.HOST_EXPORTS)


def test_every_header_is_checked():
    assert sorted(os.listdir(os.path.join(ROOT, "include"))) == ["tcx.h", "tcx_host.h"]


def test_library_exports_every_declared_symbol():
    L = tcx.load_library()
    for name in declared_symbols() + declared_symbols("tcx_host.h", "tcxh_"):
        assert hasattr(L, name), name


def test_host_abi_errors():
    """The host entry points report bad input through a status and tcxh_last_error."""
    L = tcx.load_library()
    h = ctypes.c_void_p()
    st = L.tcxh_qe_load(b"/nonexistent/in.tcxqe", ctypes.byref(h))
    assert st == tcx.TCX_ERR_ARG and not h.value
    assert b"cannot open" in L.tcxh_last_error()
    assert L.tcxh_qe_num_replicas(None, None) == tcx.TCX_ERR_ARG


def test_host_abi_rejects_truncated_and_malformed_inputs(tmp_path):
    """A flat input cut anywhere, or with a bad magic or record tag, is TCX_ERR_ARG; the whole
    file loads and expands (no GPU calls)."""
    from tc_amd import qe as Q
    qe = Q.load_fixture(os.path.join(ROOT, "tests", "golden", "qe_example14.npz"))
    path = tmp_path / "in.tcxqe"
    Q.write_flat(qe, str(path))
    data = path.read_bytes()
    L = tcx.load_library()
    cuts = [4, 8 + 20, 8 + 48 + 3, len(data) // 2, len(data) - 1]
    bad = [b"TCXQE002" + data[8:], data[:8 + 32] + b"<x9\0\0\0\0\0" + data[8 + 40:]]
    for blob in [data[:c] for c in cuts] + bad:
        path.write_bytes(blob)
        h = ctypes.c_void_p()
        assert L.tcxh_qe_load(str(path).encode(), ctypes.byref(h)) == tcx.TCX_ERR_ARG
        assert not h.value
    path.write_bytes(data)
    with Q.HostInputs(qe) as h:
        assert len(h.replicas()) > qe.nirr
    # a rotation that is not unimodular is bad input too
    import dataclasses
    rot = qe.rot_qe.copy()
    rot[1] = 2 * np.eye(3, dtype=rot.dtype)
    with pytest.raises(tcx.TcxError) as e:
        Q.HostInputs(dataclasses.replace(qe, rot_qe=rot))
    assert e.value.code == tcx.TCX_ERR_ARG and "unimodular" in str(e.value)


def test_library_is_gfx950_only():
    data = open(tcx.LIB_PATH, "rb").read()
    assert b"gfx950" in data


def test_sum_of_vaux_host_utility_matches_oracle():
    for sys_ in (S.tiny_system(), S.heg_sc_rs4_24()):
        a1, s1 = tcx.sum_of_vaux(sys_.recip, sys_.volume, sys_.kfrac)
        a2, s2 = O.sum_of_vaux(sys_.recip, sys_.volume, sys_.kfrac)
        assert abs(a1 - a2) < 1e-15 * a2
        assert np.abs(s1 - s2).max() < 1e-12 * np.abs(s2).max()


def test_sum_of_vaux_rejects_bad_arguments():
    L = tcx.load_library()
    r = np.eye(3).reshape(9)
    kf = np.zeros(3)
    out = np.zeros(1)
    alpha = np.zeros(1)
    dp = ctypes.POINTER(ctypes.c_double)
    st = L.tcx_sum_of_vaux(r.ctypes.data_as(dp), -1.0, 1, kf.ctypes.data_as(dp), alpha.ctypes.data_as(dp),
                           out.ctypes.data_as(dp))
    assert st == tcx.TCX_ERR_ARG


def test_null_context_is_an_error_not_a_crash():
    L = tcx.load_library()
    assert L.tcx_destroy(None) == tcx.TCX_ERR_ARG
    assert L.tcx_synchronize(None) == tcx.TCX_ERR_ARG
    assert L.tcx_last_error(None) == b"null context"


def test_host_shim_exports_reference_names():
    """The C++ mirror keeps the reference's calc_hamiltonian:: names (mangled C++ symbols)."""
    import subprocess
    out = subprocess.run(["nm", "-DC", tcx.LIB_PATH], capture_output=True, text=True, check=True).stdout
    for name in ("calc_hamiltonian::hf2x(", "calc_hamiltonian::tc2x(", "calc_hamiltonian::tc3a1(",
                 "calc_hamiltonian::tc3b1("):
        assert name in out, name
