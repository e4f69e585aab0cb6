"""ctypes binding of libtcx.so (the C ABI declared in include/tcx.h).

The term methods keep the reference's names (calc_hamiltonian::hf2x, tc2x,
tc3a1, tc3b1; src/include/calc_hamiltonian.hpp:54-108) and semantics: they
ACCUMULATE into the output array and raise on any non-zero status instead of
aborting (the reference's error_messages::stop -> MPI_Abort).

There is no fallback: if libtcx.so is missing or no GPU is visible, the
constructor raises.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libtcx.so")

TCX_OK = 0
TCX_ERR_ARG = 1
TCX_ERR_STATE = 2
TCX_ERR_HIP = 3
TCX_ERR_UNSUPPORTED = 4
TCX_ERR_NOMEM = 5
TCX_ERR_INTERNAL = 6

PREC_FP64 = 0
PREC_FP32 = 1
FLAG_DEVICE_PTRS = 1

METHODS = {"HF": 0, "TC": 1, "BITC": 2}

# Every entry point declared in include/tcx.h.
EXPORTS = [
    "tcx_create", "tcx_destroy", "tcx_last_error", "tcx_synchronize",
    "tcx_set_cell", "tcx_set_kpoints", "tcx_set_basis", "tcx_set_states",
    "tcx_set_jastrow_rpa", "tcx_set_jastrow_polynomial", "tcx_get_jastrow_num_rmesh", "tcx_jastrow_kernels",
    "tcx_set_div_correction", "tcx_sum_of_vaux", "tcx_set_kpoint_symmetry",
    "tcx_sum_of_vaux_weighted", "tcx_set_zero_weight_kpoints",
    "tcx_get_density", "tcx_set_density",
    "tcx_hf2x", "tcx_hf2x_noalloc", "tcx_tc2x", "tcx_tc3a1", "tcx_tc3b1", "tcx_tc3a3b3b4b6", "tcx_tc3a2a4b2b5",
    "tcx_kinetic", "tcx_hf2h", "tcx_tc2h", "tcx_apply_all",
    "tcx_davidson", "tcx_total_energy", "tcx_eig_dense", "tcx_set_filling",
    "tcx_graph_begin", "tcx_graph_end", "tcx_graph_launch", "tcx_graph_destroy",
    "tcx_grad_exchange", "tcx_grad_tc2", "tcx_kernel_gradient",
    "tcx_device_alloc", "tcx_device_free", "tcx_memcpy", "tcx_memset_zero",
    "tcx_set_launch_shape", "tcx_set_pipeline_order", "tcx_get_record_columns", "tcx_enable_timing",
    "tcx_get_timing",
]

# Every entry point declared in include/tcx_host.h (the C++ host's k-mesh expansion and SCF loop).
HOST_EXPORTS = [
    "tcxh_last_error", "tcxh_qe_load", "tcxh_qe_free", "tcxh_qe_num_replicas", "tcxh_qe_replicas",
    "tcxh_qe_rotation", "tcxh_qe_replica_npw", "tcxh_qe_replica_basis", "tcxh_scf_run", "tcxh_scf_result_sizes",
    "tcxh_scf_result_fill", "tcxh_scf_result_energy", "tcxh_scf_result_ddiff", "tcxh_scf_result_free",
]


class TcxError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"tcx error {code}: {msg}")
        self.code = code


_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)
_lib = None


def load_library():
    """Load libtcx.so; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is missing: build it with `python -m tc_amd.build` "
                           "(or __graft_entry__.build()); there is no CPU fallback")
    L = C.CDLL(LIB_PATH)
    vp = C.c_void_p
    L.tcx_create.argtypes = [C.POINTER(vp), C.c_int]
    L.tcx_destroy.argtypes = [vp]
    L.tcx_last_error.argtypes = [vp]
    L.tcx_last_error.restype = C.c_char_p
    L.tcx_synchronize.argtypes = [vp]
    L.tcx_set_cell.argtypes = [vp, _dp, C.c_double, _ip, C.c_int, C.c_int]
    L.tcx_set_kpoints.argtypes = [vp, C.c_int, _dp, C.c_int]
    L.tcx_set_basis.argtypes = [vp, _ip, _ip]
    L.tcx_set_states.argtypes = [vp, _ip, _dp, _dp, _dp]
    L.tcx_set_jastrow_rpa.argtypes = [vp, _dp, _dp]
    L.tcx_set_jastrow_polynomial.argtypes = [vp, C.c_int, C.c_int, _dp, _dp, C.c_int]
    L.tcx_get_jastrow_num_rmesh.argtypes = [vp, _ip]
    L.tcx_jastrow_kernels.argtypes = [vp, C.c_int, C.c_int, C.c_longlong, vp, vp, vp]
    L.tcx_set_div_correction.argtypes = [vp, C.c_int, C.c_double, _dp]
    L.tcx_sum_of_vaux.argtypes = [_dp, C.c_double, C.c_int, _dp, _dp, _dp]
    L.tcx_sum_of_vaux_weighted.argtypes = [_dp, C.c_double, C.c_int, _dp, _dp, _dp, _dp, _dp]
    L.tcx_set_zero_weight_kpoints.argtypes = [vp, _dp, _dp, _ip, _dp, vp, vp]
    L.tcx_set_kpoint_symmetry.argtypes = [vp, _ip, _ip, _dp, _ip]
    L.tcx_get_density.argtypes = [vp, vp]
    L.tcx_set_density.argtypes = [vp, vp]
    for name in ("tcx_hf2x", "tcx_tc2x", "tcx_tc3a1", "tcx_tc3b1", "tcx_tc3a3b3b4b6", "tcx_tc3a2a4b2b5"):
        getattr(L, name).argtypes = [vp, _ip, vp, vp, C.c_int, C.c_uint]
    L.tcx_hf2x_noalloc.argtypes = [vp, _ip, vp, vp, C.c_int]
    for name in ("tcx_kinetic", "tcx_hf2h", "tcx_tc2h"):
        getattr(L, name).argtypes = [vp, _ip, vp, vp, C.c_uint]
    L.tcx_apply_all.argtypes = [vp, _ip, vp, vp, vp, vp, C.c_int, C.c_uint]
    L.tcx_davidson.argtypes = [vp, _ip, vp, vp, vp, vp, C.c_int, C.c_int, C.c_int, C.c_int]
    L.tcx_total_energy.argtypes = [C.c_int, _ip, _dp, vp, C.c_int, C.c_int, vp]
    L.tcx_eig_dense.argtypes = [C.c_int, vp, vp, vp]
    L.tcx_grad_exchange.argtypes = [vp, vp, _dp, C.c_int]
    L.tcx_grad_tc2.argtypes = [vp, vp, vp, _dp, _dp, C.c_int]
    L.tcx_kernel_gradient.argtypes = [vp, C.c_int, C.c_int, C.c_double, vp, vp, _dp, C.c_int]
    L.tcx_graph_begin.argtypes = [vp]
    L.tcx_graph_end.argtypes = [vp, _ip]
    L.tcx_graph_launch.argtypes = [vp, C.c_int]
    L.tcx_graph_destroy.argtypes = [vp, C.c_int]
    L.tcx_set_filling.argtypes = [C.c_int, _ip, _dp, _dp, C.c_double, C.c_int, C.c_double, _dp, _dp, _ip]
    L.tcx_device_alloc.argtypes = [vp, C.c_ulonglong, C.POINTER(vp)]
    L.tcx_device_free.argtypes = [vp, vp]
    L.tcx_memcpy.argtypes = [vp, vp, vp, C.c_ulonglong, C.c_int]
    L.tcx_memset_zero.argtypes = [vp, vp, C.c_ulonglong]
    L.tcx_set_launch_shape.argtypes = [vp, C.c_int, C.c_long, C.c_long]
    L.tcx_set_pipeline_order.argtypes = [vp, C.c_int, C.c_int]
    L.tcx_get_record_columns.argtypes = [vp, C.c_int, _dp]
    L.tcx_enable_timing.argtypes = [vp, C.c_int]
    L.tcx_get_timing.argtypes = [vp, _dp, C.POINTER(C.c_long), C.POINTER(C.c_long)]
    L.tcxh_last_error.argtypes = []
    L.tcxh_last_error.restype = C.c_char_p
    L.tcxh_qe_load.argtypes = [C.c_char_p, C.POINTER(vp)]
    L.tcxh_qe_free.argtypes = [vp]
    L.tcxh_qe_free.restype = None
    L.tcxh_qe_num_replicas.argtypes = [vp, _ip]
    L.tcxh_qe_replicas.argtypes = [vp, _ip, _ip, _ip, _dp]
    L.tcxh_qe_rotation.argtypes = [vp, C.c_int, _ip]
    L.tcxh_qe_replica_npw.argtypes = [vp, C.c_int, _ip]
    L.tcxh_qe_replica_basis.argtypes = [vp, C.c_int, _ip, vp]
    L.tcxh_scf_run.argtypes = [vp, vp, C.c_int, C.c_int, C.POINTER(vp)]
    L.tcxh_scf_result_sizes.argtypes = [vp] + [_ip] * 7
    L.tcxh_scf_result_fill.argtypes = [vp, C.c_int, _dp, vp, _dp]
    L.tcxh_scf_result_energy.argtypes = [vp, C.c_int, vp]
    L.tcxh_scf_result_ddiff.argtypes = [vp, _dp]
    L.tcxh_scf_result_free.argtypes = [vp]
    L.tcxh_scf_result_free.restype = None
    _lib = L
    return L


def host_check(st, what):
    """Raise TcxError for a failed tcxh_ call, with the host's message."""
    if st != TCX_OK:
        msg = load_library().tcxh_last_error()
        raise TcxError(st, f"{what}: {msg.decode() if msg else ''}")


def _ptr(a, t=_dp):
    return a.ctypes.data_as(t)


def eig_dense(A):
    """Eigenvalues and unit-norm right eigenvectors (columns) of a small complex matrix."""
    A = np.ascontiguousarray(A, dtype=np.complex128)
    n = A.shape[0]
    w = np.zeros(n, np.complex128)
    X = np.zeros((n, n), np.complex128)
    st = load_library().tcx_eig_dense(n, A.ctypes.data, w.ctypes.data, X.ctypes.data)
    if st != TCX_OK:
        raise TcxError(st, "eig_dense failed")
    return w, X


def set_filling(nb, eig, kweight, num_electrons, smearing="fixed", width=0.01):
    """(filling, fermi_energy, num_occupied) as BlochStates::set_filling."""
    nb = np.ascontiguousarray(nb, dtype=np.int32)
    e = np.ascontiguousarray(np.real(eig), dtype=np.float64)
    kw = np.ascontiguousarray(kweight, dtype=np.float64)
    f = np.zeros(int(nb.sum()))
    mu = np.zeros(1)
    nocc = np.zeros(len(nb), np.int32)
    st = load_library().tcx_set_filling(len(nb), _ptr(nb, _ip), _ptr(e), _ptr(kw), float(num_electrons),
                                        {"fixed": 0, "gaussian": 1}[smearing], float(width), _ptr(f), _ptr(mu),
                                        _ptr(nocc, _ip))
    if st != TCX_OK:
        raise TcxError(st, "set_filling failed")
    return f, float(mu[0]), nocc


def total_energy(nb, filling, energies, num_kpoints, uses_3body):
    """(E_1, E_2, E_3, E_total) from per-band e_n (TotalEnergy::calc_total_energy)."""
    nb = np.ascontiguousarray(nb, dtype=np.int32)
    f = np.ascontiguousarray(filling, dtype=np.float64)
    e = np.ascontiguousarray(energies, dtype=np.complex128)
    out = np.zeros(4, np.complex128)
    st = load_library().tcx_total_energy(len(nb), _ptr(nb, _ip), _ptr(f), e.ctypes.data, int(num_kpoints),
                                         int(bool(uses_3body)), out.ctypes.data)
    if st != TCX_OK:
        raise TcxError(st, "total_energy failed")
    return out


def sum_of_vaux(recip, volume, kfrac):
    """alpha and sum_of_Vaux[k] (Potentials::set_sum_of_Vaux_each) from the library's host code."""
    L = load_library()
    r = np.ascontiguousarray(recip, dtype=np.float64).reshape(9)
    kf = np.ascontiguousarray(kfrac, dtype=np.float64).reshape(-1)
    nk = kf.size // 3
    out = np.zeros(nk)
    alpha = np.zeros(1)
    st = L.tcx_sum_of_vaux(_ptr(r), float(volume), nk, _ptr(kf), _ptr(alpha), _ptr(out))
    if st != TCX_OK:
        raise TcxError(st, "tcx_sum_of_vaux")
    return float(alpha[0]), out


def sum_of_vaux_weighted(recip, volume, kfrac, kweight):
    """(alpha, sum_of_Vaux[nk], sum_of_kVaux[nk, 3]) with the q-sums over the k of weight
    >= 1e-8 (Potentials::set_sum_of_Vaux_each, set_sum_of_kVaux_each), library host code."""
    L = load_library()
    r = np.ascontiguousarray(recip, dtype=np.float64).reshape(9)
    kf = np.ascontiguousarray(kfrac, dtype=np.float64).reshape(-1)
    kw = np.ascontiguousarray(kweight, dtype=np.float64)
    nk = kf.size // 3
    out, kout, alpha = np.zeros(nk), np.zeros(3 * nk), np.zeros(1)
    st = L.tcx_sum_of_vaux_weighted(_ptr(r), float(volume), nk, _ptr(kf), _ptr(kw), _ptr(alpha), _ptr(out),
                                    _ptr(kout))
    if st != TCX_OK:
        raise TcxError(st, "tcx_sum_of_vaux_weighted")
    return float(alpha[0]), out, kout.reshape(nk, 3)


class Engine:
    """One tcx_context on one GPU, loaded with a synthetic or caller-provided system."""

    def __init__(self, device: int = 0):
        self.L = load_library()
        h = C.c_void_p()
        st = self.L.tcx_create(C.byref(h), int(device))
        if st != TCX_OK:
            raise TcxError(st, "tcx_create failed (no GPU visible?)")
        self.h = h
        self._keep = []
        self.sys = None
        self.method = "HF"

    def close(self):
        if getattr(self, "h", None):
            self.L.tcx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, st, what):
        if st != TCX_OK:
            msg = self.L.tcx_last_error(self.h)
            raise TcxError(st, f"{what}: {msg.decode() if msg else ''}")

    # --------------------------------------------------------------- state
    def set_cell(self, recip, volume, grid, nspin=1, method="HF"):
        r = np.ascontiguousarray(recip, dtype=np.float64).reshape(9)
        g = np.ascontiguousarray(grid, dtype=np.int32)
        m = METHODS[method] if isinstance(method, str) else int(method)
        self.method = {v: k for k, v in METHODS.items()}[m]
        self._check(self.L.tcx_set_cell(self.h, _ptr(r), float(volume), _ptr(g, _ip), int(nspin), m), "set_cell")

    def set_kpoints(self, kfrac, num_kpoints=None):
        kf = np.ascontiguousarray(kfrac, dtype=np.float64).reshape(-1)
        nk = kf.size // 3
        self._check(self.L.tcx_set_kpoints(self.h, nk, _ptr(kf), int(num_kpoints or nk)), "set_kpoints")

    def set_basis(self, npw, gindex):
        a = np.ascontiguousarray(npw, dtype=np.int32)
        b = np.ascontiguousarray(gindex, dtype=np.int32)
        self._check(self.L.tcx_set_basis(self.h, _ptr(a, _ip), _ptr(b, _ip)), "set_basis")

    def set_states(self, nocc, filling, phi_occ, chi_occ=None):
        a = np.ascontiguousarray(nocc, dtype=np.int32)
        f = np.ascontiguousarray(filling, dtype=np.float64)
        p = np.ascontiguousarray(phi_occ, dtype=np.complex128)
        c = None if chi_occ is None else np.ascontiguousarray(chi_occ, dtype=np.complex128)
        self._check(self.L.tcx_set_states(self.h, _ptr(a, _ip), _ptr(f), p.view(np.float64).ctypes.data_as(_dp),
                                          None if c is None else c.view(np.float64).ctypes.data_as(_dp)),
                    "set_states")

    def set_jastrow_rpa(self, A, F):
        a = np.ascontiguousarray(A, dtype=np.float64).reshape(4)
        f = np.ascontiguousarray(F, dtype=np.float64).reshape(4)
        self._check(self.L.tcx_set_jastrow_rpa(self.h, _ptr(a), _ptr(f)), "set_jastrow_rpa")

    def set_jastrow_polynomial(self, deg, C=3, L=None, c=None, cusp=False):
        """Polynomial part of the Jastrow (tcx_set_jastrow_polynomial); deg 0 removes it.
        L: (2, 2), c: (2, 2, deg) c_poly_input (entry 0 of each pair is replaced by the
        cusp condition)."""
        deg = int(deg)
        Lv = np.ascontiguousarray(np.ones(4) if L is None else L, dtype=np.float64).reshape(4)
        cv = np.ascontiguousarray(np.zeros(4 * max(deg, 1)) if c is None else c, dtype=np.float64).reshape(-1)
        self._check(self.L.tcx_set_jastrow_polynomial(self.h, deg, int(C), _ptr(Lv), _ptr(cv) if deg else None,
                                                      int(bool(cusp))), "set_jastrow_polynomial")

    def jastrow_num_rmesh(self):
        n = C.c_int(-1)
        self._check(self.L.tcx_get_jastrow_num_rmesh(self.h, C.byref(n)), "get_jastrow_num_rmesh")
        return n.value

    def jastrow_kernels(self, G2, s1=0, s2=0):
        """(u(G), tc_2body(G)) at the |G|^2 values G2, evaluated on the GPU."""
        g2 = np.ascontiguousarray(G2, dtype=np.float64).reshape(-1)
        n = g2.size
        g = self.to_device(g2)
        u, v = self.device_buffer(max(8 * n, 8)), self.device_buffer(max(8 * n, 8))
        self._check(self.L.tcx_jastrow_kernels(self.h, int(s1), int(s2), n, C.c_void_p(g.ptr), C.c_void_p(u.ptr),
                                               C.c_void_p(v.ptr)), "jastrow_kernels")
        self.synchronize()
        return u.download(np.float64, n), v.download(np.float64, n)

    def set_div_correction(self, enabled, alpha=0.0, sum_vaux=None):
        s = None if sum_vaux is None else np.ascontiguousarray(sum_vaux, dtype=np.float64)
        self._check(self.L.tcx_set_div_correction(self.h, int(bool(enabled)), float(alpha),
                                                  None if s is None else _ptr(s)), "set_div_correction")

    def set_zero_weight_kpoints(self, kweight, sum_of_kvaux, nstate, filling, phi, chi=None):
        """Zero-weight k-points (BAND k-points) and the states their divergence terms
        project on (tcx_set_zero_weight_kpoints)."""
        kw = np.ascontiguousarray(kweight, dtype=np.float64)
        skv = np.ascontiguousarray(sum_of_kvaux, dtype=np.float64).reshape(-1)
        ns = np.ascontiguousarray(nstate, dtype=np.int32)
        f = np.ascontiguousarray(filling, dtype=np.float64)
        p = np.ascontiguousarray(phi, dtype=np.complex128)
        c = None if chi is None else np.ascontiguousarray(chi, dtype=np.complex128)
        self._check(self.L.tcx_set_zero_weight_kpoints(self.h, _ptr(kw), _ptr(skv), _ptr(ns, _ip), _ptr(f),
                                                       p.ctypes.data if p.size else None,
                                                       None if c is None or not c.size else c.ctypes.data),
                    "set_zero_weight_kpoints")

    def set_kpoint_symmetry(self, k_src=None, rot=None, trans=None, trev=None):
        """k-mesh symmetry (None: identity).  k_src (nk,), rot (nk, 3, 3) TC++ rotations,
        trans (nk, 3) fractional translations, trev (nk,) time-reversal flags."""
        if k_src is None:
            self._check(self.L.tcx_set_kpoint_symmetry(self.h, None, None, None, None), "set_kpoint_symmetry")
            return
        ks = np.ascontiguousarray(k_src, dtype=np.int32)
        r = np.ascontiguousarray(rot, dtype=np.int32).reshape(-1)
        t = np.ascontiguousarray(trans, dtype=np.float64).reshape(-1)
        tr = np.ascontiguousarray(trev, dtype=np.int32)
        self._check(self.L.tcx_set_kpoint_symmetry(self.h, _ptr(ks, _ip), _ptr(r, _ip), _ptr(t), _ptr(tr, _ip)),
                    "set_kpoint_symmetry")

    def get_density(self):
        """The real-space density the terms read, complex128 (2, N)."""
        if getattr(self, "sys", None) is None:
            raise TcxError(TCX_ERR_STATE, "get_density: no system loaded")
        n = int(np.prod(self.sys.n))
        out = np.zeros((2, n), np.complex128)
        self._check(self.L.tcx_get_density(self.h, out.ctypes.data), "get_density")
        return out

    def set_density(self, rho):
        """Replace the density (complex128 (2, N)) until the next set_states."""
        r = np.ascontiguousarray(rho, dtype=np.complex128)
        self._check(self.L.tcx_set_density(self.h, r.ctypes.data), "set_density")

    def load_system(self, sys, div_corr=None):
        """Push a tc_amd.system.System into the context."""
        self.sys = sys
        self.set_cell(sys.recip, sys.volume, sys.n, sys.nspin, sys.method)
        self.set_kpoints(sys.kfrac, sys.num_kpoints)
        self.set_basis(sys.npw, sys.gindex)
        self.set_states(sys.nocc, sys.filling, sys.phi_occ, sys.chi_occ)
        sym = getattr(sys, "symmetry", None)
        if sym is not None:
            self.set_kpoint_symmetry(sym["k_src"], sym["rot"], sym["trans"], sym["trev"])
        self.set_jastrow_rpa(sys.A, sys.F)
        poly = getattr(sys, "jastrow_poly", None)
        if poly is not None and int(poly["deg"]) > 0:
            self.set_jastrow_polynomial(poly["deg"], poly["C"], poly["L"], poly["c"], poly["cusp"])
        else:
            self.set_jastrow_polynomial(0)
        on = sys.div_corr if div_corr is None else div_corr
        kweight = getattr(sys, "kweight", None)
        if kweight is None:
            alpha, sv = sum_of_vaux(sys.recip, sys.volume, sys.kfrac)
            skv = None
        else:
            alpha, sv, skv = sum_of_vaux_weighted(sys.recip, sys.volume, sys.kfrac, kweight)
        if on:
            self.set_div_correction(True, alpha, sv)
        else:
            self.set_div_correction(False)
        if kweight is not None:
            zw = sys.zero_weight_states or {}
            ns = zw.get("nstate", np.zeros(sys.nk, np.int32))
            empty = np.zeros(0, np.complex128)
            self.set_zero_weight_kpoints(kweight, skv, ns, zw.get("fill", np.zeros(0)), zw.get("phi", empty),
                                         zw.get("chi") if sys.method == "BITC" else None)

    # --------------------------------------------------------------- terms
    def _term(self, name, nb, phi, out=None, precision=PREC_FP64):
        nb = np.ascontiguousarray(nb, dtype=np.int32)
        phi = np.ascontiguousarray(phi, dtype=np.complex128)
        if out is None:
            out = np.zeros_like(phi)
        st = getattr(self.L, "tcx_" + name)(self.h, _ptr(nb, _ip), phi.ctypes.data, out.ctypes.data,
                                           int(precision), 0)
        self._check(st, name)
        return out

    def _term_device(self, name, nb, phi_ptr, out_ptr, precision):
        nb = np.ascontiguousarray(nb, dtype=np.int32)
        st = getattr(self.L, "tcx_" + name)(self.h, _ptr(nb, _ip), C.c_void_p(phi_ptr), C.c_void_p(out_ptr),
                                           int(precision), FLAG_DEVICE_PTRS)
        self._check(st, name)

    def hf2x(self, nb, phi, out=None, precision=PREC_FP64):
        return self._term("hf2x", nb, phi, out, precision)

    def tc2x(self, nb, phi, out=None, precision=PREC_FP64):
        return self._term("tc2x", nb, phi, out, precision)

    def tc3a1(self, nb, phi, out=None, precision=PREC_FP64):
        return self._term("tc3a1", nb, phi, out, precision)

    def tc3b1(self, nb, phi, out=None, precision=PREC_FP64):
        return self._term("tc3b1", nb, phi, out, precision)

    def tc3a3b3b4b6(self, nb, phi, out=None, precision=PREC_FP64):
        return self._term("tc3a3b3b4b6", nb, phi, out, precision)

    def tc3a2a4b2b5(self, nb, phi, out=None, precision=PREC_FP64):
        return self._term("tc3a2a4b2b5", nb, phi, out, precision)

    def _term_fp64(self, name, nb, phi, out=None):
        nb = np.ascontiguousarray(nb, dtype=np.int32)
        phi = np.ascontiguousarray(phi, dtype=np.complex128)
        if out is None:
            out = np.zeros_like(phi)
        st = getattr(self.L, "tcx_" + name)(self.h, _ptr(nb, _ip), phi.ctypes.data, out.ctypes.data, 0)
        self._check(st, name)
        return out

    def kinetic(self, nb, phi, out=None):
        return self._term_fp64("kinetic", nb, phi, out)

    def hf2h(self, nb, phi, out=None):
        return self._term_fp64("hf2h", nb, phi, out)

    def tc2h(self, nb, phi, out=None):
        return self._term_fp64("tc2h", nb, phi, out)

    def davidson(self, nb, phi, eig, max_blocks=2, num_refresh=1, precision=PREC_FP64, phi_left=None):
        """One Diagonalization::block_davidson call: returns (phi, phi_left or None, eig, energies[:,3]).
        phi_left given (BITC): the bi-orthogonal variant (biortho_david = true) from these left bands."""
        nb = np.ascontiguousarray(nb, dtype=np.int32)
        phi = np.array(phi, dtype=np.complex128, copy=True)
        eig = np.array(eig, dtype=np.complex128, copy=True)
        bi = phi_left is not None
        if bi:
            left = np.array(phi_left, dtype=np.complex128, copy=True)
            if left.shape != phi.shape:
                raise ValueError("davidson: phi_left must have the shape of phi")
        else:
            left = np.zeros_like(phi) if self.method == "BITC" else None
        en = np.zeros((int(nb.sum()), 3), np.complex128)
        st = self.L.tcx_davidson(self.h, _ptr(nb, _ip), phi.ctypes.data, left.ctypes.data if left is not None else None,
                                 eig.ctypes.data, en.ctypes.data, int(max_blocks), int(num_refresh), int(bi),
                                 int(precision))
        self._check(st, "davidson")
        return phi, left, eig, en

    # --------------------------------------------------------------- gradients
    def grad_exchange(self, precision=PREC_FP64):
        """(E_x, dE_x/d conj(phi_occ)) for the loaded occupied set."""
        n = int(self.sys.phi_occ.size)
        g = np.zeros(n, np.complex128)
        e = np.zeros(2)
        self._check(self.L.tcx_grad_exchange(self.h, g.ctypes.data, _ptr(e), int(precision)), "grad_exchange")
        return complex(e[0], e[1]), g

    def grad_tc2(self, precision=PREC_FP64):
        """(E_2, dE_2/d conj(chi), dE_2/d phi, dE_2/dA) for the loaded occupied set."""
        n = int(self.sys.phi_occ.size)
        gl = np.zeros(n, np.complex128)
        gr = np.zeros(n, np.complex128)
        e = np.zeros(2)
        da = np.zeros(2)
        self._check(self.L.tcx_grad_tc2(self.h, gl.ctypes.data, gr.ctypes.data, _ptr(e), _ptr(da), int(precision)),
                    "grad_tc2")
        return complex(e[0], e[1]), gl, gr, complex(da[0], da[1])

    def kernel_gradient(self, term, nbins, width, precision=PREC_FP64):
        """(dE/dK bins, dE/du bins or None, sums[4]) -- term "x" (E_x, Coulomb) or "tc2" (E_2 exchange)."""
        t = {"x": 0, "tc2": 1}[term]
        gk = np.zeros(nbins, np.complex128)
        gu = np.zeros(nbins, np.complex128)
        sums = np.zeros(4)
        self._check(self.L.tcx_kernel_gradient(self.h, t, int(nbins), float(width), gk.ctypes.data,
                                               gu.ctypes.data, _ptr(sums), int(precision)), "kernel_gradient")
        return gk, (gu if t == 1 else None), sums

    # --------------------------------------------------------------- HIP graphs
    def graph_begin(self):
        self._check(self.L.tcx_graph_begin(self.h), "graph_begin")

    def graph_end(self):
        gid = np.zeros(1, np.int32)
        self._check(self.L.tcx_graph_end(self.h, _ptr(gid, _ip)), "graph_end")
        return int(gid[0])

    def graph_launch(self, gid):
        self._check(self.L.tcx_graph_launch(self.h, int(gid)), "graph_launch")

    def graph_destroy(self, gid):
        self._check(self.L.tcx_graph_destroy(self.h, int(gid)), "graph_destroy")

    def apply_all(self, nb, phi, precision=PREC_FP64, method=None):
        """H1 phi, H2 phi, H3 phi (H3 is None for HF), each zeroed first (calc_hamiltonian::all)."""
        nb = np.ascontiguousarray(nb, dtype=np.int32)
        phi = np.ascontiguousarray(phi, dtype=np.complex128)
        h1, h2 = np.empty_like(phi), np.empty_like(phi)
        h3 = np.empty_like(phi) if (method or self.method) != "HF" else None
        st = self.L.tcx_apply_all(self.h, _ptr(nb, _ip), phi.ctypes.data, h1.ctypes.data, h2.ctypes.data,
                                  h3.ctypes.data if h3 is not None else None, int(precision), 0)
        self._check(st, "apply_all")
        return h1, h2, h3

    # --------------------------------------------------------------- tuning / timing
    SHAPE_TERMS = {"hf2x": 0, "tc2x": 1, "tc3": 2, "tc3a3b3b4b6": 2, "tc3a2a4b2b5": 2}

    def set_launch_shape(self, term, targets=0, sources=0):
        """term: "hf2x", "tc2x" or "tc3" (the three-body terms, also by their names); 0 = automatic."""
        t = self.SHAPE_TERMS[term] if isinstance(term, str) else int(term)
        self._check(self.L.tcx_set_launch_shape(self.h, t, int(targets), int(sources)), "set_launch_shape")

    def set_pipeline_order(self, term, order):
        """hf2x pass order: 0 z-first, 1 plane-first (FP32-split), -1 automatic."""
        t = {"hf2x": 0}[term] if isinstance(term, str) else int(term)
        self._check(self.L.tcx_set_pipeline_order(self.h, t, int(order)), "set_pipeline_order")

    def record_columns(self, term="hf2x"):
        """Mean (kx, ky) columns per plane of the last hf2x apply's pair records."""
        t = {"hf2x": 0}[term] if isinstance(term, str) else int(term)
        out = C.c_double(0.0)
        self._check(self.L.tcx_get_record_columns(self.h, t, C.byref(out)), "get_record_columns")
        return out.value

    def enable_timing(self, on=True):
        self._check(self.L.tcx_enable_timing(self.h, int(bool(on))), "enable_timing")

    def get_timing(self):
        ms = np.zeros(4)
        la = (C.c_long * 4)()
        pa = (C.c_long * 4)()
        self._check(self.L.tcx_get_timing(self.h, _ptr(ms), la, pa), "get_timing")
        return ms, np.array(la[:]), np.array(pa[:])

    def synchronize(self):
        self._check(self.L.tcx_synchronize(self.h), "synchronize")

    def device_buffer(self, nbytes):
        return DeviceBuffer(self, nbytes)

    def to_device(self, array):
        a = np.ascontiguousarray(array)
        buf = DeviceBuffer(self, a.nbytes)
        buf.upload(a)
        return buf


class DeviceBuffer:
    """Device memory owned through the library's own HIP runtime (no torch on
    the GPU path: torch bundles a second HIP runtime)."""

    def __init__(self, eng: Engine, nbytes: int):
        self.eng = eng
        self.nbytes = int(nbytes)
        p = C.c_void_p()
        eng._check(eng.L.tcx_device_alloc(eng.h, self.nbytes, C.byref(p)), "device_alloc")
        self.ptr = p.value

    def upload(self, array):
        a = np.ascontiguousarray(array)
        assert a.nbytes <= self.nbytes
        self.eng._check(self.eng.L.tcx_memcpy(self.eng.h, C.c_void_p(self.ptr), C.c_void_p(a.ctypes.data),
                                              a.nbytes, 0), "memcpy h2d")

    def download(self, dtype, count=None):
        dt = np.dtype(dtype)
        n = self.nbytes // dt.itemsize if count is None else int(count)
        out = np.empty(n, dtype=dt)
        self.eng._check(self.eng.L.tcx_memcpy(self.eng.h, C.c_void_p(out.ctypes.data), C.c_void_p(self.ptr),
                                              out.nbytes, 1), "memcpy d2h")
        return out

    def zero(self):
        self.eng._check(self.eng.L.tcx_memset_zero(self.eng.h, C.c_void_p(self.ptr), self.nbytes), "memset")

    def free(self):
        if self.ptr:
            self.eng._check(self.eng.L.tcx_device_free(self.eng.h, C.c_void_p(self.ptr)), "device_free")
            self.ptr = None

    def __del__(self):
        try:
            if self.ptr and self.eng.h:
                self.free()
        except Exception:
            pass
