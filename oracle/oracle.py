"""ctypes binding of the FP64 CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, never by the product package.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)


class OrcProblem(C.Structure):
    _fields_ = [
        ("n", C.c_int * 3),
        ("recip", C.c_double * 9),
        ("volume", C.c_double),
        ("nspin", C.c_int),
        ("method", C.c_int),
        ("nk", C.c_int),
        ("kfrac", _dp),
        ("num_kpoints", C.c_int),
        ("npw", _ip),
        ("gindex", _ip),
        ("nocc", _ip),
        ("filling", _dp),
        ("phi_occ", _dp),
        ("chi_occ", _dp),
        ("A", C.c_double * 4),
        ("F", C.c_double * 4),
        ("jastrow_A_zero", C.c_int),
        ("div_corr", C.c_int),
        ("alpha", C.c_double),
        ("sum_vaux", _dp),
        ("density", _dp),
        ("kweight", _dp),
        ("sum_kvaux", _dp),
        ("zw_nstate", _ip),
        ("zw_fill", _dp),
        ("zw_phi", _dp),
        ("zw_chi", _dp),
        ("poly", C.c_void_p),
    ]


class OrcJastrowPoly(C.Structure):
    _fields_ = [
        ("deg", C.c_int),
        ("C", C.c_int),
        ("rpa", C.c_int),
        ("nr", C.c_int),
        ("L", C.c_double * 4),
        ("A", C.c_double * 4),
        ("F", C.c_double * 4),
        ("c_input", (C.c_double * 16) * 4),
        ("c_int", (C.c_double * 19) * 4),
        ("fourpi_L", (C.c_double * 38) * 4),
        ("exp_LF", C.c_double * 4),
        ("M", (C.c_double * 47) * 4),
        ("rmesh", _dp),
        ("integrand", _dp),
    ]


def build(force: bool = False) -> str:
    if force or not os.path.exists(LIB_PATH):
        subprocess.run(["make", "-s", "-C", HERE], check=True)
    return LIB_PATH


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        P = C.POINTER(OrcProblem)
        L.orc_fft3d.argtypes = [_dp, _ip, C.c_int]
        L.orc_dft3d_naive.argtypes = [_dp, _dp, _ip, C.c_int]
        L.orc_uk.argtypes = [P, C.c_double, C.c_int, C.c_int]
        L.orc_uk.restype = C.c_double
        L.orc_tc_2body.argtypes = [P, C.c_double, C.c_int, C.c_int]
        L.orc_tc_2body.restype = C.c_double
        L.orc_sum_of_vaux.argtypes = [_dp, C.c_double, C.c_int, _dp, _dp, _dp]
        L.orc_sum_of_vaux_w.argtypes = [_dp, C.c_double, C.c_int, _dp, _dp, _dp, _dp, _dp]
        L.orc_density.argtypes = [P, _dp]
        JP = C.POINTER(OrcJastrowPoly)
        L.orc_jastrow_poly_init.argtypes = [JP, C.c_int, C.c_int, _dp, _dp, C.c_int, C.c_int, _dp, _dp, _dp, _ip]
        L.orc_jastrow_poly_init.restype = C.c_int
        L.orc_jastrow_poly_free.argtypes = [JP]
        L.orc_J_integ.argtypes = [JP, C.c_double, C.c_int, _dp, C.c_int, C.c_int]
        L.orc_J2_integ.argtypes = [JP, C.c_double, C.c_int, _dp, C.c_int]
        for name in ("orc_J3_integ", "orc_uk_poly", "orc_tc2body_poly"):
            getattr(L, name).argtypes = [JP, C.c_double, C.c_int]
            getattr(L, name).restype = C.c_double
        L.orc_tc2body_rpapoly.argtypes = [JP, C.c_double, C.c_int, C.c_double]
        L.orc_tc2body_rpapoly.restype = C.c_double
        for name in ("orc_hf2x", "orc_tc2x", "orc_tc3a3b3b4b6", "orc_tc3a2a4b2b5"):
            getattr(L, name).argtypes = [P, _ip, _dp, _dp, C.c_int]
        for name in ("orc_tc3a1", "orc_tc3b1", "orc_kinetic", "orc_hf2h", "orc_tc2h"):
            getattr(L, name).argtypes = [P, _ip, _dp, _dp]
        _lib = L
    return _lib


def _p(a, t=_dp):
    return a.ctypes.data_as(t) if a is not None else None


def fft3d(grid: np.ndarray, sign: int) -> np.ndarray:
    """grid: complex128 array shaped (n3, n2, n1) (i1 fastest); returns a new array."""
    g = np.ascontiguousarray(grid, dtype=np.complex128).copy()
    n = np.array([g.shape[2], g.shape[1], g.shape[0]], dtype=np.int32)
    lib().orc_fft3d(g.view(np.float64).ctypes.data_as(_dp), _p(n, _ip), sign)
    return g


def dft3d_naive(grid: np.ndarray, sign: int) -> np.ndarray:
    g = np.ascontiguousarray(grid, dtype=np.complex128)
    out = np.empty_like(g)
    n = np.array([g.shape[2], g.shape[1], g.shape[0]], dtype=np.int32)
    lib().orc_dft3d_naive(g.view(np.float64).ctypes.data_as(_dp), out.view(np.float64).ctypes.data_as(_dp),
                          _p(n, _ip), sign)
    return out


def sum_of_vaux(recip, volume, kfrac):
    r = np.ascontiguousarray(recip, dtype=np.float64).reshape(9)
    kf = np.ascontiguousarray(kfrac, dtype=np.float64).reshape(-1)
    nk = len(kf) // 3
    out = np.zeros(nk)
    alpha = np.zeros(1)
    lib().orc_sum_of_vaux(_p(r), float(volume), nk, _p(kf), _p(alpha), _p(out))
    return float(alpha[0]), out


def sum_of_vaux_w(recip, volume, kfrac, kweight):
    """(alpha, sum_of_Vaux[nk], sum_of_kVaux[nk, 3]) with the q-mesh restricted to the k
    of weight >= 1e-8 (potentials_coulomb.cpp:60-254)."""
    r = np.ascontiguousarray(recip, dtype=np.float64).reshape(9)
    kf = np.ascontiguousarray(kfrac, dtype=np.float64).reshape(-1)
    kw = np.ascontiguousarray(kweight, dtype=np.float64)
    nk = len(kf) // 3
    out, kout, alpha = np.zeros(nk), np.zeros(3 * nk), np.zeros(1)
    lib().orc_sum_of_vaux_w(_p(r), float(volume), nk, _p(kf), _p(kw), _p(alpha), _p(out), _p(kout))
    return float(alpha[0]), out, kout.reshape(nk, 3)


class JastrowPoly:
    """The oracle's polynomial Jastrow (jastrow_poly.c) for one parameter set.

    spec: {"deg": int, "C": 2|3, "L": (2, 2), "c": (2, 2, deg) c_poly_input (entry 0 of each
    pair is replaced by the cusp condition), "cusp": bool}; A, F: the RPA parameters (A all
    zero: polynomial only); recip, n: the FFT grid that fixes num_rmesh_J3_integ."""

    def __init__(self, spec, A, F, recip, n):
        self.s = OrcJastrowPoly()
        deg = int(spec["deg"])
        L = np.ascontiguousarray(spec["L"], dtype=np.float64).reshape(4)
        c = np.ascontiguousarray(spec["c"], dtype=np.float64).reshape(4 * deg) if deg else np.zeros(1)
        A = np.ascontiguousarray(A, dtype=np.float64).reshape(4)
        F = np.ascontiguousarray(F, dtype=np.float64).reshape(4)
        rpa = int(np.any(np.abs(A) > 1e-8))
        r = np.ascontiguousarray(recip, dtype=np.float64).reshape(9)
        nn = np.ascontiguousarray(n, dtype=np.int32)
        rc = lib().orc_jastrow_poly_init(C.byref(self.s), deg, int(spec["C"]), _p(L), _p(c), int(spec["cusp"]), rpa,
                                         _p(A), _p(F), _p(r), _p(nn, _ip))
        if rc != 0:
            raise ValueError(f"polynomial Jastrow set-up failed ({rc})")
        if spec.get("hermitian_conj"):
            # Jastrow::take_Hermitian_conj (jastrow_functions.cpp:417-440) negates c_poly_internal_
            # (A is negated by the caller, whose RPA part reads it too)
            for p in range(4):
                for i in range(19):
                    self.s.c_int[p][i] = -self.s.c_int[p][i]

    def __del__(self):
        try:
            lib().orc_jastrow_poly_free(C.byref(self.s))
        except Exception:
            pass

    @property
    def num_rmesh(self) -> int:
        return int(self.s.nr)

    def c_input(self, p):
        return np.array(self.s.c_input[p][:self.s.deg])

    def c_internal(self, p):
        return np.array(self.s.c_int[p][:self.s.deg + self.s.C])

    def J(self, G2, p, nJ, zeroth=True):
        out = np.zeros(nJ)
        lib().orc_J_integ(C.byref(self.s), float(G2), p, _p(out), nJ, int(zeroth))
        return out

    def J2(self, G1, p, nJ2):
        out = np.zeros(nJ2)
        lib().orc_J2_integ(C.byref(self.s), float(G1), p, _p(out), nJ2)
        return out

    def J3(self, G1, p):
        return lib().orc_J3_integ(C.byref(self.s), float(G1), p)

    def uk_poly(self, G2, p):
        return lib().orc_uk_poly(C.byref(self.s), float(G2), p)

    def tc2body_poly(self, G2, p):
        return lib().orc_tc2body_poly(C.byref(self.s), float(G2), p)

    def tc2body_rpapoly(self, G2, p, rpa):
        return lib().orc_tc2body_rpapoly(C.byref(self.s), float(G2), p, float(rpa))


class Oracle:
    """Holds the ctypes problem struct (and keeps its arrays alive) for one System."""

    def __init__(self, sys, density: np.ndarray | None = None, div_corr: bool | None = None):
        self.sys = sys
        self._keep = []
        P = OrcProblem()
        P.n[:] = list(sys.n)
        P.recip[:] = list(np.asarray(sys.recip, dtype=np.float64).reshape(9))
        P.volume = sys.volume
        P.nspin = sys.nspin
        P.method = sys.method_id
        P.nk = sys.nk
        P.num_kpoints = sys.num_kpoints
        self.kfrac = self._hold(np.ascontiguousarray(sys.kfrac, dtype=np.float64).reshape(-1))
        self.npw = self._hold(np.ascontiguousarray(sys.npw, dtype=np.int32))
        self.gindex = self._hold(np.ascontiguousarray(sys.gindex, dtype=np.int32))
        self.nocc = self._hold(np.ascontiguousarray(sys.nocc, dtype=np.int32))
        self.filling = self._hold(np.ascontiguousarray(sys.filling, dtype=np.float64))
        self.phi_occ = self._hold(np.ascontiguousarray(sys.phi_occ, dtype=np.complex128))
        self.chi_occ = (self._hold(np.ascontiguousarray(sys.chi_occ, dtype=np.complex128))
                        if sys.chi_occ is not None else None)
        P.kfrac = _p(self.kfrac)
        P.npw = _p(self.npw, _ip)
        P.gindex = _p(self.gindex, _ip)
        P.nocc = _p(self.nocc, _ip)
        P.filling = _p(self.filling)
        P.phi_occ = self.phi_occ.view(np.float64).ctypes.data_as(_dp)
        P.chi_occ = self.chi_occ.view(np.float64).ctypes.data_as(_dp) if self.chi_occ is not None else None
        P.A[:] = list(np.asarray(sys.A, dtype=np.float64).reshape(4))
        P.F[:] = list(np.asarray(sys.F, dtype=np.float64).reshape(4))
        P.jastrow_A_zero = int(sys.jastrow_A_zero)
        spec = getattr(sys, "jastrow_poly", None)
        if spec is not None and int(spec["deg"]) > 0:
            self.poly = JastrowPoly(spec, sys.A, sys.F, sys.recip, sys.n)
            P.poly = C.cast(C.pointer(self.poly.s), C.c_void_p)
        P.div_corr = int(sys.div_corr if div_corr is None else div_corr)
        kweight = getattr(sys, "kweight", None)
        if kweight is None:
            alpha, svaux = sum_of_vaux(sys.recip, sys.volume, sys.kfrac)
        else:
            alpha, svaux, skv = sum_of_vaux_w(sys.recip, sys.volume, sys.kfrac, kweight)
            self.kweight = self._hold(np.ascontiguousarray(kweight, dtype=np.float64))
            self.skv = self._hold(np.ascontiguousarray(skv, dtype=np.float64).reshape(-1))
            P.kweight = _p(self.kweight)
            P.sum_kvaux = _p(self.skv)
            zw = sys.zero_weight_states
            if zw is not None:
                self.zw_n = self._hold(np.ascontiguousarray(zw["nstate"], dtype=np.int32))
                self.zw_f = self._hold(np.ascontiguousarray(zw["fill"], dtype=np.float64))
                self.zw_p = self._hold(np.ascontiguousarray(zw["phi"], dtype=np.complex128))
                P.zw_nstate = _p(self.zw_n, _ip)
                P.zw_fill = _p(self.zw_f)
                P.zw_phi = self.zw_p.view(np.float64).ctypes.data_as(_dp)
                if zw.get("chi") is not None:
                    self.zw_c = self._hold(np.ascontiguousarray(zw["chi"], dtype=np.complex128))
                    P.zw_chi = self.zw_c.view(np.float64).ctypes.data_as(_dp)
        P.alpha = alpha
        self.svaux = self._hold(svaux)
        P.sum_vaux = _p(self.svaux)
        self.P = P
        if density is None:
            density = self.density()
        if density is not None:
            self.dens = self._hold(np.ascontiguousarray(density, dtype=np.complex128))
            self.P.density = self.dens.view(np.float64).ctypes.data_as(_dp)

    def _hold(self, a):
        self._keep.append(a)
        return a

    def density(self) -> np.ndarray:
        out = np.zeros(2 * self.sys.ngrid, dtype=np.complex128)
        lib().orc_density(C.byref(self.P), out.view(np.float64).ctypes.data_as(_dp))
        return out

    def _apply(self, fname, nb, phi, hphi=None, nthreads=0):
        nb = np.ascontiguousarray(nb, dtype=np.int32)
        phi = np.ascontiguousarray(phi, dtype=np.complex128)
        out = np.zeros_like(phi) if hphi is None else hphi
        f = getattr(lib(), fname)
        args = [C.byref(self.P), _p(nb, _ip), phi.view(np.float64).ctypes.data_as(_dp),
                out.view(np.float64).ctypes.data_as(_dp)]
        if fname in ("orc_hf2x", "orc_tc2x", "orc_tc3a3b3b4b6", "orc_tc3a2a4b2b5"):
            args.append(int(nthreads))
        f(*args)
        return out

    def hf2x(self, nb, phi, nthreads=0):
        return self._apply("orc_hf2x", nb, phi, nthreads=nthreads)

    def tc2x(self, nb, phi, nthreads=0):
        return self._apply("orc_tc2x", nb, phi, nthreads=nthreads)

    def tc3a1(self, nb, phi):
        return self._apply("orc_tc3a1", nb, phi)

    def tc3b1(self, nb, phi):
        return self._apply("orc_tc3b1", nb, phi)

    def tc3a3b3b4b6(self, nb, phi, nthreads=0):
        return self._apply("orc_tc3a3b3b4b6", nb, phi, nthreads=nthreads)

    def tc3a2a4b2b5(self, nb, phi, nthreads=0):
        return self._apply("orc_tc3a2a4b2b5", nb, phi, nthreads=nthreads)

    def kinetic(self, nb, phi):
        return self._apply("orc_kinetic", nb, phi)

    def hf2h(self, nb, phi):
        return self._apply("orc_hf2h", nb, phi)

    def tc2h(self, nb, phi):
        return self._apply("orc_tc2h", nb, phi)

    def uk(self, G2, s1=0, s2=0):
        return lib().orc_uk(C.byref(self.P), float(G2), s1, s2)

    def tc_2body(self, G2, s1=0, s2=0):
        return lib().orc_tc_2body(C.byref(self.P), float(G2), s1, s2)


def fillings(nb, eig, kweight, num_electrons, smearing="fixed", width=0.01):
    """Band fillings and Fermi energy, one independent spin: BlochStates::set_filling
    (src/bloch_states_scfloop_filling_density.cpp:35-194) with Kpoints::return_band_filling
    (src/kpoints.cpp:349-387).  Returns (filling, fermi_energy, num_occupied) or raises
    ValueError where the reference stops."""
    import math
    nb = np.asarray(nb, dtype=int)
    eig = np.real(np.asarray(eig, dtype=np.complex128))
    off = np.concatenate([[0], np.cumsum(nb)])
    f = np.zeros(int(off[-1]))
    spin = 2.0
    if smearing == "fixed":                                   # :37-70
        nocc = int(round(num_electrons / 2.0))
        if abs(nocc - num_electrons / 2.0) > 1e-5:
            raise ValueError("fixed smearing needs an even electron count")
        mu = -1e5
        for k in range(len(nb)):
            if nocc > nb[k]:
                raise ValueError("num_bands_tc too small")
            if kweight[k] < 1e-8:
                continue
            for b in range(nocc):
                f[off[k] + b] = spin
                mu = max(mu, eig[off[k] + b])
    else:                                                     # :91-130
        lo, hi = -1e5, 1e5
        for _ in range(60):
            mu = 0.5 * (hi + lo)
            ne = 0.0
            for k in range(len(nb)):
                for b in range(nb[k]):
                    v = 0.0 if kweight[k] < 1e-8 else spin * 0.5 * (1.0 - math.erf((eig[off[k] + b] - mu) / width))
                    f[off[k] + b] = v
                    ne += v * kweight[k] / spin
            if ne > num_electrons:
                hi = mu
            else:
                lo = mu
    ne = sum(f[off[k]:off[k + 1]].sum() * kweight[k] / spin for k in range(len(nb)))
    if abs(ne - num_electrons) > 1e-8:                        # :174
        raise ValueError("set Fermi energy failed")
    nocc_k = np.array([int((f[off[k]:off[k + 1]] > 1e-8).sum()) for k in range(len(nb))])  # :182-194
    return f, mu, nocc_k
