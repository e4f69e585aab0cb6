"""The SCF loop of TC++ (Diagonalization::scf, src/diagonalization.cpp:154-303) on the GPU
engine, starting from QE orbitals: a Python front end to the C++ host's loop
(tcxh_scf_run, include/tcx_host.h; tc_amd/host/scf.cpp), which is the one implementation.

Per iteration the host runs, in the reference's order: set_filling from the current
eigenvalues (bloch_states_scfloop_filling_density.cpp:9-199), calc_total_energy
(total_energy.cpp:101-170) and the energy check, mix_density_matrix (:204-223) or density
mixing, set_density (:256-358) and the charge check, then block_davidson on every
irreducible k (diagonalization_davidson.cpp:7-689) with reset_phik's old bands under
density-matrix mixing (bloch_states_scfloop_phik.cpp:86-99).  This module returns what the
run printed: the fillings and Fermi energies, the energies, the density differences.
"""
from __future__ import annotations

import ctypes as C
import dataclasses

import numpy as np

from . import qe as Q
from . import tcx


@dataclasses.dataclass
class FillingStep:
    fermi: float                  # Hartree
    eigenvalues: np.ndarray       # (nirr, nb) complex, the eigenvalues the fillings came from
    filling: np.ndarray           # (nirr, nb)


@dataclasses.dataclass
class SCFRun:
    fillings: list                # FillingStep per set_filling, in order
    energies: list                # [E_1, E_2, E_3, E_total] complex per calc_total_energy
    density_difference: list      # per set_density after the first
    converged: bool
    iterations: int               # Davidson calls made
    replicas: list


@dataclasses.dataclass
class FirstIteration:
    fermi_initial: float          # Hartree
    fermi: float                  # Hartree, after iteration 1
    eigenvalues: np.ndarray       # (nirr, nb) complex, after iteration 1
    filling: np.ndarray           # (nirr, nb)
    energies: np.ndarray          # [E_1, E_2, E_3, E_total] complex
    replicas: list


def _fillings(qe: Q.QEInput, eig: np.ndarray, smearing: str, width: float):
    nb = np.full(qe.nirr, eig.shape[1], np.int32)
    f, mu, nocc = tcx.set_filling(nb, eig.reshape(-1), qe.kweight, qe.nelec, smearing, width)
    return f.reshape(qe.nirr, -1), mu, nocc


def run(engine: tcx.Engine, qe: Q.QEInput, precision=tcx.PREC_FP64, max_iterations=None) -> SCFRun:
    """Diagonalization::scf on the engine's context; max_iterations replaces the input's
    max_num_iterations.  The context holds the run's last state afterwards."""
    L = engine.L
    with Q.HostInputs(qe) as h:
        reps = h.replicas()
        res = C.c_void_p()
        st = L.tcxh_scf_run(engine.h, h.h, int(precision), int(max_iterations or 0), C.byref(res))
    engine.sys = None  # the context no longer holds the engine's last loaded system
    tcx.host_check(st, "tcxh_scf_run")
    try:
        n = [C.c_int() for _ in range(7)]
        tcx.host_check(L.tcxh_scf_result_sizes(res, *[C.byref(x) for x in n]), "tcxh_scf_result_sizes")
        nfill, nenergy, nddiff, iterations, converged, nirr, nb = (x.value for x in n)
        fills = []
        for i in range(nfill):
            mu = np.zeros(1)
            eig = np.zeros((nirr, nb), np.complex128)
            f = np.zeros((nirr, nb))
            tcx.host_check(L.tcxh_scf_result_fill(res, i, tcx._ptr(mu), eig.ctypes.data, tcx._ptr(f)),
                           "tcxh_scf_result_fill")
            fills.append(FillingStep(fermi=float(mu[0]), eigenvalues=eig, filling=f))
        energies = []
        for i in range(nenergy):
            e = np.zeros(4, np.complex128)
            tcx.host_check(L.tcxh_scf_result_energy(res, i, e.ctypes.data), "tcxh_scf_result_energy")
            energies.append(e)
        dd = np.zeros(max(nddiff, 1))
        tcx.host_check(L.tcxh_scf_result_ddiff(res, tcx._ptr(dd)), "tcxh_scf_result_ddiff")
    finally:
        L.tcxh_scf_result_free(res)
    return SCFRun(fillings=fills, energies=energies, density_difference=[float(x) for x in dd[:nddiff]],
                  converged=bool(converged), iterations=iterations, replicas=reps)


def first_iteration(engine: tcx.Engine, qe: Q.QEInput, precision=tcx.PREC_FP64) -> FirstIteration:
    """Iteration 1 only: the fillings from the QE eigenvalues, one Davidson, then the
    fillings and the total energy printed after "Iteration 1"."""
    r = run(engine, qe, precision, max_iterations=1)
    return FirstIteration(fermi_initial=r.fillings[0].fermi, fermi=r.fillings[1].fermi,
                          eigenvalues=r.fillings[1].eigenvalues, filling=r.fillings[1].filling,
                          energies=r.energies[0], replicas=r.replicas)
