"""The first SCF iteration of TC++ (Diagonalization::scf, src/diagonalization.cpp:154-303)
on the GPU engine, starting from QE orbitals.

    iter 0: set_filling from the QE eigenvalues (bloch_states_scfloop_filling_density.cpp:9-199);
            density from the QE orbitals with those fillings (no mixing in the first iteration);
            "Iteration 1": block_davidson on every irreducible k with H = calc_hamiltonian::all
            built from the full-mesh occupied set (diagonalization_davidson.cpp:7-689);
    iter 1: set_filling from the new eigenvalues, then TotalEnergy::calc_total_energy with the
            new fillings and the band energies e_n of the Davidson (total_energy.cpp:98-188).

The full mesh, the replicated orbitals and the Jastrow parameters come from tc_amd.qe.  The
Davidson runs only on the irreducible k (zero bands elsewhere); at a replica obtained with
time reversal the band energies enter conjugated, as in calc_total_energy.
"""
from __future__ import annotations

import dataclasses

import numpy as np

from . import qe as Q
from . import tcx


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


def first_iteration(engine: tcx.Engine, qe: Q.QEInput, precision=tcx.PREC_FP64) -> FirstIteration:
    setup = qe.setup
    method = setup["method"]
    smearing, width = setup["smearing_mode"], float(setup["smearing_width"])
    nb = qe.nbands

    # iteration 0: fillings from the QE eigenvalues, occupied set = replicated QE orbitals
    f0, mu0, nocc0 = _fillings(qe, qe.eig_qe, smearing, width)
    phi_irr = [np.asarray(c) for c in qe.coef]
    chi_irr = phi_irr if method == "BITC" else None  # phik_left = phik at the start (bloch_states_initialize.cpp:370)
    sys_, reps = Q.build_system(qe, phi_irr, f0, nocc0, chi_irr, method)
    engine.load_system(sys_)

    # Davidson on the irreducible k (the replica with isym 0 of each irreducible k)
    nb_full = np.zeros(sys_.nk, np.int32)
    first = {}
    for i, r in enumerate(reps):
        if r.ik not in first:
            first[r.ik] = i
            nb_full[i] = nb
    phi = np.concatenate([phi_irr[ik].reshape(-1) for ik in range(qe.nirr)])
    eig_in = qe.eig_qe.reshape(-1).astype(np.complex128)
    phi_out, left, eig, en = engine.davidson(nb_full, phi, eig_in, max_blocks=setup["max_num_blocks_david"],
                                             num_refresh=setup["num_refresh_david"], precision=precision)
    eig = eig.reshape(qe.nirr, nb)
    en = en.reshape(qe.nirr, nb, 3)

    # iteration 1: fillings from the new eigenvalues, total energy from e_n
    f1, mu1, _ = _fillings(qe, eig.real, smearing, width)
    e = np.zeros(3, np.complex128)
    for r in reps:
        w = en[r.ik] if not r.trev else np.conj(en[r.ik])
        e += np.sum(f1[r.ik][:, None] * w, axis=0) / np.array([1.0, 2.0, 3.0])
    if method == "HF":
        e[2] = 0.0
    e /= len(reps)
    energies = np.array([e[0], e[1], e[2], e.sum()])
    return FirstIteration(fermi_initial=mu0, fermi=mu1, eigenvalues=eig, filling=f1, energies=energies,
                          replicas=reps)
