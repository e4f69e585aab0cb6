"""The SCF loop of TC++ (Diagonalization::scf, src/diagonalization.cpp:154-303) on the GPU
engine, starting from QE orbitals.

Per iteration `it` (0-based), as the reference orders it:
  1. set_filling from the current eigenvalues (bloch_states_scfloop_filling_density.cpp:9-199);
     with density-matrix mixing and it > 0 the previous fillings of the occupied bands are
     kept first as filling_old (:19-31);
  2. it > 0: TotalEnergy::calc_total_energy with the new fillings and the band energies e_n
     of the last Davidson (total_energy.cpp:101-170); energy converged when
     |E - E_previous| < energy_tolerance (E_previous starts at 0); with density-matrix mixing
     filling *= beta, filling_old *= 1 - beta (mix_density_matrix, :204-223);
  3. set_density (:256-358) from the occupied set (plus the "old" bands of the previous
     iteration with filling_old under density-matrix mixing); with density mixing and it > 0
     rho = (1 - beta) rho_previous + beta rho_new; the L1 density difference
     2 sum_r |rho - rho_previous| / (N Nk); converged when the energy and the charge were
     converged in this and the previous iteration;
  4. "Iteration it+1": block_davidson on every irreducible k (diagonalization_davidson.cpp:7-689)
     with H = calc_hamiltonian::all of this occupied set and density; under density-matrix
     mixing the occupied bands before the update become the "old" bands
     (BlochStates::reset_phik, bloch_states_scfloop_phik.cpp:86-99), and the fillings are
     restored afterwards (recover_filling_for_density_matrix, :225-242).
Without convergence after max_num_iterations, a final set_filling and calc_total_energy
(and the density of step 3) close the run (:247-258).

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


def _total_energy(reps, fill, en, method):
    """calc_total_energy (total_energy.cpp:101-170): sum over the replicas of every
    irreducible k (e_n conjugated at time-reversal replicas) / Nk."""
    e = np.zeros(3, np.complex128)
    for r in reps:
        w = en[r.ik] if not r.trev else np.conj(en[r.ik])
        e += np.sum(fill[r.ik][:, None] * w, axis=0) / np.array([1.0, 2.0, 3.0])
    if method == "HF":
        e[2] = 0.0
    e /= len(reps)
    return np.array([e[0], e[1], e[2], e.sum()])


def run(engine: tcx.Engine, qe: Q.QEInput, precision=tcx.PREC_FP64, max_iterations=None) -> SCFRun:
    setup = qe.setup
    method = setup["method"]
    bitc = method == "BITC"
    smearing, width = setup["smearing_mode"], float(setup["smearing_width"])
    beta = float(setup.get("mixing_beta", 0.7))
    dm = bool(setup.get("mixes_density_matrix", False))
    etol = float(setup.get("energy_tolerance", 1e-6))
    ctol = float(setup.get("charge_tolerance", 1e-5))
    maxit = int(setup.get("max_num_iterations", 30)) if max_iterations is None else int(max_iterations)
    nb, nirr = qe.nbands, qe.nirr

    phi = [np.array(c) for c in qe.coef]                    # phik_scf
    chi = [np.array(c) for c in qe.coef] if bitc else None  # phik_left_scf = phik at the start
    eig = qe.eig_qe.astype(np.complex128)
    en = None
    fill = nocc = None
    old_phi = old_chi = fill_old = None
    rho = None
    e_prev = 0.0
    e_conv = e_conv_prev = c_conv = c_conv_prev = conv = False
    out = SCFRun(fillings=[], energies=[], density_difference=[], converged=False, iterations=0, replicas=[])

    def filling_step(keep_old):
        nonlocal fill, nocc, fill_old
        if keep_old:
            fill_old = [fill[ik][:nocc[ik]].copy() for ik in range(nirr)]
        fill, mu, nocc = _fillings(qe, eig.real, smearing, width)
        out.fillings.append(FillingStep(fermi=mu, eigenvalues=eig.copy(), filling=fill.copy()))

    def density_step(first):
        """Occupied set (+ old bands), density and mixing; returns the density difference."""
        nonlocal rho
        occ, f, left = [], [], []
        for ik in range(nirr):
            n = int(nocc[ik])
            o, ff, lt = [phi[ik][:n]], [fill[ik][:n]], [chi[ik][:n]] if bitc else []
            if old_phi is not None and fill_old is not None:
                o.append(old_phi[ik])
                ff.append(fill_old[ik])
                if bitc:
                    lt.append(old_chi[ik])
            occ.append(np.concatenate(o))
            f.append(np.concatenate(ff))
            if bitc:
                left.append(np.concatenate(lt))
        n_occ = np.array([len(x) for x in f], np.int32)
        sys_, reps = Q.build_system(qe, occ, f, n_occ, left if bitc else None, method)
        out.replicas = reps
        engine.load_system(sys_)
        new = engine.get_density()
        if first:
            rho = new
            return None, reps
        if not dm:
            new = (1.0 - beta) * rho + beta * new
            engine.set_density(new)
        diff = 2.0 * float(np.abs(new[0] - rho[0]).sum()) / (new.shape[1] * len(reps))
        rho = new
        out.density_difference.append(diff)
        return diff, reps

    reps = None
    for it in range(maxit):
        first = it == 0
        filling_step(keep_old=(not first and dm))
        if not first:
            e = _total_energy(reps, fill, en, method)
            out.energies.append(e)
            e_conv = abs(e[3] - e_prev) < etol
            e_prev = e[3]
            if dm:
                fill = fill * beta
                fill_old = [x * (1.0 - beta) for x in fill_old]
        diff, reps = density_step(first)
        if not first:
            c_conv = diff < ctol
            conv = e_conv and e_conv_prev and c_conv and c_conv_prev
            if conv:
                break
        e_conv_prev, c_conv_prev = e_conv, c_conv

        # block Davidson on the irreducible k (the replica with isym 0 of each irreducible k)
        nb_full = np.zeros(len(reps), np.int32)
        seen = set()
        for i, r in enumerate(reps):
            if r.ik not in seen:
                seen.add(r.ik)
                nb_full[i] = nb
        phi_cat = np.concatenate([phi[ik].reshape(-1) for ik in range(nirr)])
        biortho = bitc and bool(setup.get("biortho_david", False))  # diagonalization.cpp:117-122
        left_cat = np.concatenate([chi[ik].reshape(-1) for ik in range(nirr)]) if biortho else None
        phi_out, left_out, eig_out, en_out = engine.davidson(
            nb_full, phi_cat, eig.reshape(-1), max_blocks=setup["max_num_blocks_david"],
            num_refresh=setup["num_refresh_david"], precision=precision, phi_left=left_cat)
        out.iterations += 1
        if dm:  # reset_phik keeps the occupied bands before the update as the old bands
            old_phi = [phi[ik][:nocc[ik]].copy() for ik in range(nirr)]
            old_chi = [chi[ik][:nocc[ik]].copy() for ik in range(nirr)] if bitc else None
        split = np.cumsum([0] + [nb * len(qe.mill[ik]) for ik in range(nirr)])
        phi = [phi_out[split[ik]:split[ik + 1]].reshape(nb, -1) for ik in range(nirr)]
        if bitc:
            chi = [left_out[split[ik]:split[ik + 1]].reshape(nb, -1) for ik in range(nirr)]
        eig = eig_out.reshape(nirr, nb)
        en = en_out.reshape(nirr, nb, 3)
        if not first and dm:
            fill = fill / beta
            fill_old = [np.zeros_like(x) for x in fill_old]

    if not conv:
        filling_step(keep_old=dm)
        out.energies.append(_total_energy(reps, fill, en, method))
        if dm:
            fill = fill * beta
            fill_old = [x * (1.0 - beta) for x in fill_old]
        density_step(False)
    out.converged = conv
    return out


def first_iteration(engine: tcx.Engine, qe: Q.QEInput, precision=tcx.PREC_FP64) -> FirstIteration:
    """Iteration 1 only: the fillings from the QE eigenvalues, one Davidson, then the
    fillings and the total energy printed after "Iteration 1"."""
    r = run(engine, qe, precision, max_iterations=1)
    return FirstIteration(fermi_initial=r.fillings[0].fermi, fermi=r.fillings[1].fermi,
                          eigenvalues=r.fillings[1].eigenvalues, filling=r.fillings[1].filling,
                          energies=r.energies[0], replicas=r.replicas)
