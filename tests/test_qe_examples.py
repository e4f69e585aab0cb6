"""Ingestion of the reference's own end-to-end examples (test/example14/15/16: HEG HF, TC
and BITC on the fcc Si cell, fixtures in tests/golden/qe_example*.npz) -- CPU part.

Pins against numbers the reference printed (output.out):
  * the symmetry replication of the irreducible k-points (kpoints.cpp:34-110), every
    replica vector and its time-reversal flag, in the reference's order;
  * the Fermi energy of the starting fillings (set_filling on the QE eigenvalues with
    Gaussian smearing, bloch_states_scfloop_filling_density.cpp:9-199) to 1e-8 eV.
The first SCF iteration itself runs on the GPU (test_gpu_qe_examples.py)."""
import os

import numpy as np
import pytest

from tc_amd import qe as Q
from tc_amd import tcx

HERE = os.path.dirname(os.path.abspath(__file__))
EXAMPLES = [14, 15, 16]


def fixture(ex):
    return Q.load_fixture(os.path.join(HERE, "golden", f"qe_example{ex}.npz"))


@pytest.mark.parametrize("ex", EXAMPLES)
def test_replicated_kpoints_match_reference(ex):
    qe = fixture(ex)
    reps = Q.replicate_kpoints(qe)
    printed = qe.expected["equivalent_kvectors"]
    got = {}
    for r in reps:
        got.setdefault(r.ik, []).append(list(r.kcrys) + [1.0 if r.trev else 0.0])
    assert len(reps) == sum(len(p) for p in printed)
    for ik, rows in enumerate(printed):
        assert np.allclose(np.array(got[ik]), np.array(rows), atol=1e-9), ik
    # QE weights = (2 / Nk) x replicas of each irreducible k (the reference's consistency check)
    mult = np.array([len(got[ik]) for ik in range(qe.nirr)])
    assert np.allclose(qe.kweight, 2.0 * mult / len(reps))


@pytest.mark.parametrize("ex", EXAMPLES)
def test_replicated_orbitals_are_the_irreducible_ones_reindexed(ex):
    """Each replica's plane waves are distinct grid points and the replicated bands stay
    orthonormal (the phase factors are unimodular, the G map a permutation)."""
    qe = fixture(ex)
    f = np.full((qe.nirr, qe.nbands), 2.0)
    sys_, reps = Q.build_system(qe, qe.coef, f, np.full(qe.nirr, qe.nbands), None, "HF")
    off = 0
    coff = 0
    for i, r in enumerate(reps):
        n = sys_.npw[i]
        g = sys_.gindex[off:off + n]
        assert len(np.unique(g)) == n
        B = sys_.phi_occ[coff:coff + qe.nbands * n].reshape(qe.nbands, n)
        assert np.abs(B.conj() @ B.T - np.eye(qe.nbands)).max() < 1e-9
        off += n
        coff += qe.nbands * n


@pytest.mark.parametrize("ex", EXAMPLES)
def test_initial_fermi_energy_matches_reference(ex):
    qe = fixture(ex)
    nb = np.full(qe.nirr, qe.nbands, np.int32)
    f, mu, nocc = tcx.set_filling(nb, qe.eig_qe.reshape(-1), qe.kweight, qe.nelec, qe.setup["smearing_mode"],
                                  qe.setup["smearing_width"])
    print(ex, mu * Q.HT_IN_EV, qe.expected["fermi_energy_initial_eV"])
    assert abs(mu * Q.HT_IN_EV - qe.expected["fermi_energy_initial_eV"]) < 1e-8
    # electron count with the QE weights
    assert abs(np.sum(qe.kweight[:, None] * f.reshape(qe.nirr, -1)) / 2.0 - qe.nelec) < 1e-9


@pytest.mark.parametrize("ex", [14, 15, 16])
def test_host_rotations_and_replica_bases(ex):
    """The C++ host's rotations and replica plane waves against a numpy restatement of
    symmetry.cpp:17-33 and plane_wave_basis.cpp:81-164: U = (QE rotation)^-1; at a replica,
    G' = +-U^T G of every irreducible plane wave, folded onto the grid, with phase
    exp(2 pi i G'.tau).  The replica's k itself is U^T k (or its negative) modulo the
    reciprocal lattice."""
    qe = fixture(ex)
    U = Q.tc_rotations(qe)
    for i, r in enumerate(qe.rot_qe):
        assert np.array_equal(U[i] @ r.astype(np.int64), np.eye(3, dtype=np.int64)), i
    n = np.array(qe.grid)
    kcrys = [np.linalg.solve(qe.b_norm.T, k) for k in qe.kcart]
    with Q.HostInputs(qe) as h:
        reps = h.replicas()
        assert [r.index for r in reps] == list(range(len(reps)))
        for r in reps:
            gidx, ph = h.replica_basis(r)
            Gs = (-1 if r.trev else 1) * (qe.mill[r.ik].astype(np.int64) @ U[r.isym])
            w = Gs % n
            assert np.array_equal(gidx, w[:, 0] + n[0] * (w[:, 1] + n[1] * w[:, 2])), r.index
            assert np.abs(ph - np.exp(2j * np.pi * (Gs @ qe.trans[r.isym]))).max() < 1e-12, r.index
            v = (-1 if r.trev else 1) * (U[r.isym].T @ kcrys[r.ik])
            d = v - r.kcrys
            assert np.abs(d - np.round(d)).max() < 1e-4, r.index
