"""GPU block Davidson (Diagonalization::block_davidson) against exact
diagonalisation of the full per-k Hamiltonian.

The full matrix H_k at fixed occupied orbitals is assembled from the FP64
oracle by applying every term of calc_hamiltonian::all to each plane-wave
unit vector, then diagonalised by LAPACK (numpy.linalg.eig).  Checks:
  * a subspace that can span the whole basis reproduces the lowest eigenvalues
    (by real part) to 1e-9 in one call;
  * repeated calls with the reference defaults (2 blocks, 1 refresh) converge
    to them;
  * eigen-residuals ||H phi - e phi|| of the returned bands;
  * BITC: left and right bands are bi-orthonormal;
  * per-band energies e_n = <phi|H_n|phi> (HF) and eig = e_1 + e_2 (+ e_3).
"""
import numpy as np
import pytest

from oracle import oracle as O
from tc_amd import system as S
from tc_amd import tcx

pytestmark = pytest.mark.gpu

TERMS = {"HF": (("kinetic",), ("hf2h", "hf2x"), ()),
         "TC": (("kinetic",), ("tc2h", "tc2x"), ("tc3a1", "tc3b1", "tc3a2a4b2b5", "tc3a3b3b4b6"))}
TERMS["BITC"] = TERMS["TC"]


def full_hamiltonians(sys_):
    """Per-k dense H_k (columns = H e_l) and the split H1/H2/H3 operators."""
    o = O.Oracle(sys_)
    nb = np.array(sys_.npw, np.int32)
    eye = np.concatenate([np.eye(n, dtype=np.complex128).reshape(-1) for n in sys_.npw])
    parts = []
    for group in TERMS[sys_.method]:
        acc = np.zeros_like(eye)
        for t in group:
            acc += getattr(o, t)(nb, eye)
        parts.append(acc)
    bo = sys_.band_offsets(nb)
    mats = []
    for ik in range(sys_.nk):
        n = sys_.npw[ik]
        mats.append([p[bo[ik]:bo[ik + 1]].reshape(n, n).T for p in parts])
    return mats


def lowest(M, nb):
    w = np.linalg.eigvals(M)
    return w[np.argsort(w.real, kind="stable")][:nb]


def trial_bands(sys_, nb, seed):
    """Seeded random bands, orthonormal per k (the Davidson's input contract)."""
    _, phi = S.applied_bands(sys_, nb=nb, seed=seed)
    bo = sys_.band_offsets(nb)
    out = phi.copy()
    for ik in range(sys_.nk):
        blk = phi[bo[ik]:bo[ik + 1]].reshape(nb[ik], -1)
        q, _ = np.linalg.qr(blk.T)
        out[bo[ik]:bo[ik + 1]] = q.T.reshape(-1)
    return out


def system(method, stream):
    return S.tiny_system(method=method, n=(8, 8, 8), kmesh=(2, 1, 1), nbands=3, nocc=2, seed_stream=stream)


@pytest.mark.parametrize("method", ["HF", "TC", "BITC"])
def test_davidson_full_subspace_is_exact(engine, method):
    sys_ = system(method, 61)
    engine.load_system(sys_)
    mats = full_hamiltonians(sys_)
    nbk = 2
    nb = np.full(sys_.nk, nbk, np.int32)
    phi = trial_bands(sys_, nb, 62)
    blocks = int(np.ceil(max(sys_.npw) / nbk)) + 2
    phi_out, left, eig, en = engine.davidson(nb, phi, np.zeros(nb.sum(), np.complex128), max_blocks=blocks)
    bo = sys_.band_offsets(nb)
    for ik in range(sys_.nk):
        M = sum(mats[ik])
        ref = lowest(M, nbk)
        got = eig[ik * nbk:(ik + 1) * nbk]
        print(method, ik, ref, got - ref, flush=True)
        assert np.abs(got - ref).max() < 1e-9 * max(1.0, np.abs(ref).max())
        P = phi_out[bo[ik]:bo[ik + 1]].reshape(nbk, -1)
        if method != "TC":  # TC bands are re-orthonormalised after the rotation (:530-537)
            res = M @ P.T - P.T * got[None, :]
            assert np.abs(res).max() < 1e-8
        if method == "BITC":
            L = left[bo[ik]:bo[ik + 1]].reshape(nbk, -1)
            assert np.abs(L.conj() @ P.T - np.eye(nbk)).max() < 1e-9
        # eigenvalue = e_1 + e_2 + e_3 of the band (diag X_L^H H X = D)
        assert np.abs(en[ik * nbk:(ik + 1) * nbk].sum(axis=1) - got).max() < 1e-9 * max(1.0, np.abs(got).max())


def _kG2(sys_, ik):
    sl = slice(sys_.pw_offsets()[ik], sys_.pw_offsets()[ik + 1])
    kg = (sys_.kfrac[ik] + S.miller_of_index(sys_.n)[sys_.gindex[sl]]) @ sys_.recip
    return np.sum(kg * kg, axis=1)


@pytest.mark.parametrize("method", ["HF", "TC"])
def test_davidson_default_blocks_converge(engine, method):
    """Repeated calls with the reference defaults (2 blocks, 1 refresh), eigenvalues fed back.
    HF converges to the exact eigenvalues.  For TC the reference re-orthonormalises the
    Ritz vectors of the non-Hermitian H (diagonalization_davidson.cpp:530-537), so with
    these defaults only the lowest band converges; the second stops at a fixed point of
    the algorithm (5.8e-5 Ha off at k=1 here).  The numpy restatement of block_davidson
    run for the same number of calls stops at the same point, which the GPU must match."""
    from oracle import davidson as OD
    sys_ = system(method, 63)
    engine.load_system(sys_)
    mats = full_hamiltonians(sys_)
    nbk = 2
    nb = np.full(sys_.nk, nbk, np.int32)
    phi0 = trial_bands(sys_, nb, 64)
    phi = phi0
    bo = sys_.band_offsets(nb)
    eig = np.zeros(nb.sum(), np.complex128)
    calls = 120
    for it in range(calls):
        phi, _, eig, en = engine.davidson(nb, phi, eig)
    for ik in range(sys_.nk):
        ref = lowest(sum(mats[ik]), nbk)
        got = eig[ik * nbk:(ik + 1) * nbk]
        print(method, ik, got - ref, flush=True)
        if method == "HF":
            assert np.abs(got - ref).max() < 1e-7
            continue
        assert abs(got[0] - ref[0]) < 1e-7
        P, e = phi0[bo[ik]:bo[ik + 1]].reshape(nbk, -1), np.zeros(nbk, complex)
        for it in range(calls):
            P, _, e, _ = OD.block_davidson(list(mats[ik]), _kG2(sys_, ik), P, e)
        assert np.abs(got - e).max() < 1e-9, (got - e)


@pytest.mark.parametrize("method", ["HF", "TC", "BITC"])
def test_davidson_follows_the_reference_algorithm(engine, method):
    """Eigenvalues, band energies and bands of successive calls match the numpy
    restatement of block_davidson (oracle/davidson.py) run on the oracle's dense H_n."""
    from oracle import davidson as OD
    sys_ = system(method, 67)
    engine.load_system(sys_)
    mats = full_hamiltonians(sys_)
    nbk = 2
    nb = np.full(sys_.nk, nbk, np.int32)
    phi = trial_bands(sys_, nb, 68)
    eig = np.zeros(nb.sum(), np.complex128)
    bo = sys_.band_offsets(nb)
    ref_phi = [phi[bo[ik]:bo[ik + 1]].reshape(nbk, -1) for ik in range(sys_.nk)]
    ref_eig = [eig[ik * nbk:(ik + 1) * nbk] for ik in range(sys_.nk)]
    kG2 = [_kG2(sys_, ik) for ik in range(sys_.nk)]
    for it in range(6):
        phi, left, eig, en = engine.davidson(nb, phi, eig, max_blocks=3, num_refresh=2)
        for ik in range(sys_.nk):
            Hn = list(mats[ik]) + ([np.zeros_like(mats[ik][0])] if method == "HF" else [])
            rp, rl, re, ren = OD.block_davidson(Hn, kG2[ik], ref_phi[ik], ref_eig[ik], max_blocks=3, num_refresh=2,
                                                bitc=method == "BITC")
            ref_phi[ik], ref_eig[ik] = rp, re
            got = eig[ik * nbk:(ik + 1) * nbk]
            assert np.abs(got - re).max() < 1e-10 * max(1.0, np.abs(re).max()), (it, ik)
            assert np.abs(en[ik * nbk:(ik + 1) * nbk] - ren).max() < 1e-10 * max(1.0, np.abs(ren).max())
            # bands agree up to one phase per band (eigenvectors carry an arbitrary phase)
            P = phi[bo[ik]:bo[ik + 1]].reshape(nbk, -1)
            for b in range(nbk):
                ov = np.vdot(rp[b], P[b])
                assert abs(abs(ov) - np.linalg.norm(rp[b]) * np.linalg.norm(P[b])) < 1e-8
        print(method, it, eig, flush=True)


def test_davidson_hf_band_energies(engine):
    """HF: e_1 = <phi|T|phi>, e_2 = <phi|V_H + V_x|phi> for the returned unit-norm bands."""
    sys_ = system("HF", 65)
    engine.load_system(sys_)
    mats = full_hamiltonians(sys_)
    nbk = 2
    nb = np.full(sys_.nk, nbk, np.int32)
    phi = trial_bands(sys_, nb, 66)
    blocks = int(np.ceil(max(sys_.npw) / nbk)) + 2
    phi_out, _, eig, en = engine.davidson(nb, phi, np.zeros(nb.sum(), np.complex128), max_blocks=blocks)
    bo = sys_.band_offsets(nb)
    for ik in range(sys_.nk):
        P = phi_out[bo[ik]:bo[ik + 1]].reshape(nbk, -1)
        for b in range(nbk):
            for n in range(2):
                ref = np.vdot(P[b], mats[ik][n] @ P[b])
                assert abs(en[ik * nbk + b, n] - ref) < 1e-9
    tot = tcx.total_energy(nb, np.full(nb.sum(), 2.0), en, sys_.num_kpoints, False)
    assert abs(tot[3] - (tot[0] + tot[1])) < 1e-14
