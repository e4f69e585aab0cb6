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
  * per-band energies e_n = <phi|H_n|phi> (HF) and eig = e_1 + e_2 (+ e_3);
  * the bi-orthogonal variant (biortho_david = true, BITC): successive calls follow the numpy
    restatement (oracle/davidson.py) on the dense H_n and on the reference's H^dagger_n
    (u -> -u, right and left occupied orbitals swapped, density kept); a full subspace gives
    exact eigenvalues with right and left eigenvectors of H, bi-orthonormal.
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


def full_hamiltonians(sys_, density=None):
    """Per-k dense H_k (columns = H e_l) and the split H1/H2/H3 operators."""
    o = O.Oracle(sys_, density=density)
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


def hermitian_conj_hamiltonians(sys_):
    """Dense H^dagger_n as the reference applies it in the bi-orthogonal Davidson
    (Jastrow::take_Hermitian_conj and BlochStates::switch_left_right_orbitals,
    diagonalization_davidson.cpp:234-251): u -> -u, phi <-> chi, the density unchanged."""
    dens = O.Oracle(sys_).density()
    poly = sys_.jastrow_poly
    if poly is not None:
        poly = {**poly, "hermitian_conj": not poly.get("hermitian_conj", False)}
    sw = S.System(**{**sys_.__dict__, "A": -np.asarray(sys_.A), "phi_occ": sys_.chi_occ, "chi_occ": sys_.phi_occ,
                     "jastrow_poly": poly})
    return full_hamiltonians(sw, density=dens)


def test_davidson_biortho_follows_the_reference_algorithm(engine):
    """biortho_david = true: eigenvalues, band energies and both band sets of successive calls
    match block_davidson_biortho on the oracle's dense H_n and H^dagger_n.  Every call is
    compared with the restatement run on the same inputs (the GPU's previous outputs) to
    1e-10 (measured: 1.4e-12 at the first call, ~1e-15 after).  The two independent
    trajectories from the same start agree to 1e-9 over the first four calls; the
    bi-orthogonal iteration amplifies rounding differences about fivefold per call
    (measured 1.6e-12, 6e-12, 2e-11, 3.5e-11, 8e-11, then 3.5e-8 when a band jumps at
    the sixth call), so later calls are compared per call only."""
    from oracle import davidson as OD
    sys_ = system("BITC", 71)
    engine.load_system(sys_)
    mats = full_hamiltonians(sys_)
    dmats = hermitian_conj_hamiltonians(sys_)
    nbk = 2
    nb = np.full(sys_.nk, nbk, np.int32)
    phi = trial_bands(sys_, nb, 72)
    left = phi.copy()  # phik_left = phik at the start of a BITC run (bloch_states_initialize.cpp:370)
    eig = np.zeros(nb.sum(), np.complex128)
    bo = sys_.band_offsets(nb)
    ref = [[phi[bo[ik]:bo[ik + 1]].reshape(nbk, -1), left[bo[ik]:bo[ik + 1]].reshape(nbk, -1),
            eig[ik * nbk:(ik + 1) * nbk]] for ik in range(sys_.nk)]
    kG2 = [_kG2(sys_, ik) for ik in range(sys_.nk)]
    for it in range(6):
        same_in = [[phi[bo[ik]:bo[ik + 1]].reshape(nbk, -1), left[bo[ik]:bo[ik + 1]].reshape(nbk, -1),
                    eig[ik * nbk:(ik + 1) * nbk]] for ik in range(sys_.nk)]
        phi, left, eig, en = engine.davidson(nb, phi, eig, max_blocks=3, num_refresh=2, phi_left=left)
        for ik in range(sys_.nk):
            rp, rl, re, ren = OD.block_davidson_biortho(list(mats[ik]), list(dmats[ik]), kG2[ik], *same_in[ik],
                                                        max_blocks=3, num_refresh=2)
            got = eig[ik * nbk:(ik + 1) * nbk]
            print("biortho", it, ik, got, got - re, flush=True)
            assert np.abs(got - re).max() < 1e-10 * max(1.0, np.abs(re).max()), (it, ik)
            assert np.abs(en[ik * nbk:(ik + 1) * nbk] - ren).max() < 1e-10 * max(1.0, np.abs(ren).max())
            P = phi[bo[ik]:bo[ik + 1]].reshape(nbk, -1)
            L = left[bo[ik]:bo[ik + 1]].reshape(nbk, -1)
            for b in range(nbk):  # bands agree up to a phase per band (conjugate-inverse on the left)
                assert abs(abs(np.vdot(rp[b], P[b])) - np.linalg.norm(rp[b]) * np.linalg.norm(P[b])) < 1e-8
                assert abs(abs(np.vdot(rl[b], L[b])) - np.linalg.norm(rl[b]) * np.linalg.norm(L[b])) < 1e-8
                assert abs(np.linalg.norm(P[b]) - 1.0) < 1e-12
            assert np.abs(L.conj() @ P.T - np.eye(nbk)).max() < 1e-9
            # the restatement's own trajectory from the same start
            tp, tl, te, ten = OD.block_davidson_biortho(list(mats[ik]), list(dmats[ik]), kG2[ik], *ref[ik],
                                                        max_blocks=3, num_refresh=2)
            ref[ik] = [tp, tl, te]
            print("biortho trajectory", it, ik, np.abs(got - te).max(), flush=True)
            if it < 4:
                assert np.abs(got - te).max() < 1e-9 * max(1.0, np.abs(te).max()), (it, ik)


def test_davidson_biortho_full_subspace_is_exact(engine):
    """A subspace that spans the basis: the returned eigenvalues are eigenvalues of H, the
    right bands right eigenvectors, the left bands left eigenvectors (chi^H H = e chi^H),
    and <chi_a|phi_b> = delta_ab."""
    sys_ = system("BITC", 73)
    engine.load_system(sys_)
    mats = full_hamiltonians(sys_)
    nbk = 2
    nb = np.full(sys_.nk, nbk, np.int32)
    phi = trial_bands(sys_, nb, 74)
    blocks = int(np.ceil(max(sys_.npw) / nbk)) + 2
    phi_out, left, eig, en = engine.davidson(nb, phi, np.zeros(nb.sum(), np.complex128), max_blocks=blocks,
                                             phi_left=phi.copy())
    bo = sys_.band_offsets(nb)
    for ik in range(sys_.nk):
        M = sum(mats[ik])
        w = np.linalg.eigvals(M)
        got = eig[ik * nbk:(ik + 1) * nbk]
        P = phi_out[bo[ik]:bo[ik + 1]].reshape(nbk, -1)
        L = left[bo[ik]:bo[ik + 1]].reshape(nbk, -1)
        scale = max(1.0, np.abs(w).max())
        for b in range(nbk):
            assert np.abs(w - got[b]).min() < 1e-9 * scale, (ik, b, got[b])
            assert np.abs(M @ P[b] - got[b] * P[b]).max() < 1e-8 * scale
            assert np.abs(L[b].conj() @ M - got[b] * L[b].conj()).max() < 1e-8 * scale
        print("biortho full", ik, got, flush=True)
        assert np.abs(L.conj() @ P.T - np.eye(nbk)).max() < 1e-9
        assert np.abs(en[ik * nbk:(ik + 1) * nbk].sum(axis=1) - got).max() < 1e-9 * scale


def test_davidson_biortho_needs_bitc(engine):
    sys_ = system("TC", 75)
    engine.load_system(sys_)
    nb = np.full(sys_.nk, 2, np.int32)
    phi = trial_bands(sys_, nb, 76)
    with pytest.raises(tcx.TcxError) as ei:
        engine.davidson(nb, phi, np.zeros(nb.sum(), np.complex128), phi_left=phi.copy())
    assert ei.value.code == tcx.TCX_ERR_ARG
