"""Host-side pieces of the Davidson consumer (no GPU): the dense complex
eigensolver against numpy.linalg.eig (LAPACK zgeev) and the total-energy sum
of TotalEnergy::calc_total_energy (src/total_energy.cpp:98-163)."""
import numpy as np
import pytest

from tc_amd import tcx


@pytest.mark.parametrize("n", [1, 2, 3, 7, 32, 64])
@pytest.mark.parametrize("kind", ["general", "hermitian", "near_hermitian"])
def test_eig_dense_against_lapack(n, kind):
    rng = np.random.default_rng(100 + n)
    A = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    if kind != "general":
        A = 0.5 * (A + A.conj().T)
    if kind == "near_hermitian":  # the shape of a TC subspace Hamiltonian
        A = A + 1e-2 * (rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n)))
    w, X = tcx.eig_dense(A)
    scale = np.abs(A).max()
    # A X = X diag(w), unit columns
    assert np.abs(A @ X - X * w[None, :]).max() < 1e-11 * scale * n
    assert np.allclose(np.linalg.norm(X, axis=0), 1.0, atol=1e-13)
    ref = np.linalg.eigvals(A)
    assert np.abs(np.sort_complex(w) - np.sort_complex(ref)).max() < 1e-10 * scale * n
    if kind == "hermitian":
        assert np.abs(w.imag).max() < 1e-11 * scale


def test_eig_dense_triangular_and_diagonal():
    T = np.triu(np.arange(1, 26).reshape(5, 5).astype(complex))
    w, X = tcx.eig_dense(T)
    assert np.allclose(np.sort(w.real), [1, 7, 13, 19, 25])
    assert np.abs(T @ X - X * w[None, :]).max() < 1e-11
    w, X = tcx.eig_dense(np.diag([3.0, -1.0, 2.0]).astype(complex))
    assert np.allclose(np.sort(w.real), [-1, 2, 3])


def test_total_energy_sum():
    nb = np.array([2, 1], np.int32)
    f = np.array([2.0, 1.0, 2.0])
    e = np.array([[1 + 1j, 4, 9], [2, 6 - 2j, 3], [0.5, 1, 1.5j]], np.complex128)
    out = tcx.total_energy(nb, f, e, num_kpoints=2, uses_3body=True)
    e1 = (2 * e[0, 0] + 1 * e[1, 0] + 2 * e[2, 0]) / 2
    e2 = (2 * e[0, 1] + 1 * e[1, 1] + 2 * e[2, 1]) / 2 / 2
    e3 = (2 * e[0, 2] + 1 * e[1, 2] + 2 * e[2, 2]) / 3 / 2
    assert np.allclose(out, [e1, e2, e3, e1 + e2 + e3], atol=1e-15)
    out = tcx.total_energy(nb, f, e, num_kpoints=2, uses_3body=False)
    assert out[2] == 0 and np.isclose(out[3], e1 + e2)


# ---------------------------------------------------------------- fillings (no GPU)
from oracle import oracle as O


def _spectrum(seed, nk=4, nb=6, gap=0.3):
    rng = np.random.default_rng(seed)
    e = np.sort(rng.uniform(-1.0, 0.0, (nk, nb)), axis=1)
    e[:, 3:] += gap  # an insulating gap above band 3
    return np.full(nk, nb, np.int32), e.reshape(-1), np.full(nk, 2.0 / nk)


@pytest.mark.parametrize("smearing,width", [("fixed", 0.0), ("gaussian", 0.05), ("gaussian", 0.3)])
def test_fillings_match_oracle(smearing, width):
    nb, e, kw = _spectrum(7)
    f, mu, nocc = tcx.set_filling(nb, e, kw, 6.0, smearing, width if width else 0.01)
    fo, muo, nocco = O.fillings(nb, e, kw, 6.0, smearing, width if width else 0.01)
    assert np.abs(f - fo).max() < 1e-12 and abs(mu - muo) < 1e-12 and (nocc == nocco).all()
    # electrons per cell: sum_k sum_b f kweight / 2
    assert abs(sum(f.reshape(len(nb), -1).sum(axis=1) * kw / 2) - 6.0) < 1e-8


def test_fillings_fixed_and_narrow_gaussian_agree_for_a_gap():
    nb, e, kw = _spectrum(8)
    f0, mu0, n0 = tcx.set_filling(nb, e, kw, 6.0, "fixed")
    f1, mu1, n1 = tcx.set_filling(nb, e, kw, 6.0, "gaussian", 1e-3)
    assert (n0 == 3).all() and np.abs(f0 - f1).max() < 1e-12
    assert mu0 == e.reshape(len(nb), -1)[:, 2].max()          # valence-band top (:56-59)
    assert e.reshape(len(nb), -1)[:, 2].max() < mu1 < e.reshape(len(nb), -1)[:, 3].min()


def test_fillings_errors_like_the_reference():
    nb, e, kw = _spectrum(9)
    with pytest.raises(tcx.TcxError):
        tcx.set_filling(nb, e, kw, 14.0, "fixed")             # more occupied than bands
    with pytest.raises(tcx.TcxError):
        tcx.set_filling(nb, e, kw, 5.0, "fixed")              # odd electron count
    with pytest.raises(ValueError):
        O.fillings(nb, e, kw, 14.0, "fixed")
