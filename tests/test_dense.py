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
