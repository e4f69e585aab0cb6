"""Pins the FP64 CPU oracle (oracle/) before anything is compared against it.

FFT: textbook DFT (naive O(N^2) sum), numpy.fft, round trip and Parseval.
Terms: the golden fixtures written by a brute-force G-space restatement with
no FFT (tests/golden/), the plane-wave closed form of the exchange operator,
and operator invariants chosen so that a dropped term, a wrong sign or a
transposed operand fails at least one of them.
"""
import numpy as np
import pytest

from conftest import golden_names
from oracle import oracle as O
from tc_amd import system as S
from tc_amd import energy as E
import make_golden as MG


# ------------------------------------------------------------------ FFT
@pytest.mark.parametrize("shape", [(4, 6, 5), (8, 8, 8), (3, 10, 6), (6, 6, 6), (9, 5, 7)])
@pytest.mark.parametrize("sign", [-1, 1])
def test_fft_matches_naive_dft(shape, sign):
    rng = np.random.default_rng(1)
    g = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
    ref = O.dft3d_naive(g, sign)
    got = O.fft3d(g, sign)
    assert np.abs(got - ref).max() <= 1e-13 * max(1.0, np.abs(ref).max())
    # FFTW conventions: forward exp(-iGr)/N, backward exp(+iGr)
    npref = np.fft.fftn(g) / g.size if sign < 0 else np.fft.ifftn(g) * g.size
    assert np.abs(got - npref).max() <= 1e-12 * max(1.0, np.abs(npref).max())


def test_fft_roundtrip_and_parseval():
    rng = np.random.default_rng(2)
    g = rng.standard_normal((12, 10, 8)) + 1j * rng.standard_normal((12, 10, 8))
    f = O.fft3d(g, -1)
    back = O.fft3d(f, +1)
    assert np.abs(back - g).max() < 1e-13
    assert abs(np.sum(np.abs(g) ** 2) / g.size - np.sum(np.abs(f) ** 2)) < 1e-10


# ------------------------------------------------------------------ golden fixtures
@pytest.mark.parametrize("name", golden_names())
def test_oracle_matches_golden(golden, name):
    z = golden[name]
    sys_ = MG.load_system(z)
    o = O.Oracle(sys_)
    assert abs(o.P.alpha - float(z["alpha"])) < 1e-14
    assert np.abs(o.svaux - z["sum_vaux"]).max() < 1e-10 * np.abs(z["sum_vaux"]).max()
    nb, phi = z["nb"], z["phi"]
    terms = (["hf2x", "hf2h", "kinetic"] if str(z["method"]) == "HF" else
             ["tc2x", "tc2h", "tc3a1", "tc3b1", "tc3a3b3b4b6", "tc3a2a4b2b5", "kinetic"])
    for t in terms:
        got = getattr(o, t)(nb, phi)
        ref = z[t]
        assert np.abs(got - ref).max() <= 1e-12 * np.abs(ref).max(), t
    if "hf2x_nodiv" in z:
        o2 = O.Oracle(sys_, div_corr=False)
        got = o2.hf2x(nb, phi)
        assert np.abs(got - z["hf2x_nodiv"]).max() <= 1e-12 * np.abs(z["hf2x_nodiv"]).max()


# ------------------------------------------------------------------ closed forms and invariants
def plane_wave_system(method="HF", n=(8, 8, 8), kmesh=(2, 2, 1), nocc=3, lattice=None, A=None):
    """Occupied orbitals are single plane waves (the lowest |q+G|), so the
    exchange operator is diagonal with a closed form and the density is uniform."""
    sys_ = S.tiny_system(method=method, n=n, kmesh=kmesh, nbands=nocc, nocc=nocc,
                         lattice=S.sc_lattice(6.0) if lattice is None else lattice, A=A, seed_stream=77)
    po = sys_.pw_offsets()
    blocks = []
    for ik in range(sys_.nk):
        c = np.zeros((nocc, sys_.npw[ik]), dtype=np.complex128)
        for b in range(nocc):
            c[b, b] = np.exp(0.7j * b)  # unit modulus, arbitrary phase
        blocks.append(c.reshape(-1))
    sys_.phi_occ = np.concatenate(blocks)
    if method == "BITC":
        sys_.chi_occ = sys_.phi_occ.copy()
    return sys_


def test_hf2x_plane_wave_closed_form():
    sys_ = plane_wave_system()
    sys_.div_corr = False
    o = O.Oracle(sys_)
    nb = sys_.npw.copy()
    # apply to every plane wave of the sphere at every k: columns of the identity
    phi = np.concatenate([np.eye(sys_.npw[k], dtype=np.complex128).reshape(-1) for k in range(sys_.nk)])
    h = o.hf2x(nb, phi)
    cx = -4 * np.pi / (sys_.num_kpoints * sys_.volume) / 2.0
    bo = sys_.band_offsets(nb)
    fo = np.concatenate([[0], np.cumsum(sys_.nocc)])
    for ik in range(sys_.nk):
        kG = sys_.gcart_at_k(ik)
        H = h[bo[ik]:bo[ik + 1]].reshape(nb[ik], sys_.npw[ik])
        expect = np.zeros(sys_.npw[ik])
        for iq in range(sys_.nk):
            qG = sys_.gcart_at_k(iq)
            for b in range(sys_.nocc[iq]):
                d2 = np.sum((kG - qG[b]) ** 2, axis=1)
                with np.errstate(divide="ignore"):
                    expect += sys_.filling[fo[iq] + b] * np.where(d2 < 1e-8, 0.0, 1.0 / d2)
        assert np.abs(H - np.diag(cx * expect)).max() < 1e-12 * np.abs(cx * expect).max()


def _random_bands(sys_, seed):
    return S.applied_bands(sys_, nb=sys_.nocc + 1, seed=seed)


@pytest.mark.parametrize("div", [False, True])
def test_hf2x_hermitian_and_negative(div):
    sys_ = S.tiny_system(method="HF", n=(8, 8, 8), kmesh=(2, 1, 1), nbands=3, nocc=2)
    sys_.div_corr = div
    o = O.Oracle(sys_)
    nb, phi = _random_bands(sys_, 5)
    h = o.hf2x(nb, phi)
    bo = sys_.band_offsets(nb)
    for ik in range(sys_.nk):
        P = phi[bo[ik]:bo[ik + 1]].reshape(nb[ik], -1)
        H = h[bo[ik]:bo[ik + 1]].reshape(nb[ik], -1)
        M = np.conj(P) @ H.T  # M[a, b] = <a|Vx b>
        assert np.abs(M - M.conj().T).max() < 1e-13 * np.abs(M).max()
        if not div:
            assert np.linalg.eigvalsh(0.5 * (M + M.conj().T)).max() < 1e-14


def test_exchange_energy_unitary_invariance():
    """E_x is invariant under a unitary rotation of the occupied bands at one k
    (equal fillings), which fixes the pair structure phi_qi^* ... phi_qi."""
    sys_ = S.tiny_system(method="HF", n=(8, 8, 8), kmesh=(2, 1, 1), nbands=3, nocc=3)
    o = O.Oracle(sys_)
    h = o.hf2x(sys_.nocc, sys_.phi_occ)
    e0 = E.term_energy(sys_, h, 2)
    rng = np.random.default_rng(3)
    X = np.linalg.qr(rng.standard_normal((3, 3)) + 1j * rng.standard_normal((3, 3)))[0]
    off = sys_.occ_offsets()
    rot = sys_.phi_occ.copy()
    blk = rot[off[0]:off[1]].reshape(3, -1)
    rot[off[0]:off[1]] = (X.T @ blk).reshape(-1)
    sys_.phi_occ = rot
    o2 = O.Oracle(sys_)
    e1 = E.term_energy(sys_, o2.hf2x(sys_.nocc, sys_.phi_occ), 2)
    assert abs(e0 - e1) < 1e-13 * abs(e0)
    assert e0.real < 0 and abs(e0.imag) < 1e-14 * abs(e0)


def test_tc2x_reduces_to_hf2x_without_jastrow():
    sys_ = S.tiny_system(method="TC", n=(8, 8, 8), kmesh=(2, 1, 1), nbands=3, nocc=2, A=np.zeros((2, 2)))
    assert sys_.jastrow_A_zero
    o = O.Oracle(sys_)
    nb, phi = _random_bands(sys_, 6)
    a = o.tc2x(nb, phi)
    b = o.hf2x(nb, phi)
    assert np.abs(a - b).max() < 1e-13 * np.abs(b).max()


def test_tc_terms_without_jastrow():
    """A = 0: tc2h is the Hartree term hf2h and every three-body term vanishes."""
    sys_ = S.tiny_system(method="TC", n=(8, 8, 8), kmesh=(2, 1, 1), nbands=3, nocc=2, A=np.zeros((2, 2)))
    o = O.Oracle(sys_)
    nb, phi = _random_bands(sys_, 7)
    a = o.tc2h(nb, phi)
    b = o.hf2h(nb, phi)
    assert np.abs(a - b).max() < 1e-13 * np.abs(b).max()
    for t in ("tc3a3b3b4b6", "tc3a2a4b2b5"):
        assert np.abs(getattr(o, t)(nb, phi)).max() == 0.0, t


def test_three_body_exchange_needs_divergence_flag():
    """The [3b6] and [3a2] G=0 corrections are O(A^2) and switch off with the
    divergence-correction flag: the difference is non-zero and scales as lambda^2."""
    d = []
    for lam in (1.0, 2.0):
        s_on = _scaled(lam)
        nb, phi = _random_bands(s_on, 5)
        on = O.Oracle(s_on)
        off = O.Oracle(s_on, div_corr=False)
        d.append(sum(getattr(on, t)(nb, phi) - getattr(off, t)(nb, phi) for t in ("tc3a3b3b4b6", "tc3a2a4b2b5")))
    assert np.abs(d[0]).max() > 1e-6 * np.abs(d[1]).max()
    assert np.abs(d[1] - 4.0 * d[0]).max() < 1e-12 * np.abs(d[1]).max()


def test_bitc_with_equal_orbitals_is_tc():
    base = S.tiny_system(method="TC", n=(8, 8, 8), kmesh=(2, 1, 1), nbands=3, nocc=2)
    bitc = S.tiny_system(method="TC", n=(8, 8, 8), kmesh=(2, 1, 1), nbands=3, nocc=2)
    bitc.method = "BITC"
    bitc.chi_occ = bitc.phi_occ.copy()
    nb, phi = _random_bands(base, 8)
    for t in ("tc2x", "tc2h", "tc3a1", "tc3b1", "tc3a3b3b4b6", "tc3a2a4b2b5"):
        a = getattr(O.Oracle(base), t)(nb, phi)
        b = getattr(O.Oracle(bitc), t)(nb, phi)
        assert np.abs(a - b).max() < 1e-14 * max(1e-300, np.abs(a).max()), t


def _scaled(A_scale, method="TC"):
    sys_ = S.tiny_system(method=method, n=(8, 8, 8), kmesh=(2, 1, 1), nbands=3, nocc=2,
                         A=np.full((2, 2), 0.5))
    sys_.A = sys_.A * A_scale  # F held fixed: u -> lambda u
    return sys_


@pytest.mark.parametrize("term,degree", [("tc2x", 2), ("tc2h", 2), ("tc3a1", 2), ("tc3b1", 2),
                                         ("tc3a3b3b4b6", 2), ("tc3a2a4b2b5", 2)])
def test_polynomial_in_jastrow_strength(term, degree):
    """At fixed F, u -> lambda*u makes tc2x a quadratic and the three-body terms
    pure quadratics in lambda: third differences vanish, and tc3 scales as lambda^2."""
    base = _scaled(1.0)
    nb, phi = _random_bands(base, 9)
    vals = [getattr(O.Oracle(_scaled(l)), term)(nb, phi) for l in (0.5, 1.0, 1.5, 2.0)]
    third = vals[3] - 3 * vals[2] + 3 * vals[1] - vals[0]
    assert np.abs(third).max() < 1e-10 * np.abs(vals[3]).max()
    if term.startswith("tc3"):
        assert np.abs(vals[3] - 4.0 * vals[1]).max() < 1e-12 * np.abs(vals[3]).max()


def test_three_body_local_terms_vanish_for_uniform_density():
    sys_ = plane_wave_system(method="TC", A=np.full((2, 2), 0.5))
    o = O.Oracle(sys_)
    dens = o.density()
    assert np.abs(dens - dens[0]).max() < 1e-12
    nb, phi = _random_bands(sys_, 4)
    assert np.abs(o.tc3a1(nb, phi)).max() < 1e-16
    assert np.abs(o.tc3b1(nb, phi)).max() < 1e-16


def test_density_normalisation():
    """sum_r density_s / (N Nk) = electrons per spin (bloch_states_scfloop_filling_density.cpp:363-364)."""
    sys_ = S.tiny_system(method="TC", n=(8, 8, 8), kmesh=(2, 1, 1), nbands=3, nocc=2)
    d = O.Oracle(sys_).density()
    N = sys_.ngrid
    per_spin = np.sum(d[:N]).real / (N * sys_.num_kpoints)
    assert abs(per_spin - sys_.num_electrons() / 2) < 1e-12


def test_jastrow_kernels():
    """uk_RPA is the Fourier transform of (A/r)(1-exp(-r/F)); tc_2body_RPA's
    -(grad u)^2 part is checked by radial quadrature of 4 pi int r^2 u'(r)^2 j0(Gr) dr."""
    sys_ = S.tiny_system(method="TC", A=np.full((2, 2), 0.5))
    o = O.Oracle(sys_)
    A, F = 0.5, 1.0
    r = np.linspace(1e-6, 60.0, 600001)
    dr = r[1] - r[0]
    u = A / r * (1 - np.exp(-r / F))
    du = -A / r ** 2 * (1 - np.exp(-r / F)) + A / (r * F) * np.exp(-r / F)
    for G in (0.3, 1.0, 2.5):
        j0 = np.sin(G * r) / (G * r)
        ft_u = 4 * np.pi * np.sum(r ** 2 * u * j0) * dr
        # the long-range A/r tail needs the analytic remainder: FT of A/r is 4 pi A/G^2
        ft_tail = 4 * np.pi * A / G ** 2 - 4 * np.pi * np.sum(r ** 2 * (A / r) * j0) * dr
        assert abs(o.uk(G * G) - (ft_u + ft_tail)) < 2e-3 * abs(o.uk(G * G))
        grad2 = 4 * np.pi * np.sum(r ** 2 * du * du * j0) * dr
        # tc_2body = 4 pi/G^2 + nabla^2 u - (grad u)^2 with nabla^2 u -> -G^2 uk
        expect = 4 * np.pi / G ** 2 - G ** 2 * o.uk(G * G) - grad2
        assert abs(o.tc_2body(G * G) - expect) < 1e-4 * abs(expect)
    assert abs(o.uk(0.0) - (-4 * np.pi * A * F * F)) < 1e-15
    assert abs(o.tc_2body(0.0) - (-2 * np.pi * A * A / F)) < 1e-14


@pytest.mark.parametrize("method", ["TC", "BITC"])
def test_two_body_adjoint_is_minus_u_with_swapped_orbitals(method):
    """<a|H[u; phi, chi] b> = conj <b|H[-u; chi, phi] a> for tc2x + tc2h
    (src/diagonalization_davidson.cpp:234-251: take_Hermitian_conj + switch_left_right)."""
    sys_ = S.tiny_system(method=method, n=(8, 8, 8), kmesh=(2, 1, 1), nbands=3, nocc=2, seed_stream=81)
    adj = S.System(**{**sys_.__dict__, "A": -sys_.A})
    if method == "BITC":
        adj.phi_occ, adj.chi_occ = sys_.chi_occ.copy(), sys_.phi_occ.copy()
    nb, a = _random_bands(sys_, 82)
    _, b = _random_bands(sys_, 83)
    o, oa = O.Oracle(sys_), O.Oracle(adj)
    Hb = o.tc2x(nb, b) + o.tc2h(nb, b)
    Ha = oa.tc2x(nb, a) + oa.tc2h(nb, a)
    lhs, rhs = np.vdot(a, Hb), np.conj(np.vdot(b, Ha))
    assert abs(lhs - rhs) < 1e-12 * abs(lhs)


def test_hf2x_batched_restatement(golden):
    """The CPU baseline's batched SciPy restatement of hf2x (oracle/hf2x_batched.py) against
    the golden fixtures held without the divergence correction and against orc_hf2x, on
    ragged band counts (an empty k included) and two grids."""
    from oracle import hf2x_batched as HB
    for name in [n for n in golden_names() if n.startswith("hf_")]:
        z = golden[name]
        sys_ = MG.load_system(z)
        got = HB.hf2x(sys_, z["nb"], z["phi"], workers=2)
        ref = z["hf2x_nodiv"]
        assert np.abs(got - ref).max() <= 1e-12 * np.abs(ref).max(), name
    for n, km, nb in (((8, 8, 8), (2, 2, 1), [3, 0, 2, 4]), ((12, 12, 12), (2, 1, 1), [1, 4])):
        sys_ = S.tiny_system(method="HF", n=n, kmesh=km, nbands=4, nocc=3, div_corr=False)
        nb, phi = S.applied_bands(sys_, nb=nb, seed=5)
        ref = O.Oracle(sys_).hf2x(nb, phi)
        got = HB.hf2x(sys_, nb, phi, workers=3)
        assert np.abs(got - ref).max() <= 1e-13 * np.abs(ref).max(), n
