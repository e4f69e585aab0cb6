"""Gradients of the occupied-set energies (tcx_grad_exchange, tcx_grad_tc2)
against central finite differences (step 1e-4 along seeded random directions),
and the torch.autograd.Function wrapper of the test harness (CPU tensors; the
GPU work goes through libtcx).

E_x and E_2 are quartic in the orbitals (central-difference error O(h^2)) and
E_2 is quadratic in A at fixed F (central differences exact up to rounding), also with a
polynomial part on top of the RPA one (its cross terms with the RPA part are linear in A)."""
import numpy as np
import pytest

from tc_amd import system as S
from tc_amd import tcx

pytestmark = pytest.mark.gpu
H = 1e-4


def _direction(n, seed):
    rng = np.random.default_rng(seed)
    return rng.standard_normal(n) + 1j * rng.standard_normal(n)


def _with(sys_, **kw):
    return S.System(**{**sys_.__dict__, **kw})


def _poly(sys_):
    """The RPA Jastrow of sys_ plus example 32's 6th-order polynomial (RPA+polynomial form)."""
    from test_oracle_jastrow_poly import example_spec
    return _with(sys_, jastrow_poly=example_spec(32)[0])


def test_exchange_gradient_finite_differences(engine):
    sys_ = S.tiny_system(method="HF", n=(8, 8, 8), kmesh=(2, 1, 1), nbands=3, nocc=2, seed_stream=91)
    engine.load_system(sys_)
    e0, g = engine.grad_exchange()
    assert abs(e0.imag) < 1e-14 * abs(e0.real) and e0.real < 0
    for seed in (1, 2):
        d = _direction(sys_.phi_occ.size, seed)
        ep = engine.load_system(_with(sys_, phi_occ=sys_.phi_occ + H * d)) or engine.grad_exchange()[0]
        em = engine.load_system(_with(sys_, phi_occ=sys_.phi_occ - H * d)) or engine.grad_exchange()[0]
        fd = (ep - em) / (2 * H)
        an = 2.0 * np.real(np.vdot(g, d))          # dE = 2 Re <g|d> for real E, g = dE/d conj(c)
        print("E_x directional derivative", fd, an, flush=True)
        assert abs(fd - an) < 1e-6 * max(1.0, abs(an))
    engine.load_system(sys_)


@pytest.mark.parametrize("poly", [False, True], ids=["rpa", "rpa_poly"])
@pytest.mark.parametrize("method", ["TC", "BITC"])
def test_tc2_gradients_finite_differences(engine, method, poly):
    sys_ = S.tiny_system(method=method, n=(8, 8, 8), kmesh=(2, 1, 1), nbands=3, nocc=2, seed_stream=92)
    if poly:
        sys_ = _poly(sys_)
    engine.load_system(sys_)
    e0, gl, gr, dA = engine.grad_tc2()
    chi = sys_.chi_occ if method == "BITC" else sys_.phi_occ
    d = _direction(sys_.phi_occ.size, 3)

    def energy(**kw):
        engine.load_system(_with(sys_, **kw))
        return engine.grad_tc2()[0]

    if method == "BITC":   # chi and phi independent: dE = <d|g_left> along chi, sum d g_right along phi
        fd_l = (energy(chi_occ=chi + H * d) - energy(chi_occ=chi - H * d)) / (2 * H)
        fd_r = (energy(phi_occ=sys_.phi_occ + H * d) - energy(phi_occ=sys_.phi_occ - H * d)) / (2 * H)
        an_l, an_r = np.vdot(d, gl), np.sum(d * gr)
        print("BITC", fd_l, an_l, fd_r, an_r, flush=True)
        assert abs(fd_l - an_l) < 1e-6 * max(1.0, abs(an_l))
        assert abs(fd_r - an_r) < 1e-6 * max(1.0, abs(an_r))
    else:                  # TC: chi = phi moves together
        fd = (energy(phi_occ=sys_.phi_occ + H * d) - energy(phi_occ=sys_.phi_occ - H * d)) / (2 * H)
        an = np.vdot(d, gl) + np.sum(d * gr)
        print("TC", fd, an, flush=True)
        assert abs(fd - an) < 1e-6 * max(1.0, abs(an))
    # dE/dA at fixed F
    fdA = (energy(A=sys_.A + H) - energy(A=sys_.A - H)) / (2 * H)
    print("dE/dA", fdA, dA, flush=True)
    assert abs(fdA - dA) < 1e-8 * max(1.0, abs(dA))
    engine.load_system(sys_)


def test_exchange_energy_autograd_function(engine):
    """torch.autograd.Function around the C-ABI energy and gradient, checked by gradcheck."""
    import torch

    sys_ = S.tiny_system(method="HF", n=(6, 6, 6), kmesh=(2, 1, 1), nbands=2, nocc=1, seed_stream=93,
                         lattice=S.sc_lattice(5.0))

    class ExchangeEnergy(torch.autograd.Function):
        @staticmethod
        def forward(ctx, re, im):
            c = re.detach().numpy() + 1j * im.detach().numpy()
            engine.load_system(_with(sys_, phi_occ=c))
            e, g = engine.grad_exchange()
            ctx.save_for_backward(torch.from_numpy(g.real.copy()), torch.from_numpy(g.imag.copy()))
            return torch.tensor(e.real, dtype=torch.float64)

        @staticmethod
        def backward(ctx, gout):
            gre, gim = ctx.saved_tensors
            return 2.0 * gout * gre, 2.0 * gout * gim   # dE/dRe c = 2 Re g, dE/dIm c = 2 Im g

    re = torch.tensor(sys_.phi_occ.real.copy(), dtype=torch.float64, requires_grad=True)
    im = torch.tensor(sys_.phi_occ.imag.copy(), dtype=torch.float64, requires_grad=True)
    assert torch.autograd.gradcheck(ExchangeEnergy.apply, (re, im), eps=1e-6, atol=1e-7, rtol=1e-5)
    engine.load_system(sys_)


def _half_energy(sys_, bra, h):
    e, off, b = 0.0, 0, 0
    for ik in range(sys_.nk):
        n = sys_.npw[ik]
        for i in range(sys_.nocc[ik]):
            e += sys_.filling[b] * np.vdot(bra[off:off + n], h[off:off + n])
            off += n
            b += 1
    return e / (2 * sys_.num_kpoints)


def _real_space(sys_, coef):
    """phi(r) = FFT_backward (unscaled) of the scattered coefficients, per occupied orbital."""
    n1, n2, n3 = sys_.n
    po, out, off = sys_.pw_offsets(), [], 0
    for ik in range(sys_.nk):
        for i in range(sys_.nocc[ik]):
            g = np.zeros(sys_.ngrid, np.complex128)
            g[sys_.gindex[po[ik]:po[ik + 1]]] = coef[off:off + sys_.npw[ik]]
            off += sys_.npw[ik]
            out.append((ik, np.fft.ifftn(g.reshape(n3, n2, n1)) * sys_.ngrid))
    return out


def test_exchange_kernel_gradient_bins_and_euler(engine):
    sys_ = S.tiny_system(method="HF", n=(8, 8, 8), kmesh=(2, 1, 1), nbands=3, nocc=2, seed_stream=94)
    engine.load_system(sys_, div_corr=False)
    e, _ = engine.grad_exchange()
    nbins, width = 48, 0.3
    gk, _, sums = engine.kernel_gradient("x", nbins, width)
    assert abs(sums[0] - e.real) < 1e-11 * abs(e)              # sum_x v dE/dv = E_x (linear in v)
    # the same bins from numpy FFTs of the orbitals
    orbs = _real_space(sys_, sys_.phi_occ)
    f = sys_.filling
    cx = -4 * np.pi / (sys_.num_kpoints * sys_.volume) / 2 / (2 * sys_.num_kpoints)
    n1, n2, n3 = sys_.n
    m = [np.where(np.arange(n) <= n // 2, np.arange(n), np.arange(n) - n) for n in (n1, n2, n3)]
    M3, M2, M1 = np.meshgrid(m[2], m[1], m[0], indexing="ij")
    ref = np.zeros(nbins)
    for a, (ka, pa) in enumerate(orbs):
        for b, (kb, pb) in enumerate(orbs):
            rho = np.fft.fftn(np.conj(pb) * pa) / sys_.ngrid
            kq = (sys_.kfrac[ka] - sys_.kfrac[kb])
            x = (kq[0] + M1)[..., None] * sys_.recip[0] + (kq[1] + M2)[..., None] * sys_.recip[1] + \
                (kq[2] + M3)[..., None] * sys_.recip[2]
            idx = np.minimum(nbins - 1, (np.sqrt(np.sum(x * x, axis=-1)) / width).astype(int))
            np.add.at(ref, idx.reshape(-1), (cx * f[a] * f[b] * np.abs(rho) ** 2).reshape(-1))
    assert np.abs(gk.real - ref).max() < 1e-11 * np.abs(ref).max()
    assert np.abs(gk.imag).max() == 0.0


@pytest.mark.parametrize("poly", [False, True], ids=["rpa", "rpa_poly"])
@pytest.mark.parametrize("method", ["TC", "BITC"])
@pytest.mark.parametrize("prec", [tcx.PREC_FP64, tcx.PREC_FP32])
def test_tc2_kernel_gradient_euler_and_dA(engine, method, prec, poly):
    sys_ = S.tiny_system(method=method, n=(8, 8, 8), kmesh=(2, 1, 1), nbands=3, nocc=2, seed_stream=95)
    if poly:
        sys_ = _poly(sys_)
    chi = sys_.chi_occ if method == "BITC" else sys_.phi_occ

    def e2x(s):
        engine.load_system(s, div_corr=False)
        return _half_energy(s, chi, engine.tc2x(s.nocc, s.phi_occ))

    e = e2x(sys_)
    gk, gu, sums = engine.kernel_gradient("tc2", 64, 0.25, precision=prec)
    lin, dA = complex(sums[0], sums[1]), complex(sums[2], sums[3])
    tol = 1e-11 if prec == tcx.PREC_FP64 else 2e-6
    print(method, prec, e, lin - e, flush=True)
    assert abs(lin - e) < tol * abs(e)                          # sum_x (V2 dE/dV2 + u dE/du) = E_2x
    assert abs(gk.sum() + gu.sum()) > 0
    fdA = (e2x(S.System(**{**sys_.__dict__, "A": sys_.A + H})) -
           e2x(S.System(**{**sys_.__dict__, "A": sys_.A - H}))) / (2 * H)
    print("dE/dA", fdA, dA, flush=True)
    assert abs(fdA - dA) < (1e-8 if prec == tcx.PREC_FP64 else 2e-6) * abs(fdA)
