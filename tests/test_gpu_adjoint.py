"""Gradients of the occupied-set energies (tcx_grad_exchange, tcx_grad_tc2)
against central finite differences (step 1e-4 along seeded random directions),
and the torch.autograd.Function wrapper of the test harness (CPU tensors; the
GPU work goes through libtcx).

E_x and E_2 are quartic in the orbitals (central-difference error O(h^2)) and
E_2 is quadratic in A at fixed F (central differences exact up to rounding)."""
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


def test_exchange_gradient_finite_differences(engine):
    sys_ = S.tiny_system(method="HF", n=(8, 8, 8), kmesh=(2, 1, 1), nbands=3, nocc=2, config=91)
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


@pytest.mark.parametrize("method", ["TC", "BITC"])
def test_tc2_gradients_finite_differences(engine, method):
    sys_ = S.tiny_system(method=method, n=(8, 8, 8), kmesh=(2, 1, 1), nbands=3, nocc=2, config=92)
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

    sys_ = S.tiny_system(method="HF", n=(6, 6, 6), kmesh=(2, 1, 1), nbands=2, nocc=1, config=93,
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
