"""GPU parity of the TC/BITC terms (tc2x, tc3a1, tc3b1) through the C ABI.

Tolerances:
  FP64 path : max |delta| <= 1e-10 * max |ref|
  FP32 path (tc2x pair pipeline): per band ||delta||/||ref|| <= 3e-5, |delta E| <= 1e-6 Ha
  tc3a1/tc3b1 run in FP64 for either precision flag: 1e-10.
"""
import numpy as np
import pytest

from conftest import golden_names
from oracle import oracle as O
from tc_amd import energy as E
from tc_amd import system as S
from tc_amd import tcx
import make_golden as MG
from test_gpu_hf2x import per_band_rel

pytestmark = pytest.mark.gpu

TC_GOLDEN = [n for n in golden_names() if not n.startswith("hf_")]


@pytest.mark.parametrize("name", TC_GOLDEN)
@pytest.mark.parametrize("prec", [tcx.PREC_FP64, tcx.PREC_FP32])
def test_tc_terms_golden(engine, golden, name, prec):
    z = golden[name]
    sys_ = MG.load_system(z)
    engine.load_system(sys_)
    for term in ("tc2x", "tc3a1", "tc3b1"):
        got = getattr(engine, term)(z["nb"], z["phi"], precision=prec)
        ref = z[term]
        if prec == tcx.PREC_FP64 or term != "tc2x":
            assert np.abs(got - ref).max() <= 1e-10 * np.abs(ref).max(), term
        else:
            assert per_band_rel(sys_, z["nb"], got, ref).max() <= 3e-5, term


@pytest.mark.parametrize("method", ["TC", "BITC"])
def test_tc2x_against_oracle_12(engine, method):
    sys_ = S.tiny_system(method=method, n=(12, 12, 12), kmesh=(2, 2, 1), nbands=4, nocc=3, seed_stream=31)
    engine.load_system(sys_)
    nb, phi = S.applied_bands(sys_, nb=[4, 3, 2, 4], seed=32)
    o = O.Oracle(sys_)
    for term in ("tc2x", "tc3a1", "tc3b1"):
        ref = getattr(o, term)(nb, phi)
        got64 = getattr(engine, term)(nb, phi, precision=tcx.PREC_FP64)
        assert np.abs(got64 - ref).max() <= 1e-10 * np.abs(ref).max(), term
        got32 = getattr(engine, term)(nb, phi, precision=tcx.PREC_FP32)
        assert per_band_rel(sys_, nb, got32, ref).max() <= 3e-5, term


def test_tc2x_without_jastrow_equals_hf2x(engine):
    sys_ = S.tiny_system(method="TC", n=(8, 8, 8), kmesh=(2, 1, 1), nbands=3, nocc=2, A=np.zeros((2, 2)))
    engine.load_system(sys_)
    nb, phi = S.applied_bands(sys_, nb=[3, 3], seed=33)
    a = engine.tc2x(nb, phi)
    hf = S.System(**{**sys_.__dict__, "method": "HF"})
    engine.load_system(hf)
    b = engine.hf2x(nb, phi)
    assert np.abs(a - b).max() <= 1e-11 * np.abs(b).max()


def test_tc_energies_fp32_vs_oracle(engine):
    """Two- and three-body energies of the occupied bands, FP32-split vs FP64 oracle."""
    sys_ = S.tiny_system(method="BITC", n=(12, 12, 12), kmesh=(2, 2, 1), nbands=3, nocc=3, seed_stream=34)
    engine.load_system(sys_)
    o = O.Oracle(sys_)
    for term, nbody in (("tc2x", 2), ("tc3a1", 3), ("tc3b1", 3)):
        ref = getattr(o, term)(sys_.nocc, sys_.phi_occ)
        got = getattr(engine, term)(sys_.nocc, sys_.phi_occ, precision=tcx.PREC_FP32)
        e_ref = E.term_energy(sys_, ref, nbody, left=True)
        e_got = E.term_energy(sys_, got, nbody, left=True)
        assert abs(e_got - e_ref) < 1e-6, term


def test_tc_requires_jastrow_and_method(engine):
    sys_ = S.tiny_system(method="HF")
    engine.load_system(sys_)
    nb, phi = S.applied_bands(sys_, nb=[1, 1], seed=35)
    with pytest.raises(tcx.TcxError) as ei:
        engine.tc2x(nb, phi)
    assert ei.value.code == tcx.TCX_ERR_STATE


def test_density_get_set(engine):
    """tcx_get_density / tcx_set_density (BlochStates::density_): the density formed from
    the occupied set integrates to the electron count; the terms read a replaced density
    (the Hartree term hf2h is linear in it, tc3a1 is not: both follow the replacement);
    with one independent spin the second grid is ignored and mirrors the first;
    tcx_set_states forms the density from the orbitals again."""
    sys_ = S.tiny_system(method="TC", n=(12, 12, 12), kmesh=(2, 2, 1), nbands=4, nocc=3, seed_stream=35)
    nb, phi = S.applied_bands(sys_, nb=[2, 3, 1, 2], seed=7)
    engine.load_system(sys_)
    rho = engine.get_density()
    N = sys_.ngrid
    assert rho.shape == (2, N)
    assert abs(rho[0].sum().real / (N * sys_.num_kpoints) - sys_.num_electrons() / 2) < 1e-10
    assert np.array_equal(rho[0], rho[1])
    h1 = engine.hf2h(nb, phi)
    t1 = engine.tc3a1(nb, phi)
    two = np.stack([2.0 * rho[0], 5.0 * rho[0]])  # second spin grid differs: ignored
    engine.set_density(two)
    back = engine.get_density()
    assert np.array_equal(back[0], 2.0 * rho[0]) and np.array_equal(back[1], 2.0 * rho[0])
    h2 = engine.hf2h(nb, phi)
    t2 = engine.tc3a1(nb, phi)
    assert np.abs(h2 - 2.0 * h1).max() <= 1e-12 * np.abs(h1).max()
    assert np.abs(t2 - t1).max() > 1e-3 * np.abs(t1).max()  # tc3a1 reads the replaced density too
    engine.set_states(sys_.nocc, sys_.filling, sys_.phi_occ, sys_.chi_occ)
    assert np.array_equal(engine.get_density(), rho)
    assert np.abs(engine.hf2h(nb, phi) - h1).max() <= 1e-13 * np.abs(h1).max()
    with pytest.raises(tcx.TcxError):
        engine.set_density(np.full((2, N), np.nan, np.complex128))
