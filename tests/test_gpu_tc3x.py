"""GPU parity of the three-body exchange terms (tc3a3b3b4b6, tc3a2a4b2b5) and
the assembly terms (kinetic, hf2h, tc2h, all) through the C ABI.

Tolerances:
  FP64 path : max |delta| <= 1e-10 * max |ref|
  FP32 path (three-body pair records in FP32, FP64 sums):
      per band ||delta||/||ref|| <= 1e-4 and |delta E_3| <= 1e-6 Ha
  kinetic / hf2h / tc2h are FP64 only: 1e-10.
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
HF_GOLDEN = [n for n in golden_names() if n.startswith("hf_")]
THREE_BODY_X = ("tc3a3b3b4b6", "tc3a2a4b2b5")


def _close(got, ref, tol=1e-10):
    return np.abs(got - ref).max() <= tol * np.abs(ref).max()


@pytest.mark.parametrize("name", TC_GOLDEN)
@pytest.mark.parametrize("prec", [tcx.PREC_FP64, tcx.PREC_FP32])
def test_three_body_exchange_golden(engine, golden, name, prec):
    z = golden[name]
    sys_ = MG.load_system(z)
    engine.load_system(sys_)
    for term in THREE_BODY_X:
        print(f"{name} {term} prec={prec}", flush=True)
        got = getattr(engine, term)(z["nb"], z["phi"], precision=prec)
        if prec == tcx.PREC_FP64:
            assert _close(got, z[term]), term
        else:
            assert per_band_rel(sys_, z["nb"], got, z[term]).max() <= 1e-4, term


@pytest.mark.parametrize("name", TC_GOLDEN + HF_GOLDEN)
def test_assembly_terms_golden(engine, golden, name):
    z = golden[name]
    sys_ = MG.load_system(z)
    engine.load_system(sys_)
    terms = ["kinetic", "hf2h"] if name in HF_GOLDEN else ["kinetic", "tc2h"]
    for term in terms:
        print(f"{name} {term}", flush=True)
        assert _close(getattr(engine, term)(z["nb"], z["phi"]), z[term]), term


@pytest.mark.parametrize("method", ["TC", "BITC"])
def test_three_body_exchange_against_oracle_12(engine, method):
    sys_ = S.tiny_system(method=method, n=(12, 12, 12), kmesh=(2, 2, 1), nbands=4, nocc=3, seed_stream=41)
    engine.load_system(sys_)
    nb, phi = S.applied_bands(sys_, nb=[3, 2, 1, 3], seed=42)
    o = O.Oracle(sys_)
    for term in THREE_BODY_X + ("tc2h",):
        print(f"{method} {term}", flush=True)
        ref = getattr(o, term)(nb, phi)
        got64 = getattr(engine, term)(nb, phi, precision=tcx.PREC_FP64) if term != "tc2h" else engine.tc2h(nb, phi)
        assert _close(got64, ref), term
        if term != "tc2h":
            got32 = getattr(engine, term)(nb, phi, precision=tcx.PREC_FP32)
            assert per_band_rel(sys_, nb, got32, ref).max() <= 1e-4, term


@pytest.mark.parametrize("method", ["TC", "BITC"])
def test_three_body_exchange_against_oracle_6(engine, method):
    """NZ = 6 is not a multiple of 4: every z consumer splits its pencils over 2 lanes,
    k_za7_mid with the lane-swap exchange, and so do the FP64 instances of tc2x passes A
    and C and hf2x pass C."""
    sys_ = S.tiny_system(method=method, n=(6, 6, 6), kmesh=(2, 1, 1), nbands=3, nocc=2, seed_stream=53,
                         lattice=S.sc_lattice(5.0))
    engine.load_system(sys_)
    nb, phi = S.applied_bands(sys_, nb=[3, 2], seed=54)
    o = O.Oracle(sys_)
    for term in ("tc2x",) + THREE_BODY_X:
        print(f"{method} {term} 6^3", flush=True)
        ref = getattr(o, term)(nb, phi)
        assert _close(getattr(engine, term)(nb, phi, precision=tcx.PREC_FP64), ref), term
        got32 = getattr(engine, term)(nb, phi, precision=tcx.PREC_FP32)
        assert per_band_rel(sys_, nb, got32, ref).max() <= 1e-4, term
    hf = S.System(**{**sys_.__dict__, "method": "HF", "chi_occ": None})
    engine.load_system(hf)
    ref = O.Oracle(hf).hf2x(nb, phi)
    assert _close(engine.hf2x(nb, phi, precision=tcx.PREC_FP64), ref), "hf2x"
    assert per_band_rel(hf, nb, engine.hf2x(nb, phi, precision=tcx.PREC_FP32), ref).max() <= 3e-5, "hf2x"


def test_three_body_exchange_without_divergence_correction(engine):
    sys_ = S.tiny_system(method="BITC", n=(12, 12, 12), kmesh=(2, 1, 1), nbands=3, nocc=2, seed_stream=43)
    engine.load_system(sys_, div_corr=False)
    nb, phi = S.applied_bands(sys_, nb=[3, 2], seed=44)
    o = O.Oracle(sys_, div_corr=False)
    for term in THREE_BODY_X:
        assert _close(getattr(engine, term)(nb, phi), getattr(o, term)(nb, phi)), term


def test_three_body_exchange_zero_without_jastrow(engine):
    sys_ = S.tiny_system(method="TC", n=(8, 8, 8), kmesh=(2, 1, 1), nbands=3, nocc=2, A=np.zeros((2, 2)))
    engine.load_system(sys_)
    nb, phi = S.applied_bands(sys_, nb=[2, 2], seed=45)
    for term in THREE_BODY_X:
        assert np.abs(getattr(engine, term)(nb, phi)).max() == 0.0, term


def test_three_body_energy_fp32_vs_oracle(engine):
    """E_3 of the occupied bands from the two exchange terms, FP32-split vs the FP64 oracle."""
    sys_ = S.tiny_system(method="BITC", n=(12, 12, 12), kmesh=(2, 2, 1), nbands=3, nocc=3, seed_stream=46)
    engine.load_system(sys_)
    o = O.Oracle(sys_)
    for term in THREE_BODY_X:
        ref = getattr(o, term)(sys_.nocc, sys_.phi_occ)
        got = getattr(engine, term)(sys_.nocc, sys_.phi_occ, precision=tcx.PREC_FP32)
        e_ref = E.term_energy(sys_, ref, 3, left=True)
        e_got = E.term_energy(sys_, got, 3, left=True)
        print(term, e_ref, e_got - e_ref, flush=True)
        assert abs(e_got - e_ref) < 1e-6, term


@pytest.mark.parametrize("method", ["HF", "TC", "BITC"])
def test_apply_all_matches_sum_of_terms(engine, method):
    """calc_hamiltonian::all: H1 = kinetic, H2 = Hartree + exchange, H3 = the four three-body terms."""
    sys_ = S.tiny_system(method=method, n=(8, 8, 8), kmesh=(2, 1, 1), nbands=3, nocc=2, seed_stream=47)
    engine.load_system(sys_)
    nb, phi = S.applied_bands(sys_, nb=[3, 2], seed=48)
    o = O.Oracle(sys_)
    h1, h2, h3 = engine.apply_all(nb, phi, precision=tcx.PREC_FP64)
    assert _close(h1, o.kinetic(nb, phi))
    if method == "HF":
        assert h3 is None
        assert _close(h2, o.hf2h(nb, phi) + o.hf2x(nb, phi))
    else:
        assert _close(h2, o.tc2h(nb, phi) + o.tc2x(nb, phi))
        ref3 = o.tc3a1(nb, phi) + o.tc3b1(nb, phi) + o.tc3a2a4b2b5(nb, phi) + o.tc3a3b3b4b6(nb, phi)
        assert _close(h3, ref3)


def test_three_body_exchange_device_pointers(engine):
    sys_ = S.tiny_system(method="TC", n=(8, 8, 8), kmesh=(2, 1, 1), nbands=3, nocc=2, seed_stream=49)
    engine.load_system(sys_)
    nb, phi = S.applied_bands(sys_, nb=[3, 1], seed=50)
    ref = engine.tc3a2a4b2b5(nb, phi)
    d_phi = engine.to_device(phi)
    d_out = engine.device_buffer(phi.nbytes)
    d_out.zero()
    engine._term_device("tc3a2a4b2b5", nb, d_phi.ptr, d_out.ptr, tcx.PREC_FP64)
    got = d_out.download(np.complex128)
    assert _close(got, ref, 1e-13)


def test_pair_lists_follow_state_changes(engine):
    """Two systems with the same applied-band counts but different occupied sets
    in a row: the cached three-body pair lists must be rebuilt for the second."""
    for nocc, cfg in ((2, 51), (3, 52)):
        sys_ = S.tiny_system(method="TC", n=(8, 8, 8), kmesh=(2, 1, 1), nbands=3, nocc=nocc, seed_stream=cfg)
        engine.load_system(sys_)
        nb, phi = S.applied_bands(sys_, nb=[2, 2], seed=53)
        o = O.Oracle(sys_)
        for term in THREE_BODY_X:
            assert _close(getattr(engine, term)(nb, phi), getattr(o, term)(nb, phi)), (nocc, term)


@pytest.mark.parametrize("shape", [(1, 1), (2, 3), (5, 2)])
def test_three_body_launch_shape(engine, shape):
    """The three-body pair sets follow tcx_set_launch_shape (term 2) and the results do not
    depend on it beyond FP64 summation order (1e-12)."""
    sys_ = S.tiny_system(method="BITC", n=(8, 8, 8), kmesh=(2, 1, 1), nbands=3, nocc=2, seed_stream=131,
                         A=np.full((2, 2), 0.5))
    engine.load_system(sys_)
    nb = np.array([2, 1], np.int32)
    _, phi = S.applied_bands(sys_, nb=nb, seed=132)
    engine.set_launch_shape("tc3", 0, 0)
    base = {t: getattr(engine, t)(nb, phi) for t in THREE_BODY_X}
    engine.set_launch_shape("tc3", *shape)
    engine.load_system(sys_)  # fresh occupied-set caches under the new shape
    for t in THREE_BODY_X:
        got = getattr(engine, t)(nb, phi)
        err = np.abs(got - base[t]).max() / np.abs(base[t]).max()
        print(t, shape, err, flush=True)
        assert err <= 1e-12, t
    engine.set_launch_shape("tc3", 0, 0)
