"""Grid sizes 10, 18 and 30 (n1 = n2, and n3 from the same set, equal or not) through every
term against the oracle on the same seeded inputs.

These lengths reach the register FFT's radix-5 and radix-3 stages in other orders than the
other sizes, and their z pencils (not multiples of 4) split over 2 lanes in every consumer.
Tolerances as in the per-term tests:
  FP64 path : max |delta| <= 1e-10 max |ref|
  FP32 path : per band ||delta||/||ref|| <= 3e-5 (hf2x, both pass orders; tc2x) and 1e-4
              (the three-body terms, longer FP32 chains).
"""
import numpy as np
import pytest

from oracle import oracle as O
from tc_amd import system as S
from tc_amd import tcx
from test_gpu_hf2x import per_band_rel

pytestmark = pytest.mark.gpu

GRIDS = [(10, 10, 10), (18, 18, 18), (30, 30, 30), (18, 18, 30), (30, 30, 10)]


def _close(got, ref, tol=1e-10):
    return np.abs(got - ref).max() <= tol * np.abs(ref).max()


def _system(method, n, seed):
    lattice = S.fcc_lattice(6.0 * n[0] / 12.0)  # about the same cutoff per grid point
    return S.tiny_system(method=method, n=n, kmesh=(2, 1, 1), nbands=3, nocc=2, seed_stream=seed,
                         lattice=lattice)


@pytest.mark.parametrize("n", GRIDS)
def test_hf_terms(engine, n):
    sys_ = _system("HF", n, 700 + n[0] + n[2])
    engine.load_system(sys_)
    nb, phi = S.applied_bands(sys_, nb=[3, 2], seed=701)
    o = O.Oracle(sys_)
    for term in ("kinetic", "hf2h"):
        assert _close(getattr(engine, term)(nb, phi), getattr(o, term)(nb, phi)), term
    ref = o.hf2x(nb, phi)
    assert _close(engine.hf2x(nb, phi, precision=tcx.PREC_FP64), ref), "hf2x fp64"
    for order in ((0, 1) if n[0] == n[1] else (0,)):
        engine.set_pipeline_order("hf2x", order)
        rel = per_band_rel(sys_, nb, engine.hf2x(nb, phi, precision=tcx.PREC_FP32), ref).max()
        print(f"{n} hf2x fp32 order {order}: per-band {rel:.2e}", flush=True)
        assert rel <= 3e-5, order
    engine.set_pipeline_order("hf2x", -1)


@pytest.mark.parametrize("n", GRIDS)
@pytest.mark.parametrize("method", ["TC", "BITC"])
def test_tc_terms(engine, method, n):
    sys_ = _system(method, n, 710 + n[0] + n[2])
    engine.load_system(sys_)
    nb, phi = S.applied_bands(sys_, nb=[2, 1], seed=711)
    o = O.Oracle(sys_)
    for term in ("tc2x", "tc3a1", "tc3b1", "tc3a3b3b4b6", "tc3a2a4b2b5"):
        ref = getattr(o, term)(nb, phi)
        got64 = getattr(engine, term)(nb, phi, precision=tcx.PREC_FP64)
        assert _close(got64, ref), term
        if term in ("tc3a1", "tc3b1"):
            continue  # FP64-only local terms
        rel = per_band_rel(sys_, nb, getattr(engine, term)(nb, phi, precision=tcx.PREC_FP32), ref).max()
        print(f"{method} {n} {term}: fp32 per-band {rel:.2e}", flush=True)
        assert rel <= (3e-5 if term == "tc2x" else 1e-4), term
