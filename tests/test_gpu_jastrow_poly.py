"""Polynomial and RPA+polynomial Jastrow on the GPU against the oracle.

Coefficients are those of the reference's examples (tests/golden/jastrow_poly_examples.json):
example 28 (polynomial only, 16th order, with the cusp condition), example 30 (RPA +
16th-order polynomial), example 32 (RPA + 6th order, separate up-up / up-down / down-down
channels) and example 32's polynomial with the cusp condition on top of the RPA part
(which exercises the linear cross term J3_integ).

* Kernels: u(G) and tc_2body(G) evaluated on the GPU (tcx_jastrow_kernels) at |G|^2 from 0
  across both series thresholds.  The reference forms them from double products of the
  expanded coefficients and J_integ, whose terms reach 1e12 for the 16th-order examples
  while the results are of order 1; any two correct evaluations (the reference's 50-digit
  one included) may therefore differ by a few eps times those sums.  The bound used is
  1e-13 |ref| + 8 eps (sum of |terms|), with the sums taken from the oracle.
* num_rmesh_J3_integ chosen by the engine for the examples' cell and 16^3 grid: 401, as the
  reference printed.
* Terms: every TC/BITC term with these Jastrows against the oracle, FP64 (max |delta| /
  max |ref|) to 1e-9 for the 6th-order polynomial and to 1e-7 for the 16th-order ones: there
  one ulp of one J_integ value moves tc_2body by ~1e-4 (the 1e12 sums above), and the
  double-double series here and the binary128 series of the oracle round a few of the
  thousands of J values differently (the reference's 50-digit series is a third rounding);
  the FP32-split paths per band to 3e-5.
* The Hermitian conjugate (u -> -u) of the bi-orthogonal Davidson flips the polynomial too
  (Jastrow::take_Hermitian_conj, jastrow_functions.cpp:417-440).
"""
import dataclasses
import math

import numpy as np
import pytest

from oracle import oracle as O
from tc_amd import system as S
from tc_amd import tcx
from test_gpu_hf2x import per_band_rel
from test_oracle_jastrow_poly import EPS, example_spec

pytestmark = pytest.mark.gpu

CASES = {"ex28": (28, None), "ex30": (30, None), "ex32": (32, None), "ex32_cusp": (32, True)}
FP64_TOL = {"ex28": 1e-7, "ex30": 1e-7, "ex32": 1e-9, "ex32_cusp": 1e-9}


def case_spec(name):
    ex, cusp = CASES[name]
    spec, A, F, recip, grid = example_spec(ex)
    if cusp is not None:
        spec = dict(spec, cusp=cusp)
    return spec, A, F, recip, grid


def condition_bound(jp, G2, p, A):
    """8 eps times the sums of |terms| the kernels are formed from (see the module doc)."""
    c = jp.c_internal(p)
    nc = len(c)
    J = jp.J(G2, p, 2 * nc - 2)
    t = sum(abs(i * j * c[i] * c[j] * J[i + j - 1]) for i in range(1, nc) for j in range(1, nc))
    u = sum(abs(c[i] * J[i + 1]) for i in range(nc))
    s4 = 0.0
    if abs(A) > 1e-8 and jp.num_rmesh:
        J2 = jp.J2(math.sqrt(G2), p, nc - 1)
        s4 = sum(abs(2 * i * c[i] * A) * (abs(J[i - 2]) + abs(J2[i - 1]) + abs(J2[i - 2])) for i in range(2, nc))
    return 8 * EPS * (t + G2 * u + s4), 8 * EPS * u


def poly_system(name, method, seed_stream=301):
    spec, A, F, _, _ = case_spec(name)
    base = S.tiny_system(method=method, n=(8, 8, 8), kmesh=(2, 1, 1), nbands=3, nocc=2, seed_stream=seed_stream)
    return dataclasses.replace(base, A=A.reshape(2, 2), F=F.reshape(2, 2), jastrow_poly=spec)


@pytest.mark.parametrize("name", list(CASES))
def test_kernels_vs_oracle(engine, name):
    sys_ = poly_system(name, "TC")
    engine.load_system(sys_)
    o = O.Oracle(sys_)
    L = float(np.asarray(sys_.jastrow_poly["L"]).min())
    G = np.concatenate([[0.0, 1e-7, 0.01], np.linspace(0.02, 8.0, 157),
                        [math.sqrt(1000.0) / L * (1 + d) for d in (-1e-9, 1e-9)],
                        [math.sqrt(1.5) / L * (1 + d) for d in (-1e-9, 1e-9)]])
    G2 = G * G
    worst = 0.0
    for s1, s2 in ((0, 0), (0, 1), (1, 0), (1, 1)):
        p = 2 * s1 + s2
        u, v = engine.jastrow_kernels(G2, s1, s2)
        for i, g2 in enumerate(G2):
            ru, rv = o.uk(g2, s1, s2), o.tc_2body(g2, s1, s2)
            bv, bu = condition_bound(o.poly, g2, p, sys_.A.reshape(4)[p])
            assert abs(u[i] - ru) <= 1e-13 * abs(ru) + bu + 1e-300, (p, g2, u[i], ru)
            assert abs(v[i] - rv) <= 1e-13 * abs(rv) + bv + 1e-300, (p, g2, v[i], rv)
            worst = max(worst, abs(v[i] - rv) / max(abs(rv), 1e-300), abs(u[i] - ru) / max(abs(ru), 1e-300))
    print(name, "max rel kernel difference", worst, flush=True)


@pytest.mark.parametrize("ex", [30, 32])
def test_num_rmesh_matches_the_examples(engine, ex):
    spec, A, F, recip, grid = example_spec(ex)
    vol = abs(np.linalg.det(2 * math.pi * np.linalg.inv(recip).T))
    engine.set_cell(recip, vol, grid, 1, "TC")
    engine.set_jastrow_rpa(A, F)
    engine.set_jastrow_polynomial(spec["deg"], spec["C"], spec["L"], spec["c"], spec["cusp"])
    assert engine.jastrow_num_rmesh() == 401


def test_polynomial_setter_errors(engine):
    sys_ = poly_system("ex32", "TC")
    engine.load_system(sys_)
    L = np.full(4, 5.0)
    c = np.zeros(4 * 3)
    for args in ((17, 3, L, np.zeros(4 * 17), False), (3, 4, L, c, False), (0, 3, L, None, True),
                 (3, 3, -L, c, False), (3, 3, L, np.full(12, np.nan), False)):
        with pytest.raises(tcx.TcxError) as e:
            engine.set_jastrow_polynomial(*args)
        assert e.value.code == tcx.TCX_ERR_ARG
    # the previous parameters are still in force
    o = O.Oracle(sys_)
    u, v = engine.jastrow_kernels(np.array([0.0, 0.7]), 0, 1)
    assert abs(u[1] - o.uk(0.7, 0, 1)) < 1e-12 * abs(o.uk(0.7, 0, 1))


TERMS = ("tc2x", "tc3a1", "tc3b1", "tc3a3b3b4b6", "tc3a2a4b2b5", "tc2h")
FP64_ONLY = ("tc3a1", "tc3b1", "tc2h")


@pytest.mark.parametrize("name", list(CASES))
@pytest.mark.parametrize("method", ["TC", "BITC"])
@pytest.mark.parametrize("prec", [tcx.PREC_FP64, tcx.PREC_FP32], ids=["fp64", "fp32"])
def test_terms_vs_oracle(engine, name, method, prec):
    sys_ = poly_system(name, method)
    engine.load_system(sys_)
    nb = np.array([2, 1], np.int32)
    _, phi = S.applied_bands(sys_, nb=nb, seed=302)
    o = O.Oracle(sys_)
    for term in TERMS:
        ref = getattr(o, term)(nb, phi)
        if term in FP64_ONLY or prec == tcx.PREC_FP64:
            got = getattr(engine, term)(nb, phi) if term in FP64_ONLY else getattr(engine, term)(nb, phi, precision=prec)
            err = np.abs(got - ref).max() / np.abs(ref).max()
            print(f"{name} {method} {term} fp64 max rel {err:.3e}", flush=True)
            assert err <= FP64_TOL[name], term
        else:
            got = getattr(engine, term)(nb, phi, precision=prec)
            rel = per_band_rel(sys_, nb, got, ref)
            print(f"{name} {method} {term} fp32 per-band rel max {rel.max():.3e}", flush=True)
            assert rel.max() <= 3e-5, term


def test_biortho_davidson_with_polynomial(engine):
    """BITC, RPA+polynomial, bi-orthogonal Davidson: each call matches block_davidson_biortho
    on the oracle's dense H_n and H^dagger_n (u -> -u includes the polynomial) on the same
    inputs."""
    from oracle import davidson as OD
    from test_gpu_davidson import _kG2, full_hamiltonians, hermitian_conj_hamiltonians, trial_bands
    sys_ = poly_system("ex32", "BITC", seed_stream=303)
    engine.load_system(sys_)
    mats = full_hamiltonians(sys_)
    dmats = hermitian_conj_hamiltonians(sys_)
    nbk = 2
    nb = np.full(sys_.nk, nbk, np.int32)
    phi = trial_bands(sys_, nb, 304)
    left = phi.copy()
    eig = np.zeros(nb.sum(), np.complex128)
    bo = sys_.band_offsets(nb)
    kG2 = [_kG2(sys_, ik) for ik in range(sys_.nk)]
    for it in range(2):
        same_in = [[phi[bo[ik]:bo[ik + 1]].reshape(nbk, -1), left[bo[ik]:bo[ik + 1]].reshape(nbk, -1),
                    eig[ik * nbk:(ik + 1) * nbk]] for ik in range(sys_.nk)]
        phi, left, eig, en = engine.davidson(nb, phi, eig, max_blocks=3, num_refresh=2, phi_left=left)
        for ik in range(sys_.nk):
            rp, rl, re, ren = OD.block_davidson_biortho(list(mats[ik]), list(dmats[ik]), kG2[ik], *same_in[ik],
                                                        max_blocks=3, num_refresh=2)
            got = eig[ik * nbk:(ik + 1) * nbk]
            print("biortho poly", it, ik, got, got - re, flush=True)
            assert np.abs(got - re).max() < 1e-9 * max(1.0, np.abs(re).max()), (it, ik)
