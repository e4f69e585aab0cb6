"""Pins of the oracle's polynomial and RPA+polynomial Jastrow (oracle/jastrow_poly.c).

* The reference's own runs: for examples 28-33 the coefficients after the cusp condition
  (as printed, 10 digits) and num_rmesh_J3_integ (= 401 for the RPA+polynomial runs) follow
  from the examples' parameters.casl and cell (tests/golden/jastrow_poly_examples.json).
* The mathematics: u(G) is the 3D Fourier transform of u(r) and tc_2body(G) that of
  4 pi/G^2 + lap u - (grad u)^2 (jastrow_functions.cpp:53-56, :210-329), with
  u_poly(r) = -sum_i c_i r^(i-1) (1 - r/L)^C (r < L) built from the input coefficients, not
  from the expanded c_poly_internal, and u_RPA(r) = A (1 - exp(-r/F)) / r.  Each transform is
  a radial integral done here by adaptive quadrature (scipy), as are J_integ, J2_integ and
  J3_integ against their definitions; J_integ(G=0) = 4 pi L^(n+2)/(n+2) exactly.
* The terms: with u_poly scaled by lambda (no RPA part, no cusp) every TC term is a
  polynomial of degree <= 2 in lambda that changes with lambda.
"""
import json
import math
import os

import numpy as np
import pytest
from scipy import integrate

from oracle import oracle as O
from tc_amd import system as S

HERE = os.path.dirname(os.path.abspath(__file__))
EX = json.load(open(os.path.join(HERE, "golden", "jastrow_poly_examples.json")))
PAIRS = ("1-1", "1-2", "1-2", "2-2")          # [s1*2+s2] -> casl channel
PRINTED = ("up-up", "up-down", "down-up", "down-down")


def example_spec(ex):
    d = EX[str(ex)]
    poly = d["casl"]["poly"]
    deg = poly["deg"]
    c = np.zeros((4, deg))
    L = np.zeros(4)
    for p, ch in enumerate(PAIRS):
        src = ch if ch in poly["c"] else "1-1"
        c[p] = poly["c"][src]
        L[p] = poly["L"][ch if ch in poly["L"] else "1-1"]
    spec = {"deg": deg, "C": poly["C"], "L": L.reshape(2, 2), "c": c.reshape(2, 2, deg), "cusp": poly["cusp"]}
    A, F = np.zeros(4), np.ones(4)
    if d["casl"]["rpa_F"] is not None:
        Fc = d["casl"]["rpa_F"]
        for p, ch in enumerate(PAIRS):
            F[p] = Fc[ch if ch in Fc else "1-1"]
            A[p] = F[p] ** 2 / 2.0 if p in (0, 3) else F[p] ** 2   # impose_cusp(false), jastrow_initialize.cpp:80-81
    cell = np.array(d["cell"])
    recip = 2.0 * math.pi * np.linalg.inv(cell).T
    return spec, A, F, recip, d["fft_grid"]


@pytest.mark.parametrize("ex", [28, 29, 30, 31, 32, 33])
def test_cusp_coefficients_and_rmesh_match_examples(ex):
    spec, A, F, recip, grid = example_spec(ex)
    jp = O.JastrowPoly(spec, A, F, recip, grid)
    printed = EX[str(ex)]["printed"]
    for p in range(4):
        got = jp.c_input(p)
        want = np.array(printed["c_poly"][PRINTED[p]])
        np.testing.assert_allclose(got, want, rtol=6e-10, atol=0)
    if printed["num_rmesh"] is not None:
        assert jp.num_rmesh == printed["num_rmesh"]
    else:
        assert jp.num_rmesh == 0                    # polynomial only: no J3 mesh (:153)


def test_cusp_condition_values():
    """c_1 = (L/C)(c_2 - cusp) with cusp 1/4 (parallel) and 1/2 (antiparallel): u_poly'(0)
    then equals -1/4 / -1/2 (minus the cusp of J = exp(-u)) when the input carries it."""
    deg, C, L = 4, 3, 5.0
    c = np.tile(np.array([0.0, 0.03, -0.01, 0.002]), (4, 1))
    spec = {"deg": deg, "C": C, "L": np.full((2, 2), L), "c": c.reshape(2, 2, deg), "cusp": True}
    jp = O.JastrowPoly(spec, np.zeros(4), np.ones(4), np.eye(3), (4, 4, 4))
    for p, cusp in enumerate((0.25, 0.5, 0.5, 0.25)):
        ci = jp.c_input(p)
        # d/dr [sum c_i r^i (1 - r/L)^C] at 0 = c_1 - C c_0 / L; u_poly = -that polynomial
        assert abs(-(ci[1] - C * ci[0] / L) - (-cusp)) < 1e-15


def radial_ft(f, G, rmax, points=None):
    """int f(r) exp(-iG.r) d^3r = 4 pi int_0^rmax f(r) r^2 sinc(G r) dr."""
    if G < 1e-12:
        g = lambda r: 4.0 * math.pi * f(r) * r * r
    else:
        g = lambda r: 4.0 * math.pi * f(r) * r * math.sin(G * r) / G
    v, _ = integrate.quad(g, 0.0, rmax, limit=400, epsabs=0.0, epsrel=1e-13, points=points)
    return v


EPS = np.finfo(float).eps


def rounding_scale(jp, G2, p):
    """sum_ij |i j c_i c_j J[i+j-1]| and sum_i |c_i J[i+1]|: the reference forms u and
    -(grad u)^2 from double products of the expanded coefficients and J_integ, so its
    results carry rounding errors of a few eps times these sums (for the examples' 16th
    order polynomials up to 1e12 against results of order 1)."""
    c = jp.c_internal(p)
    nc = len(c)
    J = jp.J(G2, p, 2 * nc - 2)
    t = sum(abs(i * j * c[i] * c[j] * J[i + j - 1]) for i in range(1, nc) for j in range(1, nc))
    u = sum(abs(c[i] * J[i + 1]) for i in range(nc))
    return t, u


def poly_u(spec, p):
    deg, C = spec["deg"], spec["C"]
    L = float(np.asarray(spec["L"]).reshape(4)[p])
    return deg, C, L


@pytest.fixture(scope="module")
def ex30():
    spec, A, F, recip, grid = example_spec(30)
    return spec, A, F, O.JastrowPoly(spec, A, F, recip, grid)


def test_J_integ_definition(ex30):
    _, _, _, jp = ex30
    p, L = 0, 7.0
    nJ = 36
    for n in range(nJ):
        assert abs(jp.J(0.0, p, nJ)[n] / (4 * math.pi * L ** (n + 2) / (n + 2)) - 1) < 1e-14
    for G in (0.05, 0.7, 2.3, 4.45, 4.53, 7.9):      # G L crosses the Taylor threshold sqrt(1000)
        J = jp.J(G * G, p, nJ)
        for n in range(0, nJ, 5):
            ref, _ = integrate.quad(lambda r: r ** n * math.sin(G * r), 0.0, L, limit=400, epsabs=0.0, epsrel=2e-14)
            ref *= 4 * math.pi / G
            scale = 4 * math.pi * L ** (n + 2) / (n + 2)
            assert abs(J[n] - ref) <= 1e-12 * scale, (G, n)


def test_J_integ_continuous_at_taylor_threshold(ex30):
    _, _, _, jp = ex30
    L = 7.0
    G_lo = math.sqrt(1000.0 * (1 - 1e-12)) / L
    G_hi = math.sqrt(1000.0 * (1 + 1e-12)) / L
    a, b = jp.J(G_lo ** 2, 0, 36), jp.J(G_hi ** 2, 0, 36)
    scale = np.array([4 * math.pi * L ** (n + 2) / (n + 2) for n in range(36)])
    assert np.max(np.abs(a - b) / scale) < 1e-12


@pytest.mark.parametrize("ex", [28, 32])
def test_uk_poly_is_fourier_transform(ex):
    spec, A, F, recip, grid = example_spec(ex)
    spec = dict(spec)
    jp = O.JastrowPoly(spec, np.zeros(4), np.ones(4), recip, grid)
    for p in (0, 1, 3):
        deg, C, L = poly_u(spec, p)
        ci = jp.c_input(p)
        u = lambda r: -np.polynomial.polynomial.polyval(r, ci) * (1 - r / L) ** C
        for G in (0.0, 0.3, 1.1, 2.9, 6.0):
            ref = radial_ft(u, G, L)
            got = jp.uk_poly(G * G, p)
            scale = radial_ft(lambda r: abs(u(r)), 0.0, L)
            assert abs(got - ref) <= 1e-11 * scale + 8 * EPS * rounding_scale(jp, G * G, p)[1], (p, G, got, ref)


@pytest.mark.parametrize("ex", [28, 32])
def test_tc2body_poly_is_fourier_transform(ex):
    spec, A, F, recip, grid = example_spec(ex)
    jp = O.JastrowPoly(spec, np.zeros(4), np.ones(4), recip, grid)
    for p in (0, 1, 3):
        deg, C, L = poly_u(spec, p)
        ci = jp.c_input(p)
        P = np.polynomial.Polynomial(-ci) * np.polynomial.Polynomial([1.0, -1.0 / L]) ** C
        dP = P.deriv()
        for G in (0.0, 0.3, 1.1, 2.9, 6.0):
            g2 = radial_ft(lambda r: dP(r) ** 2, G, L)
            coul = 4 * math.pi / (G * G) if G > 0 else 0.0
            ref = coul - G * G * radial_ft(P, G, L) - g2
            got = jp.tc2body_poly(G * G, p)
            scale = radial_ft(lambda r: dP(r) ** 2, 0.0, L) + G * G * radial_ft(lambda r: abs(P(r)), 0.0, L) + 1e-3
            t, u = rounding_scale(jp, G * G, p)
            assert abs(got - ref) <= 1e-10 * (scale + coul) + 8 * EPS * (t + G * G * u), (p, G, got, ref)


@pytest.mark.parametrize("ex,cusp", [(30, False), (32, False), (32, True)])
def test_tc2body_rpapoly_is_fourier_transform(ex, cusp):
    """-(grad u)^2 of u = u_RPA + u_poly: the RPA x RPA part is tc_2body_RPA's closed form;
    the cross term 2 u_RPA' u_poly' and u_poly'^2 are integrated here.  The reference takes
    the linear (c_1) cross term from J3_integ's Simpson rule, accurate to ~1e-6 by
    construction (jastrow_initialize_polynomial.cpp:183-193), which bounds this comparison.
    Without the cusp condition that linear term is zero (c_1 = L c_2 / C cancels it), so the
    cusp case, which set_polynomial_parameters accepts although the casl reader does not
    pair it with RPA terms, is what exercises it."""
    spec, A, F, recip, grid = example_spec(ex)
    spec = dict(spec, cusp=cusp)
    jp = O.JastrowPoly(spec, A, F, recip, grid)
    base = dataclass_replace(S.tiny_system(method="TC", n=(6, 6, 6), kmesh=(1, 1, 1), nbands=1, nocc=1),
                             A=A.reshape(2, 2), F=F.reshape(2, 2))
    o_rpa = O.Oracle(base)
    o = O.Oracle(dataclass_replace(base, jastrow_poly=spec))
    for p in (0, 1, 3):
        s1, s2 = divmod(p, 2)
        deg, C, L = poly_u(spec, p)
        Ap, Fp = A[p], F[p]
        ci = jp.c_input(p)
        P = np.polynomial.Polynomial(-ci) * np.polynomial.Polynomial([1.0, -1.0 / L]) ** C
        dP = P.deriv()
        dR = lambda r: Ap * (-(1 - math.exp(-r / Fp)) / (r * r) + math.exp(-r / Fp) / (Fp * r))
        for G in (0.0, 0.4, 1.3, 3.1):
            cross = radial_ft(lambda r: 2 * dR(r) * dP(r), G, L)
            ref = (o_rpa.tc_2body(G * G, s1, s2) - G * G * radial_ft(P, G, L)
                   - radial_ft(lambda r: dP(r) ** 2, G, L) - cross)
            got = o.tc_2body(G * G, s1, s2)
            t, u = rounding_scale(jp, G * G, p)
            assert abs(got - ref) <= 2e-5 * max(1.0, abs(ref)) + 8 * EPS * (t + G * G * u), (p, G, got, ref)
            # u(G) of the sum is the sum of the transforms
            u_ref = o_rpa.uk(G * G, s1, s2) + radial_ft(P, G, L)
            assert abs(o.uk(G * G, s1, s2) - u_ref) <= 1e-11 * max(1.0, abs(u_ref)) + 8 * EPS * u


def test_J2_J3_integ_definition(ex30):
    _, A, F, jp = ex30
    p, L, Fp = 1, 7.0, F[1]
    for G in (0.02, 0.15, 0.9, 3.0):                 # G L crosses the J2 Taylor threshold sqrt(1.5)
        J2 = jp.J2(G, p, 18)
        for n in range(0, 18, 3):
            ref, _ = integrate.quad(lambda r: r ** n * math.sin(G * r) * math.exp(-r / Fp), 0.0, L,
                                    limit=400, epsabs=0.0, epsrel=2e-14)
            ref *= 4 * math.pi / G
            scale = 4 * math.pi * L ** (n + 1) / (n + 1)
            assert abs(J2[n] - ref) <= 1e-9 * scale, (G, n, J2[n], ref)
        j3ref, _ = integrate.quad(lambda s: math.sin(Fp * G * s) * (math.exp(-s) + (math.exp(-s) - 1) / s)
                                  if s > 0 else 0.0, 0.0, L / Fp, limit=400, epsabs=0.0, epsrel=2e-14)
        j3ref *= 4 * math.pi / G
        assert abs(jp.J3(G, p) - j3ref) < 1e-6


def dataclass_replace(obj, **kw):
    import dataclasses
    return dataclasses.replace(obj, **kw)


def poly_system(lam, method="TC"):
    base = S.tiny_system(method=method, n=(8, 8, 8), kmesh=(2, 1, 1), nbands=3, nocc=2, seed_stream=77)
    deg = 5
    c = lam * np.array([0.0, 0.05, -0.02, 0.004, -0.0003])
    spec = {"deg": deg, "C": 3, "L": np.full((2, 2), 3.0), "c": np.tile(c, (4, 1)).reshape(2, 2, deg),
            "cusp": False}
    return dataclass_replace(base, A=np.zeros((2, 2)), F=np.ones((2, 2)), jastrow_poly=spec)


@pytest.mark.parametrize("term", ["tc2x", "tc3a1", "tc3b1", "tc3a3b3b4b6", "tc3a2a4b2b5", "tc2h"])
def test_terms_quadratic_in_polynomial_scale(term):
    lams = (0.0, 1.0, 2.0, 3.0)
    outs = []
    for lam in lams:
        s = poly_system(lam)
        nb = np.array([2, 1], np.int32)
        _, phi = S.applied_bands(s, nb=nb, seed=78)
        outs.append(getattr(O.Oracle(s), term)(nb, phi))
    d3 = outs[3] - 3 * outs[2] + 3 * outs[1] - outs[0]
    scale = max(np.abs(o).max() for o in outs)
    assert np.abs(d3).max() <= 1e-11 * scale
    assert np.abs(outs[1] - outs[0]).max() > 1e-6 * scale   # the term sees the polynomial
