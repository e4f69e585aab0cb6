"""Two independent spins (num_independent_spins = 2) through the C ABI.

The states, fillings and applied bands run over (spin, k), spin-major, with the k-points
and the basis given once.  The exchange-type terms pair states of one spin only and drop
the spin-sum halving of their prefactors (src/calc_hamiltonian_hf2x.cpp:29-30, :46, :90-134;
tc2x.cpp:32-33 with u(ispin, ispin), :269-271); the Hartree term reads the sum of the two
spin densities (hf2h.cpp:30).  Pins:
  * two identical spin channels with fillings f/2 reproduce the one-spin result with
    fillings f, term by term (FP64, 1e-12 of the largest value);
  * different states per spin and a spin-dependent Jastrow: spin channel s equals the
    one-spin oracle on that channel's states with fillings 2 f_s (and u(s, s) moved to
    u(0, 0) for tc2x), the Hartree term the one-spin oracle on both channels' states with
    fillings f_s (FP64 1e-10, FP32-split per band 3e-5);
  * tc2h, which couples the spins through u(s, s'), equals per channel a sum of one-spin
    oracle applies, one per spin's states (FP64 1e-10); tc3a1 and tc3b1 equal the one-spin
    oracle on both spins' states together where the Jastrow makes that exact (FP64 1e-10);
  * the three-body terms are refused with TCX_ERR_UNSUPPORTED.
"""
import dataclasses

import numpy as np
import pytest

from oracle import oracle as O
from tc_amd import system as S
from tc_amd import tcx
from test_gpu_hf2x import per_band_rel

pytestmark = pytest.mark.gpu


def two_spin(sa, sb, fa=None, fb=None):
    """A two-spin System with the states of sa as spin 0 and of sb as spin 1 (same cell,
    k-points and basis), fillings fa / fb (default: half of each system's)."""
    assert np.array_equal(sa.npw, sb.npw) and np.array_equal(sa.gindex, sb.gindex)
    fa = sa.filling / 2 if fa is None else fa
    fb = sb.filling / 2 if fb is None else fb
    chi = None
    if sa.chi_occ is not None:
        chi = np.concatenate([sa.chi_occ, sb.chi_occ])
    return dataclasses.replace(sa, nspin=2, nocc=np.concatenate([sa.nocc, sb.nocc]).astype(np.int32),
                               filling=np.concatenate([fa, fb]), phi_occ=np.concatenate([sa.phi_occ, sb.phi_occ]),
                               chi_occ=chi)


def with_states(sys_, nocc, filling, phi, chi):
    return dataclasses.replace(sys_, nocc=np.asarray(nocc, np.int32), filling=np.asarray(filling, np.float64),
                               phi_occ=phi, chi_occ=chi)


def combined(sa, sb, fa, fb):
    """One one-spin system holding, at every k, the states of sa (fillings fa) and then those
    of sb (fillings fb): its density is the sum of the two spin densities."""
    nocc_all, f_all, phi_all, chi_all = [], [], [], []
    oa, ob = sa.occ_offsets(), sb.occ_offsets()
    fo_a = np.concatenate([[0], np.cumsum(sa.nocc)])
    fo_b = np.concatenate([[0], np.cumsum(sb.nocc)])
    for k in range(sa.nk):
        nocc_all.append(sa.nocc[k] + sb.nocc[k])
        f_all += [fa[fo_a[k]:fo_a[k + 1]], fb[fo_b[k]:fo_b[k + 1]]]
        phi_all += [sa.phi_occ[oa[k]:oa[k + 1]], sb.phi_occ[ob[k]:ob[k + 1]]]
        if sa.chi_occ is not None:
            chi_all += [sa.chi_occ[oa[k]:oa[k + 1]], sb.chi_occ[ob[k]:ob[k + 1]]]
    return with_states(sa, nocc_all, np.concatenate(f_all), np.concatenate(phi_all),
                       np.concatenate(chi_all) if chi_all else None)


def both_channels(sys_, nb, seed):
    """Applied bands: the same nb per k in both channels, different coefficients."""
    nb0, phi0 = S.applied_bands(sys_, nb=nb, seed=seed)
    nb1, phi1 = S.applied_bands(sys_, nb=nb, seed=seed + 1)
    return nb0, phi0, nb1, phi1, np.concatenate([nb0, nb1]).astype(np.int32), np.concatenate([phi0, phi1])


@pytest.mark.parametrize("method", ["HF", "TC", "BITC"])
def test_identical_channels_match_one_spin(engine, method):
    s1 = S.tiny_system(method=method, n=(8, 8, 8), kmesh=(2, 2, 1), nbands=3, nocc=2, seed_stream=151)
    nb0, phi0 = S.applied_bands(s1, nb=[3, 0, 2, 1], seed=152)
    engine.load_system(s1)
    terms = ["kinetic", "hf2h", "hf2x"] + (["tc2x", "tc2h", "tc3a1", "tc3b1"] if method != "HF" else [])
    one = {t: getattr(engine, t)(nb0, phi0) for t in terms}
    s2 = two_spin(s1, s1)
    engine.load_system(s2)
    nb2, phi2 = np.concatenate([nb0, nb0]).astype(np.int32), np.concatenate([phi0, phi0])
    n = phi0.size
    for t in terms:
        got = getattr(engine, t)(nb2, phi2)
        scale = np.abs(one[t]).max()
        for s in range(2):
            err = np.abs(got[s * n:(s + 1) * n] - one[t]).max() / scale
            print(f"{method} {t} spin {s}: {err:.2e}", flush=True)
            assert err <= 1e-12, (t, s)


@pytest.mark.parametrize("method", ["HF", "TC", "BITC"])
def test_channels_against_oracle(engine, method):
    A = np.array([[0.5, 0.4], [0.4, 0.8]]) if method != "HF" else None
    sa = S.tiny_system(method=method, n=(12, 12, 12), kmesh=(2, 1, 1), nbands=4, nocc=3, seed_stream=161, A=A)
    sb = S.tiny_system(method=method, n=(12, 12, 12), kmesh=(2, 1, 1), nbands=4, nocc=2, seed_stream=162, A=A)
    fa = np.tile([1.0, 1.0, 0.6], sa.nk)
    fb = np.tile([1.0, 0.3], sb.nk)
    s2 = two_spin(sa, sb, fa, fb)
    engine.load_system(s2)
    nb0, phi0, nb1, phi1, nb2, phi2 = both_channels(sa, [3, 2], 163)
    n0 = phi0.size
    # one-spin oracle systems: channel s with fillings 2 f_s, u(s, s) in the (0, 0) slot
    ref_sys = [with_states(sa, sa.nocc, 2 * fa, sa.phi_occ, sa.chi_occ),
               with_states(sb, sb.nocc, 2 * fb, sb.phi_occ, sb.chi_occ)]
    if method != "HF":
        A1, F1 = sb.A.copy(), sb.F.copy()
        A1[0, 0], F1[0, 0] = sb.A[1, 1], sb.F[1, 1]
        ref_sys[1] = dataclasses.replace(ref_sys[1], A=A1, F=F1)
    chans = [(nb0, phi0), (nb1, phi1)]
    terms = ["hf2x"] + (["tc2x"] if method != "HF" else [])
    for t in terms:
        refs = [getattr(O.Oracle(ref_sys[s]), t)(*chans[s]) for s in range(2)]
        for prec in (tcx.PREC_FP64, tcx.PREC_FP32):
            got = getattr(engine, t)(nb2, phi2, precision=prec)
            for s, (lo, hi) in enumerate(((0, n0), (n0, phi2.size))):
                g, r = got[lo:hi], refs[s]
                if prec == tcx.PREC_FP64:
                    err = np.abs(g - r).max() / np.abs(r).max()
                    print(f"{method} {t} fp64 spin {s}: {err:.2e}", flush=True)
                    assert err <= 1e-10, (t, s)
                else:
                    rel = per_band_rel(sa, chans[s][0], g, r)
                    print(f"{method} {t} fp32 spin {s}: per-band {rel.max():.2e}", flush=True)
                    assert rel.max() <= 3e-5, (t, s)
    # Hartree: both channels' states with their fillings in one one-spin system
    hsys = combined(sa, sb, fa, fb)
    got = engine.hf2h(nb2, phi2)
    for s, (lo, hi) in enumerate(((0, n0), (n0, phi2.size))):
        r = O.Oracle(hsys).hf2h(*chans[s])
        err = np.abs(got[lo:hi] - r).max() / np.abs(r).max()
        print(f"{method} hf2h spin {s}: {err:.2e}", flush=True)
        assert err <= 1e-10, s


@pytest.mark.parametrize("method", ["TC", "BITC"])
def test_tc2h_channels_against_oracle(engine, method):
    """tc2h with two spins (tc2h.cpp:133-171, spin_factor 1): channel s's potential is
    sum_js [tc_2body(s,js) n_js + u(s,js) G.pgp_js] and its dnu sum_js u(s,js) G n_js, each
    js term linear in the states of spin js.  A one-spin system holding spin js's states
    with their fillings and u(s, js) in every Jastrow slot gives exactly that js term, so
    channel s is the sum over js of two one-spin oracle applies (FP64, 1e-10)."""
    A = np.array([[0.5, 0.4], [0.4, 0.8]])
    sa = S.tiny_system(method=method, n=(12, 12, 12), kmesh=(2, 1, 1), nbands=4, nocc=3, seed_stream=191, A=A)
    sb = S.tiny_system(method=method, n=(12, 12, 12), kmesh=(2, 1, 1), nbands=4, nocc=2, seed_stream=192, A=A)
    fa = np.tile([1.0, 1.0, 0.6], sa.nk)
    fb = np.tile([1.0, 0.3], sb.nk)
    engine.load_system(two_spin(sa, sb, fa, fb))
    nb0, phi0, nb1, phi1, nb2, phi2 = both_channels(sa, [3, 2], 193)
    got = engine.tc2h(nb2, phi2)
    states = [(sa, fa), (sb, fb)]
    n0 = phi0.size
    for s, (nbs, phis, lo, hi) in enumerate(((nb0, phi0, 0, n0), (nb1, phi1, n0, phi2.size))):
        ref = np.zeros_like(phis)
        for js, (sys_js, f_js) in enumerate(states):
            one = dataclasses.replace(with_states(sys_js, sys_js.nocc, f_js, sys_js.phi_occ, sys_js.chi_occ),
                                      A=np.full((2, 2), sa.A[s, js]), F=np.full((2, 2), sa.F[s, js]))
            ref = ref + O.Oracle(one).tc2h(nbs, phis)
        err = np.abs(got[lo:hi] - ref).max() / np.abs(ref).max()
        print(f"{method} tc2h spin {s}: {err:.2e}", flush=True)
        assert err <= 1e-10, s


@pytest.mark.parametrize("method", ["TC", "BITC"])
def test_tc3_local_channels_against_oracle(engine, method):
    """tc3a1 and tc3b1 with two spins (tc3a1.cpp:54-84, tc3b1.cpp:52-100): dnu_s =
    sum_s' u(s,s') G n_s'; tc3a1's V_s = sum_d dnu_s,d^2, tc3b1's V_s1 = sum_s2 u(s1,s2)
    G.FFT[n_s2 dnu_s2].  With u(s, 0) = u(s, 1) = u_s, channel s of tc3a1 equals the one-spin
    oracle on both spins' states together with u_s in every Jastrow slot; with one u for
    every spin pair, so does tc3b1, which then checks that each n_s2 weights its own term
    (FP64, 1e-10)."""
    sa = S.tiny_system(method=method, n=(12, 12, 12), kmesh=(2, 1, 1), nbands=4, nocc=3, seed_stream=195)
    sb = S.tiny_system(method=method, n=(12, 12, 12), kmesh=(2, 1, 1), nbands=4, nocc=2, seed_stream=196)
    fa = np.tile([1.0, 1.0, 0.6], sa.nk)
    fb = np.tile([1.0, 0.3], sb.nk)
    nb0, phi0, nb1, phi1, nb2, phi2 = both_channels(sa, [3, 2], 197)
    n0 = phi0.size
    chans = ((nb0, phi0, 0, n0), (nb1, phi1, n0, phi2.size))
    rows = {"tc3a1": (np.array([[0.5, 0.5], [0.8, 0.8]]), np.array([[1.1, 1.1], [0.7, 0.7]])),
            "tc3b1": (np.full((2, 2), 0.6), np.full((2, 2), 0.9))}
    for term, (A, F) in rows.items():
        two = two_spin(dataclasses.replace(sa, A=A, F=F), dataclasses.replace(sb, A=A, F=F), fa, fb)
        engine.load_system(two)
        got = getattr(engine, term)(nb2, phi2)
        for s, (nbs, phis, lo, hi) in enumerate(chans):
            one = dataclasses.replace(combined(sa, sb, fa, fb), A=np.full((2, 2), A[s, 0]), F=np.full((2, 2), F[s, 0]))
            ref = getattr(O.Oracle(one), term)(nbs, phis)
            err = np.abs(got[lo:hi] - ref).max() / np.abs(ref).max()
            print(f"{method} {term} spin {s}: {err:.2e}", flush=True)
            assert err <= 1e-10, (term, s)


def test_hf_all_is_the_sum_of_its_terms(engine):
    sa = S.tiny_system(method="HF", n=(8, 8, 8), kmesh=(2, 1, 1), nbands=3, nocc=2, seed_stream=171)
    sb = S.tiny_system(method="HF", n=(8, 8, 8), kmesh=(2, 1, 1), nbands=3, nocc=1, seed_stream=172)
    engine.load_system(two_spin(sa, sb))
    _, _, _, _, nb2, phi2 = both_channels(sa, [2, 3], 173)
    h1, h2, _ = engine.apply_all(nb2, phi2)
    k = engine.kinetic(nb2, phi2)
    h = engine.hf2h(nb2, phi2) + engine.hf2x(nb2, phi2)
    assert np.abs(h1 - k).max() <= 1e-14 * np.abs(k).max()
    assert np.abs(h2 - h).max() <= 1e-12 * np.abs(h).max()


def test_spin_coupling_terms_are_refused(engine):
    s1 = S.tiny_system(method="BITC", n=(8, 8, 8), kmesh=(2, 1, 1), nbands=3, nocc=2, seed_stream=181)
    engine.load_system(two_spin(s1, s1))
    nb, phi = S.applied_bands(s1, nb=[1, 1], seed=182)
    nb2, phi2 = np.concatenate([nb, nb]).astype(np.int32), np.concatenate([phi, phi])
    for t in ("tc3a3b3b4b6", "tc3a2a4b2b5"):
        with pytest.raises(tcx.TcxError) as e:
            getattr(engine, t)(nb2, phi2)
        assert e.value.code == tcx.TCX_ERR_UNSUPPORTED, t
    # the two-body terms still run on the same context
    assert np.isfinite(engine.tc2x(nb2, phi2)).all() and np.isfinite(engine.tc2h(nb2, phi2)).all()
    engine.load_system(s1)  # one spin again
