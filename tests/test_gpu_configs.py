"""GPU parity at the grid sizes and shapes of the benchmark workloads, against the FP64 oracle.

Each test id names the workload it covers:
  C1  hf2x, fcc a=10.26, 4x4x4 k, 16 bands, 40^3 at full size (the benchmark's launch
      shapes and kernel instances): two bands, (k=0, j=0) and one seeded (k, j), against
      the oracle, in FP64 and FP32-split; E_x in FP32-split vs FP64 over all bands;
  C2  the TC/BITC terms at 40^3 on a reduced k-mesh and band count against the oracle,
      plus full-size property checks: FP32-split vs FP64 term energies, and BITC with
      chi = phi equal to TC;
  C3  reduced cells at 32^3, 36^3 and 48^3 against the oracle, and captured HIP graphs of
      the three grids replayed together, every replay against the oracle;
  C4  the gradients on a reduced 48^3 cell: dE_x/d conj(phi) against the oracle's V_x phi,
      and the BITC two-body gradients against central differences.

Tolerances (per result, written where they are checked):
  FP64 apply : max |delta| <= 1e-10 * max |ref|
  FP32-split : per band ||delta|| / ||ref|| <= 3e-5 (C1: 1e-6), term energies |dE| <= 1e-6 Ha
"""
import numpy as np
import pytest

from oracle import oracle as O
from tc_amd import energy as E
from tc_amd import system as S
from tc_amd import tcx
from test_gpu_hf2x import per_band_rel

pytestmark = pytest.mark.gpu

TERMS_TC = ("tc2x", "tc3a1", "tc3b1", "tc3a3b3b4b6", "tc3a2a4b2b5")
NBODY = {"hf2x": 2, "tc2x": 2, "tc3a1": 3, "tc3b1": 3, "tc3a3b3b4b6": 3, "tc3a2a4b2b5": 3}


def _reduced_cell(grid, method, seed_stream):
    """fcc a=10.26 at the given grid, 2x1x1 k, 2 occupied bands (the benchmark cell's
    plane-wave sphere rule and Jastrow A=0.5 for TC/BITC)."""
    A = np.full((2, 2), 0.5) if method != "HF" else None
    return S.build_system(S.fcc_lattice(10.26), (grid,) * 3, (2, 1, 1), nbands=2, nocc=2,
                          seed_stream=seed_stream, method=method, A=A,
                          name=f"{method} fcc a10.26 2x1x1k 2b {grid}^3")


def _check(sys_, nb, got, ref, prec, what, tol32=3e-5):
    if prec == tcx.PREC_FP64:
        err = np.abs(got - ref).max() / np.abs(ref).max()
        print(f"{what} fp64 max rel {err:.3e}", flush=True)
        assert err <= 1e-10, what
    else:
        rel = per_band_rel(sys_, nb, got, ref)
        print(f"{what} fp32 per-band rel max {rel.max():.3e}", flush=True)
        assert rel.max() <= tol32, what


def _one_band(sys_, nb_all, phi_all, k, j):
    """(nb, phi) of band j at k alone, cut out of a ragged block."""
    bo = sys_.band_offsets(nb_all)
    nb = np.zeros(sys_.nk, np.int32)
    nb[k] = 1
    npw = int(sys_.npw[k])
    return nb, phi_all[bo[k] + j * npw: bo[k] + (j + 1) * npw].copy()


# ------------------------------------------------------------------------------- C1
@pytest.fixture(scope="module")
def c1():
    return S.fcc_hf_40()


@pytest.mark.parametrize("prec", [tcx.PREC_FP64, tcx.PREC_FP32], ids=["fp64", "fp32"])
def test_C1_hf2x_fcc40_4x4x4k_16b_bands_vs_oracle(engine, c1, prec):
    """Full C1 apply (all 1024 bands, the benchmark's launch shapes); bands (0,0) and a seeded
    (k,j) compared with the oracle's V_x of that band alone, and their band energies."""
    from tc_amd import autotune as AT
    sys_ = c1
    nb, phi = sys_.nocc.copy(), sys_.phi_occ
    engine.load_system(sys_)
    entry = AT.load_table().get(AT.signature(sys_, nb, "hf2x", prec))
    engine.set_launch_shape("hf2x", *((entry["targets"], entry["sources"]) if entry else (0, 0)))
    got = engine.hf2x(nb, phi, precision=prec)
    engine.set_launch_shape("hf2x", 0, 0)
    rng = np.random.default_rng(11)
    picks = [(0, 0), (int(rng.integers(1, sys_.nk)), int(rng.integers(0, 16)))]
    o = O.Oracle(sys_)
    bo = sys_.band_offsets(nb)
    for k, j in picks:
        nb1, phi1 = _one_band(sys_, nb, phi, k, j)
        ref = o.hf2x(nb1, phi1)
        npw = int(sys_.npw[k])
        g = got[bo[k] + j * npw: bo[k] + (j + 1) * npw]
        tol = 1e-10 if prec == tcx.PREC_FP64 else 1e-6
        rel = np.linalg.norm(g - ref) / np.linalg.norm(ref)
        e_ref, e_got = np.vdot(phi1, ref), np.vdot(phi1, g)
        print(f"C1 band (k={k}, j={j}) prec={prec}: rel {rel:.3e}, <phi|Vx phi> {e_got.real:.12f} vs "
              f"{e_ref.real:.12f}", flush=True)
        assert rel <= tol
        assert abs(e_got - e_ref) <= tol * abs(e_ref)


def test_C1_hf2x_fcc40_4x4x4k_16b_energy_fp32_vs_fp64(engine, c1):
    sys_ = c1
    nb, phi = sys_.nocc.copy(), sys_.phi_occ
    engine.load_system(sys_)
    h64 = engine.hf2x(nb, phi, precision=tcx.PREC_FP64)
    h32 = engine.hf2x(nb, phi, precision=tcx.PREC_FP32)
    e64, e32 = E.term_energy(sys_, h64, 2), E.term_energy(sys_, h32, 2)
    rel = per_band_rel(sys_, nb, h32, h64)
    print(f"C1 E_x fp64 {e64.real:.10f} fp32 {e32.real:.10f} dE {abs(e32 - e64):.3e} band rel {rel.max():.3e}",
          flush=True)
    assert abs(e32 - e64) < 1e-6
    assert rel.max() < 1e-6


# ------------------------------------------------------------------------------- C2 at 40^3
@pytest.mark.parametrize("method", ["TC", "BITC"])
@pytest.mark.parametrize("prec", [tcx.PREC_FP64, tcx.PREC_FP32], ids=["fp64", "fp32"])
def test_C2_terms_40_reduced_vs_oracle(engine, method, prec):
    sys_ = _reduced_cell(40, method, seed_stream=41)
    engine.load_system(sys_)
    nb, phi = S.applied_bands(sys_, nb=[2, 1], seed=42)
    o = O.Oracle(sys_)
    for term in TERMS_TC:
        ref = getattr(o, term)(nb, phi)
        got = getattr(engine, term)(nb, phi, precision=prec)
        # tc3a1/tc3b1 always run in FP64
        _check(sys_, nb, got, ref, prec if term not in ("tc3a1", "tc3b1") else tcx.PREC_FP64,
               f"C2 40^3 {method} {term}")
    # energies of the occupied set through the heaviest terms, FP32-split or FP64 vs the oracle
    for term in ("tc2x", "tc3a3b3b4b6", "tc3a2a4b2b5"):
        ref = getattr(o, term)(sys_.nocc, sys_.phi_occ)
        got = getattr(engine, term)(sys_.nocc, sys_.phi_occ, precision=prec)
        e_ref = E.term_energy(sys_, ref, NBODY[term], left=True)
        e_got = E.term_energy(sys_, got, NBODY[term], left=True)
        print(f"C2 40^3 {method} {term} energy {e_got} vs {e_ref}", flush=True)
        assert abs(e_got - e_ref) < (1e-10 if prec == tcx.PREC_FP64 else 1e-6)


@pytest.fixture(scope="module")
def c2():
    return S.fcc_bitc_40()


@pytest.mark.parametrize("term", ["tc2x", "tc3a3b3b4b6", "tc3a2a4b2b5"])
def test_C2_bitc_fcc40_4x4x4k_16b_fp32_vs_fp64(engine, c2, term):
    """Full C2: FP32-split term energy within 1e-6 Ha of the FP64 apply."""
    sys_ = c2
    nb, phi = sys_.nocc.copy(), sys_.phi_occ
    engine.load_system(sys_)
    h32 = getattr(engine, term)(nb, phi, precision=tcx.PREC_FP32)
    h64 = getattr(engine, term)(nb, phi, precision=tcx.PREC_FP64)
    e32 = E.term_energy(sys_, h32, NBODY[term], left=True)
    e64 = E.term_energy(sys_, h64, NBODY[term], left=True)
    rel = per_band_rel(sys_, nb, h32, h64)
    print(f"C2 {term}: E fp64 {e64} dE {abs(e32 - e64):.3e} band rel {rel.max():.3e}", flush=True)
    assert abs(e32 - e64) < 1e-6
    assert rel.max() < 3e-5


@pytest.mark.parametrize("term", ["tc2x", "tc3a3b3b4b6", "tc3a2a4b2b5"])
def test_C2_fcc40_4x4x4k_16b_bitc_with_chi_eq_phi_is_tc(engine, c2, term):
    """Full C2 cell: BITC with the left orbitals equal to the right ones reduces to TC."""
    tc = S.System(**{**c2.__dict__, "method": "TC", "chi_occ": None})
    bitc = S.System(**{**c2.__dict__, "method": "BITC", "chi_occ": c2.phi_occ.copy()})
    nb, phi = c2.nocc.copy(), c2.phi_occ
    engine.load_system(tc)
    a = getattr(engine, term)(nb, phi, precision=tcx.PREC_FP32)
    engine.load_system(bitc)
    b = getattr(engine, term)(nb, phi, precision=tcx.PREC_FP32)
    rel = per_band_rel(c2, nb, b, a)
    print(f"C2 {term}: BITC(chi=phi) vs TC band rel {rel.max():.3e}", flush=True)
    assert rel.max() < 2e-6


# ------------------------------------------------------------------------------- C3 / C4 grids
@pytest.mark.parametrize("grid", [32, 36, 48])
@pytest.mark.parametrize("method", ["TC", "BITC"])
@pytest.mark.parametrize("prec", [tcx.PREC_FP64, tcx.PREC_FP32], ids=["fp64", "fp32"])
def test_C3_grids_reduced_vs_oracle(engine, grid, method, prec):
    sys_ = _reduced_cell(grid, method, seed_stream=43 + grid)
    engine.load_system(sys_)
    nb, phi = S.applied_bands(sys_, nb=[1, 2], seed=44)
    o = O.Oracle(sys_)
    hf = S.System(**{**sys_.__dict__, "method": "HF", "chi_occ": None, "A": np.zeros((2, 2))})
    for term in ("tc2x", "tc3a3b3b4b6", "tc3a2a4b2b5"):
        ref = getattr(o, term)(nb, phi)
        got = getattr(engine, term)(nb, phi, precision=prec)
        _check(sys_, nb, got, ref, prec, f"{grid}^3 {method} {term}")
    engine.load_system(hf)
    ref = O.Oracle(hf).hf2x(nb, phi)
    got = engine.hf2x(nb, phi, precision=prec)
    _check(hf, nb, got, ref, prec, f"{grid}^3 hf2x")


def test_C3_mixed_grid_graphs_vs_oracle():
    """The batched uneven workload in miniature: three cells at 32^3, 36^3 and 48^3 with
    uneven band counts, hf2x + tc2x FP32-split captured per cell and the three graphs
    replayed together; every replay is checked against the oracle."""
    rng = np.random.default_rng(45)
    cells = []
    for g, grid in enumerate((32, 36, 48)):
        base = S.build_system(S.fcc_lattice(10.26), (grid,) * 3, (2, 2, 1), nbands=4, nocc=4,
                              seed_stream=46 + g, method="TC", A=np.full((2, 2), 0.5))
        cells.append(S.uneven_bands(base, rng.integers(1, 5, size=base.nk)))
    runs = []
    try:
        for sys_ in cells:
            eng = tcx.Engine(0)
            eng.load_system(sys_)
            nb, phi = sys_.nocc.copy(), sys_.phi_occ
            o = O.Oracle(sys_)
            hf = S.System(**{**sys_.__dict__, "method": "HF"})
            ref = o.tc2x(nb, phi) + O.Oracle(hf).hf2x(nb, phi)
            d_phi = eng.to_device(phi)
            d_out = eng.device_buffer(phi.nbytes)

            def calls(eng=eng, nb=nb, d_phi=d_phi, d_out=d_out):
                d_out.zero()
                eng._term_device("hf2x", nb, d_phi.ptr, d_out.ptr, tcx.PREC_FP32)
                eng._term_device("tc2x", nb, d_phi.ptr, d_out.ptr, tcx.PREC_FP32)

            calls()
            eng.synchronize()
            eng.graph_begin()
            calls()
            gid = eng.graph_end()
            runs.append((sys_, eng, nb, d_phi, d_out, gid, ref))
        for rep in range(3):
            for sys_, eng, nb, d_phi, d_out, gid, ref in runs:
                d_out.upload(np.full(sys_.phi_occ.size, 3.0 - 1j))   # stale data the graph must clear
            for sys_, eng, nb, d_phi, d_out, gid, ref in runs:
                eng.graph_launch(gid)
            for sys_, eng, nb, d_phi, d_out, gid, ref in runs:
                eng.synchronize()
                got = d_out.download(np.complex128)
                rel = per_band_rel(sys_, nb, got, ref)
                print(f"replay {rep} grid {sys_.n[0]}: per-band rel {rel.max():.3e}", flush=True)
                assert rel.max() <= 3e-5
    finally:
        for sys_, eng, nb, d_phi, d_out, gid, ref in runs:
            eng.graph_destroy(gid)
            d_phi.free()
            d_out.free()
            eng.close()


def test_C4_gradients_48_reduced(engine):
    """dE_x/d conj(phi) = (f/Nk) V_x phi against the oracle, FP32-split and FP64, and the BITC
    two-body gradients (FP32-split forward) against central differences of FP64 energies."""
    hfc = _reduced_cell(48, "HF", seed_stream=47)
    engine.load_system(hfc)
    ref = O.Oracle(hfc).hf2x(hfc.nocc, hfc.phi_occ)
    fb = np.concatenate([np.repeat(hfc.filling[int(np.sum(hfc.nocc[:k])):int(np.sum(hfc.nocc[:k + 1]))],
                                   hfc.npw[k]) for k in range(hfc.nk)])
    gref = fb / hfc.num_kpoints * ref
    for prec, tol in ((tcx.PREC_FP64, 1e-10), (tcx.PREC_FP32, 3e-5)):
        e, g = engine.grad_exchange(prec)
        rel = np.abs(g - gref).max() / np.abs(gref).max()
        print(f"C4 48^3 dE_x/dphi* prec={prec} rel {rel:.3e}", flush=True)
        assert rel <= tol

    sys_ = _reduced_cell(48, "BITC", seed_stream=48)
    h = 1e-4
    rng = np.random.default_rng(49)
    d = rng.standard_normal(sys_.phi_occ.size) + 1j * rng.standard_normal(sys_.phi_occ.size)
    engine.load_system(sys_)
    _, gl, gr, dA = engine.grad_tc2(tcx.PREC_FP32)

    def energy(**kw):
        engine.load_system(S.System(**{**sys_.__dict__, **kw}))
        return engine.grad_tc2(tcx.PREC_FP64)[0]

    fd_l = (energy(chi_occ=sys_.chi_occ + h * d) - energy(chi_occ=sys_.chi_occ - h * d)) / (2 * h)
    fd_r = (energy(phi_occ=sys_.phi_occ + h * d) - energy(phi_occ=sys_.phi_occ - h * d)) / (2 * h)
    an_l, an_r = np.vdot(d, gl), np.sum(d * gr)
    fdA = (energy(A=sys_.A + h) - energy(A=sys_.A - h)) / (2 * h)
    print(f"C4 48^3 BITC: left {fd_l} vs {an_l}; right {fd_r} vs {an_r}; dA {fdA} vs {dA}", flush=True)
    # FP32-split analytic gradients against FP64 differences: FP32 product rounding ~1e-7 relative
    assert abs(fd_l - an_l) < 1e-5 * max(1.0, abs(an_l))
    assert abs(fd_r - an_r) < 1e-5 * max(1.0, abs(an_r))
    assert abs(fdA - dA) < 1e-5 * max(1.0, abs(dA))
    engine.load_system(sys_)
