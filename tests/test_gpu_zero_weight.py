"""Zero-weight k-points (BAND k-points, zero-weight SCF k-points) on the GPU against the oracle.

A 2x1x1 mesh carries the occupied orbitals; zero-weight k-points are added near a mesh
point (k0 +- 0.05 e, where the kVaux sums are large), at a generic point, and exactly on
a mesh point.  Their divergence terms project on given states (here the occupied
orbitals of a mesh point, with their fillings).  Every term, the kVaux branches of tc2x,
tc3a3b3b4b6 and tc3a2a4b2b5 among them, is compared with the oracle (itself pinned by
tests/test_oracle_zero_weight.py):
  FP64: max |delta| <= 1e-10 max |ref|;  FP32-split: per band ||delta|| / ||ref|| <= 3e-5.
A Davidson run at the zero-weight k-points (BAND mode: bands at the band k-points under the
SCF occupied set) reproduces the lowest eigenvalues of the dense H assembled from the oracle.
"""
import numpy as np
import pytest

from oracle import oracle as O
from tc_amd import system as S
from tc_amd import tcx
from test_gpu_hf2x import per_band_rel

pytestmark = pytest.mark.gpu

E_DIR = np.array([1.0, 0.3, -0.2]) / np.linalg.norm([1.0, 0.3, -0.2])
TERMS = {"HF": ("hf2x", "kinetic", "hf2h"),
         "TC": ("tc2x", "tc3a1", "tc3b1", "tc3a3b3b4b6", "tc3a2a4b2b5", "kinetic", "hf2h", "tc2h")}
TERMS["BITC"] = TERMS["TC"]
FP64_ONLY = ("tc3a1", "tc3b1", "kinetic", "hf2h", "tc2h")


def zero_weight_system(method, n=(8, 8, 8), seed_stream=101):
    base = S.tiny_system(method=method, n=n, kmesh=(2, 1, 1), nbands=3, nocc=2, seed_stream=seed_stream)
    k0, k1 = np.asarray(base.kfrac)[0], np.asarray(base.kfrac)[1]
    extra = [k0 + 0.05 * E_DIR, k0 - 0.05 * E_DIR, [0.21, -0.13, 0.07], k1]
    return S.with_zero_weight_kpoints(base, extra, [0, 0, 1, 1])


@pytest.mark.parametrize("method", ["HF", "TC", "BITC"])
@pytest.mark.parametrize("prec", [tcx.PREC_FP64, tcx.PREC_FP32], ids=["fp64", "fp32"])
def test_zero_weight_terms_vs_oracle(engine, method, prec):
    sys_ = zero_weight_system(method)
    engine.load_system(sys_)
    nb = np.array([1, 0, 2, 1, 1, 2], np.int32)
    _, phi = S.applied_bands(sys_, nb=nb, seed=102)
    o = O.Oracle(sys_)
    for term in TERMS[method]:
        ref = getattr(o, term)(nb, phi)
        if term in FP64_ONLY:
            got = getattr(engine, term)(nb, phi)
            p = tcx.PREC_FP64
        else:
            got = getattr(engine, term)(nb, phi, precision=prec)
            p = prec
        if p == tcx.PREC_FP64:
            err = np.abs(got - ref).max() / np.abs(ref).max()
            print(f"{method} {term} fp64 max rel {err:.3e}", flush=True)
            assert err <= 1e-10, term
        else:
            rel = per_band_rel(sys_, nb, got, ref)
            print(f"{method} {term} fp32 per-band rel max {rel.max():.3e}", flush=True)
            assert rel.max() <= 3e-5, term


@pytest.mark.parametrize("method", ["TC", "BITC"])
def test_zero_weight_on_a_mesh_point_equals_scf(engine, method):
    """The zero-weight copy of mesh point k1 (last extra k) gives the SCF result of k1."""
    sys_ = zero_weight_system(method)
    engine.load_system(sys_)
    rng = np.random.default_rng(103)
    v = rng.standard_normal(sys_.npw[1]) + 1j * rng.standard_normal(sys_.npw[1])
    nb = np.array([0, 1, 0, 0, 0, 1], np.int32)
    phi = np.concatenate([v, v])
    for term in ("tc2x", "tc3a3b3b4b6", "tc3a2a4b2b5"):
        h = getattr(engine, term)(nb, phi)
        n = len(v)
        err = np.abs(h[n:] - h[:n]).max() / np.abs(h[:n]).max()
        print(method, term, err, flush=True)
        assert err <= 1e-11


def test_zero_weight_davidson_band_mode(engine):
    """BAND mode: Davidson for bands at zero-weight k-points under the SCF occupied set
    (HF) converges to the lowest eigenvalues of the dense H_k assembled from the oracle's
    kinetic + hf2h + hf2x."""
    sys_ = zero_weight_system("HF")
    engine.load_system(sys_)
    o = O.Oracle(sys_)
    nbk = 2
    ks = [2, 4]
    nb = np.zeros(sys_.nk, np.int32)
    nb[ks] = nbk
    rng = np.random.default_rng(104)
    bo = sys_.band_offsets(nb)
    phi = np.zeros(int(bo[-1]), np.complex128)
    for k in ks:
        blk = rng.standard_normal((sys_.npw[k], nbk)) + 1j * rng.standard_normal((sys_.npw[k], nbk))
        q, _ = np.linalg.qr(blk)
        phi[bo[k]:bo[k + 1]] = q.T.reshape(-1)
    blocks = int(np.ceil(max(sys_.npw[k] for k in ks) / nbk)) + 2
    phi_out, _, eig, en = engine.davidson(nb, phi, np.zeros(int(nb.sum()), np.complex128), max_blocks=blocks)
    eye_nb = np.zeros(sys_.nk, np.int32)
    eye_nb[ks] = [sys_.npw[k] for k in ks]
    eye = np.concatenate([np.eye(sys_.npw[k], dtype=np.complex128).reshape(-1) for k in ks])
    H = o.kinetic(eye_nb, eye) + o.hf2h(eye_nb, eye) + o.hf2x(eye_nb, eye)
    eo = sys_.band_offsets(eye_nb)
    for i, k in enumerate(ks):
        n = sys_.npw[k]
        M = H[eo[k]:eo[k + 1]].reshape(n, n).T
        w = np.sort(np.linalg.eigvals(M).real)[:nbk]
        got = eig[i * nbk:(i + 1) * nbk].real
        print("BAND k", k, w, got - w, flush=True)
        assert np.abs(got - w).max() < 1e-9 * max(1.0, np.abs(w).max())
