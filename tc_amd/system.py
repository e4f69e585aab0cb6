"""Synthetic plane-wave systems for the pair-FFT operator apply.

Builds what TC++ would otherwise read from a Quantum ESPRESSO run: the cell
and its reciprocal vectors, a Gamma-centred Monkhorst-Pack k-mesh with
identity-only symmetry (every k irreducible, one isym, no time reversal),
the plane-wave sphere at each k as FFT-grid linear indices, and seeded
orthonormal orbitals with fillings.

Conventions follow the reference:
  * reciprocal vectors are the ROWS b_i, Cartesian G = recip.T @ m
    (src/crystal_structure.hpp:15-17, src/calc_hamiltonian_hf2x.cpp:35-36);
  * FFT-grid index ipw = i1 + n1*(i2 + n2*i3) with Miller m_d wrapped to
    (-n_d/2, n_d/2] (src/plane_wave_basis.cpp:137-140, :205-216);
  * orbitals are normalised, sum_G |c(G)|^2 = 1
    (src/bloch_states_initialize.cpp:264), orthonormalised by classical
    Gram-Schmidt (src/diagonalization_utils.cpp:7-29), and left/right sets
    for BITC by Gram_Schmidt_biortho (src/diagonalization_utils.cpp:31-77).

The plane-wave sphere is |k+G| <= G_w with G_w = min_i (n_i-1)/4 * 2*pi/|a_i|,
the rule under which the pair density (radius 2*G_w) fits the FFT grid
without aliasing.
"""
from __future__ import annotations

import dataclasses
import math

import numpy as np

FOURPI = 4.0 * math.pi


def sc_lattice(a: float) -> np.ndarray:
    return a * np.eye(3)


def fcc_lattice(a: float) -> np.ndarray:
    return 0.5 * a * np.array([[0.0, 1.0, 1.0], [1.0, 0.0, 1.0], [1.0, 1.0, 0.0]])


def reciprocal_rows(lattice: np.ndarray) -> np.ndarray:
    """Rows b_i with a_i . b_j = 2*pi*delta_ij."""
    return 2.0 * math.pi * np.linalg.inv(lattice).T


def mp_mesh(nx: int, ny: int, nz: int) -> np.ndarray:
    """Gamma-centred mesh in crystal coordinates, components in {0, 1/n, ...}."""
    pts = [(i / nx, j / ny, k / nz) for i in range(nx) for j in range(ny) for k in range(nz)]
    return np.array(pts, dtype=np.float64)


def miller_of_index(n) -> np.ndarray:
    """Miller triple for every FFT-grid linear index (plane_wave_basis.cpp:205-216)."""
    n1, n2, n3 = n
    ipw = np.arange(n1 * n2 * n3)
    i1 = ipw % n1
    i2 = (ipw // n1) % n2
    i3 = ipw // (n1 * n2)
    m1 = np.where(i1 <= n1 // 2, i1, i1 - n1)
    m2 = np.where(i2 <= n2 // 2, i2, i2 - n2)
    m3 = np.where(i3 <= n3 // 2, i3, i3 - n3)
    return np.stack([m1, m2, m3], axis=1)


def gw_rule(lattice: np.ndarray, n) -> float:
    return min((n[i] - 1) / 4.0 * 2.0 * math.pi / np.linalg.norm(lattice[i]) for i in range(3))


def gram_schmidt(vecs: np.ndarray) -> np.ndarray:
    """Classical Gram-Schmidt band by band (diagonalization_utils.cpp:7-29)."""
    out = np.array(vecs, dtype=np.complex128, copy=True)
    for b in range(out.shape[0]):
        t = out[b]
        for a in range(b):
            t = t - np.vdot(out[a], t) * out[a]
        nrm = np.linalg.norm(t)
        if nrm < 1e-8:
            raise ValueError("Gram-Schmidt: linearly dependent vector")
        out[b] = t / nrm
    return out


def gram_schmidt_biortho(right: np.ndarray, left: np.ndarray):
    """Bi-orthonormalise (diagonalization_utils.cpp:31-77): <left_a|right_b> = delta_ab,
    each right vector of unit norm."""
    R = np.array(right, dtype=np.complex128, copy=True)
    L = np.array(left, dtype=np.complex128, copy=True)
    for b in range(R.shape[0]):
        tr = R[b]
        for a in range(b):
            tr = tr - np.vdot(L[a], tr) * R[a]
        nrm = np.linalg.norm(tr)
        if nrm < 1e-10:
            raise ValueError("biorthogonal Gram-Schmidt: zero right vector")
        tr = tr / nrm
        tl = L[b]
        for a in range(b):
            tl = tl - np.vdot(R[a], tl) * L[a]
        c = np.vdot(tl, tr)
        if abs(c) < 1e-10:
            raise ValueError("biorthogonal Gram-Schmidt: zero overlap")
        tl = tl / np.conj(c)
        R[b] = tr
        L[b] = tl
    return R, L


@dataclasses.dataclass
class System:
    """One spin-unpolarised plane-wave system on an explicit k-mesh (identity
    symmetry unless `symmetry` relates its k-points).

    Ragged per-k arrays are stored concatenated over k (k-major, band-major,
    then plane wave), complex128.
    """
    lattice: np.ndarray            # rows a_i (bohr)
    recip: np.ndarray              # rows b_i (bohr^-1)
    volume: float
    n: tuple                       # FFT grid (n1, n2, n3)
    kfrac: np.ndarray              # (nk, 3)
    npw: np.ndarray                # (nk,) int32
    gindex: np.ndarray             # concat int32
    nocc: np.ndarray               # (nk,) int32: occupied bands (incl. "old" ones)
    filling: np.ndarray            # concat over k, float64
    phi_occ: np.ndarray            # concat complex128 [k][i][npw_k]
    chi_occ: np.ndarray | None     # BITC left orbitals, same layout
    nspin: int = 1
    method: str = "HF"             # HF, TC or BITC
    A: np.ndarray = dataclasses.field(default_factory=lambda: np.zeros((2, 2)))
    F: np.ndarray = dataclasses.field(default_factory=lambda: np.ones((2, 2)))
    div_corr: bool = True
    name: str = ""
    # k-mesh symmetry (tcx_set_kpoint_symmetry) or None for identity symmetry:
    # {"k_src": (nk,), "rot": (nk, 3, 3), "trans": (nk, 3), "trev": (nk,)}
    symmetry: dict | None = None
    # zero-weight k-points (BAND k-points, fake-SCF k-points): kweight[k] < 1e-8 marks
    # one; None = every k is a weighted mesh point.  Zero-weight k hold no occupied
    # orbitals (nocc = 0); the states their divergence terms project on, with their
    # return_band_filling values: {"nstate": (nk,), "fill": concat, "phi": concat
    # [k][b][npw_k], "chi": BITC left states or None}
    kweight: np.ndarray | None = None
    zero_weight_states: dict | None = None
    # polynomial Jastrow (tcx_set_jastrow_polynomial) or None: {"deg": int, "C": 2 or 3,
    # "L": (2, 2), "c": (2, 2, deg) c_poly_input with entry 0 replaced by the cusp
    # condition, "cusp": bool}; with A != 0 it is the RPA+polynomial form.  The oracle also
    # reads "hermitian_conj": True (c_poly_internal negated, as for H^dagger)
    jastrow_poly: dict | None = None

    @property
    def nk(self) -> int:
        return len(self.npw)

    @property
    def num_kpoints(self) -> int:
        """Nk of the 1/(Nk V) prefactors: the weighted k-points (Kpoints::num_kpoints)."""
        if self.kweight is None:
            return self.nk
        return int(np.sum(np.asarray(self.kweight) >= 1e-8))

    @property
    def ngrid(self) -> int:
        return int(self.n[0] * self.n[1] * self.n[2])

    @property
    def jastrow_A_zero(self) -> bool:
        return bool(np.all(np.abs(self.A) <= 1e-8))

    @property
    def method_id(self) -> int:
        return {"HF": 0, "TC": 1, "BITC": 2}[self.method]

    def pw_offsets(self) -> np.ndarray:
        return np.concatenate([[0], np.cumsum(self.npw)]).astype(np.int64)

    def occ_offsets(self) -> np.ndarray:
        """Offsets (complex numbers) of each k block of phi_occ."""
        return np.concatenate([[0], np.cumsum(self.nocc.astype(np.int64) * self.npw)]).astype(np.int64)

    def band_offsets(self, nb) -> np.ndarray:
        return np.concatenate([[0], np.cumsum(np.asarray(nb, np.int64) * self.npw)]).astype(np.int64)

    def occ_block(self, ik: int, left: bool = False) -> np.ndarray:
        off = self.occ_offsets()
        src = self.chi_occ if (left and self.chi_occ is not None) else self.phi_occ
        return src[off[ik]:off[ik + 1]].reshape(self.nocc[ik], self.npw[ik])

    def gcart_at_k(self, ik: int) -> np.ndarray:
        """Cartesian k+G of the basis at k (rows)."""
        po = self.pw_offsets()
        mill = miller_of_index(self.n)[self.gindex[po[ik]:po[ik + 1]]]
        return (self.kfrac[ik][None, :] + mill) @ self.recip

    def num_electrons(self) -> float:
        return float(self.filling.sum()) / self.num_kpoints


def build_system(lattice, n, kmesh, nbands, nocc, *, seed0=1234, seed_stream=0, method="HF",
                 A=None, F=None, div_corr=True, left_seed_offset=500, name="",
                 filling_value=2.0, gw=None) -> System:
    """Synthetic system: orbitals c(G) = complex Gaussian * exp(-|k+G|^2/2),
    Gram-Schmidt orthonormalised per k; seed = seed0 + 1000*seed_stream + ik.

    ``nbands`` orbitals are generated per k and the first ``nocc`` of them are
    occupied with ``filling_value`` (spin factor 2 included).  For BITC, an
    independent left set is drawn with seed + left_seed_offset and the pair
    is bi-orthonormalised.
    """
    lattice = np.asarray(lattice, dtype=np.float64)
    recip = reciprocal_rows(lattice)
    volume = abs(float(np.linalg.det(lattice)))
    n = tuple(int(x) for x in n)
    kfrac = mp_mesh(*kmesh)
    if gw is None:
        gw = gw_rule(lattice, n)
    mill = miller_of_index(n)
    gcart0 = mill @ recip
    npw, gidx, phis, chis = [], [], [], []
    for ik, kf in enumerate(kfrac):
        kc = kf @ recip
        kg2 = np.sum((gcart0 + kc) ** 2, axis=1)
        sel = np.nonzero(kg2 <= gw * gw + 1e-12)[0]
        # deterministic order: by |k+G|^2, ties by grid index
        order = np.lexsort((sel, np.round(kg2[sel], 10)))
        sel = sel[order]
        npw.append(len(sel))
        gidx.append(sel.astype(np.int32))
        rng = np.random.default_rng(seed0 + 1000 * seed_stream + ik)
        env = np.exp(-kg2[sel] / 2.0)
        c = (rng.standard_normal((nbands, len(sel))) + 1j * rng.standard_normal((nbands, len(sel)))) * env
        if method == "BITC":
            rngl = np.random.default_rng(seed0 + 1000 * seed_stream + ik + left_seed_offset)
            cl = (rngl.standard_normal((nbands, len(sel))) + 1j * rngl.standard_normal((nbands, len(sel)))) * env
            # left = right + perturbation keeps the pair well conditioned
            R, L = gram_schmidt_biortho(gram_schmidt(c), gram_schmidt(c) + 0.3 * gram_schmidt(cl))
            phis.append(R[:nocc])
            chis.append(L[:nocc])
        else:
            phis.append(gram_schmidt(c)[:nocc])
    nk = len(kfrac)
    Am = np.zeros((2, 2)) if A is None else np.asarray(A, dtype=np.float64)
    if F is None:
        # cusp condition, jastrow_initialize.cpp:74-75
        Fm = np.array([[math.sqrt(2 * Am[0, 0]), math.sqrt(Am[0, 1])],
                       [math.sqrt(Am[1, 0]), math.sqrt(2 * Am[1, 1])]]) if np.any(np.abs(Am) > 1e-8) else np.ones((2, 2))
    else:
        Fm = np.asarray(F, dtype=np.float64)
    return System(
        lattice=lattice, recip=recip, volume=volume, n=n, kfrac=kfrac,
        npw=np.array(npw, dtype=np.int32), gindex=np.concatenate(gidx).astype(np.int32),
        nocc=np.full(nk, nocc, dtype=np.int32),
        filling=np.full(nk * nocc, filling_value, dtype=np.float64),
        phi_occ=np.concatenate([p.reshape(-1) for p in phis]).astype(np.complex128),
        chi_occ=(np.concatenate([c.reshape(-1) for c in chis]).astype(np.complex128) if chis else None),
        method=method, A=Am, F=Fm, div_corr=div_corr, name=name)


# ---------------------------------------------------------------------------
# Named workloads
# ---------------------------------------------------------------------------

def heg_sc_rs4_24(**kw) -> System:
    """HEG, simple cubic, rs=4 bohr, 14 electrons/cell, 2x2x2 k, 8 bands (7 occupied), 24^3."""
    vol = 14 * 4.0 / 3.0 * math.pi * 4.0 ** 3
    a = vol ** (1.0 / 3.0)
    return build_system(sc_lattice(a), (24, 24, 24), (2, 2, 2), nbands=8, nocc=7, seed_stream=0,
                        name="hf2x heg-sc rs4 2x2x2k 8b(7occ) 24^3", **kw)


def fcc_hf_40(**kw) -> System:
    """fcc a=10.26 bohr, 4x4x4 k, 16 bands all occupied (f=2), 40^3."""
    return build_system(fcc_lattice(10.26), (40, 40, 40), (4, 4, 4), nbands=16, nocc=16, seed_stream=1,
                        name="hf2x fcc a10.26 4x4x4k 16b 40^3", **kw)


def fcc_bitc_40(**kw) -> System:
    """The fcc a=10.26, 4x4x4 k, 16-band, 40^3 cell with BITC, RPA Jastrow A=0.5 (F=1 from the cusp), independent left set."""
    A = np.full((2, 2), 0.5)
    return build_system(fcc_lattice(10.26), (40, 40, 40), (4, 4, 4), nbands=16, nocc=16, seed_stream=2,
                        method="BITC", A=A, name="bitc fcc a10.26 4x4x4k 16b 40^3 A0.5", **kw)


def fcc_adjoint_48(method="BITC", **kw) -> System:
    """fcc a=10.26, 3x3x3 k (27 k), 32 bands/k all occupied, 48^3; BITC with A = 0.5 by
    default (the cell of the gradient workload)."""
    A = np.full((2, 2), 0.5) if method != "HF" else None
    return build_system(fcc_lattice(10.26), (48, 48, 48), (3, 3, 3), nbands=32, nocc=32, seed_stream=4,
                        method=method, A=A, name=f"{method.lower()} fcc a10.26 3x3x3k 32b 48^3", **kw)


def uneven_bands(sys_: System, counts) -> System:
    """The same cell with only the first counts[k] occupied orbitals at k."""
    counts = np.asarray(counts, dtype=np.int32)
    po, oo = sys_.pw_offsets(), sys_.occ_offsets()
    phis, chis, fill = [], [], []
    for ik in range(sys_.nk):
        assert counts[ik] <= sys_.nocc[ik]
        blk = sys_.phi_occ[oo[ik]:oo[ik + 1]].reshape(sys_.nocc[ik], -1)[:counts[ik]]
        phis.append(blk.reshape(-1))
        if sys_.chi_occ is not None:
            chis.append(sys_.chi_occ[oo[ik]:oo[ik + 1]].reshape(sys_.nocc[ik], -1)[:counts[ik]].reshape(-1))
        fo = int(np.sum(sys_.nocc[:ik]))
        fill.append(sys_.filling[fo:fo + counts[ik]])
    return System(**{**sys_.__dict__, "nocc": counts, "filling": np.concatenate(fill),
                     "phi_occ": np.concatenate(phis),
                     "chi_occ": np.concatenate(chis) if chis else None})


def batched_uneven_systems(seed=3):
    """Three fcc a=10.26 cells, one per FFT grid 32^3 / 36^3 / 48^3, 4x4x2 k (32 k each,
    96 in all), a TC Jastrow (A = 0.5), and a seeded uniform band count in [8, 48] per k,
    all occupied."""
    rng = np.random.default_rng(1234 + 1000 * seed)
    out = []
    for g, grid in enumerate((32, 36, 48)):
        base = build_system(fcc_lattice(10.26), (grid,) * 3, (4, 4, 2), nbands=48, nocc=48, seed_stream=3 + 10 * g,
                            method="TC", A=np.full((2, 2), 0.5), name=f"tc fcc a10.26 4x4x2k {grid}^3 uneven")
        out.append(uneven_bands(base, rng.integers(8, 49, size=base.nk)))
    return out


def tiny_system(method="HF", n=(8, 8, 8), kmesh=(2, 1, 1), nbands=3, nocc=2, lattice=None,
                A=None, seed_stream=90, **kw) -> System:
    """Small systems for golden vectors and brute-force checks."""
    if lattice is None:
        lattice = fcc_lattice(6.0)
    if method in ("TC", "BITC") and A is None:
        A = np.full((2, 2), 0.5)
    return build_system(lattice, n, kmesh, nbands=nbands, nocc=nocc, seed_stream=seed_stream, method=method,
                        A=A, name=f"tiny {method} {n}", **kw)


def applied_bands(sys: System, nb=None, seed=None):
    """Bands the operator is applied to: by default the occupied orbitals
    themselves (nb = nocc), as in an SCF energy evaluation."""
    if nb is None and seed is None:
        return sys.nocc.copy(), sys.phi_occ.copy()
    nb = np.asarray(sys.nocc if nb is None else nb, dtype=np.int32)
    rng = np.random.default_rng(seed)
    blocks = []
    for ik in range(sys.nk):
        c = rng.standard_normal((nb[ik], sys.npw[ik])) + 1j * rng.standard_normal((nb[ik], sys.npw[ik]))
        c /= np.linalg.norm(c, axis=1, keepdims=True)
        blocks.append(c.reshape(-1))
    return nb, np.concatenate(blocks).astype(np.complex128)


def with_zero_weight_kpoints(sys: System, kfrac_extra, from_k) -> System:
    """`sys` plus zero-weight k-points (BAND k-points): extra k at fractional coordinates
    kfrac_extra[e], each carrying the plane-wave index list of the weighted k-point
    from_k[e] and, as the states its divergence terms project on, the occupied orbitals
    (right and left) of from_k[e] with their fillings.  The extra k hold no occupied
    orbitals, so they are no source of the q-sums."""
    kfrac_extra = np.asarray(kfrac_extra, dtype=np.float64).reshape(-1, 3)
    from_k = [int(k) for k in from_k]
    assert sys.kweight is None and len(from_k) == len(kfrac_extra)
    po, oo = sys.pw_offsets(), sys.occ_offsets()
    fo = np.concatenate([[0], np.cumsum(sys.nocc)]).astype(np.int64)
    gidx = [sys.gindex] + [sys.gindex[po[k]:po[k + 1]] for k in from_k]
    zphi = [sys.phi_occ[oo[k]:oo[k + 1]] for k in from_k]
    zchi = [sys.chi_occ[oo[k]:oo[k + 1]] for k in from_k] if sys.chi_occ is not None else None
    zfill = [sys.filling[fo[k]:fo[k + 1]] for k in from_k]
    nk0 = sys.nk
    return System(**{
        **sys.__dict__,
        "kfrac": np.concatenate([np.asarray(sys.kfrac, np.float64).reshape(-1, 3), kfrac_extra]),
        "npw": np.concatenate([sys.npw, [sys.npw[k] for k in from_k]]).astype(np.int32),
        "gindex": np.concatenate(gidx).astype(np.int32),
        "nocc": np.concatenate([sys.nocc, np.zeros(len(from_k), np.int32)]).astype(np.int32),
        "kweight": np.concatenate([np.ones(nk0), np.zeros(len(from_k))]),
        "zero_weight_states": {
            "nstate": np.concatenate([np.zeros(nk0, np.int32), [sys.nocc[k] for k in from_k]]).astype(np.int32),
            "fill": np.concatenate(zfill),
            "phi": np.concatenate(zphi),
            "chi": np.concatenate(zchi) if zchi is not None else None,
        },
        "symmetry": None,
        "name": sys.name + " + zero-weight k",
    })
