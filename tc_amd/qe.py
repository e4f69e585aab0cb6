"""Quantum ESPRESSO inputs on the reference's symmetry-replicated k-mesh.

TC++ starts an SCF run from a QE calculation: the irreducible k-points with their
weights and eigenvalues, the symmetry operations, and the plane-wave coefficients of
each irreducible k (io_qe_files_read_xml.cpp:150-241, io_qe_files_read_wfc.f90:69-140).
It then works on the full k-mesh without storing it: every k of the mesh is an
irreducible k with a rotation (and possibly time reversal), and each orbital at that k
is produced on the fly by re-indexing the irreducible coefficients
(plane_wave_basis.cpp:81-164, 250-270).  This module builds that full mesh once,
explicitly, as a `tc_amd.system.System`, so that the engine's operators see exactly the
orbitals the reference's loops see.

The expansion itself runs in the C++ host (include/tcx_host.h, tc_amd/host/scf.cpp):
  * TC++ rotation U = (QE rotation)^-1, integer, crystal coordinates (symmetry.cpp:17-33);
  * k-replicas k' = U^T k (crystal), kept when not equal to an earlier one modulo the
    reciprocal lattice, and -k' by time reversal unless QE's noinv is true
    (kpoints.cpp:34-110);
  * orbital at k': coefficient c(G) placed at G' = U^T G (G' -> -G' and c -> conj(c)
    under time reversal), times exp(2 pi i G'.t) with t the fractional translation
    (plane_wave_basis.cpp:81-164, 250-270).
This module hands it the inputs (`write_flat`) and assembles the System from its output.

The fixtures (`tests/golden/qe_example*.npz`) hold the same data as the QE files of the
reference's examples.
"""
from __future__ import annotations

import ctypes as C
import dataclasses
import json
import math
import os
import tempfile

import numpy as np

from . import system as S
from . import tcx

HT_IN_EV = 27.21138505  # src/include/header.hpp:47


@dataclasses.dataclass
class QEInput:
    alat: float
    cell: np.ndarray          # rows a_i (bohr)
    grid: tuple
    b_norm: np.ndarray        # rows b_i in 2 pi / alat
    rot_qe: np.ndarray        # (nsym, 3, 3) integer, crystal coordinates
    trans: np.ndarray         # (nsym, 3) fractional translations (crystal)
    time_reversal: bool       # noinv != true (symmetry.cpp:124)
    nelec: float
    kcart: np.ndarray         # (nirr, 3) in 2 pi / alat
    kweight: np.ndarray       # (nirr,) QE weights (sum 2 for nspin = 1)
    eig_qe: np.ndarray        # (nirr, nb) Hartree
    npw: np.ndarray           # (nirr,)
    mill: list                # per irreducible k: (npw, 3) Miller indices
    coef: list                # per irreducible k: (nb, npw) complex128
    setup: dict
    expected: dict

    @property
    def nirr(self) -> int:
        return len(self.kcart)

    @property
    def nbands(self) -> int:
        return self.eig_qe.shape[1]


def load_fixture(path: str) -> QEInput:
    z = np.load(path, allow_pickle=False)
    npw = z["npw"]
    nb = z["eig_qe"].shape[1]
    mill, coef = [], []
    m0 = c0 = 0
    for n in npw:
        mill.append(z["mill"][m0:m0 + n])
        coef.append(z["coef"][c0:c0 + nb * n].reshape(nb, n))
        m0 += n
        c0 += nb * n
    return QEInput(alat=float(z["alat"]), cell=z["cell"], grid=tuple(int(v) for v in z["grid"]),
                   b_norm=z["b_norm"], rot_qe=z["rot_qe"], trans=z["trans"],
                   time_reversal=str(z["noinv"]) != "true", nelec=float(z["nelec"]), kcart=z["kcart"],
                   kweight=z["kweight"], eig_qe=z["eig_qe"], npw=npw, mill=mill, coef=coef,
                   setup=json.loads(str(z["setup"])), expected=json.loads(str(z["expected"])))


@dataclasses.dataclass
class KReplica:
    ik: int        # irreducible k
    isym: int      # symmetry operation (TC++ index)
    trev: bool     # time reversal used
    kcrys: np.ndarray
    index: int     # position in the expanded mesh


class HostInputs:
    """The QE inputs loaded into the C++ host (tcxh_qe_load), which expands the k-mesh
    and runs the SCF loop; a context manager that frees the host's copy."""

    def __init__(self, qe: QEInput):
        self.L = tcx.load_library()
        fd, path = tempfile.mkstemp(suffix=".tcxqe")
        os.close(fd)
        h = C.c_void_p()
        try:
            write_flat(qe, path)
            st = self.L.tcxh_qe_load(path.encode(), C.byref(h))
        finally:
            os.unlink(path)
        tcx.host_check(st, "tcxh_qe_load")
        self.h = h
        self.nsym = len(qe.rot_qe)

    def close(self):
        if getattr(self, "h", None):
            self.L.tcxh_qe_free(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        self.close()

    def replicas(self) -> list:
        """Kpoints::set_kvectors_by_symmetry (kpoints.cpp:34-110), in index_all_kscf order."""
        n = C.c_int()
        tcx.host_check(self.L.tcxh_qe_num_replicas(self.h, C.byref(n)), "tcxh_qe_num_replicas")
        ik, isym, trev = (np.zeros(n.value, np.int32) for _ in range(3))
        kc = np.zeros((n.value, 3))
        tcx.host_check(self.L.tcxh_qe_replicas(self.h, tcx._ptr(ik, tcx._ip), tcx._ptr(isym, tcx._ip),
                                               tcx._ptr(trev, tcx._ip), tcx._ptr(kc)), "tcxh_qe_replicas")
        return [KReplica(int(ik[r]), int(isym[r]), bool(trev[r]), kc[r].copy(), r) for r in range(n.value)]

    def rotation(self, isym: int) -> np.ndarray:
        """U = (QE rotation)^-1, integer (symmetry.cpp:17-33)."""
        U = np.zeros(9, np.int32)
        tcx.host_check(self.L.tcxh_qe_rotation(self.h, int(isym), tcx._ptr(U, tcx._ip)), "tcxh_qe_rotation")
        return U.reshape(3, 3).astype(np.int64)

    def replica_basis(self, rep: KReplica):
        """(FFT-grid linear indices, phase factors) of the irreducible plane waves at a
        replica (plane_wave_basis.cpp:81-164)."""
        n = C.c_int()
        tcx.host_check(self.L.tcxh_qe_replica_npw(self.h, rep.index, C.byref(n)), "tcxh_qe_replica_npw")
        gidx = np.zeros(n.value, np.int32)
        ph = np.zeros(n.value, np.complex128)
        tcx.host_check(self.L.tcxh_qe_replica_basis(self.h, rep.index, tcx._ptr(gidx, tcx._ip), ph.ctypes.data),
                       "tcxh_qe_replica_basis")
        return gidx, ph


def tc_rotations(qe: QEInput) -> np.ndarray:
    """U = inverse of the QE rotation, integer (symmetry.cpp:17-33), per operation."""
    with HostInputs(qe) as h:
        return np.array([h.rotation(i) for i in range(h.nsym)])


def replicate_kpoints(qe: QEInput) -> list:
    """The expanded k-mesh (kpoints.cpp:34-110), in index_all_kscf order."""
    with HostInputs(qe) as h:
        return h.replicas()


def replica_basis(qe: QEInput, rep: KReplica):
    """(FFT-grid linear indices, phase factors) of the irreducible plane waves at a replica."""
    with HostInputs(qe) as h:
        return h.replica_basis(rep)


def jastrow_rpa(qe: QEInput, volume: float):
    """Default RPA Jastrow: A_long = A sqrt(V / (4 pi N_e)) (jastrow_initialize.cpp:28-50),
    F from the cusp condition (:62-75): F_parallel = sqrt(2A), F_antiparallel = sqrt(A)."""
    a = float(qe.setup.get("A_input", 1.0)) * math.sqrt(volume / (S.FOURPI * qe.nelec))
    A = np.full((2, 2), a)
    F = np.array([[math.sqrt(2 * a), math.sqrt(a)], [math.sqrt(a), math.sqrt(2 * a)]])
    return A, F


def build_system(qe: QEInput, phi_irr: list, filling_irr: np.ndarray, nocc_irr: np.ndarray,
                 chi_irr: list | None = None, method: str | None = None, symmetric: bool = True) -> tuple:
    """Full-mesh System whose occupied set at every replica is the irreducible set
    re-indexed; returns (System, replicas).  phi_irr[ik]: (nb, npw) coefficients,
    filling_irr (nirr, nb), nocc_irr (nirr,) occupied bands (f > 1e-8) per irreducible k.
    symmetric: carry the mesh symmetry, so that tcx_tc3a2a4b2b5 rotates its q2 sums from
    the irreducible q2 as the reference does; False treats every replica as independent."""
    with HostInputs(qe) as h:
        reps = h.replicas()
        bases = [h.replica_basis(r) for r in reps]
        U = np.array([h.rotation(i) for i in range(h.nsym)])
    recip = S.reciprocal_rows(qe.cell)
    volume = abs(float(np.linalg.det(qe.cell)))
    method = method or qe.setup["method"]
    npw, gidx_all, kfrac, nocc, fill, occ, left = [], [], [], [], [], [], []
    for r, (gidx, ph) in zip(reps, bases):
        npw.append(len(gidx))
        gidx_all.append(gidx)
        kfrac.append(r.kcrys)
        n = int(nocc_irr[r.ik])
        nocc.append(n)
        fill.append(filling_irr[r.ik][:n])
        for src, dst in ((phi_irr, occ), (chi_irr, left)):
            if src is None:
                continue
            c = np.asarray(src[r.ik])[:n]
            dst.append(((np.conj(c) if r.trev else c) * ph[None, :]).reshape(-1))
    A, F = jastrow_rpa(qe, volume) if method != "HF" else (np.zeros((2, 2)), np.ones((2, 2)))
    sys_ = S.System(lattice=np.asarray(qe.cell, dtype=np.float64), recip=recip, volume=volume, n=qe.grid,
                    kfrac=np.array(kfrac), npw=np.array(npw, np.int32), gindex=np.concatenate(gidx_all),
                    nocc=np.array(nocc, np.int32), filling=np.concatenate(fill), phi_occ=np.concatenate(occ),
                    chi_occ=np.concatenate(left) if chi_irr is not None else None, method=method, A=A, F=F,
                    div_corr=True, name=qe.setup.get("source", "qe"),
                    symmetry=mesh_symmetry(qe, reps, U) if symmetric else None)
    return sys_, reps


def mesh_symmetry(qe: QEInput, reps: list, U: np.ndarray) -> dict:
    """The symmetry relation of the replicated mesh for tcx_set_kpoint_symmetry: each
    replica built from the first replica (isym 0, the irreducible k itself) of its ik;
    U[isym] the TC++ rotations."""
    first = {}
    for i, r in enumerate(reps):
        first.setdefault(r.ik, i)
    return {"k_src": np.array([first[r.ik] for r in reps], np.int32),
            "rot": np.array([U[r.isym] for r in reps], np.int32),
            "trans": np.array([qe.trans[r.isym] for r in reps], np.float64),
            "trev": np.array([int(r.trev) for r in reps], np.int32)}


def write_flat(qe: QEInput, path: str) -> None:
    """The QE inputs and SCF settings as named little-endian records for the C++ host
    (include/tcx_host.h, read by tcxh_qe_load): after the 8-byte magic
    "TCXQE001", each record is a 32-byte name, an 8-byte numpy type tag ("<f8", "<i4",
    "<c16"), an int64 element count and the data."""
    st = qe.setup
    methods = {"HF": 0, "TC": 1, "BITC": 2}
    recs = [
        ("cell", np.asarray(qe.cell, "<f8")), ("b_norm", np.asarray(qe.b_norm, "<f8")),
        ("grid", np.asarray(qe.grid, "<i4")), ("rot_qe", np.asarray(qe.rot_qe, "<i4")),
        ("trans", np.asarray(qe.trans, "<f8")), ("time_reversal", np.asarray([int(qe.time_reversal)], "<i4")),
        ("nelec", np.asarray([qe.nelec], "<f8")), ("kcart", np.asarray(qe.kcart, "<f8")),
        ("kweight", np.asarray(qe.kweight, "<f8")), ("nb", np.asarray([qe.nbands], "<i4")),
        ("eig", np.asarray(qe.eig_qe, "<f8")), ("npw", np.asarray([len(m) for m in qe.mill], "<i4")),
        ("mill", np.concatenate([np.asarray(m, "<i4").reshape(-1) for m in qe.mill])),
        ("coef", np.concatenate([np.asarray(c, "<c16").reshape(-1) for c in qe.coef])),
        ("method", np.asarray([methods[st["method"]]], "<i4")),
        ("smearing", np.asarray([{"fixed": 0, "gaussian": 1}[st["smearing_mode"]]], "<i4")),
        ("smearing_width", np.asarray([float(st["smearing_width"])], "<f8")),
        ("mixing_beta", np.asarray([float(st.get("mixing_beta", 0.7))], "<f8")),
        ("mixes_density_matrix", np.asarray([int(bool(st.get("mixes_density_matrix", False)))], "<i4")),
        ("max_num_iterations", np.asarray([int(st.get("max_num_iterations", 30))], "<i4")),
        ("energy_tolerance", np.asarray([float(st.get("energy_tolerance", 1e-6))], "<f8")),
        ("charge_tolerance", np.asarray([float(st.get("charge_tolerance", 1e-5))], "<f8")),
        ("max_num_blocks_david", np.asarray([int(st["max_num_blocks_david"])], "<i4")),
        ("num_refresh_david", np.asarray([int(st["num_refresh_david"])], "<i4")),
        ("biortho_david", np.asarray([int(bool(st.get("biortho_david", False)))], "<i4")),
        ("A_input", np.asarray([float(st.get("A_input", 1.0))], "<f8")),
    ]
    with open(path, "wb") as f:
        f.write(b"TCXQE001")
        for name, arr in recs:
            a = np.ascontiguousarray(arr)
            f.write(name.encode().ljust(32, b"\0"))
            f.write(a.dtype.str.encode().ljust(8, b"\0"))
            f.write(np.int64(a.size).tobytes())
            f.write(a.tobytes())
