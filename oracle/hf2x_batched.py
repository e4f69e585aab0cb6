"""Batched FP64 restatement of orc_hf2x (oracle/oracle.c) with SciPy's FFT.

TEST INFRASTRUCTURE ONLY: the CPU baseline of bench.py times it, and tests/ pin it against
orc_hf2x; the product package never imports it.

The loop nest is the reference's calc_hamiltonian::hf2x (src/calc_hamiltonian_hf2x.cpp):
per applied band (k,j) and per occupied q, the pair densities conj(phi_qi) phi_kj of all
occupied i at q (:91-125), times 1/|k - q + G|^2 (0 where that is below 1e-8, :91-104),
back to real space (:127), summed with the fillings times phi_qi (:134), then FFT_forward
and the gather onto the sphere of k with -4 pi / (Nk V), halved for one independent spin
(:29-30, :143-147).  The divergence correction is left out (systems with div_corr False).

The occupied q run in parallel threads (as the C oracle's OpenMP loop over q), and the
pair transforms of one q run as one batched scipy.fft call (pocketfft, SIMD), so the
baseline's FFTs are library-class rather than the C oracle's recursive ones.  Transform conventions are those of orc_fft3d: sign +1 is exp(+2 pi i) unscaled
(ifftn with norm="forward"), sign -1 is exp(-2 pi i) / N (fftn with norm="forward").  The
grid index is i1 + n1 (i2 + n2 i3), i.e. a C-order array of shape (n3, n2, n1).
"""
from __future__ import annotations

import math
import os

import numpy as np
import scipy.fft as sf

AXES = (-3, -2, -1)


def _grid_gcart(sys_) -> np.ndarray:
    """Cartesian G of every grid point, shape (n3, n2, n1, 3)."""
    n1, n2, n3 = sys_.n
    m = [np.where(np.arange(n) <= n // 2, np.arange(n), np.arange(n) - n) for n in (n1, n2, n3)]
    m3, m2, m1 = np.meshgrid(m[2], m[1], m[0], indexing="ij")
    return (m1[..., None] * sys_.recip[0] + m2[..., None] * sys_.recip[1] + m3[..., None] * sys_.recip[2])


def _to_real(sys_, ik: int, coef: np.ndarray, workers: int) -> np.ndarray:
    """Coefficients (nb, npw_k) on the sphere of k -> real-space grids (nb, n3, n2, n1)."""
    n1, n2, n3 = sys_.n
    po = sys_.pw_offsets()
    gi = sys_.gindex[po[ik]:po[ik + 1]]
    grid = np.zeros((coef.shape[0], n1 * n2 * n3), np.complex128)
    grid[:, gi] = coef
    return sf.ifftn(grid.reshape(-1, n3, n2, n1), axes=AXES, norm="forward", workers=workers)


def hf2x(sys_, nb, phi, workers: int = 0, occ_real=None) -> np.ndarray:
    """H_x phi for the applied bands `phi` (nb[k] bands at k, concatenated over k), FP64,
    without the divergence correction.  The occupied q run in parallel on `workers` threads
    (0: every CPU this process may use), each with single-threaded batched FFTs; the q
    contributions are summed in q order.  `occ_real` may hold the occupied real-space grids
    from a previous call (list over q); they are built when None."""
    from concurrent.futures import ThreadPoolExecutor
    if sys_.kweight is not None and np.any(np.asarray(sys_.kweight) < 1e-8):
        raise ValueError("hf2x_batched: zero-weight k-points are not restated here")
    if workers <= 0:
        workers = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    nb = np.asarray(nb, dtype=np.int64)
    kc = sys_.kfrac @ sys_.recip
    G = _grid_gcart(sys_)
    factor = -4.0 * math.pi / (sys_.num_kpoints * sys_.volume)
    if sys_.nspin == 1:
        factor /= 2.0
    if occ_real is None:
        occ_real = [_to_real(sys_, iq, sys_.occ_block(iq), workers) if sys_.nocc[iq] > 0 else None
                    for iq in range(sys_.nk)]
    fo = np.concatenate([[0], np.cumsum(sys_.nocc)]).astype(np.int64)
    bo = sys_.band_offsets(nb)
    po = sys_.pw_offsets()
    out = np.zeros(int(bo[-1]), np.complex128)
    qs = [iq for iq in range(sys_.nk) if sys_.nocc[iq] > 0]

    def coulomb(ik, iq):
        x = kc[ik] - kc[iq] + G
        t = np.einsum("...d,...d->...", x, x)
        return np.where(t < 1e-8, 0.0, 1.0 / np.where(t < 1e-8, 1.0, t))

    with ThreadPoolExecutor(max_workers=workers) as pool:
        for ik in range(sys_.nk):
            if nb[ik] == 0:
                continue
            coef = np.asarray(phi[bo[ik]:bo[ik + 1]]).reshape(int(nb[ik]), int(sys_.npw[ik]))
            phik = _to_real(sys_, ik, coef, workers)
            gi = sys_.gindex[po[ik]:po[ik + 1]]
            V = dict(zip(qs, pool.map(lambda iq: coulomb(ik, iq), qs)))
            for jb in range(int(nb[ik])):
                pj = phik[jb]

                def part(iq):
                    phiq = occ_real[iq]
                    rho = sf.fftn(np.conj(phiq) * pj, axes=AXES, norm="forward", workers=1)
                    rho *= V[iq]
                    vr = sf.ifftn(rho, axes=AXES, norm="forward", workers=1)
                    vr *= phiq
                    return np.tensordot(sys_.filling[fo[iq]:fo[iq + 1]], vr, axes=1)

                hsub = np.zeros(pj.shape, np.complex128)
                for h in pool.map(part, qs):
                    hsub += h
                hG = sf.fftn(hsub, norm="forward", workers=workers).reshape(-1)
                out[bo[ik] + jb * sys_.npw[ik]:bo[ik] + (jb + 1) * sys_.npw[ik]] = factor * hG[gi]
    return out
