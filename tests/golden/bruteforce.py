"""Brute-force G-space restatement of the operator terms (no FFT anywhere).

Every product of two real-space grid functions becomes a cyclic convolution
on the FFT grid and every FFT_forward(conj(a) b) a cyclic correlation, done
with explicit index shifts (np.roll).  With the reference's conventions
(FFT_forward carries 1/N, FFT_backward is unscaled; src/plane_wave_basis.cpp:
274-289) these sums are exact, so they reproduce the FFT-based algorithm
term by term without sharing any of its code.  Used to write the golden
fixtures and to pin the oracle on tiny grids.

Formulas restated (identity symmetry, nspin = 1, SCF):
  hf2x  src/calc_hamiltonian_hf2x.cpp:29-30, 91-147, 150-201
  tc2x  src/calc_hamiltonian_tc2x.cpp:32-33, 253-401, 409-513
  tc3a1 src/calc_hamiltonian_tc3a1.cpp:26-116
  tc3b1 src/calc_hamiltonian_tc3b1.cpp:26-134
"""
from __future__ import annotations

import math

import numpy as np

FOURPI = 4.0 * math.pi


def _miller(n):
    n1, n2, n3 = n
    ipw = np.arange(n1 * n2 * n3)
    i1, i2, i3 = ipw % n1, (ipw // n1) % n2, ipw // (n1 * n2)
    return np.stack([np.where(i1 <= n1 // 2, i1, i1 - n1), np.where(i2 <= n2 // 2, i2, i2 - n2),
                     np.where(i3 <= n3 // 2, i3, i3 - n3)], axis=1)


def _grid(sys, ik, c):
    """Scatter coefficients onto the grid, returned shaped (n3, n2, n1)."""
    g = np.zeros(sys.ngrid, dtype=np.complex128)
    po = sys.pw_offsets()
    g[sys.gindex[po[ik]:po[ik + 1]]] = c
    n1, n2, n3 = sys.n
    return g.reshape(n3, n2, n1)


def _shift_of(sys, idx):
    """(s3, s2, s1) grid offsets of a linear index."""
    n1, n2, _ = sys.n
    return (idx // (n1 * n2), (idx // n1) % n2, idx % n1)


def corr(A, B):
    """R[G] = sum_{G1} conj(A[G1]) B[G1+G] on the cyclic grid (= FFT_forward(conj(a) b))."""
    R = np.zeros_like(B)
    nz = np.argwhere(A != 0)
    for (a3, a2, a1) in nz:
        R += np.conj(A[a3, a2, a1]) * np.roll(B, shift=(-a3, -a2, -a1), axis=(0, 1, 2))
    return R


def conv(A, B):
    """R[G] = sum_{G1} A[G1] B[G-G1] on the cyclic grid (= FFT_forward(a b))."""
    R = np.zeros_like(B)
    nz = np.argwhere(A != 0)
    for (a3, a2, a1) in nz:
        R += A[a3, a2, a1] * np.roll(B, shift=(a3, a2, a1), axis=(0, 1, 2))
    return R


def _kcart_grid(sys, kvec_frac):
    """Cartesian k+G of every grid index, shaped (n3, n2, n1, 3)."""
    m = _miller(sys.n)
    c = (np.asarray(kvec_frac)[None, :] + m) @ sys.recip
    n1, n2, n3 = sys.n
    return c.reshape(n3, n2, n1, 3)


def uk(sys, G2, s1, s2):
    if sys.jastrow_A_zero:
        return np.zeros_like(G2)
    A, F = sys.A[s1, s2], sys.F[s1, s2]
    with np.errstate(divide="ignore", invalid="ignore"):
        v = FOURPI * A / (G2 * (G2 * F * F + 1.0))
    return np.where(G2 < 1e-12, -FOURPI * A * F * F, v)


def tc_2body(sys, G2, s1, s2):
    with np.errstate(divide="ignore", invalid="ignore"):
        if sys.jastrow_A_zero:
            return np.where(G2 > 1e-12, FOURPI / G2, 0.0)
        A, F = sys.A[s1, s2], sys.F[s1, s2]
        F2 = F * F
        paf = math.pi * A * A / F2
        G1 = np.sqrt(G2)
        v = (FOURPI / G2 - FOURPI * A / (1.0 + F2 * G2)
             + (paf / G1) * (math.pi * F2 * G2 + (4 + 2 * F2 * G2) * np.arctan(F * G1 / 2)
                             - (4 + 4 * F2 * G2) * np.arctan(F * G1)))
    return np.where(G2 < 1e-12, -2.0 * paf * F, v)


def _div_corr(sys, alpha, svaux, ik):
    dc = 1.0 / math.sqrt(math.pi * alpha) - svaux[ik] / (sys.num_kpoints * sys.volume)
    dc *= -1.0
    if sys.nspin == 1:
        dc /= 2.0
    return dc


def _gather(sys, ik, W):
    po = sys.pw_offsets()
    return W.reshape(-1)[sys.gindex[po[ik]:po[ik + 1]]]


def hf2x(sys, nb, phi, alpha=None, svaux=None):
    out = np.zeros_like(phi)
    factor = -FOURPI / (sys.num_kpoints * sys.volume) / (2.0 if sys.nspin == 1 else 1.0)
    bo = sys.band_offsets(nb)
    fo = np.concatenate([[0], np.cumsum(sys.nocc)])
    for ik in range(sys.nk):
        npwk = sys.npw[ik]
        for jb in range(nb[ik]):
            cj = phi[bo[ik] + jb * npwk: bo[ik] + (jb + 1) * npwk]
            Cj = _grid(sys, ik, cj)
            W = np.zeros_like(Cj)
            for iq in range(sys.nk):
                kq = _kcart_grid(sys, sys.kfrac[ik] - sys.kfrac[iq])
                t = np.sum(kq * kq, axis=-1)
                with np.errstate(divide="ignore"):
                    V = np.where(t < 1e-8, 0.0, 1.0 / t)
                occ = sys.occ_block(iq)
                for ib in range(sys.nocc[iq]):
                    Cq = _grid(sys, iq, occ[ib])
                    W += sys.filling[fo[iq] + ib] * conv(Cq, V * corr(Cq, Cj))
            res = factor * _gather(sys, ik, W)
            if sys.div_corr:
                dc = _div_corr(sys, alpha, svaux, ik)
                occk = sys.occ_block(ik)
                for b in range(sys.nocc[ik]):
                    f = sys.filling[fo[ik] + b]
                    if f < 1e-8:
                        continue
                    res = res + np.vdot(occk[b], cj) * f * dc * occk[b]
            out[bo[ik] + jb * npwk: bo[ik] + (jb + 1) * npwk] = res
    return out


def tc2x(sys, nb, phi, alpha=None, svaux=None):
    out = np.zeros_like(phi)
    factor = -1.0 / (sys.num_kpoints * sys.volume) / (2.0 if sys.nspin == 1 else 1.0)
    bo = sys.band_offsets(nb)
    fo = np.concatenate([[0], np.cumsum(sys.nocc)])
    bitc = sys.method == "BITC"
    for ik in range(sys.nk):
        npwk = sys.npw[ik]
        kG = _kcart_grid(sys, sys.kfrac[ik])
        for jb in range(nb[ik]):
            cj = phi[bo[ik] + jb * npwk: bo[ik] + (jb + 1) * npwk]
            Cj = _grid(sys, ik, cj)
            gCj = [kG[..., d] * Cj for d in range(3)]
            W = np.zeros_like(Cj)
            for iq in range(sys.nk):
                kq = _kcart_grid(sys, sys.kfrac[ik] - sys.kfrac[iq])
                G2 = np.sum(kq * kq, axis=-1)
                V2 = tc_2body(sys, G2, 0, 0)
                u = uk(sys, G2, 0, 0)
                du = [kq[..., d] * u for d in range(3)]
                qG = _kcart_grid(sys, sys.kfrac[iq])
                occ = sys.occ_block(iq)
                occl = sys.occ_block(iq, left=bitc)
                for ib in range(sys.nocc[iq]):
                    f = sys.filling[fo[iq] + ib]
                    Cq = _grid(sys, iq, occ[ib])
                    Cx = _grid(sys, iq, occl[ib])
                    rho = corr(Cx, Cj)
                    for d in range(3):
                        W -= f * conv(qG[..., d] * Cq, rho * du[d])       # [2b_x1]
                    T = rho * V2                                           # [2a_x]
                    for d in range(3):
                        T = T + corr(Cx, gCj[d]) * du[d]                   # [2b_x2]
                    W += f * conv(Cq, T)
            res = factor * _gather(sys, ik, W)
            if sys.div_corr:
                dc = _div_corr(sys, alpha, svaux, ik)
                occk = sys.occ_block(ik)
                occlk = sys.occ_block(ik, left=bitc)
                for b in range(sys.nocc[ik]):
                    f = sys.filling[fo[ik] + b]
                    if f < 1e-8:
                        continue
                    res = res + np.vdot(occlk[b], cj) * dc * f * occk[b]
            out[bo[ik] + jb * npwk: bo[ik] + (jb + 1) * npwk] = res
    return out


def density_G(sys):
    """n_s(G) = FFT_forward(density_s(r)) by correlation; density /2 for nspin=1."""
    fo = np.concatenate([[0], np.cumsum(sys.nocc)])
    bitc = sys.method == "BITC"
    n1, n2, n3 = sys.n
    nG = np.zeros((n3, n2, n1), dtype=np.complex128)
    for ik in range(sys.nk):
        occ = sys.occ_block(ik)
        occl = sys.occ_block(ik, left=bitc)
        for ib in range(sys.nocc[ik]):
            nG += sys.filling[fo[ik] + ib] * corr(_grid(sys, ik, occl[ib]), _grid(sys, ik, occ[ib]))
    return nG / 2.0


def _apply_local(sys, nb, phi, VG, factor):
    out = np.zeros_like(phi)
    bo = sys.band_offsets(nb)
    for ik in range(sys.nk):
        npwk = sys.npw[ik]
        for jb in range(nb[ik]):
            cj = phi[bo[ik] + jb * npwk: bo[ik] + (jb + 1) * npwk]
            out[bo[ik] + jb * npwk: bo[ik] + (jb + 1) * npwk] = factor * _gather(sys, ik, conv(_grid(sys, ik, cj), VG))
    return out


def tc3a1(sys, nb, phi):
    NkV = sys.num_kpoints * sys.volume
    nG = density_G(sys)
    G = _kcart_grid(sys, np.zeros(3))
    G2 = np.sum(G * G, axis=-1)
    dnu = [sum(uk(sys, G2, 0, js) * G[..., d] * nG for js in range(2)) for d in range(3)]
    VG = sum(conv(dnu[d], dnu[d]) for d in range(3))
    return _apply_local(sys, nb, phi, VG, 0.5 / (NkV * NkV))


def tc3b1(sys, nb, phi):
    NkV = sys.num_kpoints * sys.volume
    nG = density_G(sys)
    G = _kcart_grid(sys, np.zeros(3))
    G2 = np.sum(G * G, axis=-1)
    X = [sum(uk(sys, G2, 0, s3) * G[..., d] * nG for s3 in range(2)) for d in range(3)]
    dnu = [conv(nG, X[d]) for d in range(3)]
    VG = sum(uk(sys, G2, 0, s2) * G[..., d] * dnu[d] for s2 in range(2) for d in range(3))
    return _apply_local(sys, nb, phi, VG, -1.0 / (NkV * NkV))
