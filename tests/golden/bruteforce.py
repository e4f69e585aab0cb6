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


_IDX = {}


def _shift_tables(shape):
    """Index tables of the cyclic grid: plus[G, G1] = G + G1 and minus[G, G1] = G - G1."""
    if shape not in _IDX:
        n3, n2, n1 = shape
        i3, i2, i1 = np.unravel_index(np.arange(n3 * n2 * n1), shape)
        def lin(a3, a2, a1):
            return (a3 % n3) * (n2 * n1) + (a2 % n2) * n1 + (a1 % n1)
        plus = lin(i3[:, None] + i3[None, :], i2[:, None] + i2[None, :], i1[:, None] + i1[None, :])
        minus = lin(i3[:, None] - i3[None, :], i2[:, None] - i2[None, :], i1[:, None] - i1[None, :])
        neg = lin(-i3, -i2, -i1)
        _IDX[shape] = (plus, minus, neg)
    return _IDX[shape]


def corr(A, B):
    """R[G] = sum_{G1} conj(A[G1]) B[G1+G] on the cyclic grid (= FFT_forward(conj(a) b))."""
    plus, _, _ = _shift_tables(B.shape)
    nz = np.flatnonzero(A)
    Bf = B.reshape(-1)
    return (Bf[plus[:, nz]] @ np.conj(A.reshape(-1)[nz])).reshape(B.shape)


def conv(A, B):
    """R[G] = sum_{G1} A[G1] B[G-G1] on the cyclic grid (= FFT_forward(a b))."""
    _, minus, _ = _shift_tables(B.shape)
    nz = np.flatnonzero(A)
    Bf = B.reshape(-1)
    return (Bf[minus[:, nz]] @ A.reshape(-1)[nz]).reshape(B.shape)


def conjneg(A):
    """Coefficients of conj(a(r)) from those of a(r): conj(A[-G])."""
    _, _, neg = _shift_tables(A.shape)
    return np.conj(A.reshape(-1)[neg]).reshape(A.shape)


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


# ---------------------------------------------------------------- assembly terms
def kinetic(sys, nb, phi):
    """src/calc_hamiltonian_kinetic.cpp:38-46: 0.5 |k+G|^2 phi."""
    out = np.zeros_like(phi)
    bo = sys.band_offsets(nb)
    po = sys.pw_offsets()
    for ik in range(sys.nk):
        npwk = sys.npw[ik]
        kG = _kcart_grid(sys, sys.kfrac[ik]).reshape(-1, 3)[sys.gindex[po[ik]:po[ik + 1]]]
        t = 0.5 * np.sum(kG * kG, axis=1)
        for jb in range(nb[ik]):
            sl = slice(bo[ik] + jb * npwk, bo[ik] + (jb + 1) * npwk)
            out[sl] = t * phi[sl]
    return out


def hf2h(sys, nb, phi):
    """src/calc_hamiltonian_hf2h.cpp:23-41: V_H(G) = 4 pi n(G)/G^2 (G = 0 dropped), 1/(Nk V)."""
    nG = 2.0 * density_G(sys)                 # density[0] + density[1]
    G = _kcart_grid(sys, np.zeros(3))
    G2 = np.sum(G * G, axis=-1)
    VG = np.zeros_like(nG)
    m = np.ones(G2.shape, bool)
    m.reshape(-1)[0] = False
    VG[m] = FOURPI * nG[m] / G2[m]
    return _apply_local(sys, nb, phi, VG, 1.0 / (sys.num_kpoints * sys.volume))


def tc2h(sys, nb, phi):
    """src/calc_hamiltonian_tc2h.cpp:50-235 (one independent spin)."""
    fo = np.concatenate([[0], np.cumsum(sys.nocc)])
    bitc = sys.method == "BITC"
    nG = density_G(sys)                        # density[0] = density[1]
    G = _kcart_grid(sys, np.zeros(3))
    G2 = np.sum(G * G, axis=-1)
    pgp = [np.zeros_like(nG) for _ in range(3)]
    for iq in range(sys.nk):
        qG = _kcart_grid(sys, sys.kfrac[iq])
        occ = sys.occ_block(iq)
        occl = sys.occ_block(iq, left=bitc)
        for ib in range(sys.nocc[iq]):
            f = sys.filling[fo[iq] + ib]
            Cq = _grid(sys, iq, occ[ib])
            Cx = _grid(sys, iq, occl[ib])
            for d in range(3):
                pgp[d] += f * corr(Cx, qG[..., d] * Cq)
    V2 = np.zeros_like(nG)
    D = [np.zeros_like(nG) for _ in range(3)]
    for js in range(2):
        V2 += tc_2body(sys, G2, 0, js) * nG
        u = uk(sys, G2, 0, js)
        for d in range(3):
            V2 += u * G[..., d] * pgp[d] / 2.0
            D[d] += u * G[..., d] * nG
    out = np.zeros_like(phi)
    bo = sys.band_offsets(nb)
    factor = 1.0 / (sys.num_kpoints * sys.volume)
    for ik in range(sys.nk):
        npwk = sys.npw[ik]
        kG = _kcart_grid(sys, sys.kfrac[ik])
        for jb in range(nb[ik]):
            sl = slice(bo[ik] + jb * npwk, bo[ik] + (jb + 1) * npwk)
            Cj = _grid(sys, ik, phi[sl])
            W = conv(Cj, V2)
            for d in range(3):
                W -= conv(kG[..., d] * Cj, D[d])
            out[sl] = factor * _gather(sys, ik, W)
    return out


# ---------------------------------------------------------------- three-body exchange terms
def _du(sys, kq_frac, s_first):
    """du[s][d] = (kq+G)_d u(kq+G) with u = uk(.., s, 0) (s_first) or uk(.., 0, s)."""
    kq = _kcart_grid(sys, kq_frac)
    G2 = np.sum(kq * kq, axis=-1)
    out = []
    for s in range(2):
        u = uk(sys, G2, s, 0) if s_first else uk(sys, G2, 0, s)
        out.append([kq[..., d] * u for d in range(3)])
    return out


def tc3a3b3b4b6(sys, nb, phi, alpha=None, svaux=None):
    """src/calc_hamiltonian_tc3a3b3b4b6.cpp:33-47, 126-151, 179-371, 393-535, 629-642, 646-770."""
    NkV = sys.num_kpoints * sys.volume
    base = 1.0 / (NkV * NkV)
    fa3, fb3, fb4, fb6 = -base / 2, base / 2, -base / 4, base / 2
    fo = np.concatenate([[0], np.cumsum(sys.nocc)])
    bitc = sys.method == "BITC"
    nG = density_G(sys)
    G = _kcart_grid(sys, np.zeros(3))
    G2 = np.sum(G * G, axis=-1)
    D = [sum(uk(sys, G2, 0, s3) * G[..., d] * nG for s3 in range(2)) for d in range(3)]
    need_div = sys.div_corr and not sys.jastrow_A_zero
    out = np.zeros_like(phi)
    bo = sys.band_offsets(nb)
    for ik in range(sys.nk):
        npwk = sys.npw[ik]
        Cjs = [_grid(sys, ik, phi[bo[ik] + jb * npwk: bo[ik] + (jb + 1) * npwk]) for jb in range(nb[ik])]
        qsum = [[np.zeros_like(nG) for _ in range(3)] for _ in range(nb[ik])]
        H = [np.zeros_like(nG) for _ in range(nb[ik])]
        for iq in range(sys.nk):
            du = _du(sys, sys.kfrac[ik] - sys.kfrac[iq], True)
            occ = sys.occ_block(iq)
            occl = sys.occ_block(iq, left=bitc)
            for jb, Cj in enumerate(Cjs):
                DCj = [conv(Cj, D[d]) for d in range(3)]        # dnu_d phi_j
                for ib in range(sys.nocc[iq]):
                    f = sys.filling[fo[iq] + ib]
                    Cq = _grid(sys, iq, occ[ib])
                    Cx = _grid(sys, iq, occl[ib])
                    rho = corr(Cx, Cj)
                    X = sum(du[0][d] * corr(Cx, DCj[d]) for d in range(3))      # [3b3]
                    H[jb] += f * fb3 * conv(Cq, X)
                    Y = np.zeros_like(nG)
                    for s2 in range(2):
                        for d in range(3):
                            P = f * rho * du[s2][d]
                            if s2 == 0:
                                qsum[jb][d] += conv(Cq, P)                        # [3a3,b4]
                            Y += du[s2][d] * conv(P, nG)                          # [3b6]
                    H[jb] += fb6 * conv(Cq, Y)
        for jb, Cj in enumerate(Cjs):
            res = _gather(sys, ik, H[jb])
            if need_div:                                                         # [3b6] div
                dcp = 1.0 / math.sqrt(math.pi * alpha) * NkV - svaux[ik]
                t1 = 0.0
                tG = np.zeros_like(nG)
                for js in range(2):
                    Aj = sys.A[0, js]
                    t1 += FOURPI * (dcp * Aj + float(uk(sys, np.zeros(1), 0, js)[0])) * Aj * nG.reshape(-1)[0]
                    t = FOURPI * nG * Aj * uk(sys, G2, 0, js)
                    t.reshape(-1)[0] = 0.0
                    tG += t
                occk = sys.occ_block(ik)
                occlk = sys.occ_block(ik, left=bitc)
                Hs = np.zeros_like(nG)
                for b in range(sys.nocc[ik]):
                    f = sys.filling[fo[ik] + b]
                    if f < 1e-8:
                        continue
                    Cb = _grid(sys, ik, occk[b])
                    pp = corr(Cj, _grid(sys, ik, occlk[b]))              # FFT(chi_b conj(phi_j))
                    Ib = np.conj(pp.reshape(-1)[0])
                    dot = np.vdot(pp, tG)
                    res = res + f * fb6 * (t1 * Ib + dot) * occk[b]
                    Hs += f * Ib * fb6 * conv(Cb, tG)
                res = res + _gather(sys, ik, Hs)
            res = res + fa3 * _gather(sys, ik, sum(conv(qsum[jb][d], D[d]) for d in range(3)))  # [3a3]
            out[bo[ik] + jb * npwk: bo[ik] + (jb + 1) * npwk] += res
        for iq in range(sys.nk):                                                 # [3b4]
            du = _du(sys, sys.kfrac[ik] - sys.kfrac[iq], True)
            occ = sys.occ_block(iq)
            occl = sys.occ_block(iq, left=bitc)
            for jb in range(nb[ik]):
                Hs = np.zeros_like(nG)
                for ib in range(sys.nocc[iq]):
                    f = sys.filling[fo[iq] + ib]
                    Cq = _grid(sys, iq, occ[ib])
                    Cx = _grid(sys, iq, occl[ib])
                    Y = sum(du[0][d] * corr(Cx, qsum[jb][d]) for d in range(3))
                    Hs += f * conv(Cq, Y)
                out[bo[ik] + jb * npwk: bo[ik] + (jb + 1) * npwk] += fb4 * _gather(sys, ik, Hs)
    return out


def tc3a2a4b2b5(sys, nb, phi, alpha=None, svaux=None):
    """src/calc_hamiltonian_tc3a2a4b2b5.cpp:34-48, 120-131, 210-443, 486-573, 644-913, 922-1034."""
    NkV = sys.num_kpoints * sys.volume
    base = 1.0 / (NkV * NkV)
    fa2, fa4, fb2, fb5 = -0.5 * base / 4, base / 4, base / 4, -base / 4
    fo = np.concatenate([[0], np.cumsum(sys.nocc)])
    bitc = sys.method == "BITC"
    G = _kcart_grid(sys, np.zeros(3))
    G2 = np.sum(G * G, axis=-1)
    uG = [uk(sys, G2, 0, s) for s in range(2)]
    need_div = sys.div_corr and not sys.jastrow_A_zero
    shape = G2.shape
    V3a2 = np.zeros(shape, np.complex128)
    B2 = [np.zeros(shape, np.complex128) for _ in range(3)]
    out = np.zeros_like(phi)
    bo = sys.band_offsets(nb)
    for iq2 in range(sys.nk):
        occ2 = sys.occ_block(iq2)
        occl2 = sys.occ_block(iq2, left=bitc)
        for ib2 in range(sys.nocc[iq2]):
            f2 = sys.filling[fo[iq2] + ib2]
            Cq2 = _grid(sys, iq2, occ2[ib2])
            Cx2 = _grid(sys, iq2, occl2[ib2])
            Q2 = np.zeros(shape, np.complex128)
            Q4 = [np.zeros(shape, np.complex128) for _ in range(3)]
            Q5 = [np.zeros(shape, np.complex128) for _ in range(3)]
            for iq1 in range(sys.nk):
                du = _du(sys, sys.kfrac[iq2] - sys.kfrac[iq1], False)
                occ1 = sys.occ_block(iq1)
                occl1 = sys.occ_block(iq1, left=bitc)
                for ib1 in range(sys.nocc[iq1]):
                    f1 = sys.filling[fo[iq1] + ib1]
                    Cq1 = _grid(sys, iq1, occ1[ib1])
                    Cx1 = _grid(sys, iq1, occl1[ib1])
                    rho = corr(Cx1, Cq2)
                    rho2 = corr(Cq1, Cx2) if bitc else rho
                    for s in range(2):
                        for d in range(3):
                            O = du[s][d] * rho                    # coefficients of IFFT(du rho)
                            O2 = du[s][d] * rho2
                            Q2 -= f1 * corr(O2, O)                # FFT(O conj(O2)) ([3a2])
                            if s == 0:
                                Q4[d] += f1 * conv(Cq1, O)        # [3a4]
                                if bitc:
                                    Q5[d] -= f1 * conjneg(conv(Cx1, O2))   # [3b5] FFT(conj(O2 chi_q1))
            if not bitc:
                Q5 = [-conjneg(Q4[d]) for d in range(3)]
            T = Q2
            if need_div:                                             # [3a2] div
                dc2 = 1.0 / math.sqrt(math.pi * alpha) * NkV - svaux[iq2]
                for ib1 in range(sys.nocc[iq2]):
                    f1 = sys.filling[fo[iq2] + ib1]
                    Cq1 = _grid(sys, iq2, occ2[ib1])
                    Cx1 = _grid(sys, iq2, occl2[ib1])
                    pa = corr(Cx1, Cq2)
                    pb = corr(Cx2, Cq1)
                    pa0, pb0 = pa.reshape(-1)[0], pb.reshape(-1)[0]
                    for s in range(2):
                        As = sys.A[0, s]
                        u0 = float(uk(sys, np.zeros(1), 0, s)[0])
                        corrG = FOURPI * As * f1 * uG[s] * (pa * pb0 + pa0 * pb)
                        corrG.reshape(-1)[0] = FOURPI * As * (As * dc2 + u0) * f1 * pa0 * pb0
                        T = T - corrG
            V3a2 += f2 * T
            for d in range(3):
                B2[d] += f2 * conv(Cq2, Q5[d])
            for ik in range(sys.nk):
                npwk = sys.npw[ik]
                du0 = _du(sys, sys.kfrac[ik] - sys.kfrac[iq2], True)[0]
                for jb in range(nb[ik]):
                    sl = slice(bo[ik] + jb * npwk, bo[ik] + (jb + 1) * npwk)
                    Cj = _grid(sys, ik, phi[sl])
                    rho = corr(Cx2, Cj) * (f2 * fa4)
                    W = sum(conv(Q4[d], du0[d] * rho) for d in range(3))          # [3a4]
                    S = sum(du0[d] * conv(Cj, Q5[d]) for d in range(3))           # [3b5]
                    W = W + fb5 * f2 * conv(Cq2, S)
                    out[sl] += _gather(sys, ik, W)
    X = sum(G[..., d] * (uG[0] + uG[1]) * B2[d] for d in range(3))                # [3b2]
    out += _apply_local(sys, nb, phi, X, fb2)
    out += _apply_local(sys, nb, phi, V3a2, fa2)                                  # [3a2]
    return out
