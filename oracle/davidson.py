"""Restatement of Diagonalization::block_davidson on dense per-k Hamiltonians -- TEST
INFRASTRUCTURE ONLY (the checker of tcx_davidson).

Follows src/diagonalization_davidson.cpp:7-689 and src/diagonalization_utils.cpp:
  Gram_Schmidt :7-29 (classical, drop below 1e-8), sort_eigen :79-118 (ascending real
  part, the same swap order), make_a_new_trial_vector :182-210 (Payne preconditioner),
  subspace H = V^H H_n V summed over n (:462-477), X_L = (X^-1)^H for BITC (:498),
  rotation V X (:507-518), re-orthonormalisation except in the BITC final loop (:530-537),
  e_n = diag(X_L^H H_n X) (:571-585).
The bi-orthogonal variant (biortho_david_ = true, block_davidson_biortho) follows the same
file's left-vector branches: V_left from phik_left (:86-103), left candidates from the
reference's H^dagger with conj(eps) (:234-251, :322-326), Gram_Schmidt_biortho
(diagonalization_utils.cpp:31-77, drops below 1e-10), subspace <V_left|H V> (:465-466),
sort_eigen_overlap (:121-167), rotations V X and V_left X_L rescaled to |V X| = 1 (:596-615).
The operator H_n (n = 1, 2, 3) of one k is given as dense matrices, so this pins the
consumer's algebra independently of the GPU H*psi."""
import numpy as np


def _gs(vs, target, t):
    for b in range(target):
        t = t - np.vdot(vs[b], t) * vs[b]
    n = np.linalg.norm(t)
    if n < 1e-8:
        return None
    return t / n


def _precond(kG2, phi, hphi, eps):
    kin = np.sum(kG2 * np.abs(phi) ** 2) / np.vdot(phi, phi).real
    x = kG2 / kin
    c = (27 + x * (18 + x * (12 + x * 8))) / (27 + x * (18 + x * (12 + x * (8 + x * 16))))
    t = -c * (hphi - eps * phi)
    n = np.linalg.norm(t)
    return np.zeros_like(t) if n < 1e-8 else t / n


def _sort(w, X):
    w, X = w.copy(), X.copy()
    n = len(w)
    for i in range(n):
        for j in range(i + 1, n):
            if w[i].real > w[j].real:
                w[[i, j]] = w[[j, i]]
                X[:, [i, j]] = X[:, [j, i]]
    return w, X


def block_davidson(Hn, kG2, phi, eig_ref, max_blocks=2, num_refresh=1, bitc=False):
    """One call at one k.  Hn: list of dense (npw x npw) H_1, H_2, H_3; phi: (nb, npw).
    Returns (phi_out, phi_left_out or None, eigenvalues[nb], energies[nb, 3])."""
    nb = phi.shape[0]
    V = []
    for b in range(nb):
        v = phi[b].copy()
        if bitc:
            v = _gs(V, b, v)
            assert v is not None, "linear independence breaks"
        V.append(v)
    for it in range(num_refresh):
        final = it == num_refresh - 1
        HV = [[], [], []]
        cnt = nb
        for blk in range(max_blocks - 1):
            first = len(V) - cnt
            for v in V[first:]:
                for n in range(3):
                    HV[n].append(Hn[n] @ v)
            nold = len(V)
            for b in range(cnt):
                h = HV[0][first + b] + HV[1][first + b] + HV[2][first + b]
                t = _gs(V, len(V), _precond(kG2, V[first + b], h, eig_ref[b]))
                if t is not None:
                    V.append(t)
            cnt = len(V) - nold
        first = len(V) - cnt
        for v in V[first:]:
            for n in range(3):
                HV[n].append(Hn[n] @ v)
        Vm = np.array(V)                     # (dim, npw)
        Hs = [Vm.conj() @ np.array(HV[n]).T for n in range(3)]
        w, X = _sort(*np.linalg.eig(Hs[0] + Hs[1] + Hs[2]))
        XL = np.linalg.inv(X).conj().T if bitc else None
        Vt = (X[:, :nb].T @ Vm)
        if not (bitc and final):
            out = []
            for b in range(nb):
                v = _gs(out, b, Vt[b])
                assert v is not None, "linear independence breaks"
                out.append(v)
            Vt = np.array(out)
        if final:
            L = XL if bitc else X
            en = np.array([[np.vdot(L[:, b], Hs[n] @ X[:, b]) for n in range(3)] for b in range(nb)])
            left = (XL[:, :nb].T @ Vm) if bitc else None
            return Vt, left, w[:nb], en
        V = [Vt[b] for b in range(nb)]


def _gs_biortho(vr, vl, t_r, t_l):
    for j in range(len(vr)):
        t_r = t_r - np.vdot(vl[j], t_r) * vr[j]
    n = np.linalg.norm(t_r)
    if n < 1e-10:
        return None
    t_r = t_r / n
    for j in range(len(vr)):
        t_l = t_l - np.vdot(vr[j], t_l) * vl[j]
    c = np.vdot(t_l, t_r)
    if abs(c) < 1e-10:
        return None
    return t_r, t_l / np.conj(c)


def _sort_overlap(w, X):
    n = len(w)
    chosen = np.zeros(n, bool)
    perm = []
    for i in range(n):
        best, js = 0.0, -1
        for j in range(n):
            if chosen[j]:
                continue
            a = abs(X[i, j]) ** 2
            if best < a:
                best, js = a, j
        chosen[js] = True
        perm.append(js)
    return w[perm], X[:, perm]


def block_davidson_biortho(Hn, Hdn, kG2, phi, phi_left, eig_ref, max_blocks=2, num_refresh=1):
    """One bi-orthogonal call at one k.  Hn / Hdn: dense H_1..3 and the reference's H^dagger
    H_1..3 (u -> -u, orbitals swapped, density kept); phi, phi_left: (nb, npw).
    Returns (phi_out, phi_left_out, eigenvalues[nb], energies[nb, 3])."""
    nb = phi.shape[0]
    V = [phi[b].copy() for b in range(nb)]
    VL = [phi_left[b].copy() for b in range(nb)]
    for it in range(num_refresh):
        final = it == num_refresh - 1
        HV = [[], [], []]
        cnt = nb
        for blk in range(max_blocks - 1):
            first = len(V) - cnt
            HLV = [[Hdn[n] @ v for v in VL[first:]] for n in range(3)]
            for v in V[first:]:
                for n in range(3):
                    HV[n].append(Hn[n] @ v)
            nold = len(V)
            for b in range(cnt):
                h = HV[0][first + b] + HV[1][first + b] + HV[2][first + b]
                hl = HLV[0][b] + HLV[1][b] + HLV[2][b]
                tr = _precond(kG2, V[first + b], h, eig_ref[b])
                tl = _precond(kG2, VL[first + b], hl, np.conj(eig_ref[b]))
                pair = _gs_biortho(V, VL, tr, tl)
                if pair is not None:
                    V.append(pair[0])
                    VL.append(pair[1])
            cnt = len(V) - nold
        first = len(V) - cnt
        for v in V[first:]:
            for n in range(3):
                HV[n].append(Hn[n] @ v)
        Vm, VLm = np.array(V), np.array(VL)
        Hs = [VLm.conj() @ np.array(HV[n]).T for n in range(3)]
        w, X = _sort_overlap(*np.linalg.eig(Hs[0] + Hs[1] + Hs[2]))
        XL = np.linalg.inv(X).conj().T
        Vt = X[:, :nb].T @ Vm
        VLt = XL[:, :nb].T @ VLm
        for b in range(nb):
            n = np.linalg.norm(Vt[b])
            assert n >= 1e-10, "linear independence breaks"
            Vt[b] /= n
            VLt[b] *= n
        if final:
            en = np.array([[np.vdot(XL[:, b], Hs[n] @ X[:, b]) for n in range(3)] for b in range(nb)])
            return Vt, VLt, w[:nb], en
        V = [Vt[b] for b in range(nb)]
        VL = [VLt[b] for b in range(nb)]
