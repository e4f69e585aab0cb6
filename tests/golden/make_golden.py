"""Write the golden fixtures tests/golden/*.npz.

Each fixture is a tiny plane-wave system (inputs) together with the H*phi
terms computed by the brute-force G-space restatement in bruteforce.py, which
uses cyclic convolutions on the FFT grid and no FFT.  The fixtures pin the
FFT-based oracle (oracle/) and, through it, the GPU engine.

Run from the repository root:  python tests/golden/make_golden.py
"""
from __future__ import annotations

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, HERE)

import bruteforce as BF  # noqa: E402
from tc_amd import system as S  # noqa: E402


def _sum_of_vaux(sys_):
    # independent restatement in numpy of potentials_coulomb.cpp:60-161
    alpha = (sys_.volume ** (1.0 / 3.0) / (2.0 * np.pi)) ** 2
    out = np.zeros(sys_.nk)
    for ik in range(sys_.nk):
        s = 0.0
        match_any = False
        for iq in range(sys_.nk):
            kq = sys_.kfrac[ik] - sys_.kfrac[iq]
            kq = kq - np.floor(kq + 1e-8)
            kq = np.where(kq > 1.0 - 1e-8, kq - 1.0, kq)
            match = np.dot(kq, kq) < 1e-8
            match_any |= match
            if not match:
                c = kq @ sys_.recip
                s += np.exp(-alpha * c @ c) / (c @ c)
            shell = 1
            while True:
                r = np.arange(-shell, shell + 1)
                g = np.stack(np.meshgrid(r, r, r, indexing="ij"), -1).reshape(-1, 3)
                g = g[np.max(np.abs(g), axis=1) == shell]
                c = (kq[None, :] + g) @ sys_.recip
                g2 = np.sum(c * c, axis=1)
                on = np.sum(np.exp(-alpha * g2) / g2)
                s += on
                if on < 1e-20:
                    break
                shell += 1
        s *= 4 * np.pi
        if match_any:
            s -= 4 * np.pi * alpha
        out[ik] = s
    return alpha, out


CASES = {
    # name: (method, grid, kmesh, nbands, nocc, lattice, applied nb, seed)
    "hf_fcc_8": ("HF", (8, 8, 8), (2, 1, 1), 3, 2, S.fcc_lattice(6.0), [3, 2], 11),
    "hf_sc_6": ("HF", (6, 6, 6), (2, 2, 1), 2, 2, S.sc_lattice(5.0), [2, 1, 2, 2], 12),
    "tc_fcc_8": ("TC", (8, 8, 8), (2, 1, 1), 3, 2, S.fcc_lattice(6.0), [3, 2], 13),
    "bitc_fcc_8": ("BITC", (8, 8, 8), (2, 1, 1), 3, 2, S.fcc_lattice(6.0), [2, 3], 14),
}


def make_case(name):
    method, grid, kmesh, nbands, nocc, lattice, nb, seed = CASES[name]
    sys_ = S.tiny_system(method=method, n=grid, kmesh=kmesh, nbands=nbands, nocc=nocc, lattice=lattice,
                         seed_stream=seed)
    nb, phi = S.applied_bands(sys_, nb=nb, seed=seed)
    alpha, sv = _sum_of_vaux(sys_)
    out = dict(
        recip=sys_.recip, volume=sys_.volume, n=np.array(sys_.n), kfrac=sys_.kfrac, npw=sys_.npw,
        gindex=sys_.gindex, nocc=sys_.nocc, filling=sys_.filling, phi_occ=sys_.phi_occ,
        chi_occ=sys_.chi_occ if sys_.chi_occ is not None else np.zeros(0, np.complex128),
        A=sys_.A, F=sys_.F, method=np.array(method), nb=nb, phi=phi, alpha=alpha, sum_vaux=sv,
    )
    out["kinetic"] = BF.kinetic(sys_, nb, phi)
    if method == "HF":
        out["hf2x"] = BF.hf2x(sys_, nb, phi, alpha, sv)
        out["hf2h"] = BF.hf2h(sys_, nb, phi)
        sys_.div_corr = False
        out["hf2x_nodiv"] = BF.hf2x(sys_, nb, phi, alpha, sv)
    else:
        out["tc2x"] = BF.tc2x(sys_, nb, phi, alpha, sv)
        out["tc2h"] = BF.tc2h(sys_, nb, phi)
        out["tc3a1"] = BF.tc3a1(sys_, nb, phi)
        out["tc3b1"] = BF.tc3b1(sys_, nb, phi)
        out["tc3a3b3b4b6"] = BF.tc3a3b3b4b6(sys_, nb, phi, alpha, sv)
        out["tc3a2a4b2b5"] = BF.tc3a2a4b2b5(sys_, nb, phi, alpha, sv)
    np.savez_compressed(os.path.join(HERE, f"{name}.npz"), **out)
    return out


def load_system(npz):
    """Rebuild a System from a fixture."""
    method = str(npz["method"])
    chi = npz["chi_occ"]
    return S.System(lattice=np.linalg.inv(npz["recip"] / (2 * np.pi)).T, recip=npz["recip"],
                    volume=float(npz["volume"]), n=tuple(int(x) for x in npz["n"]), kfrac=npz["kfrac"],
                    npw=npz["npw"], gindex=npz["gindex"], nocc=npz["nocc"], filling=npz["filling"],
                    phi_occ=npz["phi_occ"], chi_occ=chi if chi.size else None, method=method,
                    A=npz["A"], F=npz["F"], div_corr=True, name="golden")


if __name__ == "__main__":
    for name in CASES:
        make_case(name)
        print("wrote", name)
