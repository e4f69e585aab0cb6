"""Band energies and term energies from H*phi, as TotalEnergy forms them.

The reference takes the diagonal of the (rotated) subspace matrices,
e_n(b) = <chi_b|H_n phi_b>, and sums (1/Nk) sum_{k,b} f_kb e_n(b) / n for the
n-body term (src/total_energy.cpp:57-80, 101-163; identity symmetry, so no
time-reversed conjugates).  Sums are compensated (Neumaier) in FP64.
"""
from __future__ import annotations

import numpy as np


def neumaier_sum(values) -> complex:
    s = 0.0 + 0.0j
    c = 0.0 + 0.0j
    for v in np.asarray(values, dtype=np.complex128).ravel():
        for part in (1.0, 1j):
            x = (v.real if part == 1.0 else v.imag) * part
            t = s + x
            if abs(s) >= abs(x):
                c += (s - t) + x
            else:
                c += (x - t) + s
            s = t
    return s + c


def band_expectations(sys, hphi, nb=None, left=False):
    """e(k,b) = <chi_kb|H phi_kb> for the occupied bands applied in order."""
    nb = sys.nocc if nb is None else np.asarray(nb)
    bo = sys.band_offsets(nb)
    out = []
    for ik in range(sys.nk):
        bra = sys.occ_block(ik, left=left)
        h = hphi[bo[ik]:bo[ik + 1]].reshape(nb[ik], sys.npw[ik])
        out.append(np.einsum("bl,bl->b", np.conj(bra[:nb[ik]]), h))
    return out


def term_energy(sys, hphi, nbody: int, left=False) -> complex:
    """(1/Nk) sum_{k,b} f_kb <chi_kb|H_n phi_kb> / n  (total_energy.cpp:118-163)."""
    fo = np.concatenate([[0], np.cumsum(sys.nocc)])
    terms = []
    for ik, e in enumerate(band_expectations(sys, hphi, left=left)):
        f = sys.filling[fo[ik]:fo[ik] + len(e)]
        terms.append(f * e / nbody)
    return neumaier_sum(np.concatenate(terms)) / sys.num_kpoints
