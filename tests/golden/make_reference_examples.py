"""Fixtures from the reference's own end-to-end examples (HEG HF / TC / BITC on the fcc
Si cell): test/example14 (HF), test/example15 (TC), test/example16 (BITC, density-matrix
mixing) of the reference repository.

Run in a container that holds the reference at /root/reference:
    python tests/golden/make_reference_examples.py
It reads the examples' data files as data (no reference code is run or imported):
  * data-file-schema.xml  -- Quantum ESPRESSO output: cell, reciprocal vectors, FFT grid,
                             symmetry operations, irreducible k-points, weights, QE eigenvalues;
  * wfc{1,2,3}.dat_nonbin -- QE plane-wave coefficients in the text layout that
                             src/io_qe_files_read_wfc.f90:69-140 reads;
  * input.in              -- the example's TC++ input;
  * output.out            -- the printed results of the reference run (expected values).
and writes tests/golden/qe_example{14,15,16}.npz: the inputs (cell, symmetry, k-points,
Miller indices and the first num_bands_tc coefficient vectors per irreducible k) and the
expected values: the symmetry-replicated k-vectors, the Fermi energy from the QE
eigenvalues, after "Iteration 1" the Fermi energy, the eigenvalues and fillings at every
irreducible k and the 1/2/3-body and total energies, and for the whole SCF run every
printed Fermi energy / eigenvalue / filling block, total energy and density difference.
"""
from __future__ import annotations

import json
import os
import re
import sys
import xml.etree.ElementTree as ET

import numpy as np

REF = "/root/reference/test"
HERE = os.path.dirname(os.path.abspath(__file__))
EXAMPLES = {14: "HF", 15: "TC", 16: "BITC"}


def floats(text):
    return np.array([float(x) for x in text.split()])


def read_input(path):
    params = {}
    with open(path) as f:
        for line in f:
            t = line.split()
            if len(t) >= 2:
                params[t[0]] = t[1]
    return params


def read_xml(path):
    root = ET.parse(path).getroot()
    out = root.find("output")
    inp = root.find("input")
    st = out.find("atomic_structure")
    alat = float(st.attrib["alat"])
    cell = np.array([floats(st.find("cell").find(a).text) for a in ("a1", "a2", "a3")])
    bs = out.find("basis_set")
    fft = bs.find("fft_grid").attrib
    grid = [int(fft["nr1"]), int(fft["nr2"]), int(fft["nr3"])]
    b = np.array([floats(bs.find("reciprocal_lattice").find(x).text) for x in ("b1", "b2", "b3")])
    rots, trans = [], []
    for s in out.find("symmetries").findall("symmetry"):
        if s.find("equivalent_atoms") is None:  # io_qe_files_read_xml.cpp: only these are used
            continue
        rots.append(np.rint(floats(s.find("rotation").text)).astype(np.int64).reshape(3, 3))  # row-major
        trans.append(floats(s.find("fractional_translation").text))
    flags = inp.find("symmetry_flags")
    band = out.find("band_structure")
    kcart, kw, eig, npw = [], [], [], []
    for ks in band.findall("ks_energies"):
        kp = ks.find("k_point")
        kcart.append(floats(kp.text))
        kw.append(float(kp.attrib["weight"]))
        npw.append(int(ks.find("npw").text))
        eig.append(floats(ks.find("eigenvalues").text))
    return dict(alat=alat, cell=cell, grid=np.array(grid), b_norm=b, rot_qe=np.array(rots),
                trans=np.array(trans), noinv=flags.find("noinv").text.strip(),
                no_t_rev=flags.find("no_t_rev").text.strip(), nelec=float(band.find("nelec").text),
                nbnd=int(band.find("nbnd").text), kcart=np.array(kcart), kweight=np.array(kw),
                npw=np.array(npw), eig_qe=np.array(eig))


def read_wfc(path, nbands):
    """Layout of io_qe_files_read_wfc.f90:111-134 (one value group per line)."""
    with open(path) as f:
        lines = f.read().splitlines()
    ngw, igwx, npol, nbnd = (int(x) for x in lines[1].split())
    assert npol == 1
    mill = np.array([[int(x) for x in lines[3 + i].split()] for i in range(igwx)], dtype=np.int32)
    base = 3 + igwx
    coef = np.empty((nbnd, igwx), np.complex128)
    for ib in range(nbnd):
        for ig in range(igwx):
            re_, im_ = lines[base + ib * igwx + ig].split()
            coef[ib, ig] = complex(float(re_), float(im_))
    return mill, coef[:nbands]


def read_output(path, nirr):
    """Expected values printed by the reference run."""
    text = open(path).read()
    lines = text.splitlines()
    # symmetry-replicated k-vectors (kpoints.cpp: show_kvectors)
    equiv = []
    i = lines.index("  Print symmetrically-repiclated kvectors:")
    cur = None
    while i < len(lines):
        ln = lines[i]
        m = re.match(r"\s+Irreducible k-vector \(ik = (\d+)\)", ln)
        if m:
            cur = []
            equiv.append(cur)
        m = re.match(r"\s+symmetrically-equivalent k-vector \(isymk = (\d+)\):(.*)", ln)
        if m:
            cur.append([float(x) for x in m.group(2).split()] + [0.0])
        m = re.match(r"\s+time-reversal symmetry = (\w+)", ln)
        if m:
            cur[-1][3] = 1.0 if m.group(1) == "true" else 0.0
        if ln.startswith("  qes:espresso.output.band_structure.nelec"):
            break
        i += 1

    def filling_block(start):
        """Fermi energy, then per k: (eigenvalue eV, filling) rows."""
        j = start
        while "Fermi energy:" not in lines[j]:
            j += 1
        ef = float(lines[j].split(":")[1].split()[0])
        blocks = []
        while len(blocks) < nirr:
            j += 1
            if "(band index, energy (eV), filling)" in lines[j]:
                rows = []
                j += 1
                while j < len(lines) and re.match(r"\s+\d+ \S+ \S+$", lines[j]):
                    t = lines[j].split()
                    rows.append([float(t[1]), float(t[2])])
                    j += 1
                blocks.append(rows)
        return ef, blocks, j

    start = lines.index(" Start SCF calculation!")
    ef0, _, _ = filling_block(start)
    # the whole run: every set_filling print, every total energy, every density difference
    scf = {"fermi_eV": [], "eigenvalues_eV": [], "fillings": [], "energies_Ha": [], "density_difference": []}
    for i, ln in enumerate(lines):
        if i < start:
            continue
        if "Fermi energy:" in ln:
            ef, blocks, _ = filling_block(i)
            scf["fermi_eV"].append(ef)
            scf["eigenvalues_eV"].append([[r[0] for r in b] for b in blocks])
            scf["fillings"].append([[r[1] for r in b] for b in blocks])
        if ln.strip().startswith("Total energy ="):
            e = {}
            for back in lines[max(0, i - 3):i + 1]:
                for key, tag in (("e1", "(1-body terms)"), ("e2", "(2-body terms)"), ("e3", "(3-body terms)"),
                                 ("etot", "Total energy = ")):
                    if tag in back and (key != "etot" or back.strip().startswith("Total energy =")):
                        m = re.search(r"\( (\S+) , (\S+) \)", back)
                        e[key] = [float(m.group(1)), float(m.group(2))]
            scf["energies_Ha"].append(e)
        m = re.match(r"\s+density difference \(L1 norm\) from the previous loop = (\S+)", ln)
        if m:
            scf["density_difference"].append(float(m.group(1)))
    it1 = lines.index(" Iteration 1")
    ef1, blocks1, j = filling_block(it1)
    energies = {}
    for key, tag in (("e1", "(1-body terms)"), ("e2", "(2-body terms)"), ("e3", "(3-body terms)"),
                     ("etot", "Total energy = ")):
        while j < len(lines):
            if tag in lines[j] and (key != "etot" or lines[j].strip().startswith("Total energy =")):
                m = re.search(r"\( (\S+) , (\S+) \)", lines[j])
                energies[key] = [float(m.group(1)), float(m.group(2))]
                break
            if key == "e3" and lines[j].strip().startswith("Total energy ="):
                break
            j += 1
    return equiv, ef0, ef1, blocks1, energies, scf


def main():
    if not os.path.isdir(REF):
        sys.exit(f"{REF} not found")
    for ex, method in EXAMPLES.items():
        d = os.path.join(REF, f"example{ex}")
        params = read_input(os.path.join(d, "input.in"))
        x = read_xml(os.path.join(d, "data-file-schema.xml"))
        nirr = len(x["kcart"])
        nb = int(params.get("num_bands_tc", x["nbnd"]))
        mills, coefs = [], []
        for ik in range(nirr):
            m, c = read_wfc(os.path.join(d, f"wfc{ik + 1}.dat_nonbin"), nb)
            assert len(m) == x["npw"][ik]
            mills.append(m)
            coefs.append(c)
        equiv, ef0, ef1, blocks1, energies, scf = read_output(os.path.join(d, "output.out"), nirr)
        expected = {
            "fermi_energy_initial_eV": ef0,
            "fermi_energy_iter1_eV": ef1,
            "eigenvalues_iter1_eV": [[r[0] for r in b] for b in blocks1],
            "fillings_iter1": [[r[1] for r in b] for b in blocks1],
            "energies_iter1_Ha": energies,
            "equivalent_kvectors": equiv,
            "scf": scf,
        }
        setup = {
            "method": method,
            "num_bands_tc": nb,
            "smearing_mode": params.get("smearing_mode", "gaussian"),
            "smearing_width": float(params.get("smearing_width", 0.01)),
            "max_num_blocks_david": int(params.get("max_num_blocks_david", 2)),
            "num_refresh_david": int(params.get("num_refresh_david", 1)),
            # Diagonalization defaults (include/diagonalization.hpp:75-89, diagonalization.cpp:48-64)
            "max_num_iterations": int(params.get("max_num_iterations", 30)),
            "mixing_beta": float(params.get("mixing_beta", 0.7)),
            "mixes_density_matrix": params.get("mixes_density_matrix", "false").lower() == "true",
            "energy_tolerance": float(params.get("energy_tolerance", 1e-6)),
            "charge_tolerance": float(params.get("charge_tolerance", 1e-5)),
            "A_input": 1.0,  # default RPA Jastrow (no parameters.casl in the examples)
            "source": f"reference test/example{ex}",
        }
        out = os.path.join(HERE, f"qe_example{ex}.npz")
        np.savez_compressed(
            out, alat=x["alat"], cell=x["cell"], grid=x["grid"], b_norm=x["b_norm"], rot_qe=x["rot_qe"],
            trans=x["trans"], noinv=x["noinv"], no_t_rev=x["no_t_rev"], nelec=x["nelec"], kcart=x["kcart"],
            kweight=x["kweight"], eig_qe=x["eig_qe"][:, :nb], npw=x["npw"],
            mill=np.concatenate(mills), coef=np.concatenate([c.reshape(-1) for c in coefs]),
            setup=json.dumps(setup), expected=json.dumps(expected))
        print(f"wrote {out}: {method}, {nirr} irreducible k, {nb} bands, grid {x['grid'].tolist()}, "
              f"{len(x['rot_qe'])} symmetries; E_tot(iter 1) = {energies['etot'][0]}")


if __name__ == "__main__":
    main()
