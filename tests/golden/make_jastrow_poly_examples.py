"""Fixtures for the polynomial and RPA+polynomial Jastrow from the reference's examples
28-33 (Si, polynomial Jastrow in 28/29, RPA+polynomial in 30-33; 32/33 spin-polarised).

Run in a container that holds the reference at /root/reference:
    python tests/golden/make_jastrow_poly_examples.py
It reads the examples' data files as data (no reference code is run or imported):
  * parameters.casl       -- the CASINO Jastrow input (RPA F, polynomial order, cutoff C and
                             L, cusp flag, c_2 ... c_N per channel), laid out as
                             src/io_qmc_files_read_casl.cpp:120-330 reads it;
  * data-file-schema.xml  -- the cell and the FFT grid (they fix num_rmesh_J3_integ);
  * output.out            -- what the reference printed after reading them: A and F of the
                             RPA part, every channel's c_poly after the cusp condition
                             (Jastrow::print_polynomial_parameters,
                             src/jastrow_initialize_polynomial.cpp:399-424) and
                             num_rmesh_J3_integ (:426-431).
and writes tests/golden/jastrow_poly_examples.json.
"""
from __future__ import annotations

import json
import os
import re
import xml.etree.ElementTree as ET

import numpy as np

REF = "/root/reference/test"
HERE = os.path.dirname(os.path.abspath(__file__))
EXAMPLES = (28, 29, 30, 31, 32, 33)
CHANNELS = ("1-1", "1-2", "2-2")


def parse_casl(path):
    """RPA F per channel (if an RPA term is present) and the polynomial term."""
    lines = [ln.rstrip("\n") for ln in open(path)]
    terms, cur = [], None
    for ln in lines:
        if re.match(r"\s*TERM \d+:", ln):
            cur = []
            terms.append(cur)
        elif cur is not None:
            cur.append(ln)
    out = {"rpa_F": None, "poly": None}
    for t in terms:
        text = "\n".join(t)
        basis = re.search(r"e-e basis:\s*\n\s*Type:\s*(.+)", text).group(1).strip()
        if basis == "RPA":
            F = {}
            for ch in CHANNELS:
                m = re.search(rf"Channel {ch}:\s*\n\s*F:\s*\[\s*([-+0-9.Ee]+)", text)
                if m:
                    F[ch] = float(m.group(1))
            out["rpa_F"] = F
        elif basis == "natural power":
            order = int(re.search(r"Order:\s*(\d+)", text).group(1))
            C = int(re.search(r"Constants:\s*\n\s*C:\s*(\d+)", text).group(1))
            L = {}
            cut = text[text.index("e-e cutoff:"):text.index("e-e cusp:")]
            for ch in CHANNELS:
                m = re.search(rf"Channel {ch}:\s*\n\s*L:\s*\[\s*([-+0-9.Ee]+)", cut)
                if m:
                    L[ch] = float(m.group(1))
            cusp = re.search(r"e-e cusp:\s*([TF])", text).group(1) == "T"
            lin = text[text.index("Linear parameters:"):]
            coeffs = {}
            blocks = re.split(r"\n\s*Channel (\d-\d):", lin)
            for i in range(1, len(blocks), 2):
                ch, body = blocks[i], blocks[i + 1]
                c = [0.0] * order
                for m in re.finditer(r"c_(\d+):\s*\[\s*([-+0-9.Ee]+)", body):
                    c[int(m.group(1)) - 1] = float(m.group(2))
                coeffs[ch] = c
            out["poly"] = {"deg": order, "C": C, "L": L, "cusp": cusp, "c": coeffs}
    return out


def parse_output(path):
    text = open(path).read()
    res = {"c_poly": {}, "num_rmesh": None, "A": {}, "F": {}}
    for name, key in (("up_up", "1-1"), ("up_dn", "1-2"), ("dn_dn", "2-2")):
        m = re.search(rf"A_{name} = ([-+0-9.e]+)", text)
        if m:
            res["A"][key] = float(m.group(1))
        m = re.search(rf"F_{name} = ([-+0-9.e]+)", text)
        if m:
            res["F"][key] = float(m.group(1))
    for ch in ("up-up", "up-down", "down-up", "down-down"):
        m = re.search(rf"   {ch} channel:\n(.*?)Note:", text, re.S)
        vals = [float(v) for v in re.findall(r"c_poly \(r\^\d+ term = c_\d+\) = ([-+0-9.e]+)", m.group(1))]
        res["c_poly"][ch] = vals
    m = re.search(r"num_rmesh_J3_integ = (\d+)", text)
    if m:
        res["num_rmesh"] = int(m.group(1))
    return res


def read_xml(path):
    root = ET.parse(path).getroot()
    out = root.find("output")
    cell = np.array([[float(x) for x in out.find("atomic_structure").find("cell").find(a).text.split()]
                     for a in ("a1", "a2", "a3")])
    fft = out.find("basis_set").find("fft_grid").attrib
    return cell, [int(fft["nr1"]), int(fft["nr2"]), int(fft["nr3"])]


def main():
    data = {}
    for ex in EXAMPLES:
        d = os.path.join(REF, f"example{ex}")
        casl = parse_casl(os.path.join(d, "parameters.casl"))
        printed = parse_output(os.path.join(d, "output.out"))
        cell, grid = read_xml(os.path.join(d, "data-file-schema.xml"))
        data[str(ex)] = {"casl": casl, "printed": printed, "cell": cell.tolist(), "fft_grid": grid}
    with open(os.path.join(HERE, "jastrow_poly_examples.json"), "w") as f:
        json.dump(data, f, indent=1)
    print("wrote", os.path.join(HERE, "jastrow_poly_examples.json"))


if __name__ == "__main__":
    main()
