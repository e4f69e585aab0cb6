"""Autotuner for the launch shape of the pair pipelines.

A launch pairs `targets` consecutive applied bands (k,j) with `sources`
consecutive occupied orbitals (q,i).  The shape sets the pair-buffer size
(targets*sources records), the wavefronts of the accumulation pass (targets x
pencils) and how often the FP64 accumulation grids are re-read, so the best
shape depends on the grid, the band counts and the device.  The tuner times
whole applies of the real workload over a candidate grid and records the
fastest shape in tc_amd/tuned_shapes.json under a workload signature; the
bench and library users look shapes up there (falling back to the library's
automatic shape).
"""
from __future__ import annotations

import json
import os
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
TABLE = os.path.join(HERE, "tuned_shapes.json")
TERMS = {"hf2x": 0, "tc2x": 1}


def signature(sys_, nb, term, precision, device="MI355X") -> str:
    prec = "fp32" if precision == 1 else "fp64"
    n = "x".join(str(int(v)) for v in sys_.n)
    return f"{device}|{term}|{prec}|{n}|nk{sys_.nk}|nt{int(np.sum(nb))}|ns{int(np.sum(sys_.nocc))}"


def default_candidates(n_targets, n_sources, channels):
    """Shapes whose pair buffer stays within 16384 grids."""
    out = []
    for t in (16, 32, 64, 96, 128, 192, 256, 384, 512):
        if t > max(16, n_targets):
            continue
        for s in (8, 16, 32, 48, 64, 128, 256, 512, 1024):
            if s > max(8, n_sources):
                continue
            if t * s * channels > 16384 or t * s < 256:
                continue
            out.append((t, s))
    return out


def tune(eng, sys_, term, nb, d_phi, d_out, precision, candidates=None, log=print):
    """Time one full apply per candidate shape (after one untimed apply that
    builds the plan) and return (best_shape, {shape: seconds})."""
    channels = 4 if term == "tc2x" else 1
    if candidates is None:
        candidates = default_candidates(int(np.sum(nb)), int(np.sum(sys_.nocc)), channels)
    fn = getattr(eng, "_term_device")
    timings = {}
    for (t, s) in candidates:
        eng.set_launch_shape(term, t, s)
        fn(term, nb, d_phi.ptr, d_out.ptr, precision)  # plan build + warm caches
        eng.synchronize()
        t0 = time.perf_counter()
        fn(term, nb, d_phi.ptr, d_out.ptr, precision)
        eng.synchronize()
        dt = time.perf_counter() - t0
        timings[(t, s)] = dt
        log(f"tune {term} targets={t:4d} sources={s:5d}: {dt * 1e3:9.2f} ms")
    best = min(timings, key=timings.get)
    eng.set_launch_shape(term, *best)
    return best, timings


def load_table(path=TABLE) -> dict:
    if not os.path.exists(path):
        return {}
    with open(path) as f:
        return json.load(f)


def save_entry(sig, shape, seconds, path=TABLE):
    table = load_table(path)
    table[sig] = {"targets": int(shape[0]), "sources": int(shape[1]), "seconds": float(seconds)}
    with open(path, "w") as f:
        json.dump(table, f, indent=1, sort_keys=True)


def apply_tuned(eng, sys_, nb, term, precision, path=TABLE) -> tuple | None:
    """Set the tuned shape for this workload if the table has one."""
    entry = load_table(path).get(signature(sys_, nb, term, precision))
    if entry is None:
        eng.set_launch_shape(term, 0, 0)
        return None
    shape = (entry["targets"], entry["sources"])
    eng.set_launch_shape(term, *shape)
    return shape
