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
TERMS = {"hf2x": 0, "tc2x": 1, "tc3a3b3b4b6": 2, "tc3a2a4b2b5": 2}
SAMPLES = 3  # timed applies per shape; the median is kept


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


def three_body_candidates(n_targets, n_sources):
    """(targets, sources) for the three-body pair sets; the library caps the pairs per
    launch by its record budget, so large products are trimmed there."""
    out = []
    for t in (8, 16, 32, 64, 128):
        if t > max(8, n_targets):
            continue
        for s in (4, 8, 16, 32, 64):
            if s > max(4, n_sources):
                continue
            out.append((t, s))
    return out


def tune(eng, sys_, term, nb, d_phi, d_out, precision, candidates=None, log=print, samples=SAMPLES):
    """Time `samples` full applies per candidate shape (after one untimed apply that
    builds the plan and warms the caches) and return (best_shape, {shape: median s})."""
    channels = 4 if term == "tc2x" else 1
    if candidates is None:
        if TERMS[term] == 2:
            candidates = three_body_candidates(int(np.sum(nb)), int(np.sum(sys_.nocc)))
        else:
            candidates = default_candidates(int(np.sum(nb)), int(np.sum(sys_.nocc)), channels)
    fn = getattr(eng, "_term_device")
    timings = {}
    for (t, s) in candidates:
        eng.set_launch_shape(term, t, s)
        fn(term, nb, d_phi.ptr, d_out.ptr, precision)  # plan build + warm caches
        eng.synchronize()
        runs = []
        for _ in range(samples):
            t0 = time.perf_counter()
            fn(term, nb, d_phi.ptr, d_out.ptr, precision)
            eng.synchronize()
            runs.append(time.perf_counter() - t0)
        dt = float(np.median(runs))
        timings[(t, s)] = dt
        log(f"tune {term} targets={t:4d} sources={s:5d}: median {dt * 1e3:9.2f} ms of {samples} "
            f"(spread {(max(runs) - min(runs)) * 1e3:.2f} ms)")
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


def main(argv=None):
    """Tune the benchmark workloads and merge the winners into the table:
    python -m tc_amd.autotune --workloads fcc_bitc_40,batched_uneven [--out PATH]."""
    import argparse
    import sys

    from tc_amd import system as S
    from tc_amd import tcx

    ap = argparse.ArgumentParser()
    ap.add_argument("--workloads", default="fcc_bitc_40,batched_uneven")
    ap.add_argument("--out", default=TABLE)
    ap.add_argument("--samples", type=int, default=SAMPLES)
    args = ap.parse_args(argv)
    log = lambda m: print(m, flush=True)  # noqa: E731
    eng = tcx.Engine(0)

    # a short candidate list per term keeps a workload's sweep within minutes
    short = {"hf2x": [(48, 48), (96, 48), (96, 96), (192, 48), (192, 96), (384, 24)],
             "tc2x": [(16, 32), (32, 32), (32, 64), (64, 16), (64, 32), (96, 16), (128, 8), (128, 16)],
             "tc3a2a4b2b5": [(16, 16), (32, 8), (32, 16), (64, 8), (64, 16)],
             "tc3a3b3b4b6": [(16, 16), (32, 8), (32, 16), (64, 8), (64, 16)]}

    def run(sys_, terms):
        eng.load_system(sys_)
        nb = sys_.nocc.astype(np.int32)
        phi = sys_.phi_occ
        d_phi = eng.to_device(phi.view(np.float64))
        d_out = eng.device_buffer(phi.nbytes)
        for term, prec in terms:
            best, tim = tune(eng, sys_, term, nb, d_phi, d_out, prec, candidates=short[term], log=log,
                             samples=args.samples)
            sig = signature(sys_, nb, term, prec)
            save_entry(sig, best, tim[best], path=args.out)
            log(f"tuned {sig}: {best} {tim[best] * 1e3:.1f} ms")
            eng.set_launch_shape(term, 0, 0)
        d_phi.free()
        d_out.free()

    for w in args.workloads.split(","):
        if w == "fcc_bitc_40":  # BITC fcc 40^3, 4x4x4 k, 16 bands: the two-body and three-body terms
            run(S.fcc_bitc_40(), [("tc2x", tcx.PREC_FP32), ("tc2x", tcx.PREC_FP64),
                                  ("tc3a2a4b2b5", tcx.PREC_FP32), ("tc3a3b3b4b6", tcx.PREC_FP32)])
        elif w == "batched_uneven":  # the batched workload's cells at 32^3, 36^3 and 48^3
            for sys_ in S.batched_uneven_systems():
                run(sys_, [("hf2x", tcx.PREC_FP32), ("tc2x", tcx.PREC_FP32)] if sys_.method != "HF"
                    else [("hf2x", tcx.PREC_FP32)])
        else:
            sys.exit(f"unknown workload {w}")
    eng.close()


if __name__ == "__main__":
    main()
