"""Benchmark: pair-FFT operator applies/s of the Fock-exchange apply (hf2x) on one MI355X.

Workload (one "step" = one full H*phi apply of the exchange term):
  fcc a=10.26 bohr, 4x4x4 k (64 k, q over all k), 16 bands/k all occupied (f=2),
  FFT grid 40^3, plane-wave sphere ~2738 G/k, synthetic seeded orbitals;
  1,048,576 (k,j,q,i) pair applies per step.
The timed steps run the FP32-split path (FFTs and pointwise products in FP32,
band-pair sums and outputs in FP64) with inputs and outputs resident in HBM.
The same apply in plain FP64 is timed for the FP32/FP64 ratio, and the two
outputs are compared (per-band relative error, exchange-energy difference).

Output: one JSON line (see the repository's benchmark contract) with
"roofline" for the dominant kernel (HIP events on the engine's stream around
every launch inside the timed region) and "cpu_baseline" (the oracle's FP64 hf2x
restated with batched SciPy FFTs and one thread per occupied q, on a bounded sample of
the same workload, about 10 s).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "pair-FFT operator applies/s + HBM GB/s: FP32-split vs FP64 at ΔE<1e-6 Ha"
HBM_PEAK_GBS = 8000.0
CLASS_NAMES = ["pass_A_zpencil", "pass_B_xyplane", "pass_C_zpencil_acc", "orbital_and_final_fft"]


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def read_traffic(name, pairs_per_launch):
    """HBM bytes per launch of a kernel: the calibrated FETCH_SIZE + WRITE_SIZE
    bytes per pair of the committed rocprofv3 --pmc summary times the pairs of
    one launch (None when no summary is present)."""
    path = os.path.join(ROOT, "profiles", "traffic.json")
    if not os.path.exists(path):
        return None
    try:
        with open(path) as f:
            entry = json.load(f).get(name)
    except (OSError, ValueError):
        return None
    if not entry or "bytes_per_pair" not in entry:
        return None
    return float(entry["bytes_per_pair"]) * pairs_per_launch


def cpu_model():
    """Model name of the host CPU (first /proc/cpuinfo entry), or None."""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def host_cores():
    """CPUs this process may run on (the box's share), not the machine's total."""
    try:
        return len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        return os.cpu_count() or 1


def oracle_bands(sys_, nbands, threads):
    """FP64 oracle (reference loop nest, OpenMP over q) on the first `nbands` occupied bands at
    k=0 against all occupied (q,i); returns (output, seconds, pairs)."""
    from oracle import oracle as O
    o = O.Oracle(sys_)
    nb = np.zeros(sys_.nk, np.int32)
    nb[0] = nbands
    phi = sys_.occ_block(0)[:nbands].reshape(-1).copy()
    t0 = time.perf_counter()
    out = o.hf2x(nb, phi, nthreads=threads)
    dt = time.perf_counter() - t0
    return out, dt, nbands * int(sys_.nocc.sum())


def cpu_baseline(sys_, threads, target_s, max_bands, log):
    """The CPU baseline: the batched FP64 restatement of the oracle's hf2x (oracle/hf2x_batched.py:
    pocketfft through scipy.fft, the occupied q in parallel threads) on the first applied bands
    in k order.  One band calibrates the rate; the sample then takes as many bands as fit in
    about `target_s` seconds (at most `max_bands`).  The occupied orbitals' transforms (once per
    apply) are done before the timed region."""
    from oracle import hf2x_batched as HB
    occ = [HB._to_real(sys_, iq, sys_.occ_block(iq), threads) if sys_.nocc[iq] > 0 else None
           for iq in range(sys_.nk)]

    def run(nbands):
        nb = np.zeros(sys_.nk, np.int32)
        left, phis = nbands, []
        for ik in range(sys_.nk):
            nb[ik] = min(left, int(sys_.nocc[ik]))
            phis.append(sys_.occ_block(ik)[:nb[ik]].reshape(-1))
            left -= nb[ik]
        phi = np.concatenate(phis)
        t0 = time.perf_counter()
        HB.hf2x(sys_, nb, phi, workers=threads, occ_real=occ)
        return time.perf_counter() - t0, int(nbands * sys_.nocc.sum())

    dt1, _ = run(1)
    nbands = int(max(1, min(max_bands, target_s / max(dt1, 1e-3))))
    log(f"bench: CPU baseline, 1 band {dt1:.2f} s; timing {nbands} band(s) on {threads} threads")
    dt, pairs = run(nbands)
    return {"value": pairs / dt, "unit": "pairs/s", "cores": threads, "kind": "port",
            "cpu_model": cpu_model(), "machine_cpus": os.cpu_count(),
            "sample": f"hf2x FP64, batched NumPy/SciPy restatement of the oracle (pocketfft, one thread per "
                      f"occupied q), {nbands} applied bands x {pairs // nbands} occupied (q,i) = {pairs} pairs, "
                      f"{dt:.2f} s wall on {threads} threads; the occupied orbitals' transforms precede the timed "
                      "region and the divergence correction is left out; pairs/s is per pair, the same unit as value"}


def calibration_kernels(eng, log):
    """Known-size kernels that scale the HBM counters of a profiled run: k_copy reads
    and writes 1 GiB, k_zero writes 1 GiB."""
    gib = 1 << 30
    a, b = eng.device_buffer(gib), eng.device_buffer(gib)
    for _ in range(3):
        eng.L.tcx_memcpy(eng.h, b.ptr, a.ptr, gib, 2)  # k_copy: reads and writes 1 GiB
        a.zero()                                       # k_zero: writes 1 GiB
    a.free()
    b.free()
    log("bench: calibration kernels done (k_copy 1 GiB, k_zero 1 GiB)")


def copy_ceiling(eng, gib=4, reps=10):
    """Attainable HBM rate of a streaming device copy (the library's copy kernel behind
    tcx_memcpy, four 16-byte non-temporal loads in flight per lane): the best of `reps`
    synchronous copies of `gib` GiB, counting the read and the write (2 x gib GiB moved).
    The wall clock includes the launch and the synchronisation (~1% at 4 GiB)."""
    n = gib << 30
    a, b = eng.device_buffer(n), eng.device_buffer(n)
    a.zero()
    b.zero()
    best = float("inf")
    for _ in range(reps):
        t0 = time.perf_counter()
        eng._check(eng.L.tcx_memcpy(eng.h, b.ptr, a.ptr, n, 2), "memcpy d2d")
        best = min(best, time.perf_counter() - t0)
    a.free()
    b.free()
    return 2.0 * n / best / 1e9


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--fp64-steps", type=int, default=1)
    ap.add_argument("--targets", type=int, default=0, help="launch shape: targets per launch (0 = tuned/auto)")
    ap.add_argument("--sources", type=int, default=0, help="launch shape: sources per launch (0 = tuned/auto)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-threads", type=int, default=0, help="oracle threads (0: every CPU this process may use)")
    ap.add_argument("--cpu-bands", type=int, default=64, help="most applied bands in the CPU-baseline sample")
    ap.add_argument("--cpu-seconds", type=float, default=10.0, help="target wall time of the CPU-baseline sample")
    ap.add_argument("--no-tc", action="store_true",
                    help="skip the default TC legs (tc2x and the three-body terms on the BITC cell)")
    ap.add_argument("--tc-steps", type=int, default=3, help="timed FP32-split tc2x applies in the TC leg")
    ap.add_argument("--tune", action="store_true", help="run the launch-shape autotuner first")
    ap.add_argument("--quick", action="store_true",
                    help="profiling run: one FP32 apply only (no FP64 leg, parity or CPU baseline)")
    ap.add_argument("--calibrate", action="store_true",
                    help="also run a 1 GiB device copy and a 1 GiB fill (counter calibration kernels)")
    ap.add_argument("--tune-out", default=os.path.join(ROOT, "tuned_shapes.new.json"),
                    help="where --tune writes its table (tc_amd/tuned_shapes.json is the one the library reads)")
    ap.add_argument("--batched", type=int, default=0, metavar="REPLAYS",
                    help="also run the batched uneven workload (3 grids x 32 k, 8-48 bands per k, hf2x + tc2x) "
                         "as captured HIP graphs for REPLAYS replays vs sequential per-k launches; report 'batched'")
    ap.add_argument("--adjoint", action="store_true",
                    help="also time the gradients of E_x and of the BITC two-body energy on fcc 3x3x3 k, "
                         "32 bands, 48^3 (FP32-split forward) with a central-difference check; report 'adjoint'")
    ap.add_argument("--skip-headline", action="store_true",
                    help="run only the requested extra legs (--bitc/--batched/--adjoint), e.g. under a profiler")
    ap.add_argument("--davidson", action="store_true",
                    help="also time tcx_davidson (2 blocks, 1 refresh, FP32-split pair terms) on the HF "
                         "4x4x4 k, 16-band, 40^3 cell against one calc_hamiltonian::all apply; report 'davidson'")
    ap.add_argument("--bitc-terms", default="tc2x,tc3a1,tc3b1,tc3a2a4b2b5,tc3a3b3b4b6",
                    help="comma-separated terms of the --bitc leg")
    ap.add_argument("--bitc-precs", default="fp32,fp64", help="precisions of the --bitc leg (fp32, fp64)")
    ap.add_argument("--bitc-repeat", type=int, default=1,
                    help="applies per term and precision in --bitc; the last one is timed (2 = steady state)")
    ap.add_argument("--bitc", action="store_true",
                    help="also time the BITC workload (tc2x + three-body terms, fcc 4x4x4k 16b 40^3) in "
                         "FP32-split and FP64 and report its energy errors under 'bitc_terms'")
    args = ap.parse_args()
    if args.steps < 1 or args.fp64_steps < 1 or args.warmup < 0:
        ap.error("--steps and --fp64-steps must be at least 1, --warmup at least 0")

    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU only (asked for --gpus {args.gpus}); "
              "sharding over k across GPUs is a design, not part of this build.", file=sys.stderr)
        sys.exit(2)

    from tc_amd import autotune as AT
    from tc_amd import energy as E
    from tc_amd import system as S
    from tc_amd import tcx

    if args.skip_headline:
        eng = tcx.Engine(0)
        if args.calibrate:
            calibration_kernels(eng, log)
        extra = {}
        if args.bitc:
            extra["bitc_terms"] = bench_bitc_terms(eng, log, args.bitc_terms.split(","), args.bitc_precs.split(","), args.bitc_repeat)
        if args.batched:
            extra["batched"] = bench_batched(log, args.batched)
        if args.adjoint:
            extra["adjoint"] = bench_adjoint(eng, log)
        if args.davidson:
            extra["davidson"] = bench_davidson(eng, log)
        print(json.dumps(extra), flush=True)
        eng.close()
        return

    log("bench: building workload fcc 4x4x4k 16b 40^3")
    sys_ = S.fcc_hf_40()
    nb = sys_.nocc.copy()
    phi = sys_.phi_occ
    pairs_per_step = int(np.sum(nb) * np.sum(sys_.nocc))
    N = sys_.ngrid

    eng = tcx.Engine(0)
    eng.load_system(sys_)
    def shape_for(prec):
        if args.targets or args.sources:
            return (args.targets, args.sources)
        entry = AT.load_table().get(AT.signature(sys_, nb, "hf2x", prec))
        return (entry["targets"], entry["sources"]) if entry else (0, 0)
    # inputs and outputs resident in HBM (the library's own allocations)
    d_phi = eng.to_device(phi)
    d_out32 = eng.device_buffer(phi.nbytes)
    d_out64 = eng.device_buffer(phi.nbytes)
    d_scratch = eng.device_buffer(phi.nbytes)

    def step(prec, out):
        eng._term_device("hf2x", nb, d_phi.ptr, out.ptr, prec)

    if args.tune:
        os.makedirs(os.path.dirname(args.tune_out), exist_ok=True)
        for prec in (tcx.PREC_FP32, tcx.PREC_FP64):
            best, tim = AT.tune(eng, sys_, "hf2x", nb, d_phi, d_scratch, prec, log=log)
            AT.save_entry(AT.signature(sys_, nb, "hf2x", prec), best, tim[best], path=args.tune_out)
            log(f"bench: tuned {'fp32' if prec else 'fp64'} shape {best} {tim[best] * 1e3:.1f} ms")
        tuned = AT.load_table(args.tune_out)
        shape32 = tuple(tuned[AT.signature(sys_, nb, "hf2x", tcx.PREC_FP32)][k] for k in ("targets", "sources"))
        shape64 = tuple(tuned[AT.signature(sys_, nb, "hf2x", tcx.PREC_FP64)][k] for k in ("targets", "sources"))
    else:
        shape32, shape64 = shape_for(tcx.PREC_FP32), shape_for(tcx.PREC_FP64)
    log(f"bench: launch shapes fp32 {shape32} fp64 {shape64} (0 = automatic)")
    eng.set_launch_shape("hf2x", *shape32)

    if args.calibrate:
        calibration_kernels(eng, log)
    if args.quick:
        step(tcx.PREC_FP32, d_out32)
        eng.synchronize()
        log("bench: quick profiling apply done")
        eng.close()
        return

    log(f"bench: warmup {args.warmup} FP32-split step(s)")
    for _ in range(args.warmup):
        d_out32.zero()
        step(tcx.PREC_FP32, d_out32)

    log(f"bench: timing {args.steps} FP32-split step(s)")
    eng.enable_timing(True)
    eng.synchronize()
    t0 = time.perf_counter()
    for i in range(args.steps):
        step(tcx.PREC_FP32, d_out32)
        log(f"bench: step {i + 1}/{args.steps} done")
    eng.synchronize()
    t1 = time.perf_counter()
    ms, launches, kpairs = eng.get_timing()
    eng.enable_timing(False)
    elapsed = t1 - t0
    ms_per_step = elapsed * 1e3 / args.steps
    value = pairs_per_step * args.steps / elapsed

    # dominant kernel of the pair pipeline (classes 0..2) and its roofline
    dom = int(np.argmax(ms[:3]))
    avg_ms = ms[dom] / max(1, launches[dom])
    pairs_per_launch = kpairs[dom] / max(1, launches[dom])
    bytes_per_launch = 2.0 * N * 8 * pairs_per_launch  # each pass: one read + one write of a FP32 pair grid
    achieved = bytes_per_launch / (avg_ms * 1e-3) / 1e9
    pipeline_gbs = 6.0 * N * 8 * pairs_per_step / (ms_per_step * 1e-3) / 1e9
    traffic = read_traffic(CLASS_NAMES[dom], pairs_per_launch)

    ceiling = copy_ceiling(eng)
    log(f"bench: streaming copy ceiling {ceiling:.0f} GB/s")

    # FP64 apply for the ratio, then parity of the two outputs
    log(f"bench: FP64 apply x{args.fp64_steps}")
    eng.set_launch_shape("hf2x", *shape64)
    d_out64.zero()
    step(tcx.PREC_FP64, d_out64)  # also the parity reference
    eng.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.fp64_steps):
        step(tcx.PREC_FP64, d_scratch)
    eng.synchronize()
    fp64_s = (time.perf_counter() - t0) / args.fp64_steps
    value64 = pairs_per_step / fp64_s

    eng.set_launch_shape("hf2x", *shape32)
    d_out32.zero()
    step(tcx.PREC_FP32, d_out32)
    h32 = d_out32.download(np.complex128)
    h64 = d_out64.download(np.complex128)
    bo = sys_.band_offsets(nb)
    rel = []
    for ik in range(sys_.nk):
        a = h32[bo[ik]:bo[ik + 1]].reshape(nb[ik], -1)
        b = h64[bo[ik]:bo[ik + 1]].reshape(nb[ik], -1)
        rel.extend(np.linalg.norm(a - b, axis=1) / np.linalg.norm(b, axis=1))
    ex32 = E.term_energy(sys_, h32, 2)
    ex64 = E.term_energy(sys_, h64, 2)

    # parity against the FP64 CPU oracle on band (k=0, j=0)
    threads = args.cpu_threads or host_cores()
    log(f"bench: FP64 oracle on 1 band at k=0, {threads} threads")
    ref, _, _ = oracle_bands(sys_, 1, threads)
    npw0 = int(sys_.npw[0])
    ref0 = ref[:npw0]
    oracle_rel = {p: float(np.linalg.norm(h[:npw0] - ref0) / np.linalg.norm(ref0)) for p, h in (("fp64", h64), ("fp32", h32))}
    cpu = None if args.no_cpu_baseline else cpu_baseline(sys_, threads, args.cpu_seconds, args.cpu_bands, log)
    failures = []
    if not oracle_rel["fp64"] <= 1e-10:
        failures.append(f"FP64 band (0,0) differs from the oracle: {oracle_rel['fp64']:.3e}")
    if not oracle_rel["fp32"] <= 1e-6:
        failures.append(f"FP32-split band (0,0) differs from the oracle: {oracle_rel['fp32']:.3e}")
    if not abs(ex32 - ex64) < 1e-6:
        failures.append(f"|dE_x| FP32-split vs FP64 = {abs(ex32 - ex64):.3e} Ha >= 1e-6")

    out = {
        "metric": METRIC,
        "value": value,
        "unit": "pairs/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic (seeded Gaussian orbitals x exp(-|k+G|^2/2), Gram-Schmidt per k)",
        "config": {
            "workload": "hf2x fcc a=10.26 bohr, 4x4x4 k, 16 bands/k (all occupied), FFT 40^3, FP32-split",
            "pairs_per_step": pairs_per_step, "grid": list(sys_.n), "nk": sys_.nk, "bands": int(nb[0]),
            "npw_mean": float(np.mean(sys_.npw)), "div_correction": True,
            "launch_shape_fp32": list(shape32), "launch_shape_fp64": list(shape64),
        },
        "roofline": {
            "bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
            "kernel": CLASS_NAMES[dom], "avg_launch_ms": avg_ms, "pairs_per_launch": pairs_per_launch,
            "algorithmic_bytes_per_launch": bytes_per_launch,
            "copy_ceiling_GBps": ceiling, "frac_of_copy_ceiling": achieved / ceiling,
        },
        "cpu_baseline": cpu,
        "pipeline": {
            "algorithmic_GBps_6NBc_per_pair": pipeline_gbs, "frac_of_hbm": pipeline_gbs / HBM_PEAK_GBS,
            "class_ms_per_step": {CLASS_NAMES[c]: float(ms[c]) / args.steps for c in range(4)},
            "launches_per_step": {CLASS_NAMES[c]: int(launches[c]) // args.steps for c in range(4)},
        },
        "fp64": {"value": value64, "ms_per_step": fp64_s * 1e3, "fp32_over_fp64": value / value64},
        "parity_fp32_vs_fp64": {
            "per_band_rel_max": float(np.max(rel)), "per_band_rel_median": float(np.median(rel)),
            "E_x_fp64_Ha": float(ex64.real), "dE_x_Ha": float(abs(ex32 - ex64)),
        },
        "parity_vs_oracle": {"band": [0, 0], "rel_fp64": oracle_rel["fp64"], "rel_fp32": oracle_rel["fp32"],
                             "tolerance": {"fp64": 1e-10, "fp32": 1e-6}},
    }
    if not args.no_tc:
        out["tc"] = bench_tc_legs(eng, log, args.tc_steps)
        tc2 = out["tc"]["tc2x"]
        if not tc2["dE_Ha"] < 1e-6:
            failures.append(f"tc2x |dE| FP32-split vs FP64 = {tc2['dE_Ha']:.3e} Ha >= 1e-6")
        for term, e in out["tc"]["three_body"].items():
            if not e["dE_Ha"] < 1e-6:
                failures.append(f"{term} |dE| FP32-split vs FP64 = {e['dE_Ha']:.3e} Ha >= 1e-6")
        # the north-star ratio: exchange + TC two-body, FP64 GPU time over FP32-split
        out["exchange_plus_tc2x"] = {
            "fp32_ms": ms_per_step + tc2["fp32_ms"], "fp64_ms": fp64_s * 1e3 + tc2["fp64_ms"],
            "fp32_over_fp64": (fp64_s * 1e3 + tc2["fp64_ms"]) / (ms_per_step + tc2["fp32_ms"])}
    out["checks_failed"] = failures
    if args.bitc:
        out["bitc_terms"] = bench_bitc_terms(eng, log, args.bitc_terms.split(","), args.bitc_precs.split(","), args.bitc_repeat)
    if args.batched:
        out["batched"] = bench_batched(log, args.batched)
    if args.adjoint:
        out["adjoint"] = bench_adjoint(eng, log)
    if args.davidson:
        out["davidson"] = bench_davidson(eng, log)
    print(json.dumps(out), flush=True)
    for b in (d_phi, d_out32, d_out64, d_scratch):
        b.free()
    eng.close()
    if failures:
        for f in failures:
            log(f"bench: CHECK FAILED: {f}")
        sys.exit(1)


def bench_adjoint(eng, log, h=1e-3):
    """Gradients on the 3x3x3 k, 32-band, 48^3 fcc cell (746,496 pairs per apply): dE_x/d conj(phi)
    (HF cell) and dE_2/d conj(chi), dE_2/d phi, dE_2/dA (BITC cell), FP32-split forward with FP64
    accumulation (timed); each checked by a central difference (step h) along one seeded direction
    with FP64 energies (an FP32-split energy difference over 2h would be rounding-limited at
    ~1e-7 Ha / 2h); plus the kernel-gradient backward pass."""
    from tc_amd import system as S
    from tc_amd import tcx
    res = {}
    rng = np.random.default_rng(4)
    for method in ("HF", "BITC"):
        sys_ = S.fcc_adjoint_48(method)
        d = rng.standard_normal(sys_.phi_occ.size) + 1j * rng.standard_normal(sys_.phi_occ.size)
        d /= np.linalg.norm(d)
        eng.load_system(sys_)
        prec = tcx.PREC_FP32
        t0 = time.perf_counter()
        if method == "HF":
            e, g = eng.grad_exchange(prec)
            t_grad = time.perf_counter() - t0
            an = 2.0 * np.real(np.vdot(g, d))
            energy = lambda s: (eng.load_system(s), eng.grad_exchange(tcx.PREC_FP64)[0])[1].real
        else:
            e, gl, gr, dA = eng.grad_tc2(prec)
            t_grad = time.perf_counter() - t0
            an = np.sum(d * gr)  # along phi
            energy = lambda s: (eng.load_system(s), eng.grad_tc2(tcx.PREC_FP64)[0])[1]
        log(f"bench adjoint: {method} gradient {t_grad:.2f} s")
        ep = energy(S.System(**{**sys_.__dict__, "phi_occ": sys_.phi_occ + h * d}))
        em = energy(S.System(**{**sys_.__dict__, "phi_occ": sys_.phi_occ - h * d}))
        fd = (ep - em) / (2 * h)
        entry = {"pairs_per_apply": int(sys_.nocc.sum()) ** 2, "gradient_s": t_grad,
                 "energy_Ha": [float(np.real(e)), float(np.imag(e))],
                 "directional_derivative": {"analytic": [float(np.real(an)), float(np.imag(an))],
                                            "central_difference": [float(np.real(fd)), float(np.imag(fd))],
                                            "rel_diff": float(abs(fd - an) / max(abs(an), 1e-300))}}
        if method == "BITC":
            eng.load_system(sys_)
            ea = energy(S.System(**{**sys_.__dict__, "A": sys_.A + h}))
            eb = energy(S.System(**{**sys_.__dict__, "A": sys_.A - h}))
            fdA = (ea - eb) / (2 * h)
            entry["dE_dA"] = {"analytic": [dA.real, dA.imag], "central_difference": [float(np.real(fdA)), float(np.imag(fdA))],
                              "rel_diff": float(abs(fdA - dA) / max(abs(dA), 1e-300))}
        eng.load_system(sys_)
        t0 = time.perf_counter()
        gk, gu, sums = eng.kernel_gradient("x" if method == "HF" else "tc2", 256, 0.05, prec)
        entry["kernel_gradient_s"] = time.perf_counter() - t0
        entry["kernel_gradient_linear_sum"] = [float(sums[0]), float(sums[1])]
        log(f"bench adjoint: {method} kernel gradient {entry['kernel_gradient_s']:.2f} s")
        res[method] = entry
    return res


def bench_batched(log, replays):
    """Batched uneven workload: three fcc cells (32^3, 36^3, 48^3), 32 k each with 8-48 bands
    per k (all occupied), hf2x + tc2x in FP32-split.  Graph mode: one context per cell, each
    apply captured once into a HIP graph, the three graphs launched together on their own
    streams per replay.  Sequential mode: the same work as separate per-k calls (one k's
    bands at a time), each call synchronous.  Reports p50 / p99 wall time per apply."""
    from tc_amd import autotune as AT
    from tc_amd import system as S
    from tc_amd import tcx
    systems = S.batched_uneven_systems()
    ctxs = []
    pairs = 0
    for sys_ in systems:
        eng = tcx.Engine(0)
        eng.load_system(sys_)
        nb = sys_.nocc.copy()
        for term in ("hf2x", "tc2x"):  # the autotuner's shapes for this cell (automatic if none)
            AT.apply_tuned(eng, sys_, nb, term, tcx.PREC_FP32)
        d_phi = eng.to_device(sys_.phi_occ)
        d_out = eng.device_buffer(sys_.phi_occ.nbytes)
        ctxs.append((sys_, eng, nb, d_phi, d_out))
        pairs += 2 * int(nb.sum()) * int(sys_.nocc.sum())  # hf2x + tc2x pairs

    def apply(eng, nb, d_phi, d_out):
        d_out.zero()
        eng._term_device("hf2x", nb, d_phi.ptr, d_out.ptr, tcx.PREC_FP32)
        eng._term_device("tc2x", nb, d_phi.ptr, d_out.ptr, tcx.PREC_FP32)

    log("bench batched: eager warm-up of the batched applies")
    ref = []
    for sys_, eng, nb, d_phi, d_out in ctxs:
        apply(eng, nb, d_phi, d_out)
        eng.synchronize()
        ref.append(d_out.download(np.complex128))
    gids = []
    for sys_, eng, nb, d_phi, d_out in ctxs:
        eng.graph_begin()
        apply(eng, nb, d_phi, d_out)
        gids.append(eng.graph_end())
    graph_t = []
    for r in range(replays):
        t0 = time.perf_counter()
        for (sys_, eng, nb, d_phi, d_out), gid in zip(ctxs, gids):
            eng.graph_launch(gid)
        for sys_, eng, nb, d_phi, d_out in ctxs:
            eng.synchronize()
        graph_t.append(time.perf_counter() - t0)
        log(f"bench batched: graph replay {r + 1}/{replays} {graph_t[-1]:.3f} s")
    same = all(np.array_equal(d_out.download(np.complex128), rf) for (_, _, _, _, d_out), rf in zip(ctxs, ref))

    def per_k(sys_, eng, nb, d_phi, d_out):
        d_out.zero()
        for ik in range(sys_.nk):
            one = np.zeros(sys_.nk, np.int32)
            one[ik] = nb[ik]
            off = int(sys_.band_offsets(nb)[ik])
            for term in ("hf2x", "tc2x"):
                # one k's bands: phi and out pointers advanced to its block
                eng._term_device(term, one, d_phi.ptr + 16 * off, d_out.ptr + 16 * off, tcx.PREC_FP32)

    log("bench batched: sequential per-k warm-up")
    for c in ctxs:
        per_k(*c)
    seq_t = []
    for r in range(max(2, replays // 2)):
        t0 = time.perf_counter()
        for c in ctxs:
            per_k(*c)
        for c in ctxs:
            c[1].synchronize()
        seq_t.append(time.perf_counter() - t0)
        log(f"bench batched: sequential per-k apply {r + 1} {seq_t[-1]:.3f} s")
    seq_same = all(np.allclose(d_out.download(np.complex128), rf, rtol=0, atol=1e-6 * np.abs(rf).max())
                   for (_, _, _, _, d_out), rf in zip(ctxs, ref))
    for (sys_, eng, nb, d_phi, d_out), gid in zip(ctxs, gids):
        eng.graph_destroy(gid)
        d_phi.free()
        d_out.free()
        eng.close()
    q = lambda a, p: float(np.percentile(np.asarray(a), p))
    return {
        "workload": "3 fcc cells (32^3, 36^3, 48^3) x 32 k, bands per k uniform in [8,48] (seeded, all occupied), "
                    "hf2x + tc2x FP32-split",
        "bands_per_cell": [int(c[2].sum()) for c in ctxs], "pairs_per_apply": pairs,
        "graph": {"replays": replays, "p50_s": q(graph_t, 50), "p99_s": q(graph_t, 99),
                  "pairs_per_s": pairs / q(graph_t, 50), "replays_bit_identical_to_eager": bool(same)},
        "sequential_per_k": {"runs": len(seq_t), "p50_s": q(seq_t, 50), "p99_s": q(seq_t, 99),
                             "pairs_per_s": pairs / q(seq_t, 50), "matches_batched": bool(seq_same)},
        "speedup_p50": q(seq_t, 50) / q(graph_t, 50),
    }


def bench_davidson(eng, log):
    """One block-Davidson update of the 16 occupied bands of the HF 4x4x4 k, 40^3 cell
    (host pointers in and out, as TC++ calls it): wall time against one apply of
    calc_hamiltonian::all on the same bands; the difference over the two applies the
    Davidson makes is the cost of its device-resident FP64 subspace algebra."""
    from tc_amd import system as S
    from tc_amd import tcx

    sys_ = S.fcc_hf_40()
    eng.load_system(sys_)
    nb, phi = sys_.nocc.copy(), sys_.phi_occ
    h1, h2, _ = eng.apply_all(nb, phi, precision=tcx.PREC_FP32)  # warm-up (plans, workspaces)
    d_phi = eng.to_device(phi)
    d_h = [eng.device_buffer(phi.nbytes) for _ in range(2)]
    nb32 = np.ascontiguousarray(nb, np.int32)
    eng.synchronize()
    t0 = time.perf_counter()
    st = eng.L.tcx_apply_all(eng.h, tcx._ptr(nb32, tcx._ip), d_phi.ptr, d_h[0].ptr, d_h[1].ptr, None, tcx.PREC_FP32,
                             tcx.FLAG_DEVICE_PTRS)
    eng.synchronize()
    t_apply = time.perf_counter() - t0
    eng._check(st, "apply_all")
    for b in [d_phi] + d_h:
        b.free()
    bo = sys_.band_offsets(nb)
    eig = np.concatenate([np.einsum("bl,bl->b", phi[bo[k]:bo[k + 1]].reshape(nb[k], -1).conj(),
                                    (h1 + h2)[bo[k]:bo[k + 1]].reshape(nb[k], -1)) for k in range(sys_.nk)])
    log("bench: davidson warm-up call")
    phi1, _, eig1, en1 = eng.davidson(nb, phi, eig, precision=tcx.PREC_FP32)
    t0 = time.perf_counter()
    phi2, _, eig2, en2 = eng.davidson(nb, phi1, eig1, precision=tcx.PREC_FP32)
    t_dav = time.perf_counter() - t0
    log(f"bench: davidson {t_dav * 1e3:.1f} ms, apply_all {t_apply * 1e3:.1f} ms")
    return {
        "workload": "HF fcc 4x4x4 k, 16 bands/k, 40^3, max_blocks 2, num_refresh 1, FP32-split hf2x",
        "davidson_ms": t_dav * 1e3, "apply_all_device_ptrs_ms": t_apply * 1e3,
        "algebra_ms_est": (t_dav - 2 * t_apply) * 1e3,
        "sum_eig_Ha": float(np.sum(eig2.real)), "max_abs_eig_change_Ha": float(np.max(np.abs(eig2 - eig1))),
        "max_abs_eig_minus_energy_sum": float(np.max(np.abs(eig2 - en2.sum(axis=1)))),
    }


def bench_bitc_terms(eng, log, terms, precs, repeat=1):
    """BITC two- and three-body terms on the fcc 4x4x4 k, 16-band, 40^3 cell (A = 0.5, F from
    the cusp): FP32-split and FP64 applies per term on HBM-resident buffers (the last of `repeat`
    applies is timed), with the per-band error and the term energy (left orbitals) of FP32
    against FP64 when both ran."""
    from tc_amd import energy as E
    from tc_amd import system as S
    from tc_amd import tcx
    nbody_of = {"tc2x": 2, "tc3a1": 3, "tc3b1": 3, "tc3a2a4b2b5": 3, "tc3a3b3b4b6": 3}
    prec_of = {"fp32": tcx.PREC_FP32, "fp64": tcx.PREC_FP64}
    sys_ = S.fcc_bitc_40()
    nb = sys_.nocc.copy()
    phi = sys_.phi_occ
    eng.load_system(sys_)
    d_phi = eng.to_device(phi)
    d_out = eng.device_buffer(phi.nbytes)
    bo = sys_.band_offsets(nb)
    res = {"workload": "BITC fcc a=10.26, 4x4x4 k, 16 bands/k (all occupied), 40^3, RPA Jastrow A=0.5",
           "pairs_per_apply": int(np.sum(nb) * np.sum(sys_.nocc)), "timed_apply": repeat, "terms": {}}
    for term in terms:
        outs, secs = {}, {}
        for pname in precs:
            prec = prec_of[pname]
            for _ in range(repeat):
                d_out.zero()
                eng.synchronize()
                t0 = time.perf_counter()
                eng._term_device(term, nb, d_phi.ptr, d_out.ptr, prec)
                eng.synchronize()
                secs[pname] = time.perf_counter() - t0
            outs[pname] = d_out.download(np.complex128)
            log(f"bench bitc: {term} {pname} {secs[pname]:.2f} s")
        entry = {f"{p}_s": secs[p] for p in precs}
        if "fp32" in secs and term not in ("tc3a1", "tc3b1"):
            entry["fp32_pairs_per_s"] = res["pairs_per_apply"] / secs["fp32"]
        if "fp32" in outs and "fp64" in outs:
            a, b = outs["fp32"], outs["fp64"]
            rel = []
            for ik in range(sys_.nk):
                x = a[bo[ik]:bo[ik + 1]].reshape(nb[ik], -1)
                y = b[bo[ik]:bo[ik + 1]].reshape(nb[ik], -1)
                rel.extend(np.linalg.norm(x - y, axis=1) / np.maximum(np.linalg.norm(y, axis=1), 1e-300))
            e32 = E.term_energy(sys_, a, nbody_of[term], left=True)
            e64 = E.term_energy(sys_, b, nbody_of[term], left=True)
            entry.update({"fp32_over_fp64": secs["fp64"] / secs["fp32"],
                          "E_fp64_Ha": [float(e64.real), float(e64.imag)], "dE_Ha": float(abs(e32 - e64)),
                          "per_band_rel_max": float(np.max(rel))})
        res["terms"][term] = entry
    d_phi.free()
    d_out.free()
    return res


def bench_tc_legs(eng, log, tc_steps):
    """BITC cell of the TC workload (fcc a=10.26, 4x4x4 k, 16 bands, 40^3, RPA Jastrow A=0.5),
    inputs resident in HBM:
      tc2x: `tc_steps` timed FP32-split applies after one warm-up, with the HIP-event roofline
            of its dominant pass; one FP64 apply after a warm-up; energy and per-band errors;
      tc3a3b3b4b6, tc3a2a4b2b5: FP32-split first apply (occupied-set work included), repeated
            apply (occupied-set q-sums reused), FP64 first apply; energy errors.
    Algorithmic bytes per tc2x pair and pass (N points, 8 B per FP32 complex value): pass A
    reads chi_qi and writes 4 record channels (5 N 8 B), pass B reads and writes 4 channels
    (8 N 8 B), pass C reads 4 channels and the 4 grids of phi_qi and grad phi_qi (8 N 8 B)."""
    from tc_amd import energy as E
    from tc_amd import system as S
    from tc_amd import tcx
    sys_ = S.fcc_bitc_40()
    nb = sys_.nocc.copy()
    phi = sys_.phi_occ
    N = sys_.ngrid
    pairs = int(np.sum(nb) * np.sum(sys_.nocc))
    eng.load_system(sys_)
    d_phi = eng.to_device(phi)
    d_out = eng.device_buffer(phi.nbytes)
    bo = sys_.band_offsets(nb)

    def per_band(a, b):
        rel = []
        for ik in range(sys_.nk):
            x = a[bo[ik]:bo[ik + 1]].reshape(nb[ik], -1)
            y = b[bo[ik]:bo[ik + 1]].reshape(nb[ik], -1)
            rel.extend(np.linalg.norm(x - y, axis=1) / np.maximum(np.linalg.norm(y, axis=1), 1e-300))
        return float(np.max(rel))

    def apply(term, prec):
        d_out.zero()
        eng.synchronize()
        t0 = time.perf_counter()
        eng._term_device(term, nb, d_phi.ptr, d_out.ptr, prec)
        eng.synchronize()
        return time.perf_counter() - t0

    from tc_amd import autotune as AT
    res = {"workload": "BITC fcc a=10.26 bohr, 4x4x4 k, 16 bands/k (all occupied), FFT 40^3, RPA Jastrow A=0.5",
           "pairs_per_apply": pairs, "launch_shapes": {}}

    def tuned(term, prec):
        """The autotuner's shape for this term and precision (automatic if none)."""
        shape = AT.apply_tuned(eng, sys_, nb, term, prec)
        res["launch_shapes"][f"{term}_{'fp32' if prec == tcx.PREC_FP32 else 'fp64'}"] = \
            list(shape) if shape else "automatic"

    # ---- tc2x
    log("bench tc: tc2x FP32-split warm-up")
    tuned("tc2x", tcx.PREC_FP32)
    apply("tc2x", tcx.PREC_FP32)
    eng.enable_timing(True)
    t32 = []
    for i in range(tc_steps):
        t32.append(apply("tc2x", tcx.PREC_FP32))
        log(f"bench tc: tc2x FP32-split apply {i + 1}/{tc_steps} {t32[-1]:.3f} s")
    ms, launches, kpairs = eng.get_timing()
    eng.enable_timing(False)
    h32 = d_out.download(np.complex128)
    tuned("tc2x", tcx.PREC_FP64)
    apply("tc2x", tcx.PREC_FP64)
    t64 = apply("tc2x", tcx.PREC_FP64)
    log(f"bench tc: tc2x FP64 apply {t64:.3f} s")
    h64 = d_out.download(np.complex128)
    e32 = E.term_energy(sys_, h32, 2, left=True)
    e64 = E.term_energy(sys_, h64, 2, left=True)
    words = [5, 8, 8]   # grid values moved per pair in passes A, B, C
    dom = int(np.argmax(ms[:3]))
    avg_ms = ms[dom] / max(1, launches[dom])
    ppl = kpairs[dom] / max(1, launches[dom])
    bpl = words[dom] * N * 8 * ppl
    fp32_s = float(np.median(t32))
    res["tc2x"] = {
        "fp32_ms": fp32_s * 1e3, "fp64_ms": t64 * 1e3, "fp32_over_fp64": t64 / fp32_s,
        "fp32_pairs_per_s": pairs / fp32_s, "E_fp64_Ha": [float(e64.real), float(e64.imag)],
        "dE_Ha": float(abs(e32 - e64)), "per_band_rel_max": per_band(h32, h64),
        "pipeline_algorithmic_GBps_21NBc": 21 * N * 8 * pairs / fp32_s / 1e9,
        "roofline": {"bound": "hbm", "achieved": bpl / (avg_ms * 1e-3) / 1e9, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": bpl / (avg_ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
                     "traffic": read_traffic("tc2x_" + CLASS_NAMES[dom], ppl),
                     "kernel": "tc2x_" + CLASS_NAMES[dom], "avg_launch_ms": avg_ms, "pairs_per_launch": ppl,
                     "algorithmic_bytes_per_launch": bpl},
        "class_ms_per_apply": {CLASS_NAMES[c]: float(ms[c]) / tc_steps for c in range(4)},
    }
    # ---- three-body exchange terms
    res["three_body"] = {}
    for term in ("tc3a3b3b4b6", "tc3a2a4b2b5"):
        eng.load_system(sys_)             # drops the occupied-set caches: the first apply does all the work
        tuned(term, tcx.PREC_FP32)        # FP64 runs with the same pair sets
        first = apply(term, tcx.PREC_FP32)
        h32 = d_out.download(np.complex128)
        again = apply(term, tcx.PREC_FP32)
        same = bool(np.array_equal(d_out.download(np.complex128), h32))
        t64 = apply(term, tcx.PREC_FP64)
        h64 = d_out.download(np.complex128)
        e32 = E.term_energy(sys_, h32, 3, left=True)
        e64 = E.term_energy(sys_, h64, 3, left=True)
        log(f"bench tc: {term} FP32 first {first:.2f} s, repeated {again:.2f} s, FP64 first {t64:.2f} s")
        res["three_body"][term] = {
            "fp32_first_s": first, "fp32_repeated_s": again, "repeated_bit_identical": same,
            "fp64_first_s": t64, "fp32_over_fp64_first": t64 / first,
            "E_fp64_Ha": [float(e64.real), float(e64.imag)], "dE_Ha": float(abs(e32 - e64)),
            "per_band_rel_max": per_band(h32, h64)}
    d_phi.free()
    d_out.free()
    return res


if __name__ == "__main__":
    main()
