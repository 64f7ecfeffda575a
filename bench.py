"""Benchmark of the scGD0 hot path on one MI355X.

Workload (the single-GPU configuration the headline metric is quoted on): the synthetic cubic
three-band t2g model on a 24³ grid (Oh-irreducible outer states, ni = 911; inner states in the
±0.8 eV window, nf = 28152), 9 Einstein modes in 10-80 meV with seeded log-normal couplings
(g2 = 1.85 GB of FP64, resident in HBM), graded ωs of 1275 nodes, T = 100 K and 300 K at one
electron per cell, tol = 1e-6 eV.

One step = one complete scGD0 solve of every temperature to tol from Σ = -iη_init.  The value is
seconds per temperature to tol (lower is better).  The line also carries the Fan-Migdal
kernel's achieved HBM bandwidth and its roofline position (HIP events on the library's stream),
and the CPU oracle timed on a bounded sample of the same workload on this host.

    python bench.py [--steps K] [--warmup W] [--method direct|auto] [--batched] [--no-cpu]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "epspectral.jl_amd"))

HBM_PEAK_GBS = 8000.0
FP64_VALU_PEAK_TFLOPS = 78.6
FM_FLOP_PER_TERM = 60.0   # flop-equivalents of one Fan-Migdal term (2 lookups, 2 complex Im(1/z), 3 exp)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1)
    ap.add_argument("--warmup", type=int, default=0)
    ap.add_argument("--nk", type=int, default=24)
    ap.add_argument("--method", default="direct")
    ap.add_argument("--precision", default="fp64")
    ap.add_argument("--batched", action="store_true")
    ap.add_argument("--maxiter", type=int, default=100)
    ap.add_argument("--tol", type=float, default=1e-6)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (got --gpus {args.gpus}); the scGD0 path has no multi-GPU "
              "implementation in this build", file=sys.stderr)
        sys.exit(2)

    import epsx
    from epsx.models import t2g_setup

    t0 = time.time()
    setup, meta = t2g_setup(nk=args.nk)
    nm, ni, nf = setup.dims
    nT = len(meta["Tlist"])
    print(f"[bench] setup nk={args.nk} ni={ni} nf={nf} nm={nm} nω={len(setup.ωs)} nωa={len(setup.ωs_spectral)} "
          f"({time.time() - t0:.1f}s)", file=sys.stderr, flush=True)
    occ = epsx.Occupation(Tlist=meta["Tlist"], nlist=0.0, nelec=1.0, spin_degeneracy=2.0)
    ctx = epsx.Context(0)
    t0 = time.time()
    dev = epsx.Scgd0Device(setup, ctx)
    print(f"[bench] upload {time.time() - t0:.1f}s", file=sys.stderr, flush=True)
    kw = dict(η_min=0.0, η_init=5e-3, maxiter=args.maxiter, tol=args.tol, ω_acoustic=1e-6,
              method=args.method, precision=args.precision, batched=args.batched)
    import warnings
    warnings.simplefilter("ignore")
    # prime the kernels (one iteration; not a step)
    dev.loop(occ, **dict(kw, maxiter=1))
    for w in range(args.warmup):
        dev.loop(occ, **kw)
        print(f"[bench] warmup {w + 1}/{args.warmup} done", file=sys.stderr, flush=True)
    times, fm_s, fm_n, iters = [], 0.0, 0, []
    for k in range(args.steps):
        t0 = time.perf_counter()
        r = dev.loop(occ, **kw)
        times.append(time.perf_counter() - t0)
        tm = r["info"]["timings"]
        fm_s += tm["fan_migdal"]
        fm_n += tm["fan_migdal_launches"]
        iters.append(int(r["history"]["err"].shape[1]))
        print(f"[bench] step {k + 1}/{args.steps}: {times[-1]:.2f}s, iterations {iters[-1]}, "
              f"converged {r['converged'].tolist()}, μ {r['μ'].tolist()}, FM {tm['fan_migdal']:.2f}s",
              file=sys.stderr, flush=True)
    total = sum(times)
    s_per_T = total / (args.steps * nT)
    # Fan-Migdal kernel: one launch per temperature per iteration in the plain method
    fm_avg = fm_s / max(fm_n, 1)
    T_per_launch = nT if args.batched else 1
    g2_bytes = 8.0 * nm * ni * nf
    gbs = g2_bytes / fm_avg / 1e9
    terms = float(len(setup.ωs)) * nm * ni * nf * T_per_launch
    tflops = terms * FM_FLOP_PER_TERM / fm_avg / 1e12
    line = {
        "metric": "scGD0 s/temperature to tol vs plain FP64 GPU; Fan-Migdal HBM GB/s",
        "value": s_per_T,
        "unit": "s/temperature",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * total / args.steps,
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if args.precision == "fp64" else args.precision,
        "data": "synthetic",
        "config": {"workload": f"t2g {args.nk}^3 plain scGD0, 2 T, tol {args.tol} eV", "nk": args.nk, "ni": ni,
                   "nf": nf, "nmodes": nm, "nw": len(setup.ωs), "nwa": len(setup.ωs_spectral), "nT": nT,
                   "method": args.method, "precision": args.precision, "batched": bool(args.batched),
                   "iterations": iters},
        "roofline": {"bound": "hbm", "achieved": gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": gbs / HBM_PEAK_GBS, "traffic": None},
        "fm_kernel": {"avg_launch_s": fm_avg, "launches": fm_n, "terms_per_launch": terms,
                      "terms_per_s": terms / fm_avg, "valu_flop_equiv_tflops": tflops,
                      "valu_frac": tflops / FP64_VALU_PEAK_TFLOPS,
                      "algorithmic_bytes_per_launch": g2_bytes},
    }
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline(setup, meta, iters[0], args.cpu_seconds)
    print(json.dumps(line), flush=True)


def cpu_baseline(setup, meta, iterations, budget_s):
    """The oracle's Fan-Migdal sum (the reference algorithm restated in C, OpenMP over chunks of
    outer states) on a bounded sample: the first outer states of one temperature's first
    iteration, grown until the sample takes about budget_s; scaled to s/temperature to tol with
    the GPU run's iteration count."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle

    nm, ni, nf = setup.dims
    ws = np.asarray(setup.ωs)
    table = oracle.cell_search_table(ws, setup.ωs.dω_min)
    threads = min(os.cpu_count() or 1, 16)
    T, mu = float(meta["Tlist"][0]), float(setup.μ_start[0])
    Sin = np.full((len(ws), ni), -5e-3j, order="F")
    n_i = 1
    elapsed = 0.0
    while True:
        g2s = np.asfortranarray(setup.g2[:, :n_i, :])
        wq = np.asfortranarray(np.broadcast_to(setup.ωph[:, 0][:, None, None], (nm, nf, n_i)))
        ImS = np.zeros((len(ws), n_i), order="F")
        t0 = time.perf_counter()
        oracle.add_fan_migdal(ImS, g2s, wq, setup.ε_f, setup.f_to_i, setup.w_f, ws, table, Sin, mu, T, 1e-6,
                              nthreads=threads)
        elapsed = time.perf_counter() - t0
        if elapsed > budget_s / 4 or n_i * 4 > ni:
            break
        n_i *= 4
    terms = float(len(ws)) * nm * n_i * nf
    per_term = elapsed / terms
    s_per_T = per_term * len(ws) * nm * ni * nf * iterations
    return {"value": s_per_T, "unit": "s/temperature", "cores": threads, "kind": "port",
            "sample": f"oracle Fan-Migdal sum (C, OpenMP) over {n_i} of {ni} outer states x all {nf} inner states "
                      f"x {nm} modes x {len(ws)} frequencies at T = {meta['Tlist_K'][0]:g} K: {elapsed:.1f} s; "
                      f"scaled by terms to one temperature x {iterations} iterations (the Fan-Migdal sum is "
                      f">99% of the CPU iteration)",
            "terms_per_s": terms / elapsed}


if __name__ == "__main__":
    main()
