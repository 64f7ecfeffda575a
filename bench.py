"""Benchmark of the scGD0 hot path on one MI355X.

Workload (the single-GPU configuration the headline metric is quoted on): the synthetic cubic
three-band t2g model on a 24³ grid (Oh-irreducible outer states, ni = 911; inner states in the
±0.8 eV window, nf = 28152), 9 Einstein modes in 10-80 meV with seeded log-normal couplings
(g2 = 1.85 GB of FP64, resident in HBM), graded ωs of 1275 nodes, T = 100 K and 300 K at one
electron per cell, tol = 1e-6 eV.

One step = one complete scGD0 solve of every temperature to tol from Σ = -iη_init, by the
library's fastest exact FP64 method (method auto: the star-factorised Fan-Migdal sum when its
preconditions hold, all temperatures batched in a replayed hipGraph).  The value is seconds per
temperature to tol (lower is better).  Beside it:

* the plain FP64 GPU method (the reference's term-by-term Fan-Migdal sum, temperatures one after
  another) timed on a bounded sample of iterations and scaled to the same iteration count;
* the dominant kernel's roofline position (HIP events on the library's stream);
* the CPU oracle timed on a bounded sample of the same workload on this host.

    python bench.py [--steps K] [--warmup W] [--method auto|direct] [--no-batched] [--no-cpu]
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
FP64_PEAK_TFLOPS = 78.6   # FP64 vector and FP64 MFMA, dense
FP32_PEAK_TFLOPS = 157.3  # FP32 MFMA, dense
BF16_PEAK_TFLOPS = 2500.0 # BF16 MFMA, dense
FM_FLOP_PER_TERM = 60.0   # flop-equivalents of one Fan-Migdal term (2 lookups, 2 complex Im(1/z), 3 exp)
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "traffic_t2g24.json")


# the library's own FP64 GEMM tile shapes (epsx_gemm.hip SHAPES), by index
GEMM_SHAPE_NAMES = {0: "dgemm_buf_kernel 64x64", 1: "dgemm_buf_kernel 128x64", 2: "dgemm_buf_kernel 64x128",
                    3: "dgemm_buf_kernel 128x128", 4: "dgemm_buf_kernel 256x128", 5: "dgemm_buf_kernel 256x32",
                    6: "dgemm_buf_kernel 128x32", 7: "dgemm_reg_kernel 256x32 (1 wave along k)",
                    8: "dgemm_reg_kernel 256x32 (2 waves along k)",
                    9: "dgemm_pair_kernel 256x32 (16-byte loads, 4 x 2 waves)",
                    10: "dgemm_pair_kernel 256x32 (16-byte loads, 2 x 4 waves)",
                    11: "dgemm_pair_kernel 128x128 (16-byte loads, 2 x 2 waves)"}

def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--nk", type=int, default=24)
    ap.add_argument("--temps", default="100,300", help="temperatures in K")
    ap.add_argument("--method", default="auto")
    ap.add_argument("--precision", default="fp64")
    ap.add_argument("--no-batched", action="store_true", help="temperatures (scgd0) or Ωs (ladder) one by one")
    ap.add_argument("--maxiter", type=int, default=100)
    ap.add_argument("--tol", type=float, default=1e-6)
    ap.add_argument("--plain-iters", type=int, default=2, help="iterations of the plain method timed")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-configs", action="store_true",
                    help="skip the other configurations (C3 with the warm start, C4) measured beside the headline")
    ap.add_argument("--no-levels", action="store_true", help="skip the FP32 / BF16 precision-level runs")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--workload", default="scgd0", choices=("scgd0", "ladder", "warmstart", "northstar"),
                    help="scgd0: the headline solve; ladder: the Kubo ladder on the 16³ model; "
                         "warmstart: train the Σ_init network and evaluate it on held-out models")
    ap.add_argument("--models", type=int, default=64, help="warmstart: training models (12³)")
    ap.add_argument("--heldout", type=int, default=4, help="warmstart: held-out models (24³)")
    ap.add_argument("--epochs", type=int, default=40, help="warmstart: training epochs")
    ap.add_argument("--omegas", type=int, default=0, help="ladder: probe frequencies 0, 5, ... meV (default 41)")
    ap.add_argument("--coupling", default="iid", choices=("iid", "forward"),
                    help="ladder: i.i.d. couplings (the C2 model) or weighted by the momentum transfer")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="only run a 1 GiB device-to-device copy (the HBM counter calibration)")
    args = ap.parse_args()
    if args.calibrate_copy:
        return calibrate_copy()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (got --gpus {args.gpus}); the scGD0 path has no multi-GPU "
              "implementation in this build", file=sys.stderr)
        sys.exit(2)
    if args.workload == "ladder":
        return ladder_bench(args)
    if args.workload == "warmstart":
        return warmstart_bench(args)
    if args.workload == "northstar":
        return northstar_bench(args)

    import epsx
    from epsx.models import t2g_setup

    temps = tuple(float(x) for x in args.temps.split(","))
    batched = not args.no_batched
    t0 = time.time()
    setup, meta = t2g_setup(nk=args.nk, Tlist_K=temps)
    nm, ni, nf = setup.dims
    nw = len(setup.ωs)
    nT = len(meta["Tlist"])
    print(f"[bench] setup nk={args.nk} ni={ni} nf={nf} nm={nm} nω={nw} nωa={len(setup.ωs_spectral)} nT={nT} "
          f"({time.time() - t0:.1f}s)", file=sys.stderr, flush=True)
    occ = epsx.Occupation(Tlist=meta["Tlist"], nlist=0.0, nelec=1.0, spin_degeneracy=2.0)
    ctx = epsx.Context(0)
    t0 = time.time()
    dev = epsx.Scgd0Device(setup, ctx)
    print(f"[bench] upload {time.time() - t0:.1f}s", file=sys.stderr, flush=True)
    kw = dict(η_min=0.0, η_init=5e-3, maxiter=args.maxiter, tol=args.tol, ω_acoustic=1e-6,
              method=args.method, precision=args.precision, batched=batched)
    import warnings
    warnings.simplefilter("ignore")
    # prime the kernels and the GEMM tile-shape choice (three iterations; not a step)
    dev.loop(occ, **dict(kw, maxiter=3, tol=0.0))
    for w in range(args.warmup):
        dev.loop(occ, **kw)
        print(f"[bench] warmup {w + 1}/{args.warmup} done", file=sys.stderr, flush=True)
    times, iters, tsum = [], [], {}
    for k in range(args.steps):
        t0 = time.perf_counter()
        r = dev.loop(occ, **kw)
        times.append(time.perf_counter() - t0)
        tm = r["info"]["timings"]
        for key, v in tm.items():
            tsum[key] = tsum.get(key, 0) + v
        tsum["star_gemm_kernel"] = tm["star_gemm_kernel"]
        iters.append(int(r["history"]["err"].shape[1]))
        print(f"[bench] step {k + 1}/{args.steps}: {times[-1]:.3f}s, iterations {iters[-1]}, "
              f"converged {r['converged'].tolist()}, μ {r['μ'].tolist()}", file=sys.stderr, flush=True)
    method_used = r["info"]["fm_method"]
    total = sum(times)
    s_per_T = total / (args.steps * nT)
    g2_bytes = 8.0 * nm * ni * nf

    if method_used == 2:
        # dominant kernel: the GEMM ImΣ = h·G̃, nω x (nm ni) x ncol per temperature, ncol the
        # product's columns: one per multiplet (its members' ImΣ is the average the loop takes)
        # and one per other outer state (the HIP events time the second iteration of each run:
        # with a low-precision mode that is the BF16 / FP32 product, priced against its own peak);
        # one call covers every temperature of the batch (star_gemm_instances over the calls)
        ncol = ni - sum(len(m) for m in setup.multiplets) + len(setup.multiplets)
        per_launch = tsum["star_gemm_instances"] / max(tsum["star_gemm_launches"], 1)
        flops = 2.0 * nw * (nm * ni) * ncol * per_launch
        avg = tsum["star_gemm"] / max(tsum["star_gemm_launches"], 1)
        achieved = flops / avg / 1e12
        if args.precision == "fp64":
            # the kernel the library reports for the FP64 product (timing slot 7)
            kid = tsum.get("star_gemm_kernel", 0)
            peak = FP64_PEAK_TFLOPS
            if kid == 1:
                kname = "hipBLASLt FP64 GEMM (star Fan-Migdal product, strided batch with G̃ shared)"
                tr = traffic("hipblaslt_star", args.nk, nT)
            else:
                kname = (f"own FP64 MFMA GEMM {GEMM_SHAPE_NAMES.get(kid - 2, '?')} (star Fan-Migdal product, "
                         "v_mfma_f64_16x16x4f64, autotuned tile shape)")
                tr = traffic(GEMM_SHAPE_NAMES.get(kid - 2, "?").split()[0], args.nk, nT)
        else:
            # the kernel the library reports for the timed iteration's product (timing slot 7)
            kid = tsum.get("star_gemm_kernel", 0)
            peak = BF16_PEAK_TFLOPS if args.precision == "bf16" else FP32_PEAK_TFLOPS
            lp = args.precision.upper()
            kname = {1: f"hipBLASLt {lp} GEMM (star Fan-Migdal product from stored {lp} h and G̃, FP32 accumulation)",
                     30: f"lpkgemm_kernel (star Fan-Migdal product from stored {lp} operands, {lp} MFMA)",
                     31: f"lpgemm_kernel (star Fan-Migdal product, FP64 operands rounded to {lp} as staged)"}.get(
                         kid, f"own FP64 MFMA GEMM {GEMM_SHAPE_NAMES.get(kid - 2, '?')}")
            tr = None
        roof = {"bound": "mfma", "achieved": achieved, "peak": peak, "unit": "TFLOP/s",
                "frac": achieved / peak, "traffic": tr, "kernel": kname,
                "algorithmic_flop_per_launch": flops, "product_columns": ncol, "avg_launch_s": avg,
                "temperatures_per_launch": per_launch,
                "launches_timed": int(tsum["star_gemm_launches"])}
        sr = tsum["star_reduce"] / max(tsum["star_reduce_launches"], 1)
        sr_bytes = g2_bytes + 8.0 * nf + 8.0 * nm * ni * ni
        extra = {"g2_stream": {"kernel": "star_reduce_kernel (G̃ = Σ_star w_f g2, once per solve)",
                               "avg_launch_s": sr, "algorithmic_bytes": sr_bytes,
                               "achieved_GBs": sr_bytes / sr / 1e9, "frac_hbm": sr_bytes / sr / 1e9 / HBM_PEAK_GBS,
                               "traffic": traffic("star_reduce_kernel", args.nk, nT)}}
    elif method_used == 1:
        # the Einstein factorisation: per Fan-Migdal pass H' (nω x nm nf per temperature) is built and
        # multiplied by the resident g2 on the library's FP64 MFMA GEMM, 2 nω ni nm nf flop per
        # temperature; the pass time (HIP events) includes the H' builds
        fm_avg = tsum["fan_migdal"] / max(tsum["fan_migdal_launches"], 1)
        per_launch = tsum["star_gemm_instances"] / max(tsum["fan_migdal_launches"], 1)
        flops = 2.0 * nw * ni * nm * nf * per_launch
        achieved = flops / fm_avg / 1e12
        peak = {"fp64": FP64_PEAK_TFLOPS, "fp32": FP32_PEAK_TFLOPS, "bf16": BF16_PEAK_TFLOPS}[args.precision]
        roof = {"bound": "mfma", "achieved": achieved, "peak": peak, "unit": "TFLOP/s", "frac": achieved / peak,
                "traffic": None, "kernel": "Einstein Fan-Migdal pass: einstein_h_kernel + dgemm_buf_kernel (FP64 MFMA "
                "16x16x4, H' g2 batched over temperatures)", "algorithmic_flop_per_launch": flops,
                "avg_launch_s": fm_avg, "temperatures_per_launch": per_launch}
        extra = {}
    else:
        fm_avg = tsum["fan_migdal"] / max(tsum["fan_migdal_launches"], 1)
        gbs = g2_bytes / fm_avg / 1e9
        roof = {"bound": "hbm", "achieved": gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": gbs / HBM_PEAK_GBS,
                "traffic": traffic("fm_direct_kernel", args.nk, nT), "kernel": "fm_direct_kernel",
                "algorithmic_bytes_per_launch": g2_bytes, "avg_launch_s": fm_avg}
        extra = {}

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
        "config": {"workload": f"t2g {args.nk}^3 scGD0, {nT} T ({args.temps} K), tol {args.tol} eV", "nk": args.nk,
                   "ni": ni, "nf": nf, "nmodes": nm, "nw": nw, "nwa": len(setup.ωs_spectral), "nT": nT,
                   "method": ["direct", "einstein", "star"][method_used], "precision": args.precision,
                   "batched": batched, "iterations": iters},
        "roofline": roof,
    }
    line.update(extra)
    if args.precision != "fp64":
        # the error of this result against the all-FP64 solve of the same workload
        r64 = dev.loop(occ, **dict(kw, precision="fp64"))
        line["precision_error"] = {
            "vs": "FP64 solve of the same workload",
            "max_abs_dSigma_eV": float(np.max(np.abs(r["Σ"] - r64["Σ"]))),
            "max_rel_dSigma": float(np.max(np.abs(r["Σ"] - r64["Σ"])) / np.max(np.abs(r64["Σ"]))),
            "max_abs_dmu_eV": float(np.max(np.abs(r["μ"] - r64["μ"]))),
            "iterations_low_precision": int(r["info"]["n_low_precision"]),
            "iterations_total": iters[-1], "iterations_fp64_solve": int(r64["history"]["err"].shape[1])}
    if args.precision == "fp64" and not args.no_levels:
        # the precision levels on the same resident problem: BF16 / FP32 products, then FP64 to the
        # same tol (only the FP64 level freezes a temperature); errors against this FP64 result
        levels = {}
        for prec in ("fp32", "bf16"):
            kl = dict(kw, precision=prec)
            dev.loop(occ, **kl)   # warm (G̃ copies, shapes)
            tl, rl = [], None
            for k in range(2):
                t0 = time.perf_counter()
                rl = dev.loop(occ, **kl)
                tl.append(time.perf_counter() - t0)
            levels[prec] = {"value": sum(tl) / (2 * nT), "unit": "s/temperature",
                            "speedup_vs_fp64": s_per_T / (sum(tl) / (2 * nT)),
                            "iterations": int(rl["history"]["err"].shape[1]),
                            "iterations_below_fp64": int(rl["info"]["n_low_precision"]),
                            "converged": bool(rl["converged"].all()),
                            "max_abs_dSigma_eV": float(np.max(np.abs(rl["Σ"] - r["Σ"]))),
                            "max_rel_dSigma": float(np.max(np.abs(rl["Σ"] - r["Σ"])) / np.max(np.abs(r["Σ"]))),
                            "max_abs_dmu_eV": float(np.max(np.abs(rl["μ"] - r["μ"])))}
            print(f"[bench] precision {prec}: {levels[prec]}", file=sys.stderr, flush=True)
        line["precision_levels"] = levels
    if not args.no_plain:
        line["plain_fp64_gpu"] = plain_sample(dev, occ, kw, iters[0], args.plain_iters, nw * nm * ni * nf, g2_bytes)
        line["speedup_vs_plain"] = line["plain_fp64_gpu"]["value"] / s_per_T
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline(setup, meta, iters[0], args.cpu_seconds)
        print(f"[bench] cpu baseline: {line['cpu_baseline']}", file=sys.stderr, flush=True)
        line["c1_holstein"] = c1_pair(ctx)
        print(f"[bench] C1: {line['c1_holstein']}", file=sys.stderr, flush=True)
    if not args.no_configs:
        line["other_configs"] = other_configs(args, line, iters[0])
    print(json.dumps(line), flush=True)


def other_configs(args, line, iters_c2):
    """The other configurations, measured in the same run beside the headline (untimed by the
    headline's clock): C3 with the learned warm start (8 T, batched, precision levels; a network
    trained on 64 random 12³ models, 20 epochs) and C4 (the ladder, 41 Ω x 4 T, BF16 products,
    batched).  Compact summaries; `--workload northstar` / `--workload ladder` give the full lines."""
    import copy
    out = {}
    a = copy.copy(args)
    a.models, a.epochs, a.steps, a.maxiter, a.nk = 64, 20, 1, 300, 24
    plain_it = None
    if "plain_fp64_gpu" in line:
        plain_it = line["plain_fp64_gpu"]["value"] / iters_c2
    ns = northstar_bench(a, emit=False, plain_s_per_iter=plain_it)
    v = ns["variants"]
    out["c3_8T_levels_warm"] = {
        "workload": ns["config"]["workload"], "unit": "s/temperature",
        "fp64_star": ns["fp64_star_s_per_T"], "bf16_levels": v["bf16_levels"]["s_per_T"],
        "bf16_levels_warm": v["bf16_levels_warm"]["s_per_T"], "fp64_warm": v["fp64_warm"]["s_per_T"],
        "iterations": {"fp64": ns["fp64_iterations"], "bf16_levels": v["bf16_levels"]["iterations"],
                       "bf16_levels_warm": v["bf16_levels_warm"]["iterations"]},
        "max_abs_dSigma_eV_vs_fp64": v["bf16_levels_warm"]["max_abs_dSigma_eV"],
        "max_rel_dSigma_vs_fp64": v["bf16_levels_warm"]["max_rel_dSigma"],
        "plain_fp64_gpu": ns["plain_fp64_gpu"]["value"],
        "speedup_vs_plain_fp64_gpu": v["bf16_levels_warm"]["speedup_vs_plain_fp64_gpu"],
        "warm_kept": v["bf16_levels_warm"].get("warm_kept")}
    print(f"[bench] C3 + warm start: {out['c3_8T_levels_warm']}", file=sys.stderr, flush=True)
    b = copy.copy(args)
    # the forward-scattering coupling gives the ladder a vertex correction and GMRES several
    # iterations per (Ω, T) (the i.i.d. coupling converges in one)
    b.nk, b.precision, b.omegas, b.coupling, b.no_cpu, b.no_batched = 24, "bf16", 0, "forward", True, False
    ld = ladder_bench(b, emit=False)
    out["c4_ladder"] = {"workload": ld["config"]["workload"] + f", {ld['config']['coupling']} coupling",
                        "value": ld["value"], "unit": ld["unit"], "total_s": ld["ms_per_step"] / 1e3,
                        "gmres_iterations": {"total": ld["config"]["niter_total"], "max": ld["config"]["niter_max"]},
                        "vertex_correction_max_rel": ld["vertex_correction_max_rel"],
                        "gemm_kernel": ld["roofline"]["kernel"], "gemm_tflops": ld["roofline"]["achieved"],
                        "gemm_frac": ld["roofline"]["frac"], "error_vs_fp64_rtol1e-8": ld.get("error_vs_fp64_rtol1e-8")}
    print(f"[bench] C4: {out['c4_ladder']}", file=sys.stderr, flush=True)
    return out


def ladder_bench(args, emit=True):
    """C4: the Kubo ladder on the converged scGD0 result of the t2g model at 16³ (ni = 331,
    nf = 8448, nshift = 82, Oh vertex rotations): scGD0 at 100, 200, 300, 400 K to 1e-6 eV by the
    star method, then GMRES at rtol 1e-2 for all 41 probe frequencies 0:5:200 meV at all four
    temperatures, the Ωs of each temperature batched as instances of one solve (`--no-batched`: the
    reference's loop, each Ω warm-started from the previous one).  The value is seconds per (Ω, T)
    solve over the whole run; the roofline line is the shift-table GEMM W = J·B̃, 2·(16 nω)·(3 ni)·
    (6 ni·nshift) flop per instance and application.  σ is checked against an FP64 GMRES at rtol
    1e-8 at three probe frequencies.  `--coupling forward` weights the couplings by the momentum
    transfer (forward scattering), which gives the ladder a vertex correction."""
    import warnings
    import epsx
    from epsx.ladder import LadderDevice, setup_ladder
    from epsx.models import t2g_ladder_setup

    nk = 16 if args.nk == 24 else args.nk
    temps = (100.0, 200.0, 300.0, 400.0)
    t0 = time.time()
    fq = nk / 8.0 if args.coupling == "forward" else None
    setup, R, v, meta = t2g_ladder_setup(nk=nk, Tlist_K=temps, forward_q=fq)
    nm, ni, nf = setup.dims
    nw = len(setup.ωs)
    print(f"[bench] ladder setup nk={nk} ni={ni} nf={nf} nω={nw} coupling={args.coupling} ({time.time() - t0:.1f}s)",
          file=sys.stderr, flush=True)
    ctx = epsx.Context(0)
    dev = epsx.Scgd0Device(setup, ctx)
    occ = epsx.Occupation(Tlist=meta["Tlist"], nlist=0.0, nelec=1.0, spin_degeneracy=2.0)
    warnings.simplefilter("ignore")
    t0 = time.perf_counter()
    r = dev.loop(occ, η_min=0.0, η_init=5e-3, maxiter=200, tol=1e-6, ω_acoustic=1e-6, method="auto", batched=True)
    t_scgd0 = time.perf_counter() - t0
    print(f"[bench] scGD0 4 T: {t_scgd0:.2f}s, converged {r['converged'].tolist()}", file=sys.stderr, flush=True)
    L = setup_ladder(setup, R, v)
    ld = LadderDevice(dev, L)
    nO = args.omegas if args.omegas else 41
    Ωs = 0.005 * np.arange(nO)
    batched = not args.no_batched
    Tl = meta["Tlist"]
    # prime: B̃ of the first temperature and one solve in the timed precision (loads the GEMM code
    # objects and times the library's candidates)
    ld.loop(r["Σ"][:, :, :1], r["μ"][:1], Tl[:1], Ωs[:1], rtol=1e-2, maxiter=30, precision=args.precision,
            batched=batched)
    t0 = time.perf_counter()
    out = ld.loop(r["Σ"], r["μ"], Tl, Ωs, rtol=1e-2, maxiter=30, precision=args.precision, batched=batched)
    el = time.perf_counter() - t0
    tm = ctx.ladder_timings()
    print(f"[bench] ladder {nO} Ω x {len(Tl)} T: {el:.2f}s, niter {out['niter'].ravel().tolist()}",
          file=sys.stderr, flush=True)
    # the FP64 GMRES at rtol 1e-8 (maxiter 60) at three probe frequencies of the first temperature
    sub = sorted(set([0, nO // 2, nO - 1]))
    ref = ld.loop(r["Σ"][:, :, :1], r["μ"][:1], Tl[:1], Ωs[sub], rtol=1e-8, maxiter=60, batched=True)

    def relerr(a, b):
        return float(np.max(np.abs(a - b)) / np.max(np.abs(b)))
    ns = L.nshift
    flops = 2.0 * (16 * nw) * (3 * ni) * (6 * ni * ns)
    inst_apps = max(tm["instance_applications"], 1)
    achieved = flops * inst_apps / tm["shift_gemm"] / 1e12 if tm["shift_gemm"] > 0 else 0.0
    peak = {"fp64": FP64_PEAK_TFLOPS, "fp32": FP32_PEAK_TFLOPS, "bf16": BF16_PEAK_TFLOPS}[args.precision]
    nsolve = nO * len(Tl)
    line = {
        "metric": "ladder s per (Ω, T) solve at rtol 1e-2; shift-table GEMM TFLOP/s",
        "value": el / nsolve, "unit": "s/(Ω,T)", "n_gpus": 1, "steps": 1, "warmup": 1,
        "ms_per_step": 1e3 * el, "higher_is_better": False, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64" if args.precision == "fp64" else f"{args.precision} shift GEMM, f64 GMRES", "data": "synthetic",
        "config": {"workload": f"t2g {nk}^3 ladder, GMRES rtol 1e-2, Ω = 0..{5 * (nO - 1)} meV ({nO}) x 4 T "
                               f"(100-400 K), {'batched over Ω' if batched else 'sequential, warm-started'}",
                   "nk": nk, "ni": ni, "nf": nf, "nw": nw, "nshift": ns, "coupling": args.coupling,
                   "niter_total": int(out["niter"].sum()), "niter_max": int(out["niter"].max()),
                   "scgd0_s_4T": t_scgd0, "instance_applications": tm["instance_applications"],
                   "batched_applications": tm["applications"]},
        "roofline": {"bound": "mfma", "achieved": achieved, "peak": peak, "unit": "TFLOP/s",
                     "frac": achieved / peak, "traffic": None,
                     "kernel": ladder_gemm_label(args.precision),
                     "algorithmic_flop_per_instance_application": flops,
                     "gemm_s_total": tm["shift_gemm"], "apply_s_total": tm["apply"]},
        "sigma_ladder_xx_T100": out["σ_ladder"][0, 0, :, 0].tolist(),
        "sigma_bubble_xx_T100": out["σ_bubble"][0, 0, :, 0].tolist(),
        "vertex_correction_max_rel": relerr(out["σ_ladder"], out["σ_bubble"]),
        "error_vs_fp64_rtol1e-8": {"omegas_eV": Ωs[sub].tolist(),
                                   "sigma_ladder_max_rel": relerr(out["σ_ladder"][:, :, sub, :1], ref["σ_ladder"]),
                                   "sigma_bubble_max_rel": relerr(out["σ_bubble"][:, :, sub, :1], ref["σ_bubble"]),
                                   "reference_niter": ref["niter"][:, 0].tolist()},
    }
    if not args.no_cpu:
        apps_per_solve = inst_apps / nsolve
        line["cpu_baseline"] = ladder_cpu_baseline(nw, ni, ns, apps_per_solve, args.cpu_seconds)
    if emit:
        print(json.dumps(line), flush=True)
    return line


def ladder_gemm_label(precision):
    """The kernel that runs the shift-table GEMM W = J·B̃: rocBLAS (a plain GEMM), or the library's
    own MFMA kernels with EPSX_LADDER_BLAS=0."""
    if os.environ.get("EPSX_LADDER_BLAS", "1") == "0":
        return ("own FP64 MFMA GEMM, autotuned tile shape: dgemm_buf_kernel or dgemm_reg_kernel "
                "(shift-table GEMM W = J B̃, v_mfma_f64_16x16x4f64)" if precision == "fp64"
                else f"lpkgemm_kernel (shift-table GEMM W = J B̃ from stored {precision} operands, FP32 accumulation)")
    return {"fp64": "rocBLAS dgemm (shift-table GEMM W = J B̃, every shift of a column block)",
            "fp32": "rocBLAS gemm_ex FP32 (shift-table GEMM W = J B̃ from stored FP32 operands)",
            "bf16": "rocBLAS gemm_ex BF16 (shift-table GEMM W = J B̃ from stored BF16 operands, FP32 "
                    "accumulation)"}[precision]


def ladder_cpu_baseline(nw, ni, ns, apps_per_solve, budget_s):
    """The oracle's apply_ladder_kernel! (the reference algorithm restated in C, OpenMP) timed once on
    the same model at 4³, 6³ or 8³ (the same grid and shifts; the first whose application takes at
    least budget_s / 4, or 8³), scaled by the multiply-adds of its dominant product C± += X B̃
    (2 x 16 nω x 6 ni x 3 nshift ni) to the measured size, times the applications per (Ω, T) solve
    of the GPU run."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    from epsx.ladder import setup_ladder
    from epsx.models import t2g_ladder_setup

    threads = cpu_threads()
    rng = np.random.default_rng(5)
    for nk in (4, 6, 8):
        setup, R, v, meta = t2g_ladder_setup(nk=nk, Tlist_K=(100.0,))
        L = setup_ladder(setup, R, v)
        nw_s, ni_s = len(setup.ωs), setup.dims[1]
        shape = (nw_s, 4, 3, 2, ni_s)
        J = np.asfortranarray(rng.standard_normal(shape) + 1j * rng.standard_normal(shape))
        G = np.asfortranarray(rng.standard_normal((nw_s, 4, ni_s)) + 1j * rng.standard_normal((nw_s, 4, ni_s)))
        t0 = time.perf_counter()
        oracle.apply_ladder_kernel(L, J, G, G, float(meta["Tlist"][0]), 1e-6, nthreads=threads)
        el = time.perf_counter() - t0
        if el >= budget_s / 4:
            break
    macs_s = 2.0 * 16 * nw_s * 6 * ni_s * 3 * L.nshift * ni_s
    macs = 2.0 * 16 * nw * 6 * ni * 3 * ns * ni
    per_app = el * macs / macs_s
    return {"value": per_app * apps_per_solve, "unit": "s/(Ω,T)", "cores": threads, "kind": "port",
            "sample": f"oracle apply_ladder_kernel (C, OpenMP) on the {nk}³ model (ni = {ni_s}, nω = {nw_s}, "
                      f"nshift = {L.nshift}): {el:.1f} s; scaled by the multiply-adds of C± += X B̃ to ni = {ni} "
                      f"({per_app:.0f} s per application) x {apps_per_solve:g} applications per solve",
            "macs_per_s": macs_s / el}


def warmstart_bench(args):
    """The learned Σ_init warm start: a training set from scGD0 runs of `--models` seeded random
    three-band models at 12³ on 1024 uniform nodes (G0D0 -> converged pairs), the 6-layer,
    64-channel network trained in BF16 on the device, then `--heldout` held-out random models at
    24³ run to 1e-6 eV from the plain start and from the warm start.  The value is the wall time
    of the warm-started solve relative to the plain one (lower is better)."""
    import warnings
    import epsx
    from epsx.learn import WarmStartNet, make_training_set, random_t2g_model, run_warm_started

    ctx = epsx.Context(0)
    L = 1024
    warnings.simplefilter("ignore")
    t0 = time.perf_counter()
    x, t = make_training_set(ctx, args.models, nk=12, length=L, seed=2024, states_per_model=96, verbosity=1)
    t_data = time.perf_counter() - t0
    print(f"[bench] training set: {len(x)} samples from {args.models} models ({t_data:.1f}s)", file=sys.stderr, flush=True)
    net = WarmStartNet(ctx, L, seed=0)
    t0 = time.perf_counter()
    losses = net.train(x, t, epochs=args.epochs, batch=64, lr=1e-3, precision="bf16", seed=0, verbosity=1)
    t_train = time.perf_counter() - t0
    print(f"[bench] trained {args.epochs} epochs in {t_train:.1f}s, loss {losses[0]:.3e} -> {losses[-1]:.3e}",
          file=sys.stderr, flush=True)
    rng = np.random.default_rng(777)
    rows = []
    for h in range(args.heldout):
        setup, meta = random_t2g_model(rng, nk=args.nk, length=L)
        dev = epsx.Scgd0Device(setup, ctx)
        occ = epsx.Occupation(Tlist=meta["Tlist"], nlist=0.0, nelec=1.0)
        kw = dict(η_min=0.0, η_init=5e-3, ω_acoustic=1e-6, method="auto", batched=True)
        try:
            dev.loop(occ, maxiter=3, tol=0.0, **kw)  # prime (kernels and the GEMM tile-shape choice)
        except epsx.ArgumentError:
            print(f"[bench] held-out {h + 1}: no μ root for this random model, skipped", file=sys.stderr, flush=True)
            continue
        t0 = time.perf_counter()
        rc = dev.loop(occ, maxiter=300, tol=1e-6, **kw)
        t_plain = time.perf_counter() - t0
        run_warm_started(dev, setup, occ, net, maxiter=3, tol=0.0)   # prime the warm-start path
        w = run_warm_started(dev, setup, occ, net)
        dΣ = float(np.max(np.abs(w["Σ"] - rc["Σ"])))
        rows.append(dict(plain_iterations=int(rc["history"]["err"].shape[1]), warm_iterations=int(w["iterations"]),
                         fallback=bool(w["fallback"][0]), plain_s=t_plain, warm_s=w["wall"], max_dSigma=dΣ,
                         T_K=float(meta["Tlist_K"][0]), converged=bool(w["converged"].all() and rc["converged"].all())))
        print(f"[bench] held-out {h + 1}/{args.heldout}: {rows[-1]}", file=sys.stderr, flush=True)
    it_p = sum(r["plain_iterations"] for r in rows)
    it_w = sum(r["warm_iterations"] for r in rows)
    s_p = sum(r["plain_s"] for r in rows)
    s_w = sum(r["warm_s"] for r in rows)
    line = {
        "metric": "warm-started / plain scGD0 wall time to tol on held-out models",
        "value": s_w / s_p, "unit": "ratio", "n_gpus": 1, "steps": 1, "warmup": 0, "ms_per_step": 1e3 * s_w,
        "higher_is_better": False, "scaling": "weak", "vs_baseline": None, "dtype": "bf16 net, f64 scGD0",
        "data": "synthetic",
        "config": {"workload": f"warm start: {args.models} training models at 12^3, {args.heldout} held-out at "
                               f"{args.nk}^3, {L} uniform nodes, tol 1e-6 eV",
                   "samples": int(len(x)), "epochs": args.epochs, "params": net.nparam,
                   "train_s": t_train, "data_s": t_data, "loss_first": float(losses[0]), "loss_last": float(losses[-1])},
        "iterations_plain": it_p, "iterations_warm": it_w, "iterations_saved_frac": 1 - it_w / it_p,
        "heldout": rows,
    }
    print(json.dumps(line), flush=True)


def northstar_bench(args, emit=True, plain_s_per_iter=None):
    """The combined configuration: the 24³ t2g model at 8 temperatures (50, 100, ..., 400 K) to
    tol = 1e-6 eV in one batched run with the BF16 / FP32 precision levels and the learned warm
    start (the 6-layer network trained on `--models` random 12³ models on the same graded grid),
    against the plain FP64 GPU method (the direct kernel, temperatures one by one, `--plain-iters`
    iterations timed and scaled to the FP64 run's iterations) and the FP64 batched star solve,
    whose Σ and μ it is compared with.  The value is seconds per temperature to tol."""
    import warnings
    import epsx
    from epsx.learn import WarmStartNet, make_training_set
    from epsx.models import t2g_setup

    warnings.simplefilter("ignore")
    ctx = epsx.Context(0)
    temps = tuple(50.0 * (k + 1) for k in range(8))
    t0 = time.perf_counter()
    setup, meta = t2g_setup(nk=args.nk, Tlist_K=temps)
    nm, ni, nf = setup.dims
    nw = len(setup.ωs)
    print(f"[bench] setup nk={args.nk} ni={ni} nω={nw} nT={len(temps)} ({time.perf_counter() - t0:.1f}s)",
          file=sys.stderr, flush=True)
    t0 = time.perf_counter()
    x, t = make_training_set(ctx, args.models, nk=12, length=None, seed=2024, states_per_model=96, verbosity=1)
    t_data = time.perf_counter() - t0
    net = WarmStartNet(ctx, nw, seed=0)
    t0 = time.perf_counter()
    losses = net.train(x, t, epochs=args.epochs, batch=64, lr=1e-3, precision="bf16", seed=0, verbosity=1)
    t_train = time.perf_counter() - t0
    print(f"[bench] network: {len(x)} samples in {t_data:.1f}s, {args.epochs} epochs in {t_train:.1f}s, "
          f"loss {losses[0]:.3e} -> {losses[-1]:.3e}", file=sys.stderr, flush=True)
    dev = epsx.Scgd0Device(setup, ctx)
    occ = epsx.Occupation(Tlist=meta["Tlist"], nlist=0.0, nelec=1.0, spin_degeneracy=2.0)
    nT = len(occ)
    kw = dict(η_min=0.0, η_init=5e-3, maxiter=args.maxiter, tol=args.tol, ω_acoustic=1e-6, method="auto",
              batched=True)
    variants = {"fp64": dict(precision="fp64"), "fp64_warm": dict(precision="fp64", warm_start="bf16"),
                "bf16_levels": dict(precision="bf16"), "bf16_levels_warm": dict(precision="bf16", warm_start="bf16")}
    res, secs = {}, {}
    for name, extra in variants.items():
        dev.loop(occ, **dict(kw, maxiter=3, tol=0.0, **extra))   # prime (kernels, tile shapes, library candidates)
        ts = []
        for k in range(args.steps):
            t0 = time.perf_counter()
            r = dev.loop(occ, **kw, **extra)
            ts.append(time.perf_counter() - t0)
        res[name], secs[name] = r, sum(ts) / (len(ts) * nT)
        wl = r["history"]["wall"]
        print(f"[bench] {name}: {secs[name]:.4f} s/T, iterations {r['history']['err'].shape[1]}, "
              f"converged {r['converged'].all()}, first iterations {np.round(1e3 * wl[:6], 2).tolist()} ms, "
              f"the rest {1e3 * wl[6:].sum():.1f} ms", file=sys.stderr, flush=True)
    ref = res["fp64"]
    iters64 = int(ref["history"]["err"].shape[1])
    if plain_s_per_iter is None:
        plain = plain_sample(dev, occ, kw, iters64, args.plain_iters, nw * nm * ni * nf, 8.0 * nm * ni * nf)
    else:   # the per-iteration time of a plain sample on the same model, scaled to this run's iterations
        plain = {"value": plain_s_per_iter * iters64, "unit": "s/temperature",
                 "sample": f"the headline run's plain-method sample ({plain_s_per_iter:.3f} s per iteration and "
                           f"temperature), scaled to {iters64} iterations"}
    print(f"[bench] plain FP64 GPU: {plain['value']:.2f} s/T", file=sys.stderr, flush=True)
    rows = {}
    for name in ("fp64_warm", "bf16_levels", "bf16_levels_warm"):
        r = res[name]
        dS = np.abs(r["Σ"] - ref["Σ"])
        rows[name] = {"s_per_T": secs[name], "iterations": int(r["history"]["err"].shape[1]),
                      "converged_all": bool(r["converged"].all()),
                      "max_abs_dSigma_eV": float(dS.max()), "max_rel_dSigma": float(dS.max() / np.abs(ref["Σ"]).max()),
                      "max_abs_dmu_eV": float(np.abs(r["μ"] - ref["μ"]).max()),
                      "speedup_vs_plain_fp64_gpu": plain["value"] / secs[name],
                      "speedup_vs_fp64_star": secs["fp64"] / secs[name]}
        if "warm" in name:
            rows[name]["warm_kept"] = r["info"]["warm_kept"].tolist()
        rows[name]["iter_frozen"] = r["history"]["iter_frozen"].tolist()
    v = rows["bf16_levels_warm"]
    line = {
        "metric": "scGD0 s/temperature to tol: 8 T batched, BF16 levels + learned warm start, vs plain FP64 GPU",
        "value": v["s_per_T"], "unit": "s/temperature", "n_gpus": 1, "steps": args.steps, "warmup": 1,
        "ms_per_step": 1e3 * v["s_per_T"] * nT, "higher_is_better": False, "scaling": "weak", "vs_baseline": None,
        "dtype": "bf16/fp32 levels then f64", "data": "synthetic",
        "config": {"workload": f"t2g {args.nk}^3, 8 T (50-400 K), tol {args.tol} eV, batched, BF16 levels, "
                               f"warm start trained on {args.models} random 12^3 models",
                   "nomega": nw, "ni": ni, "nf": nf, "samples": int(len(x)), "epochs": args.epochs,
                   "loss_first": float(losses[0]), "loss_last": float(losses[-1]), "train_s": t_train},
        "fp64_star_s_per_T": secs["fp64"], "fp64_iterations": iters64, "plain_fp64_gpu": plain,
        "variants": rows,
    }
    if emit:
        print(json.dumps(line), flush=True)
    return line


def calibrate_copy():
    """A 1 GiB device-to-device copy, three times: reads 1 GiB and writes 1 GiB per launch.  The
    FETCH_SIZE and WRITE_SIZE counters of this run give the factor between counter readings and
    HBM bytes."""
    import torch
    n = 1 << 27
    a = torch.arange(n, dtype=torch.float64, device="cuda:0")
    b = torch.empty_like(a)
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    assert torch.equal(a, b)
    print(json.dumps({"copy_bytes_read": 8 * n, "copy_bytes_written": 8 * n, "launches": 3}), flush=True)


def traffic(kernel, nk, nT):
    """HBM bytes per launch of `kernel` at this workload, from the rocprofv3 FETCH_SIZE and
    WRITE_SIZE passes recorded under profiles/ (None when that file has no entry)."""
    try:
        with open(TRAFFIC_FILE) as f:
            tab = json.load(f)
        e = tab[kernel]
        if e["nk"] == nk and e["nT"] == nT:
            return e["bytes_per_launch"]
    except (OSError, KeyError, ValueError):
        pass
    return None


def plain_sample(dev, occ, kw, iterations, n_iter, terms_per_T, g2_bytes):
    """The plain FP64 GPU method: the direct Fan-Migdal kernel, temperatures one after another,
    the loop driven from the host.  n_iter iterations are timed and the time per iteration is
    scaled to the optimised run's iteration count (the methods are exact rewrites of one
    another and take the same iterations)."""
    nT = len(occ.Tlist)
    t0 = time.perf_counter()
    r = dev.loop(occ, **dict(kw, method="direct", batched=False, maxiter=n_iter, tol=0.0))
    el = time.perf_counter() - t0
    tm = r["info"]["timings"]
    fm_avg = tm["fan_migdal"] / max(tm["fan_migdal_launches"], 1)
    tflops = terms_per_T * FM_FLOP_PER_TERM / fm_avg / 1e12
    return {"value": el / (n_iter * nT) * iterations, "unit": "s/temperature",
            "sample": f"{n_iter} iterations x {nT} T in {el:.2f} s, scaled to {iterations} iterations",
            "fm_direct_kernel": {"avg_launch_s": fm_avg, "terms_per_launch": terms_per_T,
                                 "terms_per_s": terms_per_T / fm_avg, "valu_flop_equiv_tflops": tflops,
                                 "valu_frac": tflops / FP64_PEAK_TFLOPS, "g2_GBs": g2_bytes / fm_avg / 1e9}}


def cpu_threads():
    """Host threads for the oracle: the box's CPU share (16 for one GPU), or fewer if fewer exist."""
    return max(1, min(os.cpu_count() or 1, 16))


def cpu_baseline(setup, meta, iterations, budget_s):
    """The oracle's Fan-Migdal sum (the reference algorithm restated in C; OpenMP over 2 x threads
    chunks of outer states, as src/scgd0_kernel.jl:62-66) on a bounded sample: 2 x threads outer
    states, so that every thread has a chunk, against the first nf_s inner states, nf_s grown until
    the sample takes about budget_s / 4 or more; scaled by terms to s/temperature to tol with the GPU
    run's iteration count (the cost of a term does not depend on which states it pairs)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle

    nm, ni, nf = setup.dims
    ws = np.asarray(setup.ωs)
    table = oracle.cell_search_table(ws, setup.ωs.dω_min)
    threads = cpu_threads()
    T, mu = float(meta["Tlist"][0]), float(setup.μ_start[0])
    Sin = np.full((len(ws), ni), -5e-3j, order="F")
    n_i = min(2 * threads, ni)
    nf_s = min(256, nf)
    while True:
        g2s = np.asfortranarray(setup.g2[:, :n_i, :nf_s])
        wq = np.asfortranarray(np.broadcast_to(setup.ωph[:, 0][:, None, None], (nm, nf_s, n_i)))
        ImS = np.zeros((len(ws), n_i), order="F")
        t0 = time.perf_counter()
        oracle.add_fan_migdal(ImS, g2s, wq, setup.ε_f[:nf_s], setup.f_to_i[:nf_s], setup.w_f[:nf_s], ws, table, Sin,
                              mu, T, 1e-6, nthreads=threads)
        elapsed = time.perf_counter() - t0
        if elapsed > budget_s / 4 or nf_s >= nf:
            break
        nf_s = min(nf, 4 * nf_s)
    terms = float(len(ws)) * nm * n_i * nf_s
    per_term = elapsed / terms
    s_per_T = per_term * len(ws) * nm * ni * nf * iterations
    return {"value": s_per_T, "unit": "s/temperature", "cores": min(threads, n_i), "kind": "port",
            "sample": f"oracle Fan-Migdal sum (C, OpenMP, {threads} threads over {2 * threads} chunks of outer states) "
                      f"over {n_i} of {ni} outer states x {nf_s} of {nf} inner states x {nm} modes x {len(ws)} "
                      f"frequencies at T = {meta['Tlist_K'][0]:g} K: {elapsed:.1f} s; scaled by terms to one "
                      f"temperature x {iterations} iterations (the Fan-Migdal sum is >99% of the CPU iteration)",
            "terms_per_s": terms / elapsed}


def c1_pair(ctx):
    """The 1D Holstein chain of the reference's CPU-runnable case (nk = 128, ωs = -4:0.005:4, one
    mode ω0 = 0.2 with g0 = 0.4, η_init = 0.01, T = 0.1 and 0.3 at half filling, tol 1e-8) run to tol
    in full on the GPU (FP64, direct and star methods) and by the oracle on the host cores."""
    import warnings
    import epsx
    from epsx.grids import FrequencyGrid
    from epsx.models import bare_mu, holstein_1d_setup, holstein_energies
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle

    nk, T = 128, [0.1, 0.3]
    ωs = FrequencyGrid(-4.0, 0.005, np.arange(1601))
    εs = holstein_energies(nk, 1.0)
    μ0 = [bare_mu(εs, np.full(nk, 1.0 / nk), Ti, 1.0) for Ti in T]
    s = holstein_1d_setup(nk, 1.0, 0.2, 0.4 ** 2 / (2 * 0.2), ωs, μ0)
    occ = epsx.Occupation(Tlist=T, nlist=0.0, nelec=1.0)
    kw = dict(η_min=0.0, η_init=0.01, maxiter=100, tol=1e-8, ω_acoustic=1e-6)
    dev = epsx.Scgd0Device(s, ctx)
    out = {"workload": "Holstein 1D nk=128, 1601 nodes, T = 0.1, 0.3 t, tol 1e-8 t"}
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        for method, batched in (("direct", False), ("auto", True)):
            dev.loop(occ, **dict(kw, method=method, batched=batched, maxiter=2, tol=0.0))
            t0 = time.perf_counter()
            r = dev.loop(occ, **dict(kw, method=method, batched=batched))
            el = time.perf_counter() - t0
            out[f"gpu_{method}"] = {"s_per_T": el / 2, "iterations": r["history"]["iter_frozen"].tolist(),
                                    "converged": bool(r["converged"].all())}
    threads = cpu_threads()
    t0 = time.perf_counter()
    ref = oracle.loop_scgd0(s, T, 0.0, 1.0, nthreads=threads, **kw)
    el = time.perf_counter() - t0
    out["cpu_oracle"] = {"s_per_T": el / 2, "cores": threads, "kind": "port",
                         "iterations": ref["history"]["iter_frozen"].tolist()}
    out["max_rel_dSigma_gpu_vs_oracle"] = float(np.max(np.abs(r["Σ"] - ref["Σ"])) / np.max(np.abs(ref["Σ"])))
    out["speedup_gpu_auto_vs_cpu"] = out["cpu_oracle"]["s_per_T"] / out["gpu_auto"]["s_per_T"]
    return out


if __name__ == "__main__":
    main()
