#!/usr/bin/env python3
"""Benchmark of the AMOEBA electrostatics hot path on one MI355X.

Workload: the 96,000-atom AMOEBA water box (32,000 rigid waters on a seeded
32^3 lattice, 98.65 Å cube, PME grid 120^3 order 5, Ewald cutoff 7 Å,
mutual polarization by PCG to 1e-6 D RMS).  One "step" is one pass of the hot
path per 2-fs MD step: neighbor-list refresh check, local-frame rotation of
the multipoles, direct field (real space + PME), and the preconditioned CG
solve of the induced dipoles to convergence.  Coordinates stay resident in
HBM for the whole timed region.  ns/day = 0.1728 / seconds-per-step.

Prints ONE JSON line with the roofline of the dominant kernel (the mutual
field real-space pair kernel, timed live with HIP events on its stream) and
the CPU baseline (the FP64 oracle, one induce of the same system, on this
host's cores).
"""
import argparse
import json
import os
import sys
import time

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(REPO, "tinker-gpu_amd"))

NMOL, NSIDE, BOX, GRID, SEED = 32000, 32, 98.65, 120, 2024
KEY = ("ewald\newald-cutoff 7.0\nvdw-cutoff 9.0\npme-grid {K} {K} {K}\npme-order 5\npolar-eps 1e-6\n"
       "polar-iter 100\npolarization mutual\nneighbor-list\nlist-buffer 1.0\npcg-precond\npcg-guess\npcg-peek 1.0\n"
       "usolve-cutoff 4.5\nusolve-diag 2.0\n")
FLOP_PER_PAIR_UFIELD = 130  # algorithmic FLOPs per unique pair within the cutoff
PEAK_FP32_TFLOPS = 157.3
PEAK_FP64_TFLOPS = 78.6


def host_cores():
    try:
        return len(os.sched_getaffinity(0))
    except AttributeError:
        return os.cpu_count() or 1


def cpu_baseline(system_arrays, threads):
    """FP64 oracle: one induce() of the same 96k-atom system (bounded sample)."""
    os.environ["OMP_NUM_THREADS"] = str(threads)
    sys.path.insert(0, os.path.join(REPO, "oracle"))
    import oracle
    o = oracle.Oracle(system_arrays)
    t0 = time.perf_counter()
    ud, up, it, rms, rc = o.induce()
    dt = time.perf_counter() - t0
    return dt, it


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--precision", choices=["fp32", "fp64"], default="fp32")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--cpu-threads", type=int, default=0)
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU only (asked for --gpus {args.gpus}); "
              "multi-GPU sharding is design-only", file=sys.stderr)
        sys.exit(2)

    import t9hip
    prec = t9hip.T9_FP32 if args.precision == "fp32" else t9hip.T9_FP64
    system = t9hip.System.water_box(NMOL, NSIDE, BOX, SEED, key=KEY.format(K=GRID))
    ctx = t9hip.Context(system, prec)
    ctx.nblist_build()
    pairs = ctx.cutoff_pairs()

    def step():
        ctx.nblist_refresh()
        ctx.mpole_init()
        _, _, st = ctx.induce(copy_out=False)
        return st

    iters = []
    for _ in range(args.warmup):
        step()
    ctx.sync()
    ctx.kernel_timing(t9hip.K_UFIELD, 1)
    ctx.kernel_timing(t9hip.K_FFT, 1)
    ctx.kernel_timing(t9hip.K_DFIELD, 1)
    ctx.kernel_timing(t9hip.K_SPREAD, 1)
    ctx.kernel_timing(t9hip.K_GATHER, 1)
    ctx.kernel_timing(t9hip.K_PRECOND, 1)
    ctx.sync()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        st = step()
        iters.append(st.iterations)
    ctx.sync()
    t1 = time.perf_counter()
    kt = {}
    for name, kid in (("ufield_real", t9hip.K_UFIELD), ("fft_pass", t9hip.K_FFT), ("dfield_real", t9hip.K_DFIELD),
                      ("pme_spread", t9hip.K_SPREAD), ("pme_gather", t9hip.K_GATHER),
                      ("precond", t9hip.K_PRECOND)):
        ms, cnt = ctx.kernel_timing(kid, 0)
        kt[name] = {"total_ms": round(ms, 4), "launches": cnt, "avg_us": round(1e3 * ms / max(cnt, 1), 3)}

    sec_per_step = (t1 - t0) / args.steps
    ns_day = 0.1728 / sec_per_step
    uf = kt["ufield_real"]
    avg_s = uf["avg_us"] * 1e-6
    achieved = FLOP_PER_PAIR_UFIELD * pairs / avg_s / 1e12 if avg_s > 0 else 0.0
    peak = PEAK_FP32_TFLOPS if prec == t9hip.T9_FP32 else PEAK_FP64_TFLOPS
    traffic = None
    pmc = os.path.join(REPO, "profiles", "ufield_real_water96k_traffic.json")
    if os.path.exists(pmc):
        try:
            traffic = json.load(open(pmc)).get("hbm_bytes_per_launch")
        except (ValueError, OSError):
            traffic = None

    cpu = None
    if not args.no_cpu:
        threads = args.cpu_threads or min(16, host_cores())
        dt, cit = cpu_baseline(system.arrays(), threads)
        cpu = {"value": round(0.1728 / dt, 4), "unit": "ns/day", "cores": threads, "kind": "port",
               "sample": f"FP64 oracle, one induce() of the same 96,000-atom box ({cit} CG iterations, "
                         f"{dt:.1f} s, OpenMP {threads} threads)"}

    out = {
        "metric": "ns/day, 96k-atom AMOEBA water: FP32+learned guess vs plain FP64 CG, same tol",
        "value": round(ns_day, 3),
        "unit": "ns/day",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1e3 * sec_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32" if prec == t9hip.T9_FP32 else "f64",
        "data": "synthetic (seeded 32^3-lattice water box, random orientations)",
        "config": {
            "workload": "AMOEBA electrostatics hot path per 2-fs step on 96,000-atom water: nblist refresh + "
                        "rotpole + dfield (real+PME) + PCG induce to 1e-6 D (energy/forces, vdW, valence and "
                        "the integrator are not yet in the step)",
            "atoms": system.n, "pme_grid": GRID, "ewald_cutoff": 7.0, "list_buffer": 1.0,
            "polar_eps": 1e-6, "predictor": "direct guess", "precision": args.precision,
            "cg_iters_per_step_mean": round(sum(iters) / len(iters), 2),
            "parallelism": "single GPU",
        },
        "roofline": {
            "kernel": "ufield real-space pair kernel",
            "bound": "mfma",
            "achieved": round(achieved, 3),
            "peak": peak,
            "unit": "TFLOP/s",
            "frac": round(achieved / peak, 5),
            "traffic": traffic,
            "pairs_within_cutoff": pairs,
            "flop_per_pair": FLOP_PER_PAIR_UFIELD,
            "avg_launch_us": uf["avg_us"],
        },
        "kernels": kt,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
