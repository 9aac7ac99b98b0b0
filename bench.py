#!/usr/bin/env python3
"""Benchmark of AMOEBA water molecular dynamics on one MI355X.

Workload: the 96,000-atom AMOEBA water box (32,000 waters on a seeded 32^3
lattice, 98.65 Å cube, PME grid 120^3 order 5, Ewald cutoff 7 Å, vdW cutoff
9 Å, list buffer 2 Å, mutual polarization by PCG to 1e-6 D RMS with the ASPC
predictor), Maxwell-Boltzmann velocities at 298 K, Bussi thermostat at 298 K.
One "step" is one 2-fs RESPA step (t9_md_run): 8 inner 0.25-fs steps of the
valence terms, and the slow forces once: neighbor-list refresh, local-frame
rotation of the multipoles, direct field (real space + PME), the PCG solve of
the induced dipoles, the multipole + polarization energy and gradient (fused
real-space kernel, reciprocal terms, torques) and the buffered 14-7 vdW.
Positions, velocities and forces stay resident in HBM.
ns/day = 0.1728 / seconds-per-step.

Every leg starts from the same equilibrated state and runs untimed steps
until the predictor histories are full before its warmup, so the timed
steps are steady-state MD.  The same MD is timed in the plain FP64 form
(FP64 pair/PME/CG kernels) for the FP32-vs-FP64 comparison.  Prints ONE JSON
line with the roofline of the dominant kernel family (largest GPU time per
step, timed live with HIP events on the stream it runs on), the CPU
baseline (the FP64 oracle on this host's cores: one slow-force evaluation of
the same box on all cores, and the 648-atom single point on 1 and on all
cores), the NVE energy drift of the FP32 + learned-guess form and its
FP64-oracle errors on the 648-atom box.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(REPO, "tinker-gpu_amd"))

NMOL, NSIDE, BOX, GRID, SEED = 32000, 32, 98.65, 120, 2024
# the 648-atom box of the same construction (216 waters, K = 24)
KEY0 = ("ewald\newald-cutoff 7.0\nvdw-cutoff 9.0\npme-grid 24 24 24\npme-order 5\npolar-eps 1e-6\n"
        "polarization mutual\nneighbor-list\nlist-buffer 0.3\npcg-precond\nusolve-cutoff 4.5\nusolve-diag 2.0\n"
        "vdwterm\nbondterm\n")
KEY = ("ewald\newald-cutoff 7.0\nvdw-cutoff 9.0\npme-grid {K} {K} {K}\npme-order 5\npolar-eps 1e-6\n"
       "polar-iter 100\npolarization mutual\nneighbor-list\nlist-buffer {buf}\npcg-precond\npcg-guess\n"
       "pcg-peek 1.0\nusolve-cutoff 4.5\nusolve-diag {udiag}\npolar-predict ASPC\nvdwterm\nbondterm\n")
LIST_BUFFER = 2.0
# Diagonal factor of the sparse preconditioner: 1.5 converges ~6x per CG
# iteration on this box against ~5x at Tinker's 2.0; the FP64 leg is also
# run at 2.0 to show that its iteration count does not depend on it.
UDIAG, UDIAG_REFERENCE = 1.5, 2.0
DT, NRESPA, KELVIN, TAU = 0.002, 8, 298.0, 0.1
# Algorithmic work per unit (the units are the unique atom pairs within the
# 7 Å cutoff on the actual coordinates, or grid points / atoms for PME).
FLOP_PER_PAIR = {"ufield_real": 130, "dfield_real": 180, "emplar_real": 950}
# Untimed solves each leg runs before its warmup: the predictor's 16-deep
# dipole and field histories fill, so every warmup and timed step uses the
# steady-state guess (ASPC, or the learned guess).
PRIME = 18
PEAK_TFLOPS = {32: 157.3, 64: 78.6}
PEAK_HBM_GBS = 8000.0


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def host_cores():
    try:
        return len(os.sched_getaffinity(0))
    except AttributeError:
        return os.cpu_count() or 1


def _cpu_step(sysd, full):
    """FP64 oracle: induced dipoles by PCG from the direct guess and, when
    `full`, the multipole + polarization energy and gradient and the vdW +
    valence energy and gradient.  Returns seconds and CG iterations."""
    sys.path.insert(0, os.path.join(REPO, "oracle"))
    import oracle
    o = oracle.Oracle(sysd)
    t0 = time.perf_counter()
    ud, up, it, rms, rc = o.induce()
    if full:
        o.energy(ud, up)
        o.md_energy()
    return time.perf_counter() - t0, it


def cpu_baseline(t9hip, system, threads):
    """FP64 oracle on this host's cores (OpenMP), bounded samples: the PCG
    solve of the 96,000-atom box on `threads` threads (the ns/day value: one
    solve counted as one 2-fs step), and the 648-atom single point (solve +
    E + G + V of every term) on 1 and on `threads` threads."""
    import subprocess
    out = {}
    small = t9hip.System.water_box(216, 6, 18.643, SEED, key=KEY0)
    for label, sysd, nthr, full in (("config0_1core", small.arrays(), 1, True),
                                    ("config0_allcore", small.arrays(), threads, True),
                                    ("water96k_solve_allcore", system.arrays(), threads, False)):
        # OpenMP reads the thread count at start-up: each sample runs in a child
        # process that loads the oracle there (no GPU use in the child)
        import tempfile
        with tempfile.NamedTemporaryFile(suffix=".npz", delete=False) as f:
            np.savez(f, **{k: np.asarray(v) for k, v in sysd.items()})
            path = f.name
        env = dict(os.environ, OMP_NUM_THREADS=str(nthr), OMP_PROC_BIND="close")
        code = ("import sys, json, numpy as np; sys.path.insert(0, %r); import bench; "
                "z = np.load(%r); d = {k: (z[k][()] if z[k].ndim == 0 else z[k]) for k in z.files}; "
                "print(json.dumps(bench._cpu_step(d, %r)))" % (REPO, path, full))
        r = subprocess.run([sys.executable, "-c", code], env=env, capture_output=True, text=True, timeout=600)
        os.unlink(path)
        if r.returncode != 0:
            raise RuntimeError("CPU baseline sample failed: " + r.stderr[-2000:])
        dt, it = json.loads(r.stdout.strip().splitlines()[-1])
        out[label] = {"seconds": round(dt, 3), "cg_iterations": it, "threads": nthr}
    return out


FAMILIES = None


def equilibrate(t9hip, system, steps):
    """FP32 RESPA + Bussi from the lattice: the rigid construction geometry
    relaxes and heats the box; the thermostat brings it back to 298 K."""
    ctx = t9hip.Context(system, 32)
    ctx.md_init(KELVIN, SEED)
    st = ctx.md_run(steps, DT, NRESPA, KELVIN, TAU)
    x, v = ctx.coords(), ctx.velocities()
    ctx.close()
    return x, v, st


def run_gpu(t9hip, system, prec, steps, warmup, guess=None, start=None, timers=True, per_step_samples=20):
    global FAMILIES
    FAMILIES = (("ufield_real", t9hip.K_UFIELD), ("dfield_real", t9hip.K_DFIELD), ("emplar_real", t9hip.K_EMPLAR),
                ("fft_pass", t9hip.K_FFT), ("pme_spread", t9hip.K_SPREAD), ("pme_gather", t9hip.K_GATHER),
                ("precond", t9hip.K_PRECOND), ("vdw_pairs", t9hip.K_VDW), ("valence", t9hip.K_VALENCE))
    ctx = t9hip.Context(system, prec)
    if start is not None:
        ctx.set_coords(start[0])
    ctx.nblist_build()
    pairs = ctx.cutoff_pairs()
    if guess is not None:  # learned guess: active once the 16-deep histories are filled
        ctx.guess_set_params(guess["params"], guess["norm"])
        ctx.guess_enable(True)
    ctx.md_init(KELVIN, SEED)
    if start is not None:
        ctx.set_velocities(start[1])
    # the predictor histories fill during PRIME untimed steps; then the warmup
    ctx.md_run(PRIME, DT, NRESPA, KELVIN, TAU)
    wst = ctx.md_run(warmup, DT, NRESPA, KELVIN, TAU) if warmup > 0 else None
    ctx.sync()
    gs0 = ctx.guess_stats() if guess is not None else None
    timed = FAMILIES if timers is True else tuple(f for f in FAMILIES if f[0] in (timers or ()))
    for _, kid in timed:
        ctx.kernel_timing(kid, 1)
    ctx.sync()
    t0 = time.perf_counter()
    st = ctx.md_run(steps, DT, NRESPA, KELVIN, TAU)
    ctx.sync()
    t1 = time.perf_counter()
    kt = {}
    for name, kid in timed:
        ms, cnt = ctx.kernel_timing(kid, 0)
        kt[name] = {"total_ms_per_step": round(ms / steps, 4), "launches_per_step": round(cnt / steps, 2),
                    "avg_us": round(1e3 * ms / max(cnt, 1), 3)}
    gs = None
    if guess is not None:
        g1 = ctx.guess_stats()
        gs = {"used": g1["used"] - gs0["used"], "fallback": g1["fallback"] - gs0["fallback"],
              "last_predictor": g1["last_predictor"]}
    # CG iterations of single steps after the timed region (untimed)
    per_step = [ctx.md_run(1, DT, NRESPA, KELVIN, TAU)["iterations"] for _ in range(per_step_samples)]
    ctx.close()
    return {"sec_per_step": (t1 - t0) / steps, "iters_per_step": st["iterations"] / steps, "kernels": kt,
            "pairs": pairs, "md": st, "warmup_md": wst, "guess": gs,
            "iters_p95": float(np.percentile(per_step, 95)) if per_step else None}


def drift_leg(t9hip, system, guess, start, steps, chunk=20):
    """NVE (no thermostat) RESPA from the equilibrated state, FP32 kernels and
    the learned guess: total energy every `chunk` steps, least-squares slope
    in kcal/mol per ns per atom."""
    ctx = t9hip.Context(system, 32)
    ctx.set_coords(start[0])
    if guess is not None:
        ctx.guess_set_params(guess["params"], guess["norm"])
        ctx.guess_enable(True)
    ctx.set_velocities(start[1])
    ctx.md_run(PRIME, DT, NRESPA, 0.0, TAU)
    t, e = [], []
    for k in range(steps // chunk):
        st = ctx.md_run(chunk, DT, NRESPA, 0.0, TAU)
        t.append((k + 1) * chunk * DT * 1e-3)  # ns
        e.append(st["epot"] + st["ekin"])
    ctx.close()
    slope = float(np.polyfit(np.array(t), np.array(e), 1)[0]) if len(t) > 1 else 0.0
    return {"drift_kcal_mol_per_ns_per_atom": slope / system.n, "nve_steps": len(t) * chunk,
            "etot_first": e[0] if e else None, "etot_last": e[-1] if e else None}


def oracle_errors_config0(t9hip):
    """The 648-atom box (216 waters, K = 24, same keywords as the workload):
    FP32 GPU energy, gradient and induced dipoles against the FP64 oracle
    (the oracle is the checker here, not the thing measured)."""
    sys.path.insert(0, os.path.join(REPO, "oracle"))
    import oracle
    s = t9hip.System.water_box(216, 6, 18.643, SEED, key=KEY0)
    o = oracle.Oracle(s.arrays())
    oud, oup, it, rms, rc = o.induce()
    eo, go, vo = o.energy(oud, oup)
    em, gm, vm = o.md_energy()
    e_ref = eo["em"] + eo["ep"] + em["evdw"] + em["ebond"] + em["eangle"] + em["eurey"]
    g_ref = go + gm
    ctx = t9hip.Context(s, 32)
    ud, up, st = ctx.induce()
    eg, gg, vg = ctx.energy()
    ctx.close()
    return {"atoms": s.n, "energy_rel": abs(eg["esum"] - e_ref) / abs(e_ref),
            "force_rms_kcal_mol_A": float(np.sqrt(np.mean((gg - g_ref) ** 2))),
            "dipole_rms_debye": float(np.sqrt(np.mean((ud - oud) ** 2)) * 4.803206802),
            "virial_max_abs": float(np.abs(vg - (vo + vm)).max())}


def roofline_of(name, k, pairs, prec, natoms):
    avg_s = k["avg_us"] * 1e-6
    if avg_s <= 0:
        return None
    K3 = GRID ** 3
    esz = 8 if prec == 32 else 16  # bytes per complex grid point
    if name in FLOP_PER_PAIR:
        # FP32 / FP64 vector (VALU) arithmetic, against the vector peak
        achieved = FLOP_PER_PAIR[name] * pairs / avg_s / 1e12
        peak = PEAK_TFLOPS[prec]
        return {"bound": "valu", "achieved": round(achieved, 3), "peak": peak, "unit": "TFLOP/s",
                "frac": round(achieved / peak, 5), "unit_work": f"{FLOP_PER_PAIR[name]} FLOP x {pairs} pairs"}
    if name == "fft_pass":  # one pass reads and writes the complex grid once
        byts = 2 * esz * K3
    elif name == "pme_spread":  # grid written once + per-atom weights/bases/sources
        byts = esz * K3 + 64 * natoms
    elif name == "pme_gather":  # grid read once + per-atom inputs/outputs
        byts = esz * K3 + 64 * natoms
    else:
        return None
    achieved = byts / avg_s / 1e9
    return {"bound": "hbm", "achieved": round(achieved, 1), "peak": PEAK_HBM_GBS, "unit": "GB/s",
            "frac": round(achieved / PEAK_HBM_GBS, 5), "unit_work": f"{byts} bytes"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=30)
    ap.add_argument("--equil", type=int, default=2000, help="untimed equilibration steps before the legs")
    ap.add_argument("--no-learned", action="store_true", help="ASPC instead of the learned guess")
    ap.add_argument("--weights", default=os.path.join(REPO, "tinker-gpu_amd", "guess_water.npz"))
    ap.add_argument("--precision", choices=["fp32", "fp64"], default="fp32")
    ap.add_argument("--no-fp64", action="store_true", help="skip the plain FP64 comparison leg")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--no-drift", action="store_true", help="skip the NVE energy-drift leg")
    ap.add_argument("--drift-steps", type=int, default=200)
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: the solver runs on one GPU only (asked for --gpus {args.gpus}); "
              "the sharded multi-GPU form is described, not built", file=sys.stderr)
        sys.exit(2)

    import t9hip
    prec = 32 if args.precision == "fp32" else 64
    system = t9hip.System.water_box(NMOL, NSIDE, BOX, SEED, key=KEY.format(K=GRID, buf=LIST_BUFFER, udiag=UDIAG))
    guess = None
    if not args.no_learned and prec == 32 and os.path.exists(args.weights):
        w = np.load(args.weights)
        guess = {"params": w["params"], "norm": w["norm"]}
    x_eq, v_eq, eq = equilibrate(t9hip, system, args.equil)
    start = (x_eq, v_eq)
    # per-kernel breakdown: a separate, shorter leg with every family's timer on
    # (the events add ~0.2 ms per step, so this leg is not the measured one)
    brk = run_gpu(t9hip, system, prec, min(args.steps, 30), args.warmup, guess, start, timers=True)
    kt = brk["kernels"]
    # dominant kernel family: the largest GPU time per step; the measured leg
    # times only it
    dom = max(kt, key=lambda n: kt[n]["total_ms_per_step"])
    main_run = run_gpu(t9hip, system, prec, args.steps, args.warmup, guess, start, timers=(dom,))
    sec = main_run["sec_per_step"]
    ns_day = 0.1728 / sec
    pairs = main_run["pairs"]
    kdom = main_run["kernels"][dom]
    roof = roofline_of(dom, kdom, pairs, prec, system.n)
    for name in kt:
        r = roofline_of(name, kt[name], pairs, prec, system.n)
        if r:
            kt[name]["roofline_frac"] = r["frac"]
            kt[name]["achieved"] = f'{r["achieved"]} {r["unit"]}'
    traffic = None
    pmc = os.path.join(REPO, "profiles", f"{dom}_water96k_traffic.json")
    if os.path.exists(pmc):
        try:
            traffic = json.load(open(pmc)).get("hbm_bytes_per_launch")
        except (ValueError, OSError):
            traffic = None
    roof = dict(roof or {})
    roof.update({"kernel": dom, "traffic": traffic, "avg_launch_us": kdom["avg_us"],
                 "launches_per_step": kdom["launches_per_step"], "pairs_within_cutoff": pairs})

    fp32_aspc = None
    if guess is not None and not args.no_fp64:
        ra = run_gpu(t9hip, system, 32, args.steps, args.warmup, None, start, timers=False)
        fp32_aspc = {"ms_per_step": round(1e3 * ra["sec_per_step"], 4), "ns_day": round(0.1728 / ra["sec_per_step"], 3),
                     "cg_iters_per_step_mean": round(ra["iters_per_step"], 2), "cg_iters_per_step_p95": ra["iters_p95"]}
    fp64 = None
    if not args.no_fp64 and prec == 32:
        r64 = run_gpu(t9hip, system, 64, args.steps, args.warmup, None, start, timers=False)
        fp64 = {"ms_per_step": round(1e3 * r64["sec_per_step"], 4),
                "ns_day": round(0.1728 / r64["sec_per_step"], 3),
                "cg_iters_per_step_mean": round(r64["iters_per_step"], 2), "cg_iters_per_step_p95": r64["iters_p95"],
                "temperature_K": round(r64["md"]["temperature"], 2),
                "speedup_fp32_over_fp64": round(r64["sec_per_step"] / sec, 3)}
        sys_ref = t9hip.System.water_box(NMOL, NSIDE, BOX, SEED,
                                         key=KEY.format(K=GRID, buf=LIST_BUFFER, udiag=UDIAG_REFERENCE))
        rr = run_gpu(t9hip, sys_ref, 64, args.steps, args.warmup, None, start, timers=False)
        fp64["reference_precond"] = {"usolve_diag": UDIAG_REFERENCE, "ms_per_step": round(1e3 * rr["sec_per_step"], 4),
                                     "ns_day": round(0.1728 / rr["sec_per_step"], 3),
                                     "cg_iters_per_step_mean": round(rr["iters_per_step"], 2),
                                     "speedup_fp32_over_fp64": round(rr["sec_per_step"] / sec, 3)}

    cpu = None
    if not args.no_cpu:
        threads = args.cpu_threads or min(16, host_cores())
        smp = cpu_baseline(t9hip, system, threads)
        big = smp["water96k_solve_allcore"]
        cpu = {"value": round(0.1728 / big["seconds"], 5), "unit": "ns/day", "cores": threads, "kind": "port",
               "nproc": os.cpu_count(), "cpu_model": cpu_model(),
               "sample": f"FP64 oracle (OpenMP, {threads} threads): the induced-dipole PCG solve of the same "
                         f"96,000-atom box from the direct guess ({big['cg_iterations']} iterations, "
                         f"{big['seconds']:.1f} s) counted as one 2-fs step; the rest of the step (energy and "
                         f"gradient of every term, valence inner steps, integrator) is not in the sample, so this "
                         f"ns/day is an upper bound for the CPU.  Also the 648-atom single point (solve + energy, "
                         f"gradient and virial of every term) on 1 and on {threads} threads",
               "samples": smp}
        c1, cn = smp["config0_1core"]["seconds"], smp["config0_allcore"]["seconds"]
        cpu["one_core"] = {"config0_single_point_s": c1, "config0_ns_day_equiv": round(0.1728 / c1, 4)}
        cpu["all_core"] = {"config0_single_point_s": cn, "config0_ns_day_equiv": round(0.1728 / cn, 4)}

    drift = None
    if not args.no_drift:
        drift = drift_leg(t9hip, system, guess, start, args.drift_steps)
    oerr = None if args.no_cpu else oracle_errors_config0(t9hip)

    md = main_run["md"]
    out = {
        "metric": "ns/day, 96k-atom AMOEBA water: FP32+learned guess vs plain FP64 CG, same tol",
        "value": round(ns_day, 3),
        "unit": "ns/day",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1e3 * sec, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32" if prec == 32 else "f64",
        "data": "synthetic (seeded 32^3-lattice water box, random orientations, Maxwell-Boltzmann 298 K, "
                "equilibrated untimed before the timed steps)",
        "config": {
            "workload": "AMOEBA water MD on 96,000 atoms: 2-fs RESPA steps (8 x 0.25-fs valence inner steps; "
                        "slow forces = multipole + mutual polarization (PCG to 1e-6 D) + buffered 14-7 vdW), "
                        "Bussi thermostat 298 K",
            "atoms": system.n, "pme_grid": GRID, "ewald_cutoff": 7.0, "vdw_cutoff": 9.0,
            "list_buffer": LIST_BUFFER, "polar_eps": 1e-6, "usolve_cutoff": 4.5, "usolve_diag": UDIAG, "dt_fs": 1e3 * DT, "nrespa": NRESPA,
            "predictor": ("learned guess (ASPC fallback)" if guess is not None else "ASPC"),
            "learned_guess": main_run["guess"],
            "precision": args.precision,
            "cg_iters_per_step_mean": round(main_run["iters_per_step"], 2),
            "cg_iters_per_step_p95": main_run["iters_p95"],
            "epot_kcal_mol": md["epot"], "temperature_K": round(md["temperature"], 2),
            "list_builds": md["list_builds"],
            "equilibration": {"steps": args.equil, "temperature_K": round(eq["temperature"], 2),
                              "cg_iters_per_step_mean": round(eq["iterations"] / max(1, args.equil), 2)},
            "parallelism": "single GPU",
        },
        "roofline": roof,
        "kernels": kt,
        "kernels_note": f"from a separate {min(args.steps, 30)}-step leg with every family timed by HIP events",
        "fp32_aspc": fp32_aspc,
        "fp64_plain": fp64,
        "cpu_baseline": cpu,
        "drift": drift,
        "oracle_err_config0": oerr,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
