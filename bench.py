#!/usr/bin/env python3
"""Benchmark of the nonbonded path on one MI355X.

A step is one nonbonded pass (FP32 packing, per-cell sorts, pair/self/
exclusion kernel, FP64 energy reduction) over the resident 1,048,576-atom
20-species mixture (rc 1.2 nm, 0.6 nm cells, reference sorted window).
Positions are uploaded before the timed region; W untimed passes, then K
passes timed with HIP events on the engine's stream.  Prints one JSON line.

  value         pair interactions evaluated per second (pairs per pass /
                mean pass time)
  roofline      the pair kernel (k_nonbond): 46 algorithmic FP32 FLOP per
                evaluated pair over its mean launch time, against the FP32
                peak; traffic from the committed rocprofv3 counter summary
  cpu_baseline  the FP64 oracle (CPU restatement of the reference runner)
                on the host cores, full passes of the same system; its forces
                are the FP64 reference of force_rmse
  mlp           the surrogate pass (engine_flag_mlp) on the same system:
                trained on the GPU, timed like the table pass, its force
                error against FP64, and a table set rebuilt at the tolerance
                whose force error matches the network's, timed the same way
  nve           energy drift of 10,000 device-resident NVE steps (dt 2 fs)
                of the 32,000-atom argon system: the plain LJ table, the
                force-continuous switched table, and a trained network on
                each; and of the 100k-atom rigid water (3 ps rescaled to
                300 K, then 2,000 NVE steps of 1 fs)
  md_step       whole MD steps (shuffle, pass, kick-drift, migration) of the
                1,048,576-atom argon crystal at 100 K, dt 2 fs: host-driven
                engine_step (positions up, forces down, host integration)
                against engine_gpu_steps (particles resident on the device),
                and the resident steps with pseudo-Verlet sort reuse
"""
import argparse
import json
import os
import platform
import subprocess
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT / "mdcore-history_amd"))

FLOP_PER_PAIR = 46          # algorithmic FLOP per evaluated pair interaction
FP32_PEAK_TFLOPS = 157.3    # MI355X FP32 peak
# 64x64 layer on MFMA per network evaluation: the values and their
# r-derivatives, 2 x 64 x 64 multiply-adds each (the bf16 hi + lo split of the
# inputs runs each product twice; that precision work is not counted)
MLP_MFMA_FLOP = 2 * 2 * 64 * 64
BF16_PEAK_TFLOPS = 2500.0   # MI355X dense bf16 MFMA peak
MI355X_CUS = 256            # the pair kernel's 5-wave build runs from 64 cells per CU up
DT_PS = 0.002               # 2 fs time step for ns/day
MLP_TRAIN_STEPS = 40000     # surrogate training steps on the mixture (16,384 samples each)
METRIC = "pair-interactions/s & ns/day at 1M atoms; MLP vs table at fixed force RMSE"
TRAFFIC_FILE = ROOT / "profiles" / "k_nonbond_mixture_1m_traffic.json"
COUNTERS_FILE = ROOT / "profiles" / "k_nonbond_mixture_1m_counters.json"


_T0 = time.perf_counter()


def note(msg):
    """One progress line on stderr per finished section (the JSON line is the
    only thing on stdout)."""
    print(f"[bench {time.perf_counter() - _T0:7.1f} s] {msg}", file=sys.stderr, flush=True)


def build_workload(name):
    import systems
    if name == "mixture_1m":
        return systems.mixture(1048576), "1,048,576-atom 20-species LJ+RF mixture, rho 100/nm^3, rc 1.2 nm, 0.6 nm cells"
    if name == "water_100k":
        return systems.water(), "33,333 SPC/E-like waters (99,999 atoms), 10 nm box, LJ+RF, rc 1.0 nm, exclusions"
    if name == "argon_32k":
        return systems.argon(20, jitter=0.05), "32,000 LJ argon atoms, 11.3 nm box, rc 1.0 nm"
    raise SystemExit(f"unknown workload {name}")


def host_cpus():
    """CPU model and counts of the host: lscpu's model and CPU(s), the CPUs
    this process may run on, and the cgroup CPU quota.  The baseline runs one
    thread per CPU it may use, capped by the quota."""
    info = {"model": platform.processor() or "unknown", "host_cpus": os.cpu_count()}
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        for line in out.splitlines():
            if line.startswith("Model name"):
                info["model"] = line.split(":", 1)[1].strip()
            elif line.startswith("CPU(s):"):
                info["host_cpus"] = int(line.split(":", 1)[1])
    except Exception:
        pass
    try:
        aff = len(os.sched_getaffinity(0))
    except AttributeError:
        aff = os.cpu_count() or 1
    info["affinity_cpus"] = aff
    quota = None
    try:
        q, period = Path("/sys/fs/cgroup/cpu.max").read_text().split()[:2]
        if q != "max":
            quota = float(q) / float(period)
    except Exception:
        pass
    info["cgroup_cpu_quota"] = quota
    info["threads"] = max(1, min(aff, int(quota) if quota else aff))
    return info


def cpu_baseline_config(sysm, threads, mode, steps=10, reps=3):
    """The FP64 oracle on the host cores: 1 warm-up step, then `reps` runs of
    `steps` steps, each one nonbonded pass (sort, pair/self tasks,
    exclusions) and one kick-drift over the cells (or_advance, dt = 0 as in
    the reference's probe, so every step sees the same pairs); medians."""
    import statistics
    import numpy as np
    if str(ROOT / "oracle") not in sys.path:
        sys.path.insert(0, str(ROOT / "oracle"))
    import pyoracle as po
    orc = po.from_system(sysm)
    orc.set_motion(None, 1.0 / np.where(sysm.masses > 0, sysm.masses, 1.0))
    r = orc.nonbond(mode, threads)
    orc.advance(0.0, threads)
    t_nb, t_full = [], []
    for _ in range(reps):
        a = b = 0.0
        for _ in range(steps):
            t0 = time.perf_counter()
            r = orc.nonbond(mode, threads)
            t1 = time.perf_counter()
            orc.advance(0.0, threads)
            a += t1 - t0
            b += time.perf_counter() - t0
        t_nb.append(a / steps)
        t_full.append(b / steps)
    orc.close()
    nb, full = statistics.median(t_nb), statistics.median(t_full)
    return r, {"pairs": r["pairs"], "nonbond_ms": round(nb * 1e3, 2), "full_step_ms": round(full * 1e3, 2),
               "pairs_per_s_nonbond": r["pairs"] / nb, "pairs_per_s_full_step": r["pairs"] / full,
               "ns_per_day_full_step": DT_PS * 1e-3 * 86400.0 / full}


def gpu_vs_fp64(eng, f_gpu, r64):
    """Per-atom force relative RMSE and energy relative errors of a GPU pass
    against the FP64 oracle's pass over the same input."""
    g = eng.gpu_results()
    out = {"pairs_equal": g["pairs"] == r64["pairs"], "exclusions_equal": g["exclusions"] == r64["exclusions"],
           "force_rel_rmse": rel_rmse(f_gpu, r64["f"]),
           "epot_rel_err": abs(g["epot"] - r64["epot"]) / abs(r64["epot"])}
    if r64["exclusions"]:
        out["epot_exclusion_rel_err"] = abs(g["epot_exclusion"] - r64["epot_exclusion"]) / abs(r64["epot_exclusion"])
        tot_g, tot_o = g["epot"] + g["epot_exclusion"], r64["epot"] + r64["epot_exclusion"]
        out["epot_total_rel_err"] = abs(tot_g - tot_o) / abs(tot_o)
    return out


def config_section(name, args, threads, eng=None, f_gpu=None, ms=None):
    """One configuration: its GPU pass (timed unless given), the FP64 oracle
    timed on the host cores, and the GPU pass's error against FP64."""
    import mdcore as md
    import systems
    sysm, desc = build_workload(name)
    mode = 0 if args.window == "ref" else 1
    own = eng is None
    if own:
        flags = md.engine_flag_gpu | (md.engine_flag_projwindow if args.window == "proj" else 0)
        eng = systems.build_engine(sysm, flags)
        eng.shuffle()
        eng.sync(md.engine_gpu_sync_parts_up)
        ms = eng.gpu_time(args.steps, args.warmup)
        eng.nonbond()
        f_gpu = eng.forces()
    pairs = eng.gpu_results()["pairs"]
    r64, cpu = cpu_baseline_config(sysm, threads, mode)
    note(f"{name}: FP64 CPU baseline {cpu['nonbond_ms']} ms per pass on {threads} threads")
    out = {"workload": desc, "atoms": int(sysm.N), "pairs": int(pairs), "gpu_pass_ms": round(ms["total_ms"], 4),
           "gpu_pairs_per_s": pairs / (ms["total_ms"] * 1e-3), "cpu": cpu,
           "gpu_over_cpu_nonbond": pairs / (ms["total_ms"] * 1e-3) / cpu["pairs_per_s_nonbond"],
           "vs_fp64": gpu_vs_fp64(eng, f_gpu, r64)}
    if own:
        eng.close()
    return out, r64


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--workload", default="mixture_1m", choices=["mixture_1m", "water_100k", "argon_32k"])
    ap.add_argument("--window", default="ref", choices=["ref", "proj"])
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-mlp", action="store_true", help="skip the surrogate section")
    ap.add_argument("--mlp-fullshell", action="store_true",
                    help="time the full-shell surrogate kernel (each pair at both cells) instead of the half shell")
    ap.add_argument("--no-nve", action="store_true", help="skip the NVE drift section")
    ap.add_argument("--md-steps", type=int, default=100,
                    help="resident MD steps timed in the md_step section (0 skips it)")
    ap.add_argument("--md-only", action="store_true", help="run only the md_step section")
    ap.add_argument("--engine-steps", type=int, default=3, help="full engine_step calls timed for ns/day")
    ap.add_argument("--mixture-nve", choices=["table", "mlp"], default=None,
                    help="only run the 1,048,576-atom mixture's NVE (switched LJ + reaction-field tables): device "
                         "equilibration, then --nve-steps steps with the tables or with a surrogate trained on 2^24 "
                         "samples of the equilibrated state (its own command: it takes minutes)")
    ap.add_argument("--nve-steps", type=int, default=10000)
    ap.add_argument("--nve-dt", type=float, default=0.002, help="time step (ps) of --mixture-nve")
    ap.add_argument("--nve-window", choices=["ref", "proj"], default="ref",
                    help="pair window of --mixture-nve: the reference's sorted window, or the projected window "
                         "that admits every pair inside the cutoff")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run a plain device copy of MIB MiB (read MIB, write MIB) five times, as the "
                         "known-size reference for the FETCH_SIZE / WRITE_SIZE counter scale")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py: the nonbonded engine runs on one GPU; runs on more than one GPU are not supported "
              "(domain decomposition across GPUs is a design note, not code).", file=sys.stderr)
        sys.exit(2)
    if args.copy_calibration > 0:
        copy_calibration(args.copy_calibration)
        return

    if args.md_only:
        print(json.dumps({"md_step": md_step_section(args.md_steps)}), flush=True)
        return
    if args.mixture_nve:
        print(json.dumps({"mixture_nve": mixture_nve(args.mixture_nve == "mlp", args.nve_steps, dt=args.nve_dt,
                                                      window=args.nve_window)}),
              flush=True)
        return

    import mdcore as md
    import systems

    t_setup = time.perf_counter()
    sysm, desc = build_workload(args.workload)
    flags = md.engine_flag_gpu | (md.engine_flag_projwindow if args.window == "proj" else 0)
    eng = systems.build_engine(sysm, flags)
    eng.shuffle()
    eng.sync(md.engine_gpu_sync_parts_up)
    t_setup = time.perf_counter() - t_setup

    ms = eng.gpu_time(args.steps, args.warmup)
    res = eng.gpu_results()
    pairs = res["pairs"]
    note(f"{args.workload}: {args.steps} timed passes, {ms['total_ms']:.3f} ms per pass")
    step_s = ms["total_ms"] * 1e-3
    value = pairs / step_s

    achieved = FLOP_PER_PAIR * pairs / (ms["nonbond_ms"] * 1e-3) / 1e12
    traffic = None
    if TRAFFIC_FILE.exists():
        try:
            t = json.loads(TRAFFIC_FILE.read_text())
            if t.get("workload") == args.workload and t.get("window") == args.window:
                traffic = t.get("bytes_per_launch")
        except Exception:
            traffic = None
    valu = None
    if COUNTERS_FILE.exists():
        try:
            c = json.loads(COUNTERS_FILE.read_text())
            if c.get("workload") == args.workload and c.get("window") == args.window:
                valu = c.get("valu_issue_frac")
        except Exception:
            valu = None
    roofline = {"bound": "valu", "achieved": round(achieved, 3), "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s",
                "frac": round(achieved / FP32_PEAK_TFLOPS, 4), "traffic": traffic,
                "kernel": pair_kernel_name(sysm, md),
                "kernel_ms": round(ms["nonbond_ms"], 4),
                "valu_issue_frac": valu,
                "note": "FP32 vector ALU (VALU), not MFMA: 46 algorithmic FLOP per evaluated pair against "
                        "MI355X's 157.3 TFLOP/s FP32 peak, which counts packed FP32; this build disables packed "
                        "FP32, which halves the rate it can issue. valu_issue_frac = VALU instructions / (CUs x "
                        "cycles) from the committed counter summary: the kernel is VALU-issue-bound, with most "
                        "instructions outside the 46 FLOP of a pair evaluation"}

    out = {
        "metric": METRIC,
        "value": value,
        "unit": "pair-interactions/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms["total_ms"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic",
        "config": {"workload": f"{args.workload}: {desc}", "atoms": int(sysm.N), "pairs_per_step": int(pairs),
                   "window": args.window, "cells": int(eng_cells(sysm)), "parallelism": "single GPU"},
        "ns_per_day_pass_rate": DT_PS * 1e-3 * 86400.0 / step_s,
        "ns_per_day_pass_rate_note": "2 fs per nonbonded pass of the timed system: the rate of the pass alone, "
                                     "not of whole MD steps (md_step and the NVE sections time those)",
        "ns_per_day_engine_step": None,
        "kernels_ms": {k: round(v, 4) for k, v in ms.items()},
        "roofline": roofline,
        "setup_s": round(t_setup, 1),
    }
    f_table = None
    if not args.no_cpu_baseline or not args.no_mlp:
        eng.nonbond()
        f_table = eng.forces()
        eng.sync(md.engine_gpu_sync_parts_up)
    f64 = None
    if not args.no_cpu_baseline:
        cpus = host_cpus()
        threads = cpus["threads"]
        configs = {}
        main_cfg, r64 = config_section(args.workload, args, threads, eng, f_table, ms)
        configs[args.workload] = main_cfg
        f64 = r64["f"]
        for name in ("argon_32k", "water_100k", "mixture_1m"):
            if name != args.workload:
                configs[name] = config_section(name, args, threads)[0]
        cpu = main_cfg["cpu"]
        out["cpu_baseline"] = {
            "value": cpu["pairs_per_s_nonbond"], "unit": "pair-interactions/s", "cores": threads, "kind": "port",
            "sample": f"{args.workload}: 1 warm-up + 3 x 10 FP64 nonbonded passes (sort + pair/self + exclusions) "
                      f"of the same system, {cpu['pairs']} pairs each (equal to the GPU's: "
                      f"{cpu['pairs'] == pairs}), median; {threads} threads on {cpus['model']}",
            "host": cpus, "full_step_pairs_per_s": cpu["pairs_per_s_full_step"],
            "full_step_ns_per_day": cpu["ns_per_day_full_step"]}
        out["gpu_over_cpu"] = value / out["cpu_baseline"]["value"]
        out["force_rmse_vs_fp64"] = {"table": main_cfg["vs_fp64"]["force_rel_rmse"]}
        out["configs"] = configs
    if not args.no_mlp:
        out["mlp"] = mlp_section(sysm, eng, args, f_table, f64, value)
        note("surrogate section done")
        if f64 is not None:
            out["force_rmse_vs_fp64"]["mlp"] = out["mlp"].pop("_rmse_fp64")
        else:
            out["mlp"].pop("_rmse_fp64")
    # Full engine_step (host pack, PCIe both ways, device pass, host
    # kick-drift); last, because it moves the atoms.
    if args.engine_steps > 0:
        eng.setflags(0, md.engine_flag_mlp)
        md.check(eng.lib.engine_step(eng.e), "engine_step")  # warm-up
        t0 = time.perf_counter()
        for _ in range(args.engine_steps):
            md.check(eng.lib.engine_step(eng.e), "engine_step")
        full_s = (time.perf_counter() - t0) / args.engine_steps
        out["ns_per_day_engine_step"] = DT_PS * 1e-3 * 86400.0 / full_s
        note("engine_step section done")
    if not args.no_nve:
        out["nve"] = nve_section()
        note("NVE section done")
    eng.close()
    if args.md_steps > 0:
        out["md_step"] = md_step_section(args.md_steps)
        note("md_step section done")
    print(json.dumps(out), flush=True)


def rel_rmse(a, b):
    import numpy as np
    return float(np.sqrt(np.sum((a - b) ** 2) / np.sum(b ** 2)))


def mlp_section(sysm, eng, args, f_table, f64, table_value):
    """Train the surrogate on this system (half uniform-in-r samples, half the
    configuration's own pair distances; FP64 table targets), then time the
    surrogate pass exactly like the table pass."""
    import numpy as np
    import mdcore as md
    import surrogate as sg
    import systems
    t0 = time.perf_counter()
    dom = sg.pair_domains(eng, sysm)
    S = sg.make_samples(eng, sysm, 1 << 22, dom[0], dom[1], configs=[sysm.x], seed=5)
    t1 = time.perf_counter()
    eng.mlp_init(*dom, seed=7)
    hist = eng.mlp_train(*S, steps=MLP_TRAIN_STEPS, batch_tiles=1024, lr0=1e-2, lr1=1e-5, seed=11)
    t2 = time.perf_counter()
    eng.setflags(md.engine_flag_mlp | (md.engine_flag_fullshell if args.mlp_fullshell else 0), 0)
    eng.nonbond()
    f_mlp = eng.forces()
    eng.sync(md.engine_gpu_sync_parts_up)
    ms = eng.gpu_time(args.steps, args.warmup)
    res = eng.gpu_results()
    eng.setflags(0, md.engine_flag_mlp | md.engine_flag_fullshell)
    # share of the pairs below r_lo (evaluated from the tables), from a sample
    # of the configuration's pair distances
    _, ti, tj = sg.pair_potentials(eng, sysm)
    pid = -np.ones((sysm.max_type, sysm.max_type), np.int64)
    for k, (a, b) in enumerate(zip(ti, tj)):
        pid[a, b] = pid[b, a] = k
    r, a, b = md.sample_pairs(sysm.x, sysm.types, sysm.dim, float(sysm.cutoff), 1 << 20, 3)
    pp = pid[a, b]
    frac_table = float(np.mean(r[pp >= 0] < dom[0][pp[pp >= 0]]))
    # the default surrogate kernel is the half shell (each pair evaluated once);
    # engine_flag_fullshell evaluates every pair at both of its cells
    shells = 2.0 if args.mlp_fullshell else 1.0
    evals = shells * res["pairs"] * (1.0 - frac_table)
    achieved = MLP_MFMA_FLOP * evals / (ms["nonbond_ms"] * 1e-3) / 1e12
    out = {
        "value": res["pairs"] / (ms["total_ms"] * 1e-3), "unit": "pair-interactions/s",
        "ms_per_step": ms["total_ms"], "kernels_ms": {k: round(v, 4) for k, v in ms.items()},
        "ns_per_day": DT_PS * 1e-3 * 86400.0 / (ms["total_ms"] * 1e-3),
        "vs_table": res["pairs"] / (ms["total_ms"] * 1e-3) / table_value,
        "force_rmse_vs_table": rel_rmse(f_mlp, f_table),
        "pairs": res["pairs"], "pairs_below_r_lo_share": round(frac_table, 5),
        "training": {"samples": int(len(S[0])), "steps": MLP_TRAIN_STEPS, "batch": 1024 * 16,
                     "loss_first": float(hist[:20].mean()), "loss_last": float(hist[-50:].mean()),
                     "sample_s": round(t1 - t0, 2), "train_s": round(t2 - t1, 2)},
        "roofline": {"bound": "mfma", "achieved": round(achieved, 3), "peak": BF16_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": round(achieved / BF16_PEAK_TFLOPS, 5), "traffic": None,
                     "kernel": "k_nonbond_mlp" if args.mlp_fullshell else "k_nonbond_mlp_half + k_fold",
                     "evaluations_per_pass": evals,
                     "kernel_ms": round(ms["nonbond_ms"], 4),
                     "note": "bf16 MFMA FLOP of the 64x64 layer per network evaluation, values and "
                             "r-derivatives (16,384; the hi + lo split that doubles the products executed is not "
                             "counted); VALU work (SiLU, layer 1, head, pair search) not counted"},
        "_rmse_fp64": rel_rmse(f_mlp, f64) if f64 is not None else None,
    }
    if sysm.name == "mixture":
        out["table_at_matched_rmse"] = matched_table(sysm, f_table, out["force_rmse_vs_table"], args)
    return out


def matched_table(sysm, f_table, target, args, iters=10):
    """The table set whose force error against the default tables matches
    the network's: the tables are rebuilt at tolerances between the default
    1e-3 (error 0) and 1.2e-3, bisected on the tolerance until the error is
    within +-20 % of target.  The error is a step function of the tolerance
    (interval counts are integers); if it jumps over the band the result says
    so and gives the two tolerances that bracket it.  The matched table set
    is timed like the table pass."""
    import mdcore as md
    import systems

    def err_at(tol):
        s2 = systems.with_tolerance(sysm, tol)
        e2 = systems.build_engine(s2, md.engine_flag_gpu)
        e2.shuffle()
        e2.nonbond()
        return rel_rmse(e2.forces(), f_table), e2

    lo, hi = 1.0e-3, 1.2e-3
    e_hi, eng_hi = err_at(hi)
    eng_hi.close()
    tried = [(hi, e_hi)]
    found = None
    if e_hi < target * 0.8:
        hi = 4e-3
    for _ in range(iters):
        mid = 0.5 * (lo + hi)
        err, e2 = err_at(mid)
        tried.append((mid, err))
        if 0.8 * target <= err <= 1.2 * target:
            found = (mid, err, e2)
            break
        e2.close()
        if err < target:
            lo = mid
        else:
            hi = mid
    out = {"target_force_rmse": target, "tried": [{"tol": t, "force_rmse_vs_table": e} for t, e in tried]}
    if found is None:
        out["matched"] = False
        out["note"] = ("no table tolerance gives a force error within +-20 % of the network's: the error jumps "
                       "over the band between the last two tolerances tried")
        return out
    tol, err, e2 = found
    e2.sync(md.engine_gpu_sync_parts_up)
    ms = e2.gpu_time(args.steps, args.warmup)
    pairs = e2.gpu_results()["pairs"]
    e2.close()
    out.update({"matched": True, "tol": tol, "force_rmse_vs_table": err, "value": pairs / (ms["total_ms"] * 1e-3),
                "unit": "pair-interactions/s", "ms_per_step": ms["total_ms"]})
    return out


def nve_section(steps=10000, dt=0.002, equil=1000):
    """Energy drift of the 32,000-atom argon system (fcc start at 100 K) over
    `steps` device-resident NVE steps after `equil` steps (surrogate.nve_drift):
      table         the configuration's plain LJ 12-6 table (not shifted:
                    pairs crossing the cutoff jump in energy);
      mlp_switched  the surrogate trained on that table, its energy switched
                    off to zero between 0.9 nm and the cutoff;
      table_switch  the force-continuous LJ126_switch table (switch from
                    0.9 nm), the same function as mlp_switched's target;
      mlp           the surrogate trained on the LJ126_switch table, no switch
                    of its own;
    and the 100k-atom rigid water."""
    import dataclasses
    import mdcore as md
    import surrogate as sg
    import systems
    base = systems.argon(20)
    m = base.meta
    switched = dataclasses.replace(
        base, pots=[(1, 1, md.LJ126_switch(0.275, 1.0, m["A"], m["B"], 0.9, 1e-3))])
    out = {"system": "argon_32k: 32,000 LJ atoms, fcc start at 100 K, rc 1.0 nm", "steps": steps, "dt_ps": dt,
           "equilibration_steps": equil}
    for key, sysm, mlp, sw in (("table", base, False, None), ("mlp_switched", base, True, 0.9),
                               ("table_switch", switched, False, None), ("mlp", switched, True, None)):
        out[key] = sg.nve_drift(sysm, mlp=mlp, switch=sw, equil=equil, steps=steps, dt=dt)
        note(f"NVE argon {key}: drift {out[key]['drift_rel_per_ns']:.2e}/ns")
    out["water_rigid"] = water_nve()
    return out


def mixture_nve(mlp, steps, dt=0.002, switch=1.0, T=300.0, chunk=500, samples=1 << 24, window="ref"):
    """NVE of the 1,048,576-atom mixture with force-continuous tables (LJ
    switched off from `switch` to the 1.2 nm cutoff, conducting-boundary
    reaction field): device-resident equilibration with velocity rescaling
    (the reference drivers' scheme, examples/argon/argon.c:286-299) up to
    dt, then `steps` NVE steps.  window "ref" evaluates the pairs of the
    reference's sorted window (runner_dopair.c:455-478, which leaves out a
    few pairs inside the cutoff at this cell size), "proj" every pair inside
    the cutoff.  With mlp the NVE steps evaluate the pairs
    with a surrogate trained on `samples` samples (half uniform in r, half
    the equilibrated configuration's pair distances); its force and energy
    error against the table pass and the FP64 oracle are taken on the
    equilibrated configuration first.  Drift: least-squares slope of
    PE(x_n) + (KE(v_{n-1/2}) + KE(v_{n+1/2}))/2, per ns, relative to |<E>|."""
    import dataclasses
    import numpy as np
    import mdcore as md
    import surrogate as sg
    import systems
    t0 = time.perf_counter()
    sysm = systems.mixture(1048576, switch=switch)
    flags = md.engine_flag_gpu | (md.engine_flag_projwindow if window == "proj" else 0)
    eng = systems.build_engine(sysm, flags)
    stages = systems.MIXTURE_EQUILIBRATION + ((dt, 400, 100),)
    eq = systems.equilibrate(eng, sysm, T=T, stages=stages,
                             log=lambda d, tc, ep: None)
    note(f"mixture equilibrated: {eq[-1]}")
    out = {"system": "mixture_1m: 1,048,576 atoms, 20 species, LJ switched from "
                     f"{switch} nm + reaction field, rc 1.2 nm, 0.6 nm cells",
           "equilibration": [{"dt_ps": d, "T_K": tc, "epot": ep} for d, tc, ep in eq], "dt_ps": dt, "steps": steps,
           "window": window,
           "pass": "surrogate" if mlp else "table"}
    if mlp:
        s_eq = dataclasses.replace(sysm, x=eng.positions(), v=eng.velocities())
        eng.close()
        eng = systems.build_engine(s_eq, flags)
        eng.setdt(dt)
        eng.shuffle()
        eng.nonbond()
        f_table = eng.forces()
        r_table = eng.gpu_results()
        t1 = time.perf_counter()
        dom = sg.pair_domains(eng, s_eq)
        S = sg.make_samples(eng, s_eq, samples, dom[0], dom[1], configs=[s_eq.x], seed=5)
        eng.mlp_init(*dom, seed=7)
        hist = eng.mlp_train(*S, steps=MLP_TRAIN_STEPS, batch_tiles=1024, lr0=1e-2, lr1=1e-5, seed=11)
        note(f"surrogate trained: loss {hist[:20].mean():.3e} -> {hist[-50:].mean():.3e}")
        eng.setflags(md.engine_flag_mlp, 0)
        eng.nonbond()
        f_mlp = eng.forces()
        r_mlp = eng.gpu_results()
        if str(ROOT / "oracle") not in sys.path:
            sys.path.insert(0, str(ROOT / "oracle"))
        import pyoracle as po
        orc = po.from_system(s_eq)
        o = orc.nonbond(po.OR_WINDOW_PROJ if window == "proj" else po.OR_WINDOW_REF, host_cpus()["threads"])
        orc.close()
        out["training"] = {"samples": int(len(S[0])), "steps": MLP_TRAIN_STEPS, "batch": 16384,
                           "loss_first": float(hist[:20].mean()), "loss_last": float(hist[-50:].mean()),
                           "s": round(time.perf_counter() - t1, 1)}
        out["force_rel_rmse"] = {"mlp_vs_table": rel_rmse(f_mlp, f_table), "mlp_vs_fp64": rel_rmse(f_mlp, o["f"]),
                                 "table_vs_fp64": rel_rmse(f_table, o["f"])}
        out["epot_rel_err"] = {"mlp_vs_fp64": abs(r_mlp["epot"] - o["epot"]) / abs(o["epot"]),
                               "table_vs_fp64": abs(r_table["epot"] - o["epot"]) / abs(o["epot"])}
        out["pairs_equal"] = r_mlp["pairs"] == r_table["pairs"] == o["pairs"]
        note(f"surrogate error: {out['force_rel_rmse']}")
    else:
        eng.setdt(dt)
    eps, eks = [], []
    t2 = time.perf_counter()
    for k in range(0, steps, chunk):
        ep, ek = eng.gpu_steps(min(chunk, steps - k))
        eps.append(ep)
        eks.append(ek)
        note(f"mixture NVE ({out['pass']}): {k + len(ep)} steps, E {ep[-1] + ek[-1]:.6e}")
    wall = time.perf_counter() - t2
    eng.close()
    ep, ek = np.concatenate(eps), np.concatenate(eks)
    et = ep[1:] + 0.5 * (ek[:-1] + ek[1:])
    rel, absd = sg.drift_per_ns(et, dt)
    h = len(et) // 2
    out.update({"drift_rel_per_ns": rel, "drift_kJ_mol_per_ns": absd,
                "drift_rel_per_ns_halves": [sg.drift_per_ns(et[:h], dt)[0], sg.drift_per_ns(et[h:], dt)[0]],
                "e_mean": float(et.mean()), "e_std": float(et.std()),
                "temperature_K": float(2 * ek.mean() / ((3 * sysm.N - 3) * systems.KB)),
                "ms_per_step": round(wall / steps * 1e3, 3),
                "ns_per_day": dt * 1e-3 * 86400.0 / (wall / steps), "wall_s": round(time.perf_counter() - t0, 1)})
    return out


def md_step_section(nsteps, host_steps=5, dt=0.002):
    """ms per whole MD step of the 1,048,576-atom argon crystal (fcc 64^3,
    100 K): host-driven engine_step against engine_gpu_steps, with 1.0 nm
    cells and re-sorting every step, and with pseudo-Verlet sort reuse in
    1.13 nm cells (cell option 1.1 nm: 0.13 nm skin), and those cells
    re-sorted every step."""
    out = md_step_run(nsteps, host_steps, dt, 1.0, False)
    note(f"md_step argon 1M: {out['resident_ms_per_step']} ms per resident step")
    out["sort_reuse"] = md_step_run(nsteps, 0, dt, 1.1, True)
    out["same_cells_no_reuse"] = md_step_run(nsteps, 0, dt, 1.1, False)
    return out


def md_step_run(nsteps, host_steps, dt, cell, verlet):
    """The resident time is the wall time of one engine_gpu_steps call over
    nsteps steps, including its one upload and one download; a shorter call
    separates the per-step cost from the per-call one."""
    import dataclasses
    import numpy as np
    import mdcore as md
    import systems
    sysm = dataclasses.replace(systems.argon(64, T=100.0), L=np.full(3, cell))
    flags = md.engine_flag_gpu | (md.engine_flag_verlet_pseudo if verlet else 0)
    eng = systems.build_engine(sysm, flags)
    ncell = int(np.prod(np.floor(sysm.dim / sysm.L)))
    eng.setdt(dt)
    eng.step()
    host_ms = None
    if host_steps > 0:
        t0 = time.perf_counter()
        for _ in range(host_steps):
            eng.step()
        host_ms = (time.perf_counter() - t0) / host_steps * 1e3
    eng.gpu_steps(3)
    t0 = time.perf_counter()
    eng.gpu_steps(10)
    t_short = time.perf_counter() - t0
    t0 = time.perf_counter()
    eng.gpu_steps(nsteps)
    t_long = time.perf_counter() - t0
    rebuilds = None
    if verlet:
        # rebuild steps of one more call, counted step by step
        rebuilds = 0
        for _ in range(20):
            eng.gpu_steps(1)
            rebuilds += eng.maxdx(ncell)[0] == 0.0
    eng.close()
    res_ms = t_long / nsteps * 1e3
    # two call lengths separate the per-step cost from the per-call upload/download
    step_ms = (t_long - t_short) / (nsteps - 10) * 1e3 if nsteps > 10 else res_ms
    call_ms = t_short * 1e3 - 10 * step_ms
    h = sysm.dim[0] / np.floor(sysm.dim[0] / cell)
    out = {"system": f"argon_1m: 1,048,576 LJ atoms, fcc 64^3 at 100 K, rc 1.0 nm, {ncell} cells of {h:.3f} nm",
           "dt_ps": dt, "steps": nsteps, "pseudo_verlet": verlet,
           "resident_ms_per_step": round(res_ms, 3),
           "resident_ms_per_step_marginal": round(step_ms, 3), "resident_ms_per_call": round(call_ms, 1),
           "ns_per_day_resident": dt * 1e-3 * 86400.0 / (res_ms * 1e-3),
           "ns_per_day_resident_marginal": dt * 1e-3 * 86400.0 / (step_ms * 1e-3)}
    if host_ms is not None:
        out["host_driven_ms_per_step"] = round(host_ms, 3)
        out["ns_per_day_host_driven"] = dt * 1e-3 * 86400.0 / (host_ms * 1e-3)
    if verlet:
        out["skin_nm"] = round(h - sysm.cutoff, 4)
        out["rebuilds_in_20_steps"] = int(rebuilds)
    return out


def water_nve(dt=0.001, equil=3000, steps=2000, rescale_every=25, T=300.0):
    """Rigid water (P-SHAKE on the device) through engine_gpu_steps: 33,333
    molecules, 1.0 nm cells with the reference window.  The lattice start releases
    enough energy to heat the box to ~1100 K, so `equil` steps rescale the
    velocities to T every `rescale_every` steps (engine_setvelocities, as the
    reference's drivers rescale p->v); then `steps` NVE steps give the energy
    drift, and the largest constraint residue is taken at the end."""
    import dataclasses
    import numpy as np
    import mdcore as md
    import surrogate as sg
    import systems
    sysm = systems.water(rigid=True, cell=1.0)
    sysm = dataclasses.replace(sysm, v=md.gen_maxwell(sysm.masses[sysm.types], systems.KB * T, 11))
    c = np.array(sysm.rigid)
    dof = 3 * sysm.N - len(c) - 3
    eng = systems.build_engine(sysm, md.engine_flag_gpu)
    eng.setdt(dt)
    t0 = time.perf_counter()
    for _ in range(equil // rescale_every):
        _, ek = eng.gpu_steps(rescale_every)
        eng.set_velocities(eng.velocities() * np.sqrt(T / (2 * ek[-1] / (dof * systems.KB))))
    ep, ek = eng.gpu_steps(steps)
    wall = time.perf_counter() - t0
    x = eng.positions()
    eng.close()
    i, j, d = c[:, 0].astype(int), c[:, 1].astype(int), c[:, 2]
    dv = x[i] - x[j]
    dv -= sysm.dim * np.round(dv / sysm.dim)
    res = np.abs(d * d - (dv * dv).sum(1)).max()
    et = ep + ek
    rel, absd = sg.drift_per_ns(et, dt)
    return {"system": "water_100k rigid: 33,333 molecules (99,999 atoms), 10 nm box, LJ+RF, rc 1.0 nm, "
                      "1.0 nm cells, reference window, P-SHAKE",
            "dt_ps": dt, "equilibration_steps": equil, "rescale_every": rescale_every, "steps": steps,
            "drift_rel_per_ns": rel, "drift_kJ_mol_per_ns": absd, "drift_kJ_mol_per_ns_per_dof": absd / dof,
            "e_mean": float(et.mean()), "e_std": float(et.std()),
            "temperature_K": float(2 * ek.mean() / (dof * systems.KB)),
            "max_constraint_residue_nm2": float(res), "ms_per_step": round(wall / (equil + steps) * 1e3, 3)}


def copy_calibration(mib):
    """A plain streaming copy (torch elementwise kernel, float32): each launch
    reads and writes exactly mib MiB."""
    import torch
    n = mib * (1 << 20) // 4
    a = torch.rand(n, device="cuda:0")
    b = torch.empty_like(a)
    for _ in range(5):
        torch.mul(a, 1.0, out=b)
    torch.cuda.synchronize()
    print(json.dumps({"copy_calibration_bytes": n * 4, "launches": 5}), flush=True)


def eng_cells(sysm):
    import numpy as np
    return int(np.prod(np.floor(sysm.dim / sysm.L)))


def pair_kernel_name(sysm, md):
    """The pair kernel each upload picks: the half shell (with k_fold) for
    engine_gpu_half_mean or more atoms per cell on average in fewer cells
    than the device has CUs, otherwise the full shell with paired visits,
    followed by k_fold when the owners are split (that occupancy in fewer
    than 32 cells per CU); the 5-wave build from 64 cells per CU."""
    cells = eng_cells(sysm)
    wide = cells >= 64 * MI355X_CUS
    dense = sysm.N >= md.engine_gpu_half_mean * cells
    if dense and cells < MI355X_CUS:
        return ("k_nonbond_half_wide" if wide else "k_nonbond_half") + " + k_fold"
    name = "k_nonbond_paired_wide" if wide else "k_nonbond_paired"
    if dense and cells < 32 * MI355X_CUS:
        name += " + k_fold"
    return name


if __name__ == "__main__":
    main()
