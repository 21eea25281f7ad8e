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
                on the host cores, one full pass of the same system
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
DT_PS = 0.002               # 2 fs time step for ns/day
METRIC = "pair-interactions/s & ns/day at 1M atoms; MLP vs table at fixed force RMSE"
TRAFFIC_FILE = ROOT / "profiles" / "k_nonbond_mixture_1m_traffic.json"


def build_workload(name):
    import systems
    if name == "mixture_1m":
        return systems.mixture(1048576), "1,048,576-atom 20-species LJ+RF mixture, rho 100/nm^3, rc 1.2 nm, 0.6 nm cells"
    if name == "water_100k":
        return systems.water(), "33,333 SPC/E-like waters (99,999 atoms), 10 nm box, LJ+RF, rc 1.0 nm, exclusions"
    if name == "argon_32k":
        return systems.argon(20, jitter=0.05), "32,000 LJ argon atoms, 11.3 nm box, rc 1.0 nm"
    raise SystemExit(f"unknown workload {name}")


def cpu_baseline(sysm, pairs_gpu, mode):
    """FP64 oracle on the host cores: full passes of the same system."""
    if str(ROOT / "oracle") not in sys.path:
        sys.path.insert(0, str(ROOT / "oracle"))
    import pyoracle as po
    try:
        ncpu = len(os.sched_getaffinity(0))
    except AttributeError:
        ncpu = os.cpu_count() or 1
    threads = max(1, min(16, ncpu))
    model = platform.processor() or "unknown"
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        for line in out.splitlines():
            if line.startswith("Model name"):
                model = line.split(":", 1)[1].strip()
    except Exception:
        pass
    orc = po.from_system(sysm)
    orc.nonbond(mode, threads)  # warm-up: first-touch of the private buffers
    t0 = time.perf_counter()
    reps = 0
    pairs = 0
    while True:
        r = orc.nonbond(mode, threads)
        pairs += r["pairs"]
        reps += 1
        if time.perf_counter() - t0 > 10.0 or reps >= 5:
            break
    dt = time.perf_counter() - t0
    orc.close()
    return {"value": pairs / dt, "unit": "pair-interactions/s", "cores": threads, "kind": "port",
            "sample": f"{reps} full FP64 nonbonded passes (sort + pair/self + exclusions) of the same system, "
                      f"{r['pairs']} pairs each, {threads} threads on {model}; pair count equals the GPU's: "
                      f"{r['pairs'] == pairs_gpu}"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--workload", default="mixture_1m", choices=["mixture_1m", "water_100k", "argon_32k"])
    ap.add_argument("--window", default="ref", choices=["ref", "proj"])
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--engine-steps", type=int, default=3, help="full engine_step calls timed for ns/day")
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
    roofline = {"bound": "mfma", "achieved": round(achieved, 3), "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s",
                "frac": round(achieved / FP32_PEAK_TFLOPS, 4), "traffic": traffic,
                "kernel": "k_nonbond", "kernel_ms": round(ms["nonbond_ms"], 4),
                "note": "FP32 vector pipe (VALU); packed FP32 disabled, so 78.65 TFLOP/s is the issue ceiling"}

    # Full engine_step (host pack, PCIe both ways, device pass, host kick-drift).
    ns_day_full = None
    if args.engine_steps > 0:
        md.check(eng.lib.engine_step(eng.e), "engine_step")  # warm-up
        t0 = time.perf_counter()
        for _ in range(args.engine_steps):
            md.check(eng.lib.engine_step(eng.e), "engine_step")
        full_s = (time.perf_counter() - t0) / args.engine_steps
        ns_day_full = DT_PS * 1e-3 * 86400.0 / full_s

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
        "ns_per_day": DT_PS * 1e-3 * 86400.0 / step_s,
        "ns_per_day_engine_step": ns_day_full,
        "kernels_ms": {k: round(v, 4) for k, v in ms.items()},
        "roofline": roofline,
        "setup_s": round(t_setup, 1),
    }
    if not args.no_cpu_baseline:
        base = cpu_baseline(sysm, pairs, 0 if args.window == "ref" else 1)
        out["cpu_baseline"] = base
        out["gpu_over_cpu"] = value / base["value"]
    eng.close()
    print(json.dumps(out), flush=True)


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


if __name__ == "__main__":
    main()
