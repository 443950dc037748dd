#!/usr/bin/env python3
"""Benchmark of the gravity + KDK time-step path on one MI355X.

Workload (the headline configuration): cold uniform sphere, N = 1,048,576
equal masses (total M = 1) uniform in a unit ball, at rest, G = 1, opening
angle 0.5, base dt = 5e-4 with NEXT's speed-adaptive step, in reference mode
(NEXT's FP64 algorithm, bit-compatible tree).  A "step" is one
computeAdaptiveDt + one KDK Step (begrun.cpp:97-98), with the particle state
resident in HBM.

Prints one JSON line: ms per KDK step, the walk kernel's roofline (FP64
VALU; algorithmic flops 19 per particle-particle and 55 per particle-cell
interaction, counted on the device), the per-phase split, and NEXT's own CPU
code timed on this host on a bounded sample of the same workload.

    python bench.py [--steps K] [--warmup W] [--gpus 1]
"""
from __future__ import annotations

import argparse
import json
import os
import subprocess
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

from next_amd import ic  # noqa: E402

FP64_PEAK_TFLOPS = 78.6   # MI355X FP64 vector (and matrix) peak
HBM_PEAK_GBS = 8000.0
FLOPS_PP, FLOPS_PC = 19, 55


def cpu_threads() -> int:
    n = os.cpu_count() or 1
    return max(1, min(16, n))


def cpu_baseline(p: ic.Particles, base_dt: float, theta: float, steps: int) -> dict:
    """NEXT's own Step (oracle/_ref/ref_harness, compiled from the reference
    headers) or, where that binary is absent, the oracle restatement, timed on
    this host for `steps` steps of the same workload (no warm-up)."""
    threads = cpu_threads()
    harness = os.path.join(ROOT, "oracle", "_ref", "ref_harness")
    sample = f"{steps} KDK step(s) of the full N={len(p)} workload, theta={theta}, {threads} OpenMP threads"
    if os.path.exists(harness) and theta == 0.5:
        with tempfile.TemporaryDirectory() as td:
            path = os.path.join(td, "ic.bin")
            with open(path, "wb") as f:
                np.array([len(p)], np.int64).tofile(f)
                np.concatenate([p.x, p.y, p.z, p.vx, p.vy, p.vz, p.m]).astype(np.float64).tofile(f)
                p.type.astype(np.int32).tofile(f)
            env = dict(os.environ, OMP_NUM_THREADS=str(threads))
            r = subprocess.run([harness, "time", path, str(threads), repr(base_dt), "0", str(steps)],
                               capture_output=True, text=True, env=env, check=True)
            out = json.loads(r.stdout.strip().splitlines()[-1])
        return {"value": out["ms_median"], "unit": "ms per KDK step", "cores": threads, "kind": "reference",
                "sample": sample}
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as orc  # the oracle: test infrastructure, used here only as the CPU baseline

    st = np.stack([p.x, p.y, p.z, p.vx, p.vy, p.vz])
    t0 = time.perf_counter()
    orc.run(st, p.m, p.type, base_dt, steps, theta=theta, threads=threads)
    ms = (time.perf_counter() - t0) * 1e3 / steps
    return {"value": ms, "unit": "ms per KDK step", "cores": threads, "kind": "port", "sample": sample}


def load_traffic() -> dict | None:
    path = os.path.join(ROOT, "profiles", "walk_traffic.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        return json.load(f)


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n", type=int, default=1 << 20)
    ap.add_argument("--theta", type=float, default=0.5)
    ap.add_argument("--base-dt", type=float, default=5e-4)
    ap.add_argument("--cpu-steps", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    args = ap.parse_args()

    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print("bench.py: this engine runs one particle system per GPU and is measured on one GPU only "
              f"(--gpus {args.gpus}, WORLD_SIZE {world} requested); see DESIGN.md for how it would shard.",
              file=sys.stderr)
        return 2

    from next_amd.nbx import Engine

    p = ic.uniform_sphere(args.n, seed=2)
    eng = Engine(theta=args.theta)
    eng.upload(p)
    eng.run(args.base_dt, args.warmup)

    eng.set_profiling(1)      # HIP events around each phase, on the engine's stream
    eng.reset_timing()
    t0 = time.perf_counter()
    eng.run(args.base_dt, args.steps)   # synchronous: returns after the stream drains
    wall_ms = (time.perf_counter() - t0) * 1e3
    tim = eng.timing()

    eng.set_profiling(2)      # one more step with device-side interaction counters
    eng.reset_timing()
    eng.run(args.base_dt, 1)
    cnt = eng.timing()
    eng.close()

    k = max(tim["steps"], 1)
    ms_per_step = wall_ms / args.steps
    walk_ms = tim["walk_ms"] / k
    flops = FLOPS_PP * cnt["n_pp"] + FLOPS_PC * cnt["n_pc"]
    achieved = flops / (walk_ms * 1e-3) / 1e12 if walk_ms > 0 else 0.0
    traffic = load_traffic()

    out = {
        "metric": "ms per KDK step at N=1M, theta=0.5",
        "value": round(ms_per_step, 4),
        "unit": "ms/step",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"cold uniform sphere N={args.n}, unit radius, M=1, v=0, G=1, theta={args.theta}, "
                               f"base dt={args.base_dt} with NEXT's adaptive dt; reference-mode FP64 "
                               "(NEXT's one-body-per-leaf octree, softening and quadrupole)",
                   "n": args.n, "theta": args.theta, "base_dt": args.base_dt, "mode": "reference-fp64"},
        "roofline": {"bound": "valu", "kernel": "k_walk (FP64, fused closing half kick)",
                     "achieved": round(achieved, 4), "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": round(achieved / FP64_PEAK_TFLOPS, 5),
                     "traffic": traffic["bytes_per_launch"] if traffic else None,
                     "walk_ms": round(walk_ms, 4), "flops_per_launch": flops,
                     "n_pp": cnt["n_pp"], "n_pc": cnt["n_pc"]},
        "interactions_per_s": (cnt["n_pp"] + cnt["n_pc"]) / (walk_ms * 1e-3) if walk_ms > 0 else 0.0,
        "walk_counters": {"n_open": cnt["n_open"], "wave_iterations": cnt["n_iter"],
                          "lane_efficiency": round((cnt["n_pp"] + cnt["n_pc"] + cnt["n_open"])
                                                   / max(64 * cnt["n_iter"], 1), 4)},
        "phases_ms": {key: round(tim[key] / k, 4) for key in
                      ("kick_ms", "keys_ms", "sort_ms", "build_ms", "upward_ms", "walk_ms", "total_ms")},
        "tree": {"nodes": tim["nodes"], "depth": tim["depth"]},
    }
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(p, args.base_dt, args.theta, args.cpu_steps)
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())
