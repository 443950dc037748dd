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
interaction, counted on the device; HBM traffic from the committed PMC
summary), the per-phase split with the HBM rate of the streaming phases, the
mixed-precision physics mode on the same workload (step time, force error
against the FP64 direct sum, energy drift, the FP32 walk's FLOP/s against
the FP32 peak), a bounded sample of the
batched-ensemble and serving workloads (physics mode), and NEXT's own CPU
code timed on this host on a bounded sample of the headline workload.

    python bench.py [--steps K] [--warmup W] [--gpus 1] [--no-extra] [--no-cpu]
"""
from __future__ import annotations

import argparse
import json
import os
import subprocess
import sys
import tempfile
import threading
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

from next_amd import ic  # noqa: E402

FP64_PEAK_TFLOPS = 78.6   # MI355X FP64 vector (and matrix) peak
FP32_PEAK_TFLOPS = 157.3  # MI355X FP32 vector peak (packed FP32; unpacked issue gives half)
HBM_PEAK_GBS = 8000.0
FLOPS_PP, FLOPS_PC = 19, 55
# The physics walk is counted by the same convention (FMA = 2, rsqrt = 1,
# potential excluded): particle-particle 3 (d) + 6 (r^2 + eps^2) + 1 + 1 (r^-2)
# + 2 (m r^-3) + 6 (accumulate) = 19; particle-cell with the traceless
# quadrupole 19 + 15 (Q d) + 5 (d.Q.d) + 2 (r^-5, r^-7) + 2 (coefficient) + 12
# = 55.  k_phys_walk as written shares some products and issues about 51.
PHYS_FLOPS_PP, PHYS_FLOPS_PC = FLOPS_PP, FLOPS_PC
REBUILD_EVERY = 4         # ensembles: full tree build every 4th step, refits between


def log(msg: str) -> None:
    print(f"bench: {msg}", file=sys.stderr, flush=True)


def cpu_threads() -> int:
    """Host threads for NEXT's CPU code: the cores this process may use (its
    affinity mask), capped by OMP_NUM_THREADS when the host sets it (the GPU
    box allots a fixed share of its cores to each GPU)."""
    try:
        n = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        n = os.cpu_count() or 1
    env = os.environ.get("OMP_NUM_THREADS", "")
    if env.isdigit() and int(env) > 0:
        n = min(n, int(env))
    return max(1, n)


def host_cpu() -> dict:
    """Model, sockets and core counts of this host (lscpu), and the cores
    this process may use."""
    info = {}
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=20).stdout
        for line in out.splitlines():
            key, _, val = line.partition(":")
            info[key.strip()] = val.strip()
    except (OSError, subprocess.SubprocessError):
        pass
    def num(key):
        v = info.get(key, "")
        return int(v) if v.isdigit() else None
    sockets, per_socket, tpc = num("Socket(s)"), num("Core(s) per socket"), num("Thread(s) per core")
    try:
        usable = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        usable = None
    return {"model": info.get("Model name"), "sockets": sockets, "cores_per_socket": per_socket,
            "threads_per_core": tpc, "physical_cores": sockets * per_socket if sockets and per_socket else None,
            "logical_cpus": os.cpu_count(), "usable_cpus": usable,
            "omp_num_threads_env": os.environ.get("OMP_NUM_THREADS")}


def _harness_time(p: ic.Particles, base_dt: float, threads: int, warmup: int, steps: int) -> dict:
    harness = os.path.join(ROOT, "oracle", "_ref", "ref_harness")
    with tempfile.TemporaryDirectory() as td:
        path = os.path.join(td, "ic.bin")
        with open(path, "wb") as f:
            np.array([len(p)], np.int64).tofile(f)
            np.concatenate([p.x, p.y, p.z, p.vx, p.vy, p.vz, p.m]).astype(np.float64).tofile(f)
            p.type.astype(np.int32).tofile(f)
        env = dict(os.environ, OMP_NUM_THREADS=str(threads))
        r = subprocess.run([harness, "time", path, str(threads), repr(base_dt), str(warmup), str(steps)],
                           capture_output=True, text=True, env=env, check=True, timeout=600)
    return json.loads(r.stdout.strip().splitlines()[-1])


def cpu_baseline(p: ic.Particles, base_dt: float, theta: float, steps: int, c1: bool = True) -> dict:
    """NEXT's own Step (oracle/_ref/ref_harness, compiled from the reference
    headers) or, where that binary is absent, the oracle restatement, timed on
    this host: 1 warm-up step and `steps` timed steps (median) of the same
    workload, OpenMP over the cores this process may use; and C1 (Plummer
    N = 16,384, base dt 1e-3) for all of its 100 steps."""
    threads = cpu_threads()
    harness = os.path.join(ROOT, "oracle", "_ref", "ref_harness")
    sample = (f"1 warm-up + {steps} timed KDK steps (median) of the full N={len(p)} workload, theta={theta}, "
              f"{threads} OpenMP threads")
    host = host_cpu()
    if os.path.exists(harness) and theta == 0.5:
        out = _harness_time(p, base_dt, threads, 1, steps)
        res = {"value": out["ms_median"], "unit": "ms per KDK step", "cores": threads, "kind": "reference",
               "sample": sample, "host": host}
        if c1:
            q = ic.plummer(16384, seed=1)
            o1 = _harness_time(q, 1e-3, threads, 0, 100)
            res["c1_plummer16k_100_steps"] = {"total_s": round(o1["ms_total"] / 1e3, 3),
                                              "ms_per_step_median": o1["ms_median"], "threads": threads}
        return res
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as orc  # the oracle: test infrastructure, used here only as the CPU baseline

    st = np.stack([p.x, p.y, p.z, p.vx, p.vy, p.vz])
    st, _ = orc.run(st, p.m, p.type, base_dt, 1, theta=theta, threads=threads)   # warm-up
    ms = []
    for _ in range(steps):
        t0 = time.perf_counter()
        st, _ = orc.run(st, p.m, p.type, base_dt, 1, theta=theta, threads=threads)
        ms.append((time.perf_counter() - t0) * 1e3)
    return {"value": float(np.median(ms)), "unit": "ms per KDK step", "cores": threads, "kind": "port",
            "sample": sample, "host": host}


def c1_plummer(steps: int = 100) -> dict:
    """The Plummer configuration (N = 16,384, seed 1, virial ratio 0.5,
    truncated at 10 a) in reference mode with NEXT's adaptive dt (base dt
    1e-3): its 100 KDK steps after 4 warm-up steps, one by one and replayed
    from a captured graph."""
    from next_amd.nbx import Engine

    p = ic.plummer(16384, seed=1)
    ms = {}
    for graph in (0, 1):
        with Engine(theta=0.5, graph=graph) as e:
            e.upload(p)
            e.run(1e-3, 4)
            e.reset_timing()
            t0 = time.perf_counter()
            e.run(1e-3, steps)
            ms[graph] = (time.perf_counter() - t0) * 1e3 / steps
            t = e.timing()
    return {"workload": "Plummer N=16384 (seed 1), reference mode FP64, theta 0.5, base dt 1e-3, adaptive dt",
            "steps": steps, "ms_per_step": round(ms[1], 4), "step_by_step_ms_per_step": round(ms[0], 4),
            "host_waits_per_step": t["host_waits"] / steps}


def ensemble(nsys: int, steps: int, warmup: int, ref_body_steps_per_s: float = None) -> dict:
    """Batched ensemble: nsys Plummer spheres (a = 1, M = 1, G = 1, eps = 0.01;
    samples of the distribution function, not rescaled to an exact virial
    ratio), N_i log-uniform in [1024, 65536] from seed i, theta = 0.5,
    dt = 1e-3, FP32 walk with FP64 state, all advanced by one set of launches
    per step, the tree rebuilt every REBUILD_EVERY steps and refitted between;
    leaf size, group size and walk list size chosen by the autotuner first."""
    from next_amd.nbx import MODE_PHYSICS, PREC_MIXED, Engine

    sizes = ic.ensemble_sizes(nsys)
    systems = [ic.plummer(n, seed=i, virial=None) for i, n in enumerate(sizes)]
    cands = [(leaf, group, lst) for leaf in (8, 16, 32) for group in (32, 64) for lst in (128, 256)]
    with Engine(theta=0.5, mode=MODE_PHYSICS, precision=PREC_MIXED, eps=0.01, leaf_size=16, adaptive_dt=False,
                rebuild_every=REBUILD_EVERY) as e:
        e.upload_batch(systems)
        best, trial_ms = e.autotune(cands, 1e-3, steps=4)   # leaves the state as it was
        # warm-up for both legs: one captured period after the steps up to the
        # next rebuild, so the graph leg captures before its timed run, which
        # then starts at a rebuild step and replays whole periods
        warm = max(warmup, REBUILD_EVERY - 1 + 2 * REBUILD_EVERY)
        e.run(1e-3, warm)
        t0 = time.perf_counter()
        e.run(1e-3, steps)
        dt_eager = time.perf_counter() - t0
        tim_eager = e.timing()
        # the same steps, from the same start, replayed from a captured graph
        # (nbx_set_graph; the systems evolve, so later steps would cost more)
        e.upload_batch(systems)
        e.set_graph(1)
        e.run(1e-3, warm)                                  # first forces and the capture
        e.reset_timing()
        t0 = time.perf_counter()
        e.run(1e-3, steps)
        dt = time.perf_counter() - t0
        tg = e.timing()
    total = int(sum(sizes))
    target = ensemble_target(systems, best, ref_body_steps_per_s=ref_body_steps_per_s)
    return {"systems": nsys, "bodies": total, "steps": steps, "ensemble_steps_per_s": steps / dt,
            "ms_per_step": dt * 1e3 / steps, "body_steps_per_s": total * steps / dt,
            "captured": {"graph_launches_per_step": tg["graph_launches"] / steps,
                         "kernels_per_step": tg["graph_nodes"] / max(tg["graph_period"], 1),
                         "host_waits_per_step": tg["host_waits"] / steps, "graph_aborts": tg["graph_aborts"],
                         "steps_from_graph": tg["graph_steps"]},
            "step_by_step_ms_per_step": dt_eager * 1e3 / steps,
            "autotuned": {"leaf_size": best[0], "group_size": best[1], "walk_list": best[2],
                          "trial_ms_per_step": {f"{lf}/{g}/{ls}": round(float(m), 3)
                                                for (lf, g, ls), m in zip(cands, trial_ms)}},
            "tree_rebuild_every": REBUILD_EVERY,
            "workload": f"{nsys} Plummer spheres (DF samples), N log-uniform [1024, 65536], theta 0.5 (bmax criterion), "
                        "eps 0.01, dt 1e-3, FP32 walk / FP64 state",
            "at_target_accuracy": target}


def ensemble_target(systems, best, theta: float = 0.8, alpha: float = 5e-4, energy_steps: int = 200,
                    ref_body_steps_per_s: float = None, group: int = 64) -> dict:
    """The same ensemble with an opening criterion whose median force error
    against the FP64 direct sum is at most 1e-4 (relative criterion alpha
    behind a geometric guard theta, groups of at most `group` targets, the
    autotuner's leaf and list sizes; alpha is the largest of 4e-4, 5e-4 and
    6e-4 that keeps the median below 1e-4 with groups of 64: 8.1e-5,
    9.3e-5, 1.04e-4): force error on 20,000 sampled bodies
    after one step, ms per step, per-system |dE/E| (FP64 direct-sum
    energies) over energy_steps steps, and the per-body throughput against
    NEXT's plain FP64 algorithm (reference mode) run on the same systems one
    at a time."""
    from next_amd.nbx import MODE_PHYSICS, PREC_MIXED, Engine

    with Engine(mode=MODE_PHYSICS, precision=PREC_MIXED, eps=0.01, theta=theta, leaf_size=best[0],
                group_size=group, adaptive_dt=False, alpha=alpha, rebuild_every=REBUILD_EVERY,
                walk_list=best[2]) as e:
        e.upload_batch(systems)
        e.run(1e-3, 1)                 # the criterion needs the previous accelerations
        a = e.accel()
        a_ref, _ = e.direct()
        total = len(a_ref)
        sub = np.random.Generator(np.random.PCG64(0)).choice(total, min(total, 20000), replace=False)
        d = np.linalg.norm(a_ref[sub], axis=1)
        err = np.linalg.norm(a[sub] - a_ref[sub], axis=1) / np.where(d > 0, d, 1.0)
        e.run(1e-3, 2)
        t0 = time.perf_counter()                 # two whole rebuild periods
        e.run(1e-3, 2 * REBUILD_EVERY)
        ms = (time.perf_counter() - t0) / (2 * REBUILD_EVERY) * 1e3
        e.upload_batch(systems)
        en0 = e.energy_systems(direct=True)
        e.run(1e-3, energy_steps)
        en1 = e.energy_systems(direct=True)
    E0, E1 = en0[:, 0] + en0[:, 1], en1[:, 0] + en1[:, 1]
    drift = np.abs(E1 - E0) / np.abs(E0)
    log("ensemble at theta %g, alpha %g: %.2f ms/step, median force error %.2e, max |dE/E| %.2e"
        % (theta, alpha, ms, float(np.median(err)), float(drift.max())))
    seq = 0.0
    with Engine(theta=0.5) as e:       # reference mode: NEXT's FP64 algorithm, adaptive dt
        for p in systems:
            e.upload(p)
            e.run(1e-3, 1)
            t0 = time.perf_counter()
            e.run(1e-3, 3)
            seq += (time.perf_counter() - t0) / 3
    log("plain FP64 path, systems one at a time: %.1f ms per ensemble step" % (seq * 1e3))
    with Engine(theta=0.5) as e:       # the same, all 512 systems batched in one context
        e.upload_batch(systems)
        e.run(1e-3, 1)
        t0 = time.perf_counter()
        e.run(1e-3, 3)
        batched = (time.perf_counter() - t0) / 3
    log("plain FP64 path, systems batched: %.1f ms per ensemble step" % (batched * 1e3))
    return {"theta": theta, "alpha": alpha, "leaf_size": best[0], "group_size": group, "walk_list": best[2],
            "median_rel_force_err": float(np.median(err)),
            "p99_rel_force_err": float(np.percentile(err, 99)), "ms_per_step": round(ms, 3),
            "body_steps_per_s": total / (ms * 1e-3),
            "energy": {"steps": energy_steps, "median_abs_dE_over_E": float(np.median(drift)),
                       "max_abs_dE_over_E": float(drift.max())},
            "plain_fp64_one_system_at_a_time": {"ms_per_ensemble_step": round(seq * 1e3, 3),
                                                "body_steps_per_s": total / seq,
                                                "path": "reference mode (NEXT's FP64 algorithm), theta 0.5"},
            "vs_plain_fp64_path": round(seq * 1e3 / ms, 2),
            "plain_fp64_batched": {"ms_per_ensemble_step": round(batched * 1e3, 3), "body_steps_per_s": total / batched,
                                   "path": "reference mode, the same systems in one context"},
            "vs_plain_fp64_batched": round(batched * 1e3 / ms, 3),
            # per-body step throughput against the single-system FP64 path on the 1M sphere (the headline)
            "vs_plain_fp64_1M_single_system": (round(total / (ms * 1e-3) / ref_body_steps_per_s, 3)
                                               if ref_body_steps_per_s else None)}


def physics(n: int, ref_ms: float, thetas=(0.32, 0.5), energy_steps: int = 200, fp64_walk: bool = False,
            leaf: int = 16, group: int = 64, walk_list: int = 128) -> dict:
    """Mixed-precision physics mode (FP32 walk, FP64 state) on the headline
    workload: cold uniform sphere of n bodies, eps 0.005, fixed dt 5e-4.  Per
    theta: ms per KDK step, force error against the FP64 direct sum on 20,000
    sampled bodies, and per-body throughput relative to the reference-mode
    step (ref_ms); |dE/E| (FP64 direct-sum energies) over energy_steps steps
    at the first theta.  fp64_walk: the same with the FP64 walk; leaf,
    group, walk_list: bucket size, targets per walk group, LDS list entries."""
    from next_amd.nbx import MODE_PHYSICS, PREC_FP64, PREC_MIXED, Engine

    walk_prec = PREC_FP64 if fp64_walk else PREC_MIXED

    p = ic.uniform_sphere(n, seed=2)
    with Engine(mode=MODE_PHYSICS, theta=0.5, eps=0.005, precision=PREC_FP64) as e:
        e.upload(p)
        a_ref, _ = e.direct()
    sub = np.random.Generator(np.random.PCG64(0)).choice(n, min(n, 20000), replace=False)
    rows, energy = [], None
    for j, theta in enumerate(thetas):
        with Engine(mode=MODE_PHYSICS, theta=theta, eps=0.005, precision=walk_prec, leaf_size=leaf, group_size=group,
                    adaptive_dt=False, walk_list=walk_list) as e:
            e.upload(p)
            a = e.accel()
            d = np.linalg.norm(a_ref[sub], axis=1)
            err = np.linalg.norm(a[sub] - a_ref[sub], axis=1) / np.where(d > 0, d, 1.0)
            e.run(5e-4, 2)
            t0 = time.perf_counter()
            e.run(5e-4, 10)
            ms = (time.perf_counter() - t0) * 1e2
            # the same with the tree rebuilt every REBUILD_EVERY steps and refitted between
            e.set_rebuild_interval(REBUILD_EVERY)
            t0 = time.perf_counter()
            e.run(5e-4, 3 * REBUILD_EVERY)
            ms_refit = (time.perf_counter() - t0) * 1e3 / (3 * REBUILD_EVERY)
            e.set_rebuild_interval(1)
            # the walk kernel alone: HIP-event time, then device-counted interactions
            e.set_profiling(1)
            e.reset_timing()
            e.run(5e-4, 3)
            tim = e.timing()
            walk_ms = tim["walk_ms"] / tim["steps"]
            e.set_profiling(2)
            e.reset_timing()
            e.run(5e-4, 1)
            cnt = e.timing()
            e.set_profiling(0)
            tflops = (PHYS_FLOPS_PP * cnt["n_pp"] + PHYS_FLOPS_PC * cnt["n_pc"]) / (walk_ms * 1e-3) / 1e12
            peak = FP64_PEAK_TFLOPS if fp64_walk else FP32_PEAK_TFLOPS
            # HBM bytes per launch from the committed PMC summary of this walk (FP32, theta 0.5)
            prof = load_profile("phys_walk_traffic.json") if theta == 0.5 and not fp64_walk else None
            rows.append({"theta": theta, "ms_per_step": round(ms, 4), "median_rel_force_err": float(np.median(err)),
                         "p99_rel_force_err": float(np.percentile(err, 99)), "body_steps_per_s": n / (ms * 1e-3),
                         "vs_reference_path": round(ref_ms / ms, 3) if ref_ms else None,
                         "ms_per_step_tree_rebuilt_every_%d" % REBUILD_EVERY: round(ms_refit, 4),
                         "walk": {"ms": round(walk_ms, 4), "pp_per_body": round(cnt["n_pp"] / n, 1),
                                  "pc_per_body": round(cnt["n_pc"] / n, 1),
                                  "interactions_per_s": (cnt["n_pp"] + cnt["n_pc"]) / (walk_ms * 1e-3),
                                  "achieved": round(tflops, 2),
                                  "peak": peak, "unit": "TFLOP/s", "frac": round(tflops / peak, 4),
                                  "flops_per_pp_pc": [PHYS_FLOPS_PP, PHYS_FLOPS_PC],
                                  "traffic": prof["bytes_per_launch"] if prof else None}})
            if j == 0 and energy_steps > 0:
                e.upload(p)
                e0 = e.energy(direct=True)
                e.run(5e-4, energy_steps)
                e1 = e.energy(direct=True)
                E0, E1 = e0[0] + e0[1], e1[0] + e1[1]
                energy = {"theta": theta, "steps": energy_steps, "E0": E0, "E1": E1,
                          "abs_dE_over_E": abs(E1 - E0) / abs(E0)}
        log("physics theta %g: %.3f ms/step, median force error %.2e" % (theta, ms, rows[-1]["median_rel_force_err"]))
    return {"workload": f"cold uniform sphere N={n}, eps 0.005, fixed dt 5e-4, leaf {leaf}, group {group}, "
                        f"walk list {walk_list}, "
                        f"{'FP64' if fp64_walk else 'FP32'} walk / FP64 state",
            "by_theta": rows, "energy": energy}


def merger(n_each: int = 1 << 21, theta: float = 0.5, steps: int = 500, sample_steps: int = 20) -> dict:
    """Two-cluster merger (two Plummer spheres of n_each bodies, a = 1, M =
    0.5 each, at (+-5, 0, 0) with bulk velocities (-+0.3, +-0.1, 0)), G = 1,
    eps 0.01, dt 1e-3, physics mode with FP32 forces and FP64 state: force
    error against the FP64 direct sum on 20,000 sampled bodies, ms per step
    over sample_steps steps, then the full run of `steps` steps (wall time and
    |dE/E| from FP64 direct-sum energies)."""
    from next_amd.nbx import MODE_PHYSICS, PREC_MIXED, Engine

    p = ic.two_cluster(n_each, seed=3)
    n = len(p)
    with Engine(mode=MODE_PHYSICS, theta=theta, eps=0.01, precision=PREC_MIXED, leaf_size=16, group_size=64,
                adaptive_dt=False) as e:
        e.upload(p)
        a_ref, _ = e.direct()
        a = e.accel()
        sub = np.random.Generator(np.random.PCG64(0)).choice(n, 20000, replace=False)
        d = np.linalg.norm(a_ref[sub], axis=1)
        err = np.linalg.norm(a[sub] - a_ref[sub], axis=1) / np.where(d > 0, d, 1.0)
        log("merger: force error median %.2e" % np.median(err))
        e.run(1e-3, 2)
        t0 = time.perf_counter()
        e.run(1e-3, sample_steps)
        ms = (time.perf_counter() - t0) * 1e3 / sample_steps
        log("merger: %.3f ms/step" % ms)
        e.upload(p)
        e0 = e.energy(direct=True)
        t0 = time.perf_counter()
        done = 0
        while done < steps:                      # in chunks, so progress is visible
            k = min(100, steps - done)
            e.run(1e-3, k)
            done += k
            log("merger: %d of %d steps" % (done, steps))
        wall = time.perf_counter() - t0
        e1 = e.energy(direct=True)
    E0, E1 = e0[0] + e0[1], e1[0] + e1[1]
    return {"workload": f"two Plummer clusters, 2 x {n_each} bodies, theta {theta}, eps 0.01, dt 1e-3, "
                        "leaf 16, group 64, FP32 walk / FP64 state",
            "median_rel_force_err": float(np.median(err)), "p99_rel_force_err": float(np.percentile(err, 99)),
            "ms_per_step": round(ms, 3), "body_steps_per_s": n / (ms * 1e-3),
            "full_run": {"steps": steps, "wall_s": round(wall, 2), "abs_dE_over_E": abs(E1 - E0) / abs(E0)}}


def serving(nreq: int, rate: float, seed: int = 5, n_max: int = 131072, steps=(20, 60)) -> dict:
    """Serving sample: Poisson arrivals at `rate` per second of independent
    Plummer or cold-sphere runs, N log-uniform in [4096, n_max], step counts
    uniform in `steps`, theta from {0.4, 0.5, 0.7}, FP32 or FP64 walk, 200 GB
    budget."""
    from next_amd.nbx import MODE_PHYSICS, PREC_FP64, PREC_MIXED, Server

    rng = np.random.Generator(np.random.PCG64(seed))
    reqs = []
    for k in range(nreq):
        n = int(round(np.exp(rng.uniform(np.log(4096), np.log(n_max)))))
        p = (ic.plummer(n, seed=1000 + k, virial=None) if rng.uniform() < 0.5
             else ic.uniform_sphere(n, seed=1000 + k))
        reqs.append(dict(p=p, nsteps=int(rng.integers(steps[0], steps[1] + 1)), theta=float(rng.choice([0.4, 0.5, 0.7])),
                         precision=PREC_MIXED if rng.uniform() < 0.5 else PREC_FP64,
                         gap=float(rng.exponential(1.0 / rate))))
        if (k + 1) % 50 == 0:
            log("serving at %g/s: %d of %d requests generated" % (rate, k + 1, nreq))
    lat, ids = [], []
    peak = {"device_bytes": 0, "bytes_in_use": 0}
    with Server(budget_bytes=200 << 30) as srv:
        done = threading.Event()

        def watch():                     # device memory held by the batches, and the admission estimate
            while not done.is_set():
                cur = srv.stats()
                for key in peak:
                    peak[key] = max(peak[key], cur[key])
                time.sleep(0.005)

        watcher = threading.Thread(target=watch)
        watcher.start()
        t0 = time.perf_counter()
        for r in reqs:
            time.sleep(r["gap"])
            ids.append(srv.submit(r["p"], r["nsteps"], 1e-3, mode=MODE_PHYSICS, precision=r["precision"],
                                  theta=r["theta"], eps=0.01, leaf_size=16))
        for k, i in enumerate(ids):
            _, _, l_ms = srv.wait(i)
            lat.append(l_ms)
            if (k + 1) % 25 == 0:
                log("serving at %g/s: %d of %d requests done" % (rate, k + 1, nreq))
        wall = time.perf_counter() - t0
        done.set()
        watcher.join()
        st = srv.stats()
    return {"requests": nreq, "arrival_rate_per_s": rate, "p50_latency_ms": float(np.percentile(lat, 50)),
            "p99_latency_ms": float(np.percentile(lat, 99)), "body_steps_per_s": st["body_steps"] / wall,
            "batches": st["batches"], "peak_device_gb": round(peak["device_bytes"] / 1e9, 3),
            "peak_admitted_estimate_gb": round(peak["bytes_in_use"] / 1e9, 3),
            "workload": f"Poisson arrivals; Plummer (DF sample) or cold sphere, N log-uniform [4096, {n_max}], "
                        f"{steps[0]}-{steps[1]} steps, theta {{0.4, 0.5, 0.7}}, FP32/FP64 walk, dt 1e-3, eps 0.01, "
                        "200 GB budget"}


# Algorithmic HBM bytes per body of the step's phases (state: x y z vx vy
# vz m as FP64, type id sys as int32; keys 8 B, sort values 4 B):
#   kick + drift: read x..vz, a, sys (76 B), write x..vz (48 B);
#   bbox + keys: bbox reads x y z (24 B); keys read x y z sys (28 B) and write
#     key and index (12 B);
#   sort + gather: the digit count reads the keys (8 B); each of the P radix
#     passes reads keys and values (12 B) and writes them (12 B); the gather
#     reads the permutation (4 B) and the 68-byte state and writes it (68 B).
#     P = 3 * depth / 8 rounded up, at most 8 (the sort covers the digits
#     the last build's depth needs).  A build sorted by the nearly-sorted
#     path instead reads the keys for their descents (8 B), then reads them
#     (8 B) and writes keys and permutation (12 B); one redone after it
#     pays for both;
#   node layout: lcp reads key and system (12 B) and writes lcp and node
#     count (8 B); the count's scan reads it twice and writes it (12 B); the
#     layout kernel reads lcp, preorder offset, x y z m and system (44 B) and
#     writes the leaf's 80-byte record and 32-byte walk head (112 B), plus
#     two 4-byte list entries per internal node;
#   upward pass: per internal node the two list entries (8 B), the first
#     body's system (4 B) and the record and head it writes (112 + 32 B); per
#     node but the roots, the 64-byte sector of its record its parent reads.
HBM_FIXED_BYTES = {"kick_ms": 124, "keys_ms": 64}


def phase_bytes(n: int, nodes: int, depth: int, builds: int = 0, patched: int = 0, redos: int = 0) -> dict:
    """Bytes per body of each phase (the comment above) for a tree of
    `nodes` nodes (leaves included) and `depth` levels over n bodies; the
    sort's averaged over `builds` builds of which `patched` took the
    nearly-sorted path and `redos` of those were redone with the radix sort
    (builds 0: every build radix-sorted)."""
    passes = min(8, -(-3 * depth // 8)) if 0 < depth < 21 else 8
    internal = max(nodes - n, 0)
    per_int = internal / n if n else 0.0
    out = dict(HBM_FIXED_BYTES)
    radix, patch = 8 + 24 * passes, 8 + 8 + 12
    if builds > 0:
        sort = ((builds - patched) * radix + patched * patch + redos * radix) / builds
    else:
        sort = radix
    out["sort_ms"] = round(sort + 140, 1)
    out["build_ms"] = round(12 + 8 + 12 + 44 + 112 + 8 * per_int, 1)
    out["upward_ms"] = round(per_int * (8 + 4 + 144) + max(nodes - 1, 0) / n * 64, 1) if n else 0.0
    return out


def phase_rooflines(n: int, tim: dict, k: int) -> dict:
    per = phase_bytes(n, int(tim.get("nodes", 0)), int(tim.get("depth", 0)), int(tim.get("steps", 0)),
                      int(tim.get("sort_patched", 0)), int(tim.get("sort_redos", 0)))
    out = {}
    for key, per_body in per.items():
        if key not in tim:
            continue
        ms = tim[key] / k
        gbs = per_body * n / (ms * 1e-3) / 1e9 if ms > 0 else 0.0
        out[key[:-3]] = {"bytes_per_body": per_body, "GB/s": round(gbs, 1), "frac": round(gbs / HBM_PEAK_GBS, 4)}
    return out


def load_profile(name: str) -> dict | None:
    """A committed rocprofv3 summary of a walk kernel (HBM traffic or VALU
    issue counters), or None."""
    path = os.path.join(ROOT, "profiles", name)
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
    ap.add_argument("--no-extra", action="store_true", help="skip the C1, physics, ensemble and serving samples")
    ap.add_argument("--no-graph", action="store_true", help="skip the captured (graph-replayed) run of the headline")
    ap.add_argument("--ensemble-systems", type=int, default=512)
    ap.add_argument("--serve-requests", type=int, default=24)
    ap.add_argument("--only", choices=["physics", "ensemble", "serving", "merger"],
                    help="run just this sample (for profiling) and print its JSON")
    ap.add_argument("--thetas", default="0.32,0.5", help="physics sample: opening angles (comma-separated)")
    ap.add_argument("--walk-precision", choices=["fp32", "fp64"], default="fp32", help="physics sample: walk precision")
    ap.add_argument("--leaf", type=int, default=16, help="physics sample: bucket size")
    ap.add_argument("--group", type=int, default=64, help="physics sample: targets per walk group")
    ap.add_argument("--walk-list", type=int, default=128, help="physics sample: LDS list entries of the walk")
    ap.add_argument("--serve-config", action="store_true",
                    help="serving sample with the configuration's distributions (N up to 2^20, 50-500 steps)")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="only copy a 256 MiB device buffer 3 times (scale for HBM byte counters) and exit")
    args = ap.parse_args()
    if args.calibrate_copy:
        import torch
        a = torch.ones(1 << 26, device="cuda")
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes": a.numel() * 4, "copies": 3}))
        return 0

    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print("bench.py: this engine runs one particle system per GPU and is measured on one GPU only "
              f"(--gpus {args.gpus}, WORLD_SIZE {world} requested); multi-GPU sharding is not built.",
              file=sys.stderr)
        return 2

    from next_amd.nbx import Engine

    if args.only:
        if args.only == "physics":
            res = physics(args.n, None, thetas=[float(t) for t in args.thetas.split(",")], energy_steps=0,
                          fp64_walk=args.walk_precision == "fp64", leaf=args.leaf, group=args.group,
                          walk_list=args.walk_list)
        elif args.only == "ensemble":
            res = ensemble(args.ensemble_systems, steps=16, warmup=1)
        elif args.only == "merger":
            res = merger()
        else:
            full = dict(n_max=1 << 20, steps=(50, 500)) if args.serve_config else {}
            res = [serving(args.serve_requests, rate, **full) for rate in (50.0, 200.0)]
        print(json.dumps(res))
        return 0

    p = ic.uniform_sphere(args.n, seed=2)
    eng = Engine(theta=args.theta)
    eng.upload(p)
    eng.run(args.base_dt, args.warmup)

    # the value: K steps one by one (nbx_run's default: each tree build reads
    # its node counts back to the host while the layout kernel runs)
    t0 = time.perf_counter()
    eng.run(args.base_dt, args.steps)   # synchronous: returns after the stream drains
    wall_ms = (time.perf_counter() - t0) * 1e3
    # beside it: K steps replayed from a captured graph (nbx_set_graph; bit
    # for bit the same arithmetic, one host wait per run), and the same K
    # steps one by one.  Captured builds take the nearly-sorted path once two
    # eager builds in a row have, so that leg runs 2 steps one by one before
    # the capture, and its window starts 2 steps later.  The collapse makes
    # later steps dearer, so every leg starts from the uploaded state.
    captured_ms, tg, window_ms = None, None, None
    cap_start = 2 + max(args.warmup, 2)
    if not args.no_graph:
        eng.upload(p)
        eng.run(args.base_dt, 2)
        eng.set_graph(1)
        eng.run(args.base_dt, max(args.warmup, 2))   # the capture
        eng.reset_timing()
        t0 = time.perf_counter()
        eng.run(args.base_dt, args.steps)
        captured_ms = (time.perf_counter() - t0) * 1e3
        tg = eng.timing()
        eng.set_graph(0)
        eng.upload(p)
        eng.run(args.base_dt, cap_start)
        t0 = time.perf_counter()
        eng.run(args.base_dt, args.steps)
        window_ms = (time.perf_counter() - t0) * 1e3

    eng.upload(p)
    eng.run(args.base_dt, args.warmup)
    eng.set_profiling(1)      # HIP events around each phase, on the engine's stream (one by one)
    eng.reset_timing()
    eng.run(args.base_dt, args.steps)
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
    traffic = load_profile("walk_traffic.json")
    valu = load_profile("walk_valu.json")

    out = {
        "metric": "ms per KDK step at N=1M, \u03b8=0.5; ensemble steps/s + p50/p99 latency",
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
                     "n_pp": cnt["n_pp"], "n_pc": cnt["n_pc"],
                     "valu_issue": ({"busy_frac": valu["derived"]["valu_busy_frac"],
                                     "instructions_per_slot": valu["derived"]["valu_instructions_per_slot"],
                                     "salu_per_cu_cycle": valu["derived"]["salu_instructions_per_cu_cycle"],
                                     "source": "profiles/walk_valu.json"} if valu else None)},
        "interactions_per_s": (cnt["n_pp"] + cnt["n_pc"]) / (walk_ms * 1e-3) if walk_ms > 0 else 0.0,
        "walk_counters": {"n_open": cnt["n_open"], "traversal_wave_steps": cnt["n_iter"],
                          "traversal_lane_efficiency": round((cnt["n_pp"] + cnt["n_pc"] + cnt["n_open"])
                                                             / max(64 * cnt["n_iter"], 1), 4),
                          "evaluation_wave_steps": cnt["n_eval"],
                          "evaluation_lane_efficiency": round((cnt["n_pp"] + cnt["n_pc"])
                                                              / max(64 * cnt["n_eval"], 1), 4)},
        "phases_ms": {key: round(tim[key] / k, 4) for key in
                      ("kick_ms", "keys_ms", "sort_ms", "build_ms", "upward_ms", "walk_ms", "total_ms")},
        "captured": ({"ms_per_step": round(captured_ms / args.steps, 4),
                      "step_by_step_ms_per_step": round(window_ms / args.steps, 4),
                      "first_step": cap_start,
                      "graph_launches_per_step": tg["graph_launches"] / args.steps,
                      "kernels_per_step": tg["graph_nodes"] / max(tg["graph_period"], 1),
                      "host_waits_per_step": tg["host_waits"] / args.steps, "graph_aborts": tg["graph_aborts"],
                      "steps_from_graph": tg["graph_steps"]} if tg else None),
        "tree": {"nodes": tim["nodes"], "depth": tim["depth"], "builds": k,
                 "builds_nearly_sorted": tim["sort_patched"], "builds_redone_radix": tim["sort_redos"]},
        "phases_hbm": phase_rooflines(args.n, tim, k),
    }
    log("headline: %.3f ms/step (captured %s, one by one over the same steps %s)"
        % (ms_per_step, "%.3f" % (captured_ms / args.steps) if tg else "-",
           "%.3f" % (window_ms / args.steps) if tg else "-"))
    if not args.no_extra:
        out["c1"] = c1_plummer()
        log("C1: %.3f ms/step captured, %.3f step by step" % (out["c1"]["ms_per_step"],
                                                            out["c1"]["step_by_step_ms_per_step"]))
        out["physics"] = physics(args.n, ms_per_step)
        out["ensemble"] = ensemble(args.ensemble_systems, steps=16, warmup=1,
                                   ref_body_steps_per_s=args.n / (ms_per_step * 1e-3))
        log("ensemble: %.1f ms/step" % out["ensemble"]["ms_per_step"])
        out["serving"] = []
        full = dict(n_max=1 << 20, steps=(50, 500)) if args.serve_config else {}
        for rate in (50.0, 200.0):
            out["serving"].append(serving(args.serve_requests, rate, **full))
            log("serving at %g/s: p50 %.1f ms" % (rate, out["serving"][-1]["p50_latency_ms"]))
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(p, args.base_dt, args.theta, args.cpu_steps)
    out["child_processes_at_exit"] = reap_children()
    print(json.dumps(out))
    return 0


def reap_children() -> list:
    """Every process this one started has ended: any child still present
    (none is expected; the CPU baseline runs its harness to completion) is
    named, terminated and waited for."""
    try:
        import psutil
    except ImportError:
        return []
    left = psutil.Process().children(recursive=True)
    names = [f"{c.pid}:{c.name()}" for c in left if c.is_running()]
    for c in left:
        try:
            c.terminate()
        except psutil.Error:
            pass
    psutil.wait_procs(left, timeout=10)
    if names:
        log("child processes left at exit (terminated): " + ", ".join(names))
    return names


if __name__ == "__main__":
    sys.exit(main())
