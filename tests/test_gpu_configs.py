"""The benchmark configurations at their own sizes, on the GPU.

C2  cold uniform sphere, N = 2^20: reference mode (NEXT's FP64 algorithm)
    at theta 0.3 and 0.7 against the oracle's reference-order walk on 4,096
    sampled targets (NEXT fixes theta at step.h:32; the oracle drives
    bhAccel's theta, octree.h:133-179), and the FP32 physics walk at the
    setting that reaches a median force error of 1e-4 against the FP64 direct
    sum, with its energy drift over 200 steps.
C3  two Plummer clusters of 2^21 bodies each: keys bit-exact for every body,
    reference-mode accelerations on sampled targets against the oracle, and
    the FP32 physics walk's force error against the oracle's FP64 direct sum
    on sampled targets.
C4  512 Plummer spheres (N log-uniform in [1024, 65536]) in one context:
    sampled systems equal their runs alone, and the force error against the
    FP64 direct sum (per system, all bodies).
C5  a request stream with N log-uniform in [4096, 2^20] and 50-80 steps,
    arriving while others run: every result equals the same run alone.

Tolerances: integer work bit-exact; reference-mode accelerations median
<= 1e-12 and 99th percentile <= 1e-10 relative; physics-mode force errors
as stated per test (they measure the multipole truncation, not a
disagreement with the oracle's arithmetic).
"""
import time

import numpy as np
import pytest

import oracle_lib as orc
from next_amd import ic
from next_amd.nbx import MODE_PHYSICS, MODE_REFERENCE, PREC_FP64, PREC_MIXED, Engine, Server

pytestmark = pytest.mark.gpu

MED_TOL, P99_TOL = 1e-12, 1e-10


def rel_err(a, b):
    d = np.linalg.norm(b, axis=1)
    return np.linalg.norm(a - b, axis=1) / np.where(d > 0, d, 1.0)


def sample(n, k, seed=0):
    return np.random.Generator(np.random.PCG64(seed)).choice(n, k, replace=False).astype(np.int64)


def assert_ref_close(a, b, what):
    e = rel_err(a, b)
    med, p99 = float(np.median(e)), float(np.percentile(e, 99))
    print(f"{what}: median rel err {med:.3e}, p99 {p99:.3e}, max {e.max():.3e}", flush=True)
    assert med <= MED_TOL and p99 <= P99_TOL, (what, med, p99)


@pytest.fixture(scope="module")
def cold1m():
    return ic.uniform_sphere(1 << 20, seed=3)


@pytest.mark.parametrize("theta", [0.3, 0.7])
def test_c2_reference_mode_theta(cold1m, theta):
    p = cold1m
    with Engine(theta=theta) as e:
        e.upload(p)
        a = e.accel()
    tg = sample(len(p), 4096)
    ao = orc.accel_subset(p.x, p.y, p.z, p.m, p.type, theta, tg)
    assert_ref_close(a[tg], ao, f"C2 cold sphere 1M theta={theta}")


def test_c2_fp32_walk_force_error_and_energy():
    """The mixed-precision leg of C2: FP32 walk, FP64 state, eps 0.005, at the
    opening angle (bmax criterion) that reaches a median force error of 1e-4
    against the FP64 direct sum; |dE/E| over 200 steps of dt 5e-4."""
    p = ic.uniform_sphere(1 << 20, seed=2)
    with Engine(mode=MODE_PHYSICS, theta=0.32, eps=0.005, precision=PREC_MIXED, leaf_size=16, group_size=64,
                adaptive_dt=False) as e:
        e.upload(p)
        a = e.accel()
        ad, _ = e.direct()
        err = rel_err(a, ad)
        med, p99 = float(np.median(err)), float(np.percentile(err, 99))
        print(f"C2 FP32 walk theta 0.32: median {med:.3e} p99 {p99:.3e}", flush=True)
        assert med <= 1e-4 and p99 <= 1e-3
        e0 = e.energy(direct=True)
        e.run(5e-4, 200)
        e1 = e.energy(direct=True)
    E0, E1 = e0[0] + e0[1], e1[0] + e1[1]
    drift = abs(E1 - E0) / abs(E0)
    print(f"C2 FP32 walk: |dE/E| over 200 steps {drift:.2e}", flush=True)
    assert drift <= 1e-4


@pytest.fixture(scope="module")
def merger():
    return ic.two_cluster(1 << 21, seed=3)


def test_c3_keys_and_reference_accel(merger):
    p = merger
    with Engine(theta=0.5) as e:
        e.upload(p)
        root, hi, lo = e.debug_keys()
        a = e.accel()
    ohi, olo = orc.keys(p.x, p.y, p.z, levels=42)
    assert np.array_equal(root, orc.root(p.x, p.y, p.z))
    assert np.array_equal(hi, ohi) and np.array_equal(lo, olo)
    tg = sample(len(p), 4096)
    ao = orc.accel_subset(p.x, p.y, p.z, p.m, p.type, 0.5, tg)
    assert_ref_close(a[tg], ao, "C3 two clusters 4M theta=0.5")


def test_c3_fp32_walk_force_error(merger):
    p = merger
    with Engine(mode=MODE_PHYSICS, theta=0.5, eps=0.01, precision=PREC_MIXED, leaf_size=16, group_size=64,
                adaptive_dt=False) as e:
        e.upload(p)
        a = e.accel()
    tg = sample(len(p), 2048, seed=1)
    ad, _ = orc.direct_subset(p.x, p.y, p.z, p.m, 0.01, tg)
    err = rel_err(a[tg], ad)
    med, p99 = float(np.median(err)), float(np.percentile(err, 99))
    print(f"C3 FP32 walk theta 0.5: median {med:.3e} p99 {p99:.3e}", flush=True)
    assert med <= 5e-4 and p99 <= 5e-3


@pytest.fixture(scope="module")
def ensemble():
    sizes = ic.ensemble_sizes(512)
    return [ic.plummer(n, seed=i, virial=None) for i, n in enumerate(sizes)]


@pytest.mark.parametrize("rebuild_every", [1, 4])
def test_c4_ensemble_systems_equal_alone(ensemble, rebuild_every):
    """All 512 systems step together; four sampled systems (the smallest, the
    largest and two others) equal their runs alone, bit for bit, with the
    tree rebuilt every step or every 4th step and refitted between."""
    sizes = [len(p) for p in ensemble]
    picks = sorted({int(np.argmin(sizes)), int(np.argmax(sizes)), 7, 300})
    kw = dict(mode=MODE_PHYSICS, theta=0.5, eps=0.01, precision=PREC_MIXED, leaf_size=16, group_size=64,
              adaptive_dt=False, rebuild_every=rebuild_every)
    with Engine(**kw) as e:
        e.upload_batch(ensemble)
        e.run(1e-3, 5)
        got = {s: e.download_system(s) for s in picks}
    for s in picks:
        with Engine(**kw) as e:
            e.upload(ensemble[s])
            e.run(1e-3, 5)
            assert np.array_equal(got[s], e.download()), s


@pytest.mark.parametrize("theta,alpha,med_tol", [(0.5, 0.0, 3e-4), (0.8, 5e-4, 1e-4)])
def test_c4_ensemble_force_error(ensemble, theta, alpha, med_tol):
    """Force error of the batched FP32 walk against the FP64 direct sum of
    every system: the geometric criterion at theta 0.5, and the relative
    criterion (alpha, behind a geometric guard) at the target accuracy."""
    with Engine(mode=MODE_PHYSICS, theta=theta, eps=0.01, precision=PREC_MIXED, leaf_size=16, group_size=64,
                adaptive_dt=False, alpha=alpha) as e:
        e.upload_batch(ensemble)
        if alpha:
            e.run(1e-3, 1)            # the relative criterion needs the previous accelerations
        a = e.accel()
        ad, _ = e.direct()
    err = rel_err(a, ad)
    med, p99 = float(np.median(err)), float(np.percentile(err, 99))
    print(f"C4 ensemble theta {theta} alpha {alpha}: median {med:.3e} p99 {p99:.3e}", flush=True)
    assert med <= med_tol and p99 <= 20 * med_tol


def test_c5_stream_results_equal_alone():
    """Requests of the serving configuration's sizes (N log-uniform in
    [4096, 2^20], 50-80 steps, theta from {0.4, 0.5, 0.7}, FP32 or FP64 walk,
    Plummer or cold sphere) arrive as a Poisson stream at 200 per second and
    overlap on the device; each result equals the same run alone."""
    rng = np.random.Generator(np.random.PCG64(17))
    reqs = []
    for k in range(6):
        n = int(round(np.exp(rng.uniform(np.log(4096), np.log(1 << 20)))))
        if k == 0:
            n = 1 << 20                                # one request at the top of the range
        p = ic.plummer(n, seed=500 + k, virial=None) if k % 2 else ic.uniform_sphere(n, seed=500 + k)
        reqs.append(dict(p=p, nsteps=int(rng.integers(50, 81)), theta=float(rng.choice([0.4, 0.5, 0.7])),
                         precision=PREC_MIXED if rng.uniform() < 0.5 else PREC_FP64,
                         gap=float(rng.exponential(1 / 200.0))))
    print("C5 requests:", [(len(r["p"]), r["nsteps"], r["theta"], r["precision"]) for r in reqs], flush=True)
    with Server(budget_bytes=200 << 30) as srv:
        ids = []
        for r in reqs:
            time.sleep(r["gap"])
            ids.append(srv.submit(r["p"], r["nsteps"], 1e-3, mode=MODE_PHYSICS, precision=r["precision"],
                                  theta=r["theta"], eps=0.01, leaf_size=16))
        results = [srv.wait(i) for i in ids]
    for r, (state, q_ms, lat_ms) in zip(reqs, results):
        assert lat_ms >= q_ms >= 0
        with Engine(theta=r["theta"], mode=MODE_PHYSICS, precision=r["precision"], eps=0.01, leaf_size=16,
                    adaptive_dt=False) as e:
            e.upload(r["p"])
            e.run(1e-3, r["nsteps"])
            assert np.array_equal(state, e.download()), len(r["p"])


def test_c5_reference_mode_request_equals_alone():
    """A reference-mode (NEXT's FP64 algorithm, adaptive dt) request beside
    physics-mode ones gives NEXT's trajectory of a context of its own."""
    big = ic.uniform_sphere(300000, seed=41)
    other = ic.plummer(50000, seed=42, virial=None)
    with Server(budget_bytes=200 << 30) as srv:
        i1 = srv.submit(other, 60, 1e-3, mode=MODE_PHYSICS, precision=PREC_MIXED, theta=0.5, eps=0.01, leaf_size=16)
        i0 = srv.submit(big, 50, 1e-3, mode=MODE_REFERENCE, precision=PREC_FP64, theta=0.5, adaptive_dt=True)
        s0, _, _ = srv.wait(i0)
        srv.wait(i1)
    with Engine(theta=0.5, adaptive_dt=True) as e:
        e.upload(big)
        e.run(1e-3, 50)
        assert np.array_equal(s0, e.download())
