"""GPU parity of the reference-mode path against NEXT's FP64 results.

Each test drives libnbx.so through the C ABI (next_amd.nbx) on the device and
compares with the golden vectors written by NEXT's own code and with the CPU
oracle on the same seeded inputs.

Tolerances: keys, permutation, topology, node geometry, masses, centres of
mass and quadrupoles are bit-exact (integer work, or FP64 arithmetic with the
same operation order and no libm call).  Accelerations go through cbrt/pow,
which differ between glibc and the device math library by an ulp or so: the
bar is median relative error <= 1e-12 and 99th percentile <= 1e-10.
"""
import numpy as np
import pytest

import oracle_lib as orc
from conftest import golden
from next_amd import ic
from next_amd.nbx import Engine, NbxError

pytestmark = pytest.mark.gpu

HAND = ["hand_n1", "hand_n2", "hand_n3", "hand_oct8", "hand_ties", "hand_chain", "hand_deep",
        "hand_zeromass", "hand_dm"]
TREE_CASES = HAND + ["plummer4k"]
MED_TOL, P99_TOL = 1e-12, 1e-10


def particles_of(g):
    return ic.Particles(*(np.array(g[k]) for k in ("x", "y", "z", "vx", "vy", "vz", "m", "type")))


def rel_err(a, b):
    num = np.linalg.norm(a - b, axis=1)
    den = np.linalg.norm(b, axis=1)
    den = np.where(den > 0, den, 1.0)
    return num / den


def assert_acc_close(a, b, what=""):
    e = rel_err(a, b)
    if e.size == 0:
        return
    med, p99 = float(np.median(e)), float(np.percentile(e, 99))
    print(f"{what}: median rel err {med:.3e}, p99 {p99:.3e}, max {e.max():.3e}")
    assert med <= MED_TOL and p99 <= P99_TOL, (what, med, p99)


@pytest.fixture(scope="module")
def eng():
    e = Engine(theta=0.5)
    yield e
    e.close()


@pytest.mark.parametrize("case", TREE_CASES)
def test_keys_bit_exact(eng, case):
    g = golden(case)
    eng.upload(particles_of(g))
    root, hi, lo = eng.debug_keys()
    assert np.array_equal(root, g["root"])
    assert np.array_equal(hi, g["keys_hi"])
    assert np.array_equal(lo, g["keys_lo"])


@pytest.mark.parametrize("case", TREE_CASES)
def test_tree_bit_exact(eng, case):
    g = golden(case)
    eng.upload(particles_of(g))
    t = eng.debug_tree()
    for k in ("depth", "leaf", "body", "cmask"):
        assert np.array_equal(t[k], g["node_" + k]), k
    assert np.array_equal(t["geo"], g["node_geo"])
    assert np.array_equal(t["mass"], g["node_mass"])
    assert np.array_equal(t["com"], g["node_com"])
    assert np.array_equal(t["q"], g["node_q"])
    assert np.array_equal(t["perm"], g["perm"])


@pytest.mark.parametrize("case", TREE_CASES + ["c1_plummer16k"])
def test_accel_vs_reference(case):
    g = golden(case)
    for k, th in enumerate(g["thetas"]):
        with Engine(theta=float(th)) as e:
            e.upload(particles_of(g))
            a = e.accel()
        assert_acc_close(a, g[f"acc_{k}"], f"{case} theta={th}")


def test_uniform_64k_keys_and_perm(eng):
    g = golden("uniform64k_keys")
    eng.upload(particles_of(g))
    root, hi, _ = eng.debug_keys()
    assert np.array_equal(root, g["root"])
    assert np.array_equal(hi, g["keys_hi"])
    assert np.array_equal(eng.debug_tree()["perm"], g["perm"])


def test_adaptive_dt_matches_reference():
    g = golden("c1_steps")
    with Engine() as e:
        e.upload(particles_of(g))
        dt = e.adaptive_dt(1e-3)
    assert dt == g["dts_10"][0]
    assert dt == orc.adaptive_dt(g["vx"], g["vy"], g["vz"], 1e-3)


@pytest.mark.parametrize("nsteps", [10, 100])
def test_multistep_trajectory(nsteps):
    """nbx_run (dt and KDK on the device) against NEXT's own 10/100-step run."""
    g = golden("c1_steps")
    with Engine() as e:
        e.upload(particles_of(g))
        dts = e.run(float(g["base_dt"][0]), nsteps)
        st = e.download()
    ref = g[f"state_{nsteps}"]
    dref = g[f"dts_{nsteps}"]
    assert np.max(np.abs(dts - dref) / dref) <= 1e-12
    for k in range(6):
        scale = np.max(np.abs(ref[k]))
        err = np.abs(st[k] - ref[k]) / scale
        print(f"state[{k}] after {nsteps}: max err {err.max():.3e}")
        assert np.median(err) <= 1e-13 and err.max() <= 1e-9


def test_step_api_matches_run():
    """The two-call form (adaptive_dt + step, as the driver loop makes them)
    gives the same state as nbx_run."""
    g = golden("c1_steps")
    p = particles_of(g)
    with Engine() as a, Engine() as b:
        a.upload(p)
        b.upload(p)
        dts = a.run(1e-3, 3)
        for s in range(3):
            dt = b.adaptive_dt(1e-3)
            assert dt == dts[s]
            b.step(dt)
        assert np.array_equal(a.download(), b.download())


def test_large_tree_vs_oracle():
    """Full node table at N = 262,144 (cold sphere) against the oracle's
    insertion-built tree."""
    p = ic.uniform_sphere(262144, seed=5)
    with Engine() as e:
        e.upload(p)
        t = e.debug_tree()
    o = orc.tree(p.x, p.y, p.z, p.m)
    for k in ("depth", "leaf", "body", "cmask", "perm"):
        assert np.array_equal(t[k], o[k]), k
    for k in ("geo", "mass", "com", "q"):
        assert np.array_equal(t[k], o[k]), k


def test_tree_tracks_a_deepening_collapse():
    """Each build sorts only the key digits the previous build's depth needs
    and orders the runs of bodies that agree on them in place; while a cold
    sphere collapses and its tree deepens from step to step, every step's
    tree must still be the oracle's insertion tree of the same positions,
    bit for bit."""
    p = ic.uniform_sphere(32768, seed=9)
    depths = []
    with Engine(theta=0.5) as e:
        e.upload(p)
        for steps in (1, 4, 10, 25):
            e.run(2e-2, steps)
            x, y, z, vx, vy, vz = e.download()
            t = e.debug_tree()
            o = orc.tree(x, y, z, p.m)
            for k in ("depth", "leaf", "body", "cmask", "perm"):
                assert np.array_equal(t[k], o[k]), (steps, k)
            for k in ("geo", "mass", "com", "q"):
                assert np.array_equal(t[k], o[k]), (steps, k)
            depths.append(int(t["depth"].max()))
        tm = e.timing()
    print("tree depths", depths, "builds sorted by the nearly-sorted path", tm["sort_patched"],
          "redone", tm["sort_redos"])
    assert depths[-1] >= depths[0] + 2      # the tree deepened: the runs of the truncated sorts were fixed


def test_tree_from_the_nearly_sorted_path():
    """The headline regime at 65,536 bodies (cold sphere, base dt 5e-4): after
    the first build the bodies stay in nearly the key order of the next one,
    the builds take the nearly-sorted path, and the tree after 20 steps is the
    oracle's insertion tree of the same positions, bit for bit."""
    p = ic.uniform_sphere(65536, seed=12)
    with Engine(theta=0.5) as e:
        e.upload(p)
        e.run(5e-4, 20)
        tm = e.timing()
        x, y, z, vx, vy, vz = e.download()
        t = e.debug_tree()
    print("builds sorted by the nearly-sorted path", tm["sort_patched"], "redone", tm["sort_redos"])
    assert tm["sort_patched"] - tm["sort_redos"] >= 15
    o = orc.tree(x, y, z, p.m)
    for k in ("depth", "leaf", "body", "cmask", "perm"):
        assert np.array_equal(t[k], o[k]), k
    for k in ("geo", "mass", "com", "q"):
        assert np.array_equal(t[k], o[k]), k


def test_sort_redone_when_the_bodies_moved_too_much():
    """Three clusters of light bodies (negligible forces); in one step the
    first two trade places, so in the last build's order the new keys have a
    few descents but the rest is out of order.  The nearly-sorted path must
    refuse it, the build must be redone with the radix sort, and the tree
    must be the oracle's, bit for bit."""
    rng = np.random.Generator(np.random.PCG64(21))
    k = 3000
    ball = rng.normal(size=(3 * k, 3)) * 0.05
    ctr = np.repeat(np.array([[-1.0, 0.0, 0.0], [1.0, 0.0, 0.0], [0.0, 1.0, 0.3]]), k, axis=0)
    dt = 1e-3
    vel = np.repeat(np.array([[2.0 / dt, 0.0, 0.0], [-2.0 / dt, 0.0, 0.0], [0.0, 0.0, 0.0]]), k, axis=0)
    cols = np.column_stack([ball + ctr, vel, np.full(3 * k, 1e-12)])
    p = ic.Particles.from_columns(cols, np.zeros(3 * k, np.int32))
    with Engine(theta=0.5, adaptive_dt=False) as e:
        e.upload(p)
        e.step(dt)                                   # first tree (radix sort), then the trade of places
        tm = e.timing()
        x, y, z, vx, vy, vz = e.download()
        t = e.debug_tree()
    assert abs(float(np.mean(x[:k])) - 1.0) < 0.05 and abs(float(np.mean(x[k:2 * k])) + 1.0) < 0.05
    print("patched", tm["sort_patched"], "redone", tm["sort_redos"])
    assert tm["sort_patched"] == 1 and tm["sort_redos"] == 1
    o = orc.tree(x, y, z, p.m)
    for key in ("depth", "leaf", "body", "cmask", "perm"):
        assert np.array_equal(t[key], o[key]), key
    for key in ("geo", "mass", "com", "q"):
        assert np.array_equal(t[key], o[key]), key


def test_million_body_keys_and_sampled_accel():
    """The headline workload size: 1,048,576-body cold sphere.  Keys are
    bit-exact for every body; accelerations are checked on 4,096 sampled
    targets against the oracle's reference-order walk."""
    p = ic.uniform_sphere(1 << 20, seed=3)
    with Engine(theta=0.5) as e:
        e.upload(p)
        root, hi, lo = e.debug_keys()
        a = e.accel()
    ohi, olo = orc.keys(p.x, p.y, p.z, levels=42)
    assert np.array_equal(root, orc.root(p.x, p.y, p.z))
    assert np.array_equal(hi, ohi) and np.array_equal(lo, olo)
    rng = np.random.Generator(np.random.PCG64(0))
    tg = rng.choice(len(p), 4096, replace=False)
    ao = orc.accel_subset(p.x, p.y, p.z, p.m, p.type, 0.5, tg)
    assert_acc_close(a[tg], ao, "1M cold sphere theta=0.5")


def test_edge_cases():
    with Engine() as e:
        e.upload(ic.Particles.from_columns(np.zeros((0, 7)), np.zeros(0, np.int32)))
        assert e.run(1e-3, 2).tolist() == [1e-3, 1e-3]
        one = ic.Particles.from_columns(np.array([[0.1, 0.2, 0.3, 1.0, 0.0, 0.0, 1.0]]), np.zeros(1, np.int32))
        e.upload(one)
        assert np.array_equal(e.accel(), np.zeros((1, 3)))
        dt = e.adaptive_dt(1e-3)
        assert dt == 1e-3 / 2.0
        e.step(dt)
        st = e.download()
        assert st[0, 0] == 0.1 + 1.0 * dt
        bad = ic.uniform_sphere(100, seed=1)
        bad.x[7] = np.nan
        e.upload(bad)
        with pytest.raises(NbxError) as ei:
            e.accel()
        assert ei.value.status == -2
        same = ic.uniform_sphere(50, seed=1)
        same.x[3], same.y[3], same.z[3] = same.x[9], same.y[9], same.z[9]
        e.upload(same)
        with pytest.raises(NbxError) as ei:
            e.accel()
        assert ei.value.status == -3
