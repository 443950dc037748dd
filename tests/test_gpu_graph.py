"""Captured runs (nbx_set_graph): nbx_run's steps replayed from a HIP graph
whose tree builds never read counts back to the host, against the same run
step by step (bit for bit), and the fallback when a build does not fit the
sizes the graph was captured with.  The reference's loop body is
begrun.cpp:96-99 around Step (step.h:25-204)."""
import numpy as np
import pytest

from next_amd import ic
from next_amd.nbx import MODE_PHYSICS, PREC_MIXED, Engine

pytestmark = pytest.mark.gpu


def run_pair(p_or_systems, nsteps, dt, **kw):
    """The same run step by step and captured: (state, dts, timing) of each."""
    out = []
    for graph in (0, 1):
        with Engine(graph=graph, **kw) as e:
            if isinstance(p_or_systems, list):
                e.upload_batch(p_or_systems)
            else:
                e.upload(p_or_systems)
            dts = e.run(dt, nsteps)
            out.append((e.download(), dts, e.timing()))
    return out


@pytest.mark.parametrize("nsteps", [10, 11])
def test_reference_mode_captured_equals_step_by_step(nsteps):
    p = ic.plummer(16384, seed=21, type_fraction_dm=0.1)
    (s0, d0, _), (s1, d1, t1) = run_pair(p, nsteps, 1e-3, theta=0.5)
    print(f"graph steps {t1['graph_steps']} launches {t1['graph_launches']} kernels/period {t1['graph_nodes']} "
          f"period {t1['graph_period']} aborts {t1['graph_aborts']} host waits {t1['host_waits']}", flush=True)
    assert np.array_equal(d0, d1)
    assert np.array_equal(s0, s1)
    assert t1["graph_period"] == 2 and t1["graph_steps"] == 10 and t1["graph_launches"] == 5
    assert t1["graph_aborts"] == 0 and t1["graph_nodes"] > 0


@pytest.mark.parametrize("rebuild_every", [1, 4])
def test_physics_batch_captured_equals_step_by_step(rebuild_every):
    sizes = [700, 3000, 1500, 64, 5000, 900, 2048, 333]
    systems = [ic.plummer(n, seed=40 + k, virial=None) for k, n in enumerate(sizes)]
    kw = dict(mode=MODE_PHYSICS, precision=PREC_MIXED, theta=0.5, eps=0.01, leaf_size=16, group_size=64,
              adaptive_dt=False, rebuild_every=rebuild_every)
    (s0, _, _), (s1, _, t1) = run_pair(systems, 19, 1e-3, **kw)
    print(f"rebuild {rebuild_every}: graph steps {t1['graph_steps']} aborts {t1['graph_aborts']}", flush=True)
    assert np.array_equal(s0, s1)
    # period 2k: with k = 4 the first 3 steps (refits) run one by one up to a
    # rebuild, then 2 periods of 8; with k = 1, 9 periods of 2 and 1 step
    assert t1["graph_aborts"] == 0
    assert t1["graph_steps"] == (16 if rebuild_every == 4 else 18)


def test_fallback_when_a_build_does_not_fit():
    """Without headroom (graph mode 2) the bounded builds are sized exactly
    for the tree the run starts with; two bodies closing in deepen the tree,
    a build aborts, and the run is redone step by step: same result."""
    p = ic.plummer(4000, seed=5)
    pair = ic.Particles.from_columns(np.array([[3.0, 0.0, 0.0, 0.5, 0.0, 0.0, 1e-4],
                                               [3.004, 1e-4, 0.0, -0.5, 0.0, 0.0, 1e-4]]), np.zeros(2, np.int32))
    q = ic.Particles.concat([p, pair])
    results = []
    for graph in (0, 2):
        with Engine(theta=0.5, adaptive_dt=False, graph=graph) as e:
            e.upload(q)
            e.run(1e-3, 8)
            results.append((e.download(), e.timing()))
    (s0, _), (s1, t1) = results
    print(f"aborts {t1['graph_aborts']} graph steps {t1['graph_steps']}", flush=True)
    assert t1["graph_aborts"] >= 1
    assert np.array_equal(s0, s1)


def test_capture_is_reused_across_runs():
    p = ic.uniform_sphere(50000, seed=8)
    with Engine(theta=0.5, graph=1) as a, Engine(theta=0.5) as b:
        a.upload(p)
        b.upload(p)
        for _ in range(3):
            assert np.array_equal(a.run(5e-4, 4), b.run(5e-4, 4))
        t = a.timing()
        assert t["graph_steps"] == 12 and t["graph_aborts"] == 0
        assert np.array_equal(a.download(), b.download())


def test_captured_nearly_sorted_refusal_redoes_the_run():
    """Captured builds take the nearly-sorted path once two eager builds in a
    row have.  Three clusters of light bodies barely move for three tiny
    steps, then two of them trade places inside a captured run: the build's
    nearly-sorted path refuses, the run aborts and is redone step by step,
    and the result equals the same steps run one by one."""
    rng = np.random.Generator(np.random.PCG64(23))
    k = 2000
    ball = rng.normal(size=(3 * k, 3)) * 0.05
    ctr = np.repeat(np.array([[-1.0, 0.0, 0.0], [1.0, 0.0, 0.0], [0.0, 1.0, 0.3]]), k, axis=0)
    dt = 1e-3
    vel = np.repeat(np.array([[2.0 / dt, 0.0, 0.0], [-2.0 / dt, 0.0, 0.0], [0.0, 0.0, 0.0]]), k, axis=0)
    p = ic.Particles.from_columns(np.column_stack([ball + ctr, vel, np.full(3 * k, 1e-12)]),
                                  np.zeros(3 * k, np.int32))
    out = []
    for graph in (0, 1):
        with Engine(theta=0.5, adaptive_dt=False) as e:
            e.upload(p)
            e.run(1e-9, 3)                               # one by one: the bodies stay in key order
            e.set_graph(graph)
            e.reset_timing()
            e.run(dt, 2)                                 # the clusters trade places in the first step
            out.append((e.download(), e.timing()))
    (s0, _), (s1, t1) = out
    print("aborts", t1["graph_aborts"], "graph steps", t1["graph_steps"], "redone sorts", t1["sort_redos"])
    assert t1["graph_aborts"] == 1 and t1["graph_steps"] == 0
    assert np.array_equal(s0, s1)
