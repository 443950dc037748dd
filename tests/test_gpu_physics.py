"""Physics mode on the GPU: FP64 direct sum against the oracle, the
quadrupole term against brute force, force errors of the FP32 and FP64 group
walks against the FP64 direct sum, and energy / momentum behaviour.

Tolerances are stated per test; force errors are relative, |a - a_ref| /
|a_ref| per body, reported as median and 99th percentile."""
import numpy as np
import pytest

import oracle_lib as orc
from next_amd import ic
from next_amd.nbx import MODE_PHYSICS, PREC_FP64, PREC_MIXED, Engine, NbxError

pytestmark = pytest.mark.gpu


def rel_err(a, b):
    d = np.linalg.norm(b, axis=1)
    return np.linalg.norm(a - b, axis=1) / np.where(d > 0, d, 1.0)


def phys(theta=0.5, eps=0.01, precision=PREC_MIXED, leaf=16, adaptive=False):
    return Engine(theta=theta, mode=MODE_PHYSICS, precision=precision, eps=eps, leaf_size=leaf,
                  adaptive_dt=adaptive)


def test_direct_sum_matches_oracle():
    p = ic.plummer(3000, seed=21, virial=None)
    with phys(eps=0.01) as e:
        e.upload(p)
        a, pot = e.direct()
    ao, po = orc.direct(p.x, p.y, p.z, p.m, 0.01)
    assert np.max(rel_err(a, ao)) <= 1e-12
    assert np.max(np.abs(pot - po) / np.abs(po)) <= 1e-12


def test_quadrupole_term_against_brute_force():
    """A compact cluster seen from afar is accepted as one cell: with the
    quadrupole (right sign) the error is third order in size/distance; without
    it, or with the sign flipped, it is second order."""
    rng = np.random.Generator(np.random.PCG64(4))
    cl = rng.uniform(-0.3, 0.3, (64, 3)) * np.array([1.0, 0.4, 0.2])
    far = np.array([[4.0, 2.5, 1.0]])
    pos = np.concatenate([cl, far])
    m = np.concatenate([rng.uniform(0.5, 1.5, 64), [1e-9]])
    p = ic.Particles.from_columns(np.concatenate([pos, np.zeros_like(pos), m[:, None]], 1),
                                  np.zeros(65, np.int32))
    with phys(theta=0.5, eps=1e-6, precision=PREC_FP64, leaf=4) as e:
        e.upload(p)
        a = e.accel()[-1]
        ad, _ = e.direct()
    ref = ad[-1]
    mono = (m[:64, None] * (cl - far)).sum(0) / np.linalg.norm(cl.mean(0) - far) ** 3
    err = np.linalg.norm(a - ref) / np.linalg.norm(ref)
    err_mono = np.linalg.norm(mono - ref) / np.linalg.norm(ref)
    print(f"cell far field: quadrupole err {err:.2e}, monopole-only err {err_mono:.2e}")
    assert err < 0.2 * err_mono and err < 2e-3


@pytest.mark.parametrize("precision,theta,med_tol,p99_tol", [
    (PREC_FP64, 0.3, 6e-5, 6e-4),
    (PREC_MIXED, 0.3, 6e-5, 6e-4),
    (PREC_MIXED, 0.5, 6e-4, 6e-3),
    (PREC_MIXED, 0.7, 4e-3, 3e-2),
])
def test_group_walk_force_error(precision, theta, med_tol, p99_tol):
    p = ic.plummer(32768, seed=31)
    with phys(theta=theta, eps=0.01, precision=precision) as e:
        e.upload(p)
        a = e.accel()
        ad, _ = e.direct()
    err = rel_err(a, ad)
    med, p99 = float(np.median(err)), float(np.percentile(err, 99))
    print(f"prec={precision} theta={theta}: median {med:.2e} p99 {p99:.2e}")
    assert med <= med_tol and p99 <= p99_tol


@pytest.mark.parametrize("precision", [PREC_MIXED, PREC_FP64])
def test_walk_list_sizes_agree(precision):
    """The walk's LDS list size changes only how the accepted cells and
    bodies are batched: every size gives the default's force error against
    the FP64 direct sum, and forces that differ from the default's by FP32
    (FP64) rounding of the per-flush sums only."""
    p = ic.plummer(32768, seed=33)
    got = {}
    with phys(theta=0.5, eps=0.01, precision=precision) as e:
        e.upload(p)
        ad, _ = e.direct()
        for lst in (96, 128, 192, 256):
            e.set_walk_list(lst)
            got[lst] = e.accel()
    base = got[128]
    med0 = float(np.median(rel_err(base, ad)))
    tol = 1e-5 if precision == PREC_MIXED else 1e-12
    for lst, a in got.items():
        med = float(np.median(rel_err(a, ad)))
        diff = float(np.max(rel_err(a, base)))
        print(f"list {lst}: median {med:.3e} (128: {med0:.3e}), max diff to 128 {diff:.2e}")
        assert abs(med - med0) <= 0.02 * med0 and diff <= tol


def test_energy_and_momentum_over_steps():
    p = ic.plummer(16384, seed=1)
    with phys(theta=0.5, eps=0.01, precision=PREC_MIXED) as e:
        e.upload(p)
        e0 = e.energy(direct=True)
        e.run(1e-3, 100)
        e1 = e.energy(direct=True)
    E0, E1 = e0[0] + e0[1], e1[0] + e1[1]
    dE = abs(E1 - E0) / abs(E0)
    dp = np.linalg.norm(e1[2:] - e0[2:])
    print(f"100 steps: |dE/E| {dE:.2e}, |dP| {dp:.2e}")
    assert dE <= 1e-4 and dp <= 1e-5


def test_tree_potential_close_to_direct():
    p = ic.plummer(8192, seed=5)
    with phys(theta=0.35, eps=0.01, precision=PREC_FP64) as e:
        e.upload(p)
        pt = e.potential()
        _, pd = e.direct()
    err = np.abs(pt - pd) / np.abs(pd)
    print(f"tree potential: median {np.median(err):.2e} max {err.max():.2e}")
    assert np.median(err) <= 1e-5 and err.max() <= 1e-4


def test_physics_mode_rejects_bad_settings():
    for kw in (dict(theta=1.2), dict(eps=0.0), dict(leaf=0), dict(leaf=65)):
        args = dict(theta=0.5, eps=0.01, leaf=16)
        args.update(kw)
        with pytest.raises(NbxError):
            phys(**args)


def test_relative_opening_criterion():
    """alpha > 0: after the first (geometric) evaluation, cells that pass the
    bmax/theta test must also satisfy m bmax^3 <= alpha |a|min d^5.  On a
    Plummer sphere with a weak geometric guard (theta 0.8) and alpha 1e-3 the
    force error stays at the level of the geometric theta 0.4 walk (measured
    median 1.4e-4 and p99 5.6e-4 on the 512-system ensemble) with fewer
    interactions; alpha = 0 leaves the geometric walk unchanged bit for bit."""
    p = ic.plummer(20000, seed=31, virial=None)
    counts = {}
    for alpha, theta in ((0.0, 0.8), (1e-3, 0.8)):
        with phys(theta=theta) as e:
            e.set_alpha(alpha)
            e.upload(p)
            e.run(1e-3, 1)
            a = e.accel()
            ad, _ = e.direct()
            err = rel_err(a, ad)
            e.set_profiling(2)
            e.reset_timing()
            e.run(1e-3, 1)
            t = e.timing()
            counts[alpha] = t["n_pp"] + t["n_pc"]
        if alpha > 0:
            assert np.median(err) <= 3e-4 and np.percentile(err, 99) <= 2e-3
        else:
            with phys(theta=theta) as e:
                e.upload(p)
                e.run(1e-3, 1)
                assert np.array_equal(e.accel(), a)
    assert counts[1e-3] > counts[0.0]           # the relative test only ever rejects more cells
    with phys() as e:
        for bad in (-1.0, float("inf"), float("nan")):
            with pytest.raises(NbxError):
                e.set_alpha(bad)
    with Engine(theta=0.5) as e:                # reference mode: NEXT's geometric criterion only
        with pytest.raises(NbxError):
            e.set_alpha(1e-3)


def test_tree_refit_between_rebuilds():
    """rebuild_every = 4: three of every four steps refit the tree (same body
    order and topology, moments and bmax from the current positions).  The
    force error against the FP64 direct sum stays that of a rebuild every
    step, and the trajectories agree to the level of the force error."""
    p = ic.plummer(20000, seed=33, virial=None)
    out = {}
    for k in (1, 4):
        with Engine(theta=0.5, mode=MODE_PHYSICS, precision=PREC_MIXED, eps=0.01, leaf_size=16, adaptive_dt=False,
                    rebuild_every=k) as e:
            e.upload(p)
            e.run(1e-3, 7)                         # ends on a refit for k = 4
            a = e.accel()
            ad, _ = e.direct()
            out[k] = (np.median(rel_err(a, ad)), np.stack(e.download()))
    assert out[4][0] <= 1.2 * out[1][0]
    # forces differ at the multipole-error level (~1e-4 of |a|, |a| up to ~10
    # in the core) over 7 steps of 1e-3: velocities by up to ~1e-5, positions
    # by up to ~1e-7
    d = np.abs(out[4][1] - out[1][1])
    assert d[:3].max() <= 5e-7 and d[3:].max() <= 5e-5
    with phys() as e:
        for bad in (0, 1001):
            with pytest.raises(NbxError):
                e.set_rebuild_interval(bad)
    with Engine(theta=0.5) as e:                    # reference mode rebuilds every step, as NEXT does
        with pytest.raises(NbxError):
            e.set_rebuild_interval(2)


@pytest.mark.parametrize("alpha,every", [(0.0, 1), (1e-3, 3)])
def test_diagnostics_leave_trajectory_unchanged(alpha, every):
    """Energies (tree and direct) and the tree potential computed between
    steps keep the forces cached for the next step: the run with them ends
    bit for bit where the run without them does, also with the relative
    criterion (forces that depend on the previous step's |a|) and refits."""
    p = ic.plummer(5000, seed=41, virial=None)
    out = []
    for diag in (False, True):
        with Engine(theta=0.5, mode=MODE_PHYSICS, precision=PREC_MIXED, eps=0.01, leaf_size=16, adaptive_dt=False,
                    rebuild_every=every) as e:
            e.set_alpha(alpha)
            e.upload(p)
            for _ in range(5):
                e.run(1e-3, 1)
                if diag:
                    e.energy(direct=False)
                    e.energy(direct=True)
                    e.potential()
            out.append(np.stack(e.download()))
    assert np.array_equal(out[0], out[1])


def test_debug_tree_after_a_refit():
    """After a step that refitted the tree, the inspection hooks describe the
    tree of the current positions: the same table as a fresh context's."""
    p = ic.plummer(4000, seed=43, virial=None)
    with Engine(theta=0.5, mode=MODE_PHYSICS, precision=PREC_MIXED, eps=0.01, leaf_size=16, adaptive_dt=False,
                rebuild_every=4) as e:
        e.upload(p)
        e.run(1e-2, 2)                                # the second step refits
        got = e.debug_tree()
        x, y, z, vx, vy, vz = e.download()
    q = ic.Particles.from_columns(np.stack([x, y, z, vx, vy, vz, p.m], 1), p.type)
    with phys() as f:
        f.upload(q)
        want = f.debug_tree()
    assert got.keys() == want.keys()
    for k in want:
        assert np.array_equal(got[k], want[k]), k


def subtree_ranges(depth, leaf):
    """Body range [lo, hi) in leaf order of every node of a preorder table."""
    n_nodes = len(depth)
    before = np.concatenate([[0], np.cumsum(leaf)])           # leaves before node i
    end = np.empty(n_nodes, np.int64)
    stack = []
    for i in range(n_nodes):
        while stack and depth[stack[-1]] >= depth[i]:
            end[stack.pop()] = i
        stack.append(i)
    for j in stack:
        end[j] = n_nodes
    return before[:-1], before[end]


@pytest.mark.parametrize("gsize", [64, 32])
def test_tree_groups_pack_siblings(gsize):
    """Physics-mode walk groups: they cover every body once, hold at most
    group_size bodies each, and each is a run of consecutive children (of at
    most group_size bodies) of one node of more than group_size bodies, or a
    whole small root; so no group straddles distant cells."""
    p = ic.plummer(20000, seed=9, virial=None)
    with Engine(theta=0.5, mode=MODE_PHYSICS, precision=PREC_MIXED, eps=0.01, leaf_size=16,
                group_size=gsize, adaptive_dt=False) as e:
        e.upload(p)
        first, count, sys_ = e.debug_groups()
        t = e.debug_tree()
    order = np.argsort(first)
    f, c = first[order], count[order]
    assert f[0] == 0 and np.all(f[1:] == f[:-1] + c[:-1]) and f[-1] + c[-1] == len(p)
    assert c.min() >= 1 and c.max() <= gsize and np.all(sys_ == 0)
    lo, hi = subtree_ranges(t["depth"], t["leaf"])
    size = hi - lo
    by_lo = {}
    for node in np.nonzero(size <= gsize)[0]:
        by_lo.setdefault(int(lo[node]), []).append(node)
    parent = np.full(len(lo), -1)
    stack = []
    for i, d in enumerate(t["depth"]):
        while stack and t["depth"][stack[-1]] >= d:
            stack.pop()
        parent[i] = stack[-1] if stack else -1
        stack.append(i)
    for gf, gc in zip(f, c):
        # the children covering the group: maximal nodes of <= gsize bodies starting at each point
        pos, par = int(gf), None
        while pos < gf + gc:
            nodes = [k for k in by_lo.get(pos, []) if parent[k] < 0 or size[parent[k]] > gsize]
            assert len(nodes) == 1, (gf, gc, pos)
            k = nodes[0]
            pk = parent[k]
            if par is None:
                par = pk
            assert pk == par, "group spans two packing nodes"
            pos = int(hi[k])
        assert pos == gf + gc
    # compact in space: the widest group is far narrower than the widest run of 64 Morton-ordered bodies
    xyz = np.stack([p.x, p.y, p.z], 1)[t["perm"]]
    ext = np.array([np.ptp(xyz[a:a + b], axis=0).max() for a, b in zip(f, c)])
    runs = np.array([np.ptp(xyz[a:a + 64], axis=0).max() for a in range(0, len(p) - 63, 64)])
    print(f"group extent max {ext.max():.3f} (64-runs {runs.max():.3f}), mean size {c.mean():.1f}")
    assert ext.max() < runs.max()
