"""Checks of the oracle that the mathematics fixes, independent of the golden
vectors: closed forms, invariants and brute-force equivalences chosen so that
a dropped term, a wrong sign or a transposed index fails one of them."""
import numpy as np
import pytest

import oracle_lib as orc
from next_amd import ic


def lcp_digits(a, b, words=21):
    x = int(a) ^ int(b)
    return words if x == 0 else (63 - x.bit_length()) // 3


def test_single_body_feels_nothing():
    a = orc.accel([0.3], [0.1], [-0.2], [1.0], [0], 0.5)
    assert np.array_equal(a, np.zeros((1, 3)))
    t = orc.tree([0.3], [0.1], [-0.2], [1.0])
    assert t["depth"].tolist() == [0] and t["leaf"].tolist() == [1] and t["body"].tolist() == [0]


def test_equal_pair_obeys_newtons_third_law():
    """Two equal masses sit in leaves of the same depth, so their softening
    is the same and the forces are equal and opposite, along the line."""
    x, y, z = [0.1, -0.4], [0.2, 0.35], [-0.3, 0.1]
    a = orc.accel(x, y, z, [0.5, 0.5], [0, 0], 0.5)
    assert np.array_equal(a[0], -a[1])
    d = np.array([x[1] - x[0], y[1] - y[0], z[1] - z[0]])
    assert np.allclose(np.cross(a[0], d), 0, atol=1e-15)
    assert a[0] @ d > 0   # attractive


def test_pair_force_magnitude_limits():
    """Far apart, softening is negligible against the distance: |a| -> m/r^2."""
    x, y, z = [0.0, 1000.0], [0.0, 0.0], [0.0, 0.0]
    a = orc.accel(x, y, z, [1.0, 1.0], [0, 0], 0.5)
    assert a[0, 0] == pytest.approx(1.0 / 1000.0 ** 2, rel=1e-9)


def test_theta_zero_is_brute_force_over_leaves():
    """At theta = 0 no internal node passes s/d < theta, so the walk sums
    every other body as a leaf (depth-dependent leaf softening)."""
    p = ic.plummer(300, seed=4, type_fraction_dm=0.0, virial=None)
    t = orc.tree(p.x, p.y, p.z, p.m)
    leafsize = np.zeros(len(p))
    sel = t["leaf"] == 1
    leafsize[t["body"][sel]] = t["geo"][sel, 3]
    a = orc.accel(p.x, p.y, p.z, p.m, p.type, 0.0)
    pos = np.stack([p.x, p.y, p.z], 1)
    d = pos[None, :, :] - pos[:, None, :]
    r2 = (d ** 2).sum(-1)
    dist = np.sqrt(r2 + 1e-20)
    eps = np.maximum((leafsize[None, :] * 0.015 + np.cbrt(p.m)[None, :] * 0.002) / (1 + 10 * dist), 1e-4)
    w = p.m[None, :] / (r2 + eps ** 2) ** 1.5
    np.fill_diagonal(w, 0.0)
    ab = (w[:, :, None] * d).sum(1)
    assert np.allclose(a, ab, rtol=1e-12, atol=1e-14)


def test_tree_follows_the_prefix_rule():
    """The insertion-built tree is the canonical one: in sorted key order, an
    internal node starts at body i for each depth in (lcp[i-1], lcp[i]], and
    each leaf sits one level below its longest common prefix with a sorted
    neighbour."""
    p = ic.uniform_sphere(5000, seed=9)
    hi, _ = orc.keys(p.x, p.y, p.z, levels=21)
    order = np.argsort(hi, kind="stable")
    ks = hi[order]
    lcp = np.array([lcp_digits(ks[i], ks[i + 1]) for i in range(len(ks) - 1)])
    t = orc.tree(p.x, p.y, p.z, p.m)
    assert np.array_equal(t["perm"], order)
    ext = np.concatenate([[-1], lcp, [-1]])
    assert (t["leaf"] == 0).sum() == np.maximum(0, ext[1:] - ext[:-1]).sum()
    leafdepth = np.zeros(len(p), int)
    sel = t["leaf"] == 1
    leafdepth[t["body"][sel]] = t["depth"][sel]
    assert np.array_equal(leafdepth[order], 1 + np.maximum(ext[:-1], ext[1:]))


def test_multipole_closed_forms():
    """Every internal node: mass = sum of its bodies, centre of mass = their
    weighted mean, and the one-level softened quadrupole has trace
    -0.03 s^2 m (the 3 r r^T - r^2 I part is traceless; the 0.01 s^2 term adds
    -3 * 0.01 s^2 per unit mass).  A node with a single child has
    Q = -0.01 s^2 m I."""
    p = ic.plummer(800, seed=5, virial=None)
    t = orc.tree(p.x, p.y, p.z, p.m)
    pos = np.stack([p.x, p.y, p.z], 1)
    depth, leaf = t["depth"], t["leaf"]
    k = len(depth)
    for n in range(k):
        if leaf[n]:
            continue
        end = n + 1
        while end < k and depth[end] > depth[n]:
            end += 1
        bodies = t["body"][n:end][leaf[n:end] == 1]
        m = p.m[bodies].sum()
        assert t["mass"][n] == pytest.approx(m, rel=1e-14)
        assert np.allclose(t["com"][n], (p.m[bodies, None] * pos[bodies]).sum(0) / m, rtol=1e-12, atol=1e-15)
        s = t["geo"][n, 3]
        q = t["q"][n]
        assert q[0] + q[1] + q[2] == pytest.approx(-0.03 * s * s * t["mass"][n], rel=1e-9, abs=1e-300)
        nchild = bin(int(t["cmask"][n])).count("1")
        if nchild == 1:
            assert np.allclose(q[:3], -0.01 * s * s * t["mass"][n], rtol=1e-12) and np.all(q[3:] == 0)


def test_child_centres_follow_the_recurrence():
    p = ic.uniform_sphere(500, seed=6)
    t = orc.tree(p.x, p.y, p.z, p.m)
    geo, depth = t["geo"], t["depth"]
    stack = []
    for n in range(len(depth)):
        while stack and depth[stack[-1]] >= depth[n]:
            stack.pop()
        if stack:
            par = geo[stack[-1]]
            hs = par[3] * 0.5
            assert geo[n, 3] == hs
            assert np.all(np.abs(np.abs(geo[n, :3] - par[:3]) - hs) <= 1e-15 * max(1.0, np.abs(par[:3]).max()))
        stack.append(n)


def test_input_order_does_not_matter():
    """Reordering the bodies reorders the accelerations bit for bit."""
    p = ic.plummer(2000, seed=8, type_fraction_dm=0.2)
    a = orc.accel(p.x, p.y, p.z, p.m, p.type, 0.5)
    perm = np.random.Generator(np.random.PCG64(1)).permutation(len(p))
    b = orc.accel(p.x[perm], p.y[perm], p.z[perm], p.m[perm], p.type[perm], 0.5)
    assert np.array_equal(a[perm], b)


def test_thread_count_does_not_matter():
    p = ic.plummer(1500, seed=12, type_fraction_dm=0.1)
    assert np.array_equal(orc.accel(p.x, p.y, p.z, p.m, p.type, 0.5, threads=1),
                          orc.accel(p.x, p.y, p.z, p.m, p.type, 0.5, threads=7))


@pytest.mark.parametrize("v,expect", [
    (0.0, 1e-3),                           # at rest: the base step
    (1.0, 1e-3 / 2.0),                     # base / (1 + |v|)
    (3.0, 1e-3 / 4.0),
    (200.0, 1e-3 * 0.01),                  # clamped below at 1 % of base
    (1e9, 1e-3 * 0.01),                    # speed capped at 1e4, then clamped
    (1e-9, 1e-3),                          # below 1e-8: the base step
])
def test_adaptive_dt_closed_forms(v, expect):
    vx = np.array([0.0, v, -0.5 * v])
    z = np.zeros(3)
    assert orc.adaptive_dt(vx, z, z, 1e-3) == expect


def test_adaptive_dt_ignores_nan_speeds():
    vx = np.array([np.nan, 1.0])
    z = np.zeros(2)
    assert orc.adaptive_dt(vx, z, z, 1e-3) == 1e-3 / 2.0


def test_zero_mass_bodies_exert_nothing():
    p = ic.plummer(200, seed=13, virial=None)
    a = orc.accel(p.x, p.y, p.z, p.m, p.type, 0.5)
    q = p.copy()
    extra = ic.Particles.from_columns(np.array([[0.05, -0.02, 0.01, 0, 0, 0, 0.0]]), np.zeros(1, np.int32))
    q = ic.Particles.concat([q, extra])
    b = orc.accel(q.x, q.y, q.z, q.m, q.type, 0.0)
    a0 = orc.accel(p.x, p.y, p.z, p.m, p.type, 0.0)
    # at theta 0 only leaves act, and the massless leaf adds nothing
    assert np.allclose(b[:-1], a0, rtol=1e-13, atol=1e-16)
    assert a.shape == (200, 3)


def test_direct_sum_two_body_and_momentum():
    eps = 0.01
    a, pot = orc.direct([0.0, 0.3], [0.0, 0.4], [0.0, 0.0], [2.0, 1.0], eps)
    r2 = 0.25 + eps * eps
    assert a[0, 0] == pytest.approx(1.0 * 0.3 / r2 ** 1.5, rel=1e-14)
    assert a[1, 1] == pytest.approx(-2.0 * 0.4 / r2 ** 1.5, rel=1e-14)
    assert pot[0] == pytest.approx(-1.0 / np.sqrt(r2), rel=1e-14)
    p = ic.plummer(500, seed=2)
    a, _ = orc.direct(p.x, p.y, p.z, p.m, 0.01)
    assert np.abs((p.m[:, None] * a).sum(0)).max() < 1e-12


def test_direct_subset_equals_full_sum():
    """The sampled direct sum (used for force errors of large systems) gives
    the full sum's values on its targets, bit for bit."""
    p = ic.plummer(700, seed=9)
    a, pot = orc.direct(p.x, p.y, p.z, p.m, 0.02)
    tg = np.array([0, 5, 699, 123, 5], np.int64)
    b, pb = orc.direct_subset(p.x, p.y, p.z, p.m, 0.02, tg)
    assert np.array_equal(b, a[tg]) and np.array_equal(pb, pot[tg])
