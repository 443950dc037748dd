"""Seeded initial-condition generators."""
import math

import numpy as np

from next_amd import ic


def test_plummer_is_reproducible_and_centred():
    a = ic.plummer(4096, seed=1)
    b = ic.plummer(4096, seed=1)
    for u, v in zip(a.arrays(), b.arrays()):
        assert np.array_equal(u, v)
    assert len(a) == 4096 and np.all(a.m == 1.0 / 4096)
    for arr in (a.x, a.y, a.z, a.vx, a.vy, a.vz):
        assert abs(arr.mean()) < 1e-15
    r = np.sqrt(a.x ** 2 + a.y ** 2 + a.z ** 2)
    assert r.max() <= 10.5


def test_plummer_virial_ratio():
    p = ic.plummer(2048, seed=3)
    t = 0.5 * float((p.m * (p.vx ** 2 + p.vy ** 2 + p.vz ** 2)).sum())
    w = ic.potential_energy(p.x, p.y, p.z, p.m)
    assert abs(t / abs(w) - 0.5) < 1e-12


def test_plummer_density_profile():
    """Half of the (untruncated) Plummer mass lies inside r = a / sqrt(2^{2/3} - 1)."""
    p = ic.plummer(20000, seed=5, virial=None)
    r = np.sqrt(p.x ** 2 + p.y ** 2 + p.z ** 2)
    rh = 1.0 / math.sqrt(2 ** (2 / 3) - 1)
    assert abs((r < rh).mean() - 0.5) < 0.02


def test_uniform_sphere():
    p = ic.uniform_sphere(10000, seed=2)
    r = np.sqrt(p.x ** 2 + p.y ** 2 + p.z ** 2)
    assert r.max() <= 1.0
    assert abs((r < 0.5 ** (1 / 3)).mean() - 0.5) < 0.02
    assert np.all(p.vx == 0) and np.isclose(p.m.sum(), 1.0)


def test_two_cluster():
    p = ic.two_cluster(1000, seed=7)
    assert len(p) == 2000
    assert abs(p.x[:1000].mean() - 5.0) < 0.2 and abs(p.x[1000:].mean() + 5.0) < 0.2
    assert abs(p.vx[:1000].mean() + 0.3) < 1e-12 and abs(p.vy[1000:].mean() + 0.1) < 1e-12
    assert np.isclose(p.m.sum(), 1.0)


def test_ensemble_sizes():
    s = ic.ensemble_sizes(512)
    assert len(s) == 512 and min(s) >= 1024 and max(s) <= 65536
    assert s == ic.ensemble_sizes(512)
    geo = math.exp(np.mean(np.log(s)))
    assert 6000 < geo < 10000     # log-uniform on [2^10, 2^16]: geometric mean 2^13
