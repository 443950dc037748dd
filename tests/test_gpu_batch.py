"""Several independent systems in one context: every system's trajectory is
bit-identical to the same system run in a context of its own, in reference
mode and in physics mode, including systems added and removed mid-run."""
import numpy as np
import pytest

from next_amd import ic
from next_amd.nbx import MODE_PHYSICS, PREC_MIXED, Engine, NbxError

pytestmark = pytest.mark.gpu


def systems():
    a = ic.plummer(3000, seed=1, type_fraction_dm=0.2)
    b = ic.uniform_sphere(5000, seed=2)
    b.vx += 0.05
    c = ic.plummer(700, seed=3)
    c.x += 3.0
    return [a, b, c]


def make(mode, adaptive):
    if mode == "reference":
        return Engine(theta=0.5, adaptive_dt=adaptive)
    return Engine(theta=0.5, mode=MODE_PHYSICS, precision=PREC_MIXED, eps=0.01, leaf_size=16, adaptive_dt=adaptive)


@pytest.mark.parametrize("mode", ["reference", "physics"])
@pytest.mark.parametrize("adaptive", [True, False])
def test_batch_equals_alone(mode, adaptive):
    sy = systems()
    alone = []
    for p in sy:
        with make(mode, adaptive) as e:
            e.upload(p)
            e.run(1e-3, 5)
            alone.append(e.download())
    with make(mode, adaptive) as e:
        e.upload_batch(sy)
        assert e.num_systems() == 3
        e.run(1e-3, 5)
        for s, ref in enumerate(alone):
            assert np.array_equal(e.download_system(s), ref), s
        allst = e.download()
    assert np.array_equal(allst, np.concatenate(alone, axis=1))


@pytest.mark.parametrize("mode", ["reference", "physics"])
def test_add_and_remove_mid_run(mode):
    a, b, c = systems()

    def alone(p, nsteps):
        with make(mode, True) as e:
            e.upload(p)
            e.run(1e-3, nsteps)
            return e.download()

    def check_order(e, sizes):
        """Device slots hold each system's bodies contiguously, in system
        order, and their upload indices are a permutation of that system's."""
        ids, sys_ = e.debug_order()
        off = 0
        for s, k in enumerate(sizes):
            assert np.all(sys_[off:off + k] == s), s
            assert np.array_equal(np.sort(ids[off:off + k]), np.arange(off, off + k, dtype=np.int32)), s
            off += k
        assert off == len(ids)

    with make(mode, True) as e:
        e.upload_batch([a, b])
        check_order(e, [len(a), len(b)])
        e.run(1e-3, 3)
        check_order(e, [len(a), len(b)])
        sc = e.add_system(c)
        assert sc == 2
        check_order(e, [len(a), len(b), len(c)])
        e.run(1e-3, 2)
        check_order(e, [len(a), len(b), len(c)])
        out_a = e.remove_system(0)           # a: 5 steps; b moves to index 0, c to 1
        assert e.num_systems() == 2
        check_order(e, [len(b), len(c)])
        e.run(1e-3, 1)
        check_order(e, [len(b), len(c)])
        out_b = e.download_system(0)         # b: 6 steps
        out_c = e.download_system(1)         # c: 3 steps
    assert np.array_equal(out_a, alone(a, 5))
    assert np.array_equal(out_b, alone(b, 6))
    assert np.array_equal(out_c, alone(c, 3))


def test_energy_per_system():
    sy = systems()
    with make("physics", False) as e:
        e.upload_batch(sy)
        per = e.energy_systems(direct=True)
    for s, p in enumerate(sy):
        with make("physics", False) as e:
            e.upload(p)
            one = e.energy(direct=True)
        assert np.allclose(per[s], one, rtol=1e-12, atol=1e-15)


def test_many_small_systems():
    sizes = ic.ensemble_sizes(64, lo=64, hi=2048)
    sy = [ic.plummer(n, seed=100 + i) for i, n in enumerate(sizes)]
    with make("physics", False) as e:
        e.upload_batch(sy)
        e.run(1e-3, 3)
        got = [e.download_system(s) for s in (0, 17, 63)]
    for s, g in zip((0, 17, 63), got):
        with make("physics", False) as e:
            e.upload(sy[s])
            e.run(1e-3, 3)
            assert np.array_equal(g, e.download())


def test_autotune_restores_state_and_applies_choice():
    """nbx_autotune times candidate (leaf, group) sizes on the current
    systems, leaves the state as it found it, and the steps that follow
    equal a run made with the chosen sizes from the start."""
    sy = [ic.plummer(n, seed=40 + i) for i, n in enumerate((3000, 1200, 6000))]
    cands = [(8, 64), (16, 64), (16, 32), (32, 64)]
    with make("physics", False) as e:
        e.upload_batch(sy)
        before = e.download()
        best, ms = e.autotune(cands, 1e-3, steps=2)
        assert np.array_equal(e.download(), before)
        assert best in cands and len(ms) == len(cands) and np.all(ms > 0)
        e.run(1e-3, 3)
        after = e.download()
    with Engine(theta=0.5, mode=MODE_PHYSICS, precision=PREC_MIXED, eps=0.01, leaf_size=best[0],
                group_size=best[1], adaptive_dt=False) as e:
        e.upload_batch(sy)
        e.run(1e-3, 3)
        assert np.array_equal(e.download(), after)


def test_autotune_lists_restores_state_and_applies_choice():
    """nbx_autotune_lists times (leaf, group, walk list) candidates the same
    way: the state is left as found, and the steps that follow equal a run
    made with the chosen leaf, group and list sizes from the start."""
    sy = [ic.plummer(n, seed=50 + i) for i, n in enumerate((5000, 2500))]
    cands = [(16, 64, 128), (16, 64, 256), (16, 64, 96), (32, 64, 192)]
    with make("physics", False) as e:
        e.upload_batch(sy)
        before = e.download()
        best, ms = e.autotune(cands, 1e-3, steps=2)
        assert np.array_equal(e.download(), before)
        assert best in cands and len(ms) == len(cands) and np.all(ms > 0)
        e.run(1e-3, 3)
        after = e.download()
    with Engine(theta=0.5, mode=MODE_PHYSICS, precision=PREC_MIXED, eps=0.01, leaf_size=best[0],
                group_size=best[1], adaptive_dt=False) as e:
        e.set_walk_list(best[2])
        e.upload_batch(sy)
        e.run(1e-3, 3)
        assert np.array_equal(e.download(), after)


def test_walk_list_setting_checks():
    """nbx_set_walk_list takes 96, 128, 192 or 256 entries, in physics mode only."""
    with make("physics", False) as e:
        for bad in (0, 64, 100, 512):
            with pytest.raises(NbxError):
                e.set_walk_list(bad)
        e.set_walk_list(192)
    with make("reference", False) as e:
        with pytest.raises(NbxError):
            e.set_walk_list(128)


@pytest.mark.parametrize("mode", ["reference", "physics"])
def test_reused_context_matches_fresh(mode):
    """A context whose buffers were sized by earlier, smaller and larger
    systems gives the same trajectory as a fresh one: the node layout that
    runs ahead of the node-count read-back into the last build's buffers is
    kept only when the tree fits them, and redone otherwise."""
    small = ic.plummer(2000, seed=11)
    big = ic.plummer(40000, seed=12, type_fraction_dm=0.1)
    with make(mode, True) as e:
        for p in (small, big, small, big):
            e.upload(p)
            e.run(1e-3, 3)
            got = e.download()
            with make(mode, True) as f:
                f.upload(p)
                f.run(1e-3, 3)
                assert np.array_equal(got, f.download())


@pytest.mark.parametrize("mode", ["reference", "physics"])
def test_trim_frees_buffers_and_context_stays_usable(mode):
    """nbx_trim frees an empty context's device buffers (and refuses one that
    holds bodies); the trimmed context then runs like a fresh one."""
    p = ic.plummer(30000, seed=13)
    with make(mode, True) as e:
        e.upload(p)
        e.run(1e-3, 2)
        held = e.device_bytes()
        with pytest.raises(NbxError):
            e.trim()
        e.upload(ic.Particles.from_columns(np.zeros((0, 7)), np.zeros(0, np.int32)))
        e.trim()
        assert e.device_bytes() < held // 100
        e.upload(p)
        e.run(1e-3, 3)
        got = e.download()
    with make(mode, True) as f:
        f.upload(p)
        f.run(1e-3, 3)
        assert np.array_equal(got, f.download())
