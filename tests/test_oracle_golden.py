"""The CPU oracle against the golden vectors written by NEXT's own FP64 code.

The fixtures come from oracle/_ref/ref_harness (compiled from the reference
headers, tests/golden/make_golden.py).  The oracle is built with the same
compiler, libm and -ffp-contract=off, so every quantity must match bit for bit.
"""
import numpy as np
import pytest

import oracle_lib as orc
from conftest import golden

HAND = ["hand_n1", "hand_n2", "hand_n3", "hand_oct8", "hand_ties", "hand_chain", "hand_deep",
        "hand_zeromass", "hand_dm"]
TREE_CASES = HAND + ["plummer4k"]


def _xyz(g):
    return g["x"], g["y"], g["z"]


@pytest.mark.parametrize("case", TREE_CASES + ["c1_plummer16k", "uniform64k_keys"])
def test_root_and_keys(case):
    g = golden(case)
    assert np.array_equal(orc.root(*_xyz(g)), g["root"])
    levels = 42 if case in TREE_CASES else 21
    hi, lo = orc.keys(*_xyz(g), levels=levels)
    assert np.array_equal(hi, g["keys_hi"])
    assert np.array_equal(lo, g["keys_lo"])


@pytest.mark.parametrize("case", TREE_CASES)
def test_tree_topology_and_multipoles(case):
    g = golden(case)
    t = orc.tree(*_xyz(g), g["m"])
    assert np.array_equal(t["depth"], g["node_depth"])
    assert np.array_equal(t["leaf"], g["node_leaf"])
    assert np.array_equal(t["body"], g["node_body"])
    assert np.array_equal(t["cmask"], g["node_cmask"])
    assert np.array_equal(t["geo"], g["node_geo"])
    assert np.array_equal(t["mass"], g["node_mass"])
    assert np.array_equal(t["com"], g["node_com"])
    assert np.array_equal(t["q"], g["node_q"])
    assert np.array_equal(t["perm"], g["perm"])


@pytest.mark.parametrize("case", TREE_CASES + ["c1_plummer16k"])
def test_accelerations(case):
    g = golden(case)
    for k, th in enumerate(g["thetas"]):
        a = orc.accel(*_xyz(g), g["m"], g["type"], float(th))
        assert np.array_equal(a, g[f"acc_{k}"]), f"theta={th}"


def test_permutation_is_sorted_key_order():
    """The DFS leaf order of the reference equals ascending key order (the
    keys of distinct bodies are distinct)."""
    for case in TREE_CASES + ["c1_plummer16k", "uniform64k_keys"]:
        g = golden(case)
        order = np.lexsort((g["keys_lo"], g["keys_hi"]))
        assert np.array_equal(order, g["perm"]), case


def test_multistep_state_and_dt_sequence():
    g = golden("c1_steps")
    state0 = np.stack([g[k] for k in ("x", "y", "z", "vx", "vy", "vz")])
    for c in g["checkpoints"]:
        st, dts = orc.run(state0, g["m"], g["type"], float(g["base_dt"][0]), int(c))
        assert np.array_equal(dts, g[f"dts_{c}"])
        assert np.array_equal(st, g[f"state_{c}"])
