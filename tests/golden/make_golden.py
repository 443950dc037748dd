"""Write the golden vectors under tests/golden/ from NEXT's own FP64 code.

Runs oracle/_ref/ref_harness (built by ``make -C oracle`` from the reference
headers where they lie) on seeded inputs and stores inputs and outputs as
``.npz`` fixtures (plain arrays, no pickles).  Only needed when the fixtures
change; the committed ``.npz`` files are what the tests read.

    python tests/golden/make_golden.py
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from next_amd import ic  # noqa: E402

HARNESS = os.path.join(ROOT, "oracle", "_ref", "ref_harness")


def _run(*args):
    subprocess.run([HARNESS, *map(str, args)], check=True)


def _load(d, name, dtype):
    return np.fromfile(os.path.join(d, name), dtype=dtype)


def _inputs(p: ic.Particles):
    return dict(x=p.x, y=p.y, z=p.z, vx=p.vx, vy=p.vy, vz=p.vz, m=p.m, type=p.type)


def tree_case(name: str, p: ic.Particles, thetas=(0.5,), levels: int = 42, keep_tree=True,
              keep_acc=True):
    out = dict(_inputs(p))
    with tempfile.TemporaryDirectory() as td:
        icp = os.path.join(td, "ic.txt")
        ic.write_ic(icp, p)
        for k, th in enumerate(thetas):
            _run("tree", icp, repr(th), levels, td)
            if keep_acc:
                out[f"acc_{k}"] = _load(td, "acc.f64", np.float64).reshape(-1, 3)
        out["thetas"] = np.array(thetas, dtype=np.float64)
        out["root"] = _load(td, "root.f64", np.float64)
        out["keys_hi"] = _load(td, "keys_hi.u64", np.uint64)
        out["keys_lo"] = _load(td, "keys_lo.u64", np.uint64)
        out["leafdepth"] = _load(td, "leafdepth.i32", np.int32).astype(np.int16)
        out["perm"] = _load(td, "perm.i32", np.int32)
        if keep_tree:
            out["node_depth"] = _load(td, "node_depth.i32", np.int32).astype(np.int16)
            out["node_leaf"] = _load(td, "node_leaf.i32", np.int32).astype(np.int8)
            out["node_body"] = _load(td, "node_body.i32", np.int32)
            out["node_cmask"] = _load(td, "node_cmask.i32", np.int32).astype(np.uint8)
            out["node_geo"] = _load(td, "node_geo.f64", np.float64).reshape(-1, 4)
            out["node_mass"] = _load(td, "node_mass.f64", np.float64)
            out["node_com"] = _load(td, "node_com.f64", np.float64).reshape(-1, 3)
            out["node_q"] = _load(td, "node_q.f64", np.float64).reshape(-1, 6)
    np.savez_compressed(os.path.join(HERE, name + ".npz"), **out)
    print(f"{name}: n={len(p)} nodes={out.get('node_depth', np.zeros(0)).size}")


def steps_case(name: str, p: ic.Particles, base_dt: float, checkpoints, threads: int = 8):
    out = dict(_inputs(p))
    out["base_dt"] = np.array([base_dt])
    with tempfile.TemporaryDirectory() as td:
        icp = os.path.join(td, "ic.txt")
        ic.write_ic(icp, p)
        for c in checkpoints:
            _run("steps", icp, repr(base_dt), c, threads, td)
            out[f"state_{c}"] = _load(td, "state.f64", np.float64).reshape(6, -1)
            out[f"dts_{c}"] = _load(td, "dts.f64", np.float64)
    out["checkpoints"] = np.array(checkpoints, dtype=np.int32)
    np.savez_compressed(os.path.join(HERE, name + ".npz"), **out)
    print(f"{name}: n={len(p)} checkpoints={checkpoints}")


def writer_case(name: str, p: ic.Particles):
    with tempfile.TemporaryDirectory() as td:
        icp = os.path.join(td, "ic.txt")
        ic.write_ic(icp, p)
        ic.write_ic(os.path.join(HERE, name + ".ic.txt"), p)
        for fmt in ("vtu", "vtk"):
            _run(fmt, icp, os.path.join(HERE, f"{name}.{fmt}"))
    print(f"{name}: writers")


H5 = os.path.join(ROOT, "oracle", "_ref", "ref_h5")


def hdf5_case(name: str):
    """NEXT's SaveHDF5 of the writer case (file written as dump_0.hdf5, the
    name its XDMF sidecar records), and NEXT's loader reading that file back
    (written out with SaveVTK)."""
    with tempfile.TemporaryDirectory() as td:
        subprocess.run([H5, "h5save", os.path.join(HERE, name + ".ic.txt"), "dump_0.hdf5"], check=True, cwd=td,
                       stderr=subprocess.DEVNULL)
        shutil.copyfile(os.path.join(td, "dump_0.hdf5"), os.path.join(HERE, name + ".hdf5"))
        shutil.copyfile(os.path.join(td, "dump_0.xdmf"), os.path.join(HERE, name + ".xdmf"))
        subprocess.run([H5, "h5load", os.path.join(HERE, name + ".hdf5"), os.path.join(HERE, name + ".h5load.vtk")],
                       check=True, stderr=subprocess.DEVNULL)
    print(f"{name}: hdf5")


NEXT_REF = os.path.join(ROOT, "oracle", "_ref", "next_ref")


def startup_case(name: str, threads: int = 4):
    """The start-up lines NEXT's own driver prints (begrun.cpp:60-88): banner,
    thread count, precision and particle count, for writers_small with
    `threads` threads.  The driver loops until it reads 'q', so it runs in a
    scratch directory for a moment and is stopped; only the lines before its
    first dump line are kept."""
    with tempfile.TemporaryDirectory() as td:
        try:
            r = subprocess.run([NEXT_REF, os.path.join(HERE, name + ".ic.txt"), str(threads), "0.001", "1000", "vtu"],
                               cwd=td, stdin=subprocess.DEVNULL, capture_output=True, timeout=2)
            out = r.stdout
        except subprocess.TimeoutExpired as e:
            out = e.stdout or b""
    keep = out[: out.index(b"[Dump 0]")]
    with open(os.path.join(HERE, f"{name}.startup{threads}.txt"), "wb") as f:
        f.write(keep)
    print(f"{name}: start-up lines ({len(keep)} bytes)")


def hand_cases():
    """Small inputs that exercise single branches of the reference."""
    def mk(pos, m=None, types=None, vel=None):
        pos = np.asarray(pos, dtype=np.float64).reshape(-1, 3)
        n = pos.shape[0]
        m = np.full(n, 1.0 / max(n, 1)) if m is None else np.asarray(m, dtype=np.float64)
        vel = np.zeros((n, 3)) if vel is None else np.asarray(vel, dtype=np.float64)
        types = np.zeros(n, dtype=np.int32) if types is None else np.asarray(types, dtype=np.int32)
        return ic.Particles.from_columns(np.concatenate([pos, vel, m[:, None]], axis=1), types)

    cases = {}
    cases["hand_n1"] = mk([[0.3, -0.2, 0.1]])
    cases["hand_n2"] = mk([[0.0, 0.0, 0.0], [1.0, 0.5, -0.25]], m=[0.75, 0.25])
    cases["hand_n3"] = mk([[0.1, 0.2, 0.3], [-0.4, 0.5, 0.9], [0.7, -0.8, 0.05]], m=[0.2, 0.3, 0.5])
    corners = [[sx, sy, sz] for sz in (-1, 1) for sy in (-1, 1) for sx in (-1, 1)]
    cases["hand_oct8"] = mk(np.array(corners) * np.array([1.0, 0.7, 0.4]),
                            m=np.linspace(0.05, 0.2, 8))
    # bodies on the root's centre planes: octant ties resolve low
    cases["hand_ties"] = mk([[-1, -1, -1], [1, 1, 1], [0, 0, 0], [0, 0.5, -0.5], [0.5, 0, 0],
                             [0, 0, 1], [-0.5, 0.5, 0], [0.25, -0.25, 0.0]])
    # a close pair far from the rest: a chain of one-child nodes
    cases["hand_chain"] = mk([[-1, -1, -1], [1, 1, 1], [0.3, 0.3, 0.3], [0.3 + 1e-4, 0.3, 0.3 - 2e-5],
                              [-0.6, 0.2, 0.9]])
    # the pair shares more than 21 digits: keys need the second word
    cases["hand_deep"] = mk([[-1, -1, -1], [1, 1, 1], [0.3, 0.3, 0.3], [0.3 + 3e-9, 0.3, 0.3],
                             [0.1, -0.7, 0.2]])
    cases["hand_zeromass"] = mk([[0.1, 0.1, 0.1], [-0.5, 0.4, 0.2], [0.6, -0.3, -0.9], [0.0, 0.9, 0.0]],
                                m=[0.5, 0.0, 0.3, 0.2])
    rng = np.random.Generator(np.random.PCG64(7))
    pos = rng.uniform(-1, 1, (64, 3))
    cases["hand_dm"] = mk(pos, m=rng.uniform(0.5, 1.5, 64) / 64,
                          types=(rng.uniform(0, 1, 64) < 0.5).astype(np.int32))
    return cases


def main():
    if not os.path.exists(HARNESS):
        subprocess.run(["make", "-C", os.path.join(ROOT, "oracle")], check=True)
    for name, p in hand_cases().items():
        tree_case(name, p, thetas=(0.0, 0.5))
    tree_case("plummer4k", ic.plummer(4096, seed=11, type_fraction_dm=0.1), thetas=(0.3, 0.5, 0.7))
    c1 = ic.plummer(16384, seed=1)
    tree_case("c1_plummer16k", c1, thetas=(0.5,), levels=21, keep_tree=False)
    steps_case("c1_steps", c1, base_dt=1e-3, checkpoints=[10, 100])
    u = ic.uniform_sphere(65536, seed=2)
    tree_case("uniform64k_keys", u, thetas=(0.7,), levels=21, keep_tree=False, keep_acc=False)
    small = ic.plummer(5, seed=3, virial=None)
    small.type[1] = 1
    writer_case("writers_small", small)
    if os.path.exists(H5):
        hdf5_case("writers_small")
    if os.path.exists(NEXT_REF):
        startup_case("writers_small")


if __name__ == "__main__":
    if sys.argv[1:] == ["hdf5"]:
        hdf5_case("writers_small")
    elif sys.argv[1:] == ["startup"]:
        startup_case("writers_small")
    else:
        main()
