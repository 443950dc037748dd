"""ctypes access to the CPU oracle (oracle/liboracle.so) for the tests.

The oracle is the checker; nothing in the product imports this module.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "oracle", "liboracle.so")

_lib = None

f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
i32p = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
u64p = np.ctypeslib.ndpointer(np.uint64, flags="C_CONTIGUOUS")


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            subprocess.run(["make", "-C", os.path.join(ROOT, "oracle"), "oracle"], check=True)
        L = C.CDLL(LIB)
        L.orc_root.argtypes = [C.c_int64, f64p, f64p, f64p, f64p]
        L.orc_keys.argtypes = [C.c_int64, f64p, f64p, f64p, f64p, C.c_int, u64p, u64p]
        L.orc_tree_build.argtypes = [C.c_int64, f64p, f64p, f64p, f64p]
        L.orc_tree_build.restype = C.c_void_p
        L.orc_tree_error.argtypes = [C.c_void_p]
        L.orc_tree_nodes.argtypes = [C.c_void_p]
        L.orc_tree_nodes.restype = C.c_int64
        L.orc_tree_max_depth.argtypes = [C.c_void_p]
        L.orc_tree_export.argtypes = [C.c_void_p, i32p, i32p, i32p, i32p, f64p, f64p, f64p, f64p, i32p]
        L.orc_tree_accel.argtypes = [C.c_void_p, C.c_int64, f64p, f64p, f64p, f64p, i32p, C.c_double,
                                     f64p, f64p, f64p, C.c_int]
        L.orc_tree_free.argtypes = [C.c_void_p]
        L.orc_tree_accel_subset.argtypes = [C.c_void_p, f64p, f64p, f64p, f64p, i32p, C.c_double, C.c_int64,
                                            np.ctypeslib.ndpointer(np.int64, flags="C_CONTIGUOUS"), f64p, f64p,
                                            f64p, C.c_int]
        L.orc_adaptive_dt.argtypes = [C.c_int64, f64p, f64p, f64p, C.c_double]
        L.orc_adaptive_dt.restype = C.c_double
        L.orc_step.argtypes = [C.c_int64] + [f64p] * 7 + [i32p, C.c_double, C.c_double, C.c_int]
        L.orc_run.argtypes = [C.c_int64] + [f64p] * 7 + [i32p, C.c_double, C.c_int, C.c_double, C.c_int,
                                                       f64p]
        L.orc_direct.argtypes = [C.c_int64, f64p, f64p, f64p, f64p, C.c_double, f64p, f64p, f64p, f64p,
                                 C.c_int]
        L.orc_direct_subset.argtypes = [C.c_int64, f64p, f64p, f64p, f64p, C.c_double, C.c_int64,
                                        np.ctypeslib.ndpointer(np.int64, flags="C_CONTIGUOUS"), f64p, f64p, f64p,
                                        f64p, C.c_int]
        _lib = L
    return _lib


def _c(a, dt):
    return np.ascontiguousarray(a, dtype=dt)


def root(x, y, z):
    out = np.zeros(4)
    lib().orc_root(len(x), _c(x, np.float64), _c(y, np.float64), _c(z, np.float64), out)
    return out


def keys(x, y, z, levels=42):
    r = root(x, y, z)
    n = len(x)
    hi = np.zeros(n, np.uint64)
    lo = np.zeros(n, np.uint64)
    lib().orc_keys(n, _c(x, np.float64), _c(y, np.float64), _c(z, np.float64), r, levels, hi, lo)
    return hi, lo


def tree(x, y, z, m):
    """Preorder node table of the one-body-per-leaf octree."""
    L = lib()
    n = len(x)
    h = L.orc_tree_build(n, _c(x, np.float64), _c(y, np.float64), _c(z, np.float64), _c(m, np.float64))
    try:
        if L.orc_tree_error(h):
            raise RuntimeError("oracle: coincident bodies (depth cap)")
        k = L.orc_tree_nodes(h) if n > 0 else 0
        t = dict(depth=np.zeros(k, np.int32), leaf=np.zeros(k, np.int32), body=np.zeros(k, np.int32),
                 cmask=np.zeros(k, np.int32), geo=np.zeros((k, 4)), mass=np.zeros(k),
                 com=np.zeros((k, 3)), q=np.zeros((k, 6)), perm=np.zeros(n, np.int32))
        if n > 0:
            L.orc_tree_export(h, t["depth"], t["leaf"], t["body"], t["cmask"], t["geo"], t["mass"],
                              t["com"], t["q"], t["perm"])
        return t
    finally:
        L.orc_tree_free(h)


def accel(x, y, z, m, typ, theta, threads=8):
    L = lib()
    n = len(x)
    x, y, z, m = (_c(a, np.float64) for a in (x, y, z, m))
    typ = _c(typ, np.int32)
    out = [np.zeros(n) for _ in range(3)]
    h = L.orc_tree_build(n, x, y, z, m)
    try:
        if L.orc_tree_error(h):
            raise RuntimeError("oracle: coincident bodies (depth cap)")
        if n > 0:
            L.orc_tree_accel(h, n, x, y, z, m, typ, theta, *out, threads)
    finally:
        L.orc_tree_free(h)
    return np.stack(out, axis=1)


def accel_subset(x, y, z, m, typ, theta, targets, threads=8):
    """Reference-order walk for the given target indices only."""
    L = lib()
    x, y, z, m = (_c(a, np.float64) for a in (x, y, z, m))
    typ = _c(typ, np.int32)
    tg = _c(targets, np.int64)
    out = [np.zeros(len(tg)) for _ in range(3)]
    h = L.orc_tree_build(len(x), x, y, z, m)
    try:
        if L.orc_tree_error(h):
            raise RuntimeError("oracle: coincident bodies (depth cap)")
        L.orc_tree_accel_subset(h, x, y, z, m, typ, theta, len(tg), tg, *out, threads)
    finally:
        L.orc_tree_free(h)
    return np.stack(out, axis=1)


def adaptive_dt(vx, vy, vz, base):
    return lib().orc_adaptive_dt(len(vx), _c(vx, np.float64), _c(vy, np.float64), _c(vz, np.float64),
                                 base)


def run(state, m, typ, base_dt, nsteps, theta=0.5, threads=8):
    """Advance a copy of state (6 x n: x y z vx vy vz) nsteps; returns
    (state, dts)."""
    st = [np.array(a, dtype=np.float64, copy=True) for a in state]
    dts = np.zeros(max(nsteps, 1))
    rc = lib().orc_run(len(st[0]), *st, _c(m, np.float64), _c(typ, np.int32), base_dt, nsteps, theta,
                       threads, dts)
    if rc:
        raise RuntimeError("oracle: coincident bodies (depth cap)")
    return np.stack(st), dts[:nsteps]


def direct(x, y, z, m, eps, threads=8):
    n = len(x)
    out = [np.zeros(n) for _ in range(4)]
    lib().orc_direct(n, _c(x, np.float64), _c(y, np.float64), _c(z, np.float64), _c(m, np.float64), eps,
                     *out, threads)
    return np.stack(out[:3], axis=1), out[3]


def direct_subset(x, y, z, m, eps, targets, threads=8):
    """FP64 direct sum on the given targets only (all bodies act on them)."""
    tg = _c(targets, np.int64)
    out = [np.zeros(len(tg)) for _ in range(4)]
    lib().orc_direct_subset(len(x), _c(x, np.float64), _c(y, np.float64), _c(z, np.float64), _c(m, np.float64),
                            eps, len(tg), tg, *out, threads)
    return np.stack(out[:3], axis=1), out[3]
