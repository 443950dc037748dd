"""Python binding of the C ABI in include/nbx.h (ctypes over next_amd/libnbx.so).

The calls mirror NEXT's path (begrun.cpp:96-99): ``adaptive_dt`` is
computeAdaptiveDt (adaptive.h:23), ``step`` is Step (step.h:25); ``run``
does both ``nsteps`` times on the device.  The library must be built
(``python -m next_amd.build``); there is no fallback.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from .ic import Particles

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libnbx.so")

STATUS = {0: "OK", -1: "NBX_ERR_ARG", -2: "NBX_ERR_NONFINITE", -3: "NBX_ERR_DEPTH",
          -4: "NBX_ERR_NOMEM", -5: "NBX_ERR_HIP", -6: "NBX_ERR_NODEVICE", -7: "NBX_ERR_STATE"}

MODE_REFERENCE, MODE_PHYSICS = 0, 1
PREC_FP64, PREC_MIXED = 0, 1

EXPORTS = ["nbx_config_default", "nbx_create", "nbx_destroy", "nbx_last_error", "nbx_set_theta", "nbx_set_alpha",
           "nbx_set_walk_list", "nbx_autotune_lists",
           "nbx_set_rebuild_interval", "nbx_trim", "nbx_device_bytes",
           "nbx_upload",
           "nbx_download", "nbx_size", "nbx_adaptive_dt", "nbx_step", "nbx_run", "nbx_accel", "nbx_get_timing",
           "nbx_reset_timing", "nbx_set_profiling", "nbx_debug_keys", "nbx_debug_tree_size", "nbx_debug_tree",
           "nbx_debug_groups", "nbx_debug_order", "nbx_autotune",
           "nbx_potential", "nbx_direct", "nbx_energy", "nbx_energy_systems", "nbx_upload_batch", "nbx_add_system",
           "nbx_download_system", "nbx_remove_system", "nbx_num_systems", "nbx_system_size", "nbx_server_create",
           "nbx_server_destroy", "nbx_server_submit", "nbx_server_poll", "nbx_server_wait", "nbx_server_stats",
           "nbx_server_create_with_ops", "nbx_accel_noalloc", "nbx_set_graph"]


class NbxError(RuntimeError):
    def __init__(self, status: int, msg: str):
        super().__init__(f"{STATUS.get(status, status)}: {msg}")
        self.status = status


class Config(C.Structure):
    _fields_ = [("device", C.c_int), ("mode", C.c_int), ("precision", C.c_int), ("theta", C.c_double),
                ("eps", C.c_double), ("leaf_size", C.c_int), ("group_size", C.c_int), ("adaptive_dt", C.c_int)]


class Timing(C.Structure):
    _fields_ = [("steps", C.c_int64), ("keys_ms", C.c_double), ("sort_ms", C.c_double), ("build_ms", C.c_double),
                ("upward_ms", C.c_double), ("walk_ms", C.c_double), ("kick_ms", C.c_double),
                ("total_ms", C.c_double), ("nodes", C.c_int64), ("depth", C.c_int32), ("n_pp", C.c_int64),
                ("n_pc", C.c_int64), ("n_open", C.c_int64), ("n_iter", C.c_int64),
                ("n_eval", C.c_int64), ("graph_steps", C.c_int64), ("graph_launches", C.c_int64),
                ("graph_nodes", C.c_int64), ("graph_period", C.c_int64), ("graph_aborts", C.c_int64),
                ("host_waits", C.c_int64), ("sort_patched", C.c_int64), ("sort_redos", C.c_int64)]


class Request(C.Structure):
    _fields_ = [("n", C.c_int64)] + [(k, C.c_void_p) for k in ("x", "y", "z", "vx", "vy", "vz", "m", "type")] + [
        ("nsteps", C.c_int32), ("dt", C.c_double), ("mode", C.c_int), ("precision", C.c_int), ("theta", C.c_double),
        ("eps", C.c_double), ("leaf_size", C.c_int), ("adaptive_dt", C.c_int)]


class RequestTiming(C.Structure):
    _fields_ = [("queue_ms", C.c_double), ("latency_ms", C.c_double)]


class ServerStats(C.Structure):
    _fields_ = [("queued", C.c_int64), ("running", C.c_int64), ("batches", C.c_int64), ("bytes_in_use", C.c_uint64),
                ("body_steps", C.c_int64), ("batch_steps", C.c_int64), ("device_bytes", C.c_uint64)]


_lib = None
f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
i32p = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
u64p = np.ctypeslib.ndpointer(np.uint64, flags="C_CONTIGUOUS")
VP = C.c_void_p


def lib():
    """Load libnbx.so (raises if it is not built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is not built: run `python -m next_amd.build`")
    L = C.CDLL(LIB_PATH)
    L.nbx_config_default.argtypes = [C.POINTER(Config)]
    L.nbx_config_default.restype = None
    L.nbx_create.argtypes = [C.POINTER(VP), C.POINTER(Config)]
    L.nbx_destroy.argtypes = [VP]
    L.nbx_destroy.restype = None
    L.nbx_last_error.argtypes = [VP]
    L.nbx_last_error.restype = C.c_char_p
    L.nbx_set_theta.argtypes = [VP, C.c_double]
    L.nbx_set_alpha.argtypes = [VP, C.c_double]
    L.nbx_set_walk_list.argtypes = [VP, C.c_int32]
    L.nbx_set_rebuild_interval.argtypes = [VP, C.c_int32]
    L.nbx_set_graph.argtypes = [VP, C.c_int]
    L.nbx_trim.argtypes = [VP]
    L.nbx_device_bytes.argtypes = [VP]
    L.nbx_device_bytes.restype = C.c_uint64
    L.nbx_upload.argtypes = [VP, C.c_int64] + [f64p] * 7 + [i32p]
    L.nbx_download.argtypes = [VP] + [f64p] * 6
    L.nbx_size.argtypes = [VP]
    L.nbx_size.restype = C.c_int64
    L.nbx_adaptive_dt.argtypes = [VP, C.c_double, C.POINTER(C.c_double)]
    L.nbx_step.argtypes = [VP, C.c_double]
    L.nbx_run.argtypes = [VP, C.c_double, C.c_int32, f64p]
    L.nbx_accel.argtypes = [VP, f64p, f64p, f64p]
    L.nbx_get_timing.argtypes = [VP, C.POINTER(Timing)]
    L.nbx_reset_timing.argtypes = [VP]
    L.nbx_set_profiling.argtypes = [VP, C.c_int]
    L.nbx_debug_keys.argtypes = [VP, f64p, u64p, u64p]
    L.nbx_debug_tree_size.argtypes = [VP, C.POINTER(C.c_int64)]
    L.nbx_debug_tree.argtypes = [VP, i32p, i32p, i32p, i32p, f64p, f64p, f64p, f64p, i32p]
    L.nbx_debug_groups.argtypes = [VP, C.c_int64, i32p, i32p, i32p, C.POINTER(C.c_int64)]
    L.nbx_debug_order.argtypes = [VP, i32p, i32p]
    L.nbx_autotune.argtypes = [VP, i32p, i32p, C.c_int32, C.c_double, C.c_int32, f64p]
    L.nbx_autotune_lists.argtypes = [VP, i32p, i32p, i32p, C.c_int32, C.c_double, C.c_int32, f64p]
    L.nbx_potential.argtypes = [VP, f64p]
    L.nbx_direct.argtypes = [VP, VP, VP, VP, VP]
    L.nbx_energy.argtypes = [VP, C.c_int, f64p]
    L.nbx_energy_systems.argtypes = [VP, C.c_int, f64p]
    i64p = np.ctypeslib.ndpointer(np.int64, flags="C_CONTIGUOUS")
    L.nbx_upload_batch.argtypes = [VP, C.c_int32, i64p] + [f64p] * 7 + [i32p]
    L.nbx_add_system.argtypes = [VP, C.c_int64] + [f64p] * 7 + [i32p, C.POINTER(C.c_int32)]
    L.nbx_download_system.argtypes = [VP, C.c_int32] + [f64p] * 6
    L.nbx_remove_system.argtypes = [VP, C.c_int32] + [f64p] * 6
    L.nbx_num_systems.argtypes = [VP]
    L.nbx_num_systems.restype = C.c_int32
    L.nbx_system_size.argtypes = [VP, C.c_int32]
    L.nbx_system_size.restype = C.c_int64
    L.nbx_server_create.argtypes = [C.POINTER(VP), C.c_int, C.c_uint64]
    L.nbx_server_destroy.argtypes = [VP]
    L.nbx_server_destroy.restype = None
    L.nbx_server_submit.argtypes = [VP, C.POINTER(Request), C.POINTER(C.c_int64)]
    L.nbx_server_poll.argtypes = [VP, C.c_int64]
    L.nbx_server_wait.argtypes = [VP, C.c_int64] + [f64p] * 6 + [C.POINTER(RequestTiming)]
    L.nbx_server_stats.argtypes = [VP, C.POINTER(ServerStats)]
    L.nbx_accel_noalloc.argtypes = [VP, C.c_int64] + [VP] * 5 + [C.c_double] + [VP] * 3
    _lib = L
    return L


def accel_noalloc(n: int, x: int, y: int, z: int, m: int, type_: int, theta: float, ax: int, ay: int, az: int,
                  stream: int = 0):
    """nbx_accel_noalloc: reference-mode accelerations of n bodies held in
    device memory (x, y, z, m: float64, type_: int32; device addresses as
    ints) into device arrays ax, ay, az, on `stream` (a HIP stream handle;
    0 = the null stream) of the current device, with no device allocation."""
    st = lib().nbx_accel_noalloc(stream or None, n, x, y, z, m, type_, theta, ax, ay, az)
    if st != 0:
        raise NbxError(st, "nbx_accel_noalloc")


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


class Engine:
    """One device-resident particle system (an nbx_ctx)."""

    def __init__(self, theta: float = 0.5, device: int = 0, mode: int = MODE_REFERENCE,
                 precision: int = PREC_FP64, eps: float = 0.01, leaf_size: int = 16, group_size: int = 64,
                 adaptive_dt: bool = True, alpha: float = 0.0, rebuild_every: int = 1, graph: int = 0,
                 walk_list: int = 0):
        L = lib()
        cfg = Config()
        L.nbx_config_default(C.byref(cfg))
        cfg.device, cfg.mode, cfg.precision = device, mode, precision
        cfg.theta, cfg.eps, cfg.leaf_size, cfg.group_size = theta, eps, leaf_size, group_size
        cfg.adaptive_dt = 1 if adaptive_dt else 0
        h = VP()
        st = L.nbx_create(C.byref(h), C.byref(cfg))
        if st != 0:
            raise NbxError(st, "nbx_create failed")
        self._h = h
        self.n = 0
        if alpha:
            self.set_alpha(alpha)
        if rebuild_every != 1:
            self.set_rebuild_interval(rebuild_every)
        if walk_list:
            self.set_walk_list(walk_list)
        if graph:
            self.set_graph(graph)

    def close(self):
        if getattr(self, "_h", None):
            lib().nbx_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def _check(self, st):
        if st != 0:
            raise NbxError(st, lib().nbx_last_error(self._h).decode())

    def set_graph(self, on: int):
        """nbx_set_graph: 1 replays nbx_run's steps from a captured HIP graph,
        2 the same without headroom in the bounded builds, 0 step by step."""
        self._check(lib().nbx_set_graph(self._h, int(on)))

    def set_theta(self, theta: float):
        self._check(lib().nbx_set_theta(self._h, theta))

    def set_alpha(self, alpha: float):
        """Physics mode: relative opening criterion (0 turns it off)."""
        self._check(lib().nbx_set_alpha(self._h, alpha))

    def set_walk_list(self, entries: int):
        """Physics mode: entries of the walk's LDS lists (96, 128, 192 or 256)."""
        self._check(lib().nbx_set_walk_list(self._h, entries))

    def set_rebuild_interval(self, k: int):
        """Physics mode: full tree build every k-th step, refits between."""
        self._check(lib().nbx_set_rebuild_interval(self._h, k))

    def trim(self):
        """Free the device buffers of a context that holds no bodies."""
        self._check(lib().nbx_trim(self._h))

    def device_bytes(self) -> int:
        """Device memory the context holds."""
        return int(lib().nbx_device_bytes(self._h))

    def upload(self, p: Particles):
        n = len(p)
        arrs = [_f64(a) for a in (p.x, p.y, p.z, p.vx, p.vy, p.vz, p.m)]
        t = np.ascontiguousarray(p.type, dtype=np.int32)
        if n == 0:
            arrs = [np.zeros(1) for _ in range(7)]
            t = np.zeros(1, np.int32)
        self._check(lib().nbx_upload(self._h, n, *arrs, t))
        self.n = n

    def upload_batch(self, systems):
        """Replace the contents with several independent systems."""
        counts = np.array([len(p) for p in systems], dtype=np.int64)
        allp = Particles.concat(systems)
        arrs = [_f64(a) for a in (allp.x, allp.y, allp.z, allp.vx, allp.vy, allp.vz, allp.m)]
        self._check(lib().nbx_upload_batch(self._h, len(systems), counts, *arrs,
                                           np.ascontiguousarray(allp.type, dtype=np.int32)))
        self.n = int(counts.sum())

    def add_system(self, p: Particles) -> int:
        s = C.c_int32()
        arrs = [_f64(a) for a in (p.x, p.y, p.z, p.vx, p.vy, p.vz, p.m)]
        self._check(lib().nbx_add_system(self._h, len(p), *arrs, np.ascontiguousarray(p.type, dtype=np.int32),
                                         C.byref(s)))
        self.n += len(p)
        return s.value

    def num_systems(self) -> int:
        return lib().nbx_num_systems(self._h)

    def system_size(self, s: int) -> int:
        return lib().nbx_system_size(self._h, s)

    def download_system(self, s: int):
        k = self.system_size(s)
        out = [np.zeros(max(k, 1)) for _ in range(6)]
        self._check(lib().nbx_download_system(self._h, s, *out))
        return np.stack([o[:k] for o in out])

    def remove_system(self, s: int):
        k = self.system_size(s)
        out = [np.zeros(max(k, 1)) for _ in range(6)]
        self._check(lib().nbx_remove_system(self._h, s, *out))
        self.n -= k
        return np.stack([o[:k] for o in out])

    def energy_systems(self, direct: bool = False):
        out = np.zeros(5 * max(self.num_systems(), 1))
        self._check(lib().nbx_energy_systems(self._h, 1 if direct else 0, out))
        return out[: 5 * self.num_systems()].reshape(-1, 5)

    def download(self):
        out = [np.zeros(max(self.n, 1)) for _ in range(6)]
        self._check(lib().nbx_download(self._h, *out))
        return np.stack([o[: self.n] for o in out])

    def adaptive_dt(self, base_dt: float) -> float:
        dt = C.c_double()
        self._check(lib().nbx_adaptive_dt(self._h, base_dt, C.byref(dt)))
        return dt.value

    def step(self, dt: float):
        self._check(lib().nbx_step(self._h, dt))

    def run(self, base_dt: float, nsteps: int):
        dts = np.zeros(max(nsteps, 1))
        self._check(lib().nbx_run(self._h, base_dt, nsteps, dts))
        return dts[:nsteps]

    def accel(self):
        out = [np.zeros(max(self.n, 1)) for _ in range(3)]
        self._check(lib().nbx_accel(self._h, *out))
        return np.stack([o[: self.n] for o in out], axis=1)

    def set_profiling(self, flags: int):
        self._check(lib().nbx_set_profiling(self._h, flags))

    def reset_timing(self):
        self._check(lib().nbx_reset_timing(self._h))

    def timing(self) -> dict:
        t = Timing()
        self._check(lib().nbx_get_timing(self._h, C.byref(t)))
        return {k: getattr(t, k) for k, _ in Timing._fields_}

    def potential(self):
        out = np.zeros(max(self.n, 1))
        self._check(lib().nbx_potential(self._h, out))
        return out[: self.n]

    def direct(self):
        """FP64 direct-sum accelerations and potentials (Plummer eps)."""
        out = [np.zeros(max(self.n, 1)) for _ in range(4)]
        ptrs = [o.ctypes.data_as(C.c_void_p) for o in out]
        self._check(lib().nbx_direct(self._h, *ptrs))
        return np.stack([o[: self.n] for o in out[:3]], axis=1), out[3][: self.n]

    def energy(self, direct: bool = False):
        """(kinetic, potential, px, py, pz) of the current state."""
        out = np.zeros(5)
        self._check(lib().nbx_energy(self._h, 1 if direct else 0, out))
        return out

    def debug_keys(self):
        root = np.zeros(4)
        n = max(self.n, 1)
        hi = np.zeros(n, np.uint64)
        lo = np.zeros(n, np.uint64)
        self._check(lib().nbx_debug_keys(self._h, root, hi, lo))
        return root, hi[: self.n], lo[: self.n]

    def debug_tree(self) -> dict:
        k = C.c_int64()
        self._check(lib().nbx_debug_tree_size(self._h, C.byref(k)))
        k = k.value
        kk = max(k, 1)
        t = dict(depth=np.zeros(kk, np.int32), leaf=np.zeros(kk, np.int32), body=np.zeros(kk, np.int32),
                 cmask=np.zeros(kk, np.int32), geo=np.zeros((kk, 4)), mass=np.zeros(kk), com=np.zeros((kk, 3)),
                 q=np.zeros((kk, 6)), perm=np.zeros(max(self.n, 1), np.int32))
        self._check(lib().nbx_debug_tree(self._h, t["depth"], t["leaf"], t["body"], t["cmask"], t["geo"],
                                         t["mass"], t["com"], t["q"], t["perm"]))
        out = {key: v[:k] for key, v in t.items() if key != "perm"}
        out["perm"] = t["perm"][: self.n]
        return out

    def debug_order(self):
        """(upload index, system) of the body in each device slot (key order)."""
        n = max(self.n, 1)
        ids, sys_ = np.zeros(n, np.int32), np.zeros(n, np.int32)
        self._check(lib().nbx_debug_order(self._h, ids, sys_))
        return ids[: self.n], sys_[: self.n]

    def autotune(self, candidates, dt: float, steps: int = 2):
        """Time each (leaf_size, group_size) or (leaf_size, group_size,
        walk_list) candidate on the current systems (physics mode), keep the
        fastest; returns (best, ms per step of each)."""
        leaf = np.array([c[0] for c in candidates], np.int32)
        group = np.array([c[1] for c in candidates], np.int32)
        ms = np.zeros(len(candidates))
        if any(len(c) > 2 for c in candidates):
            lst = np.array([c[2] for c in candidates], np.int32)
            self._check(lib().nbx_autotune_lists(self._h, leaf, group, lst, len(candidates), dt, steps, ms))
            best = int(np.argmin(ms))
            return (int(leaf[best]), int(group[best]), int(lst[best])), ms
        self._check(lib().nbx_autotune(self._h, leaf, group, len(candidates), dt, steps, ms))
        best = int(np.argmin(ms))
        return (int(leaf[best]), int(group[best])), ms

    def debug_groups(self):
        """Walk groups (first body in leaf order, count, system) of the current positions."""
        k = C.c_int64()
        z = np.zeros(1, np.int32)
        self._check(lib().nbx_debug_groups(self._h, 0, z, z, z, C.byref(k)))
        n = k.value
        first, count, sys_ = (np.zeros(max(n, 1), np.int32) for _ in range(3))
        self._check(lib().nbx_debug_groups(self._h, n, first, count, sys_, C.byref(k)))
        return first[:n], count[:n], sys_[:n]


class Server:
    """Continuous-batching request loop (nbx_server_* in include/nbx.h)."""

    def __init__(self, budget_bytes: int, device: int = 0):
        h = VP()
        st = lib().nbx_server_create(C.byref(h), device, int(budget_bytes))
        if st != 0:
            raise NbxError(st, "nbx_server_create failed")
        self._h = h
        self._n = {}

    def close(self):
        if getattr(self, "_h", None):
            lib().nbx_server_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def submit(self, p: Particles, nsteps: int, dt: float, mode: int = MODE_PHYSICS, precision: int = PREC_MIXED,
               theta: float = 0.5, eps: float = 0.01, leaf_size: int = 16, adaptive_dt: bool = False) -> int:
        arrs = [_f64(a) for a in (p.x, p.y, p.z, p.vx, p.vy, p.vz, p.m)]
        t = np.ascontiguousarray(p.type, dtype=np.int32)
        r = Request(len(p), *[a.ctypes.data_as(C.c_void_p) for a in arrs], t.ctypes.data_as(C.c_void_p), nsteps, dt,
                    mode, precision, theta, eps, leaf_size, 1 if adaptive_dt else 0)
        rid = C.c_int64()
        st = lib().nbx_server_submit(self._h, C.byref(r), C.byref(rid))
        if st != 0:
            raise NbxError(st, "submit refused")
        self._n[rid.value] = len(p)
        return rid.value

    def poll(self, rid: int) -> int:
        return lib().nbx_server_poll(self._h, rid)

    def wait(self, rid: int):
        """(state 6 x n, queue_ms, latency_ms)."""
        n = self._n.pop(rid)
        out = [np.zeros(max(n, 1)) for _ in range(6)]
        tm = RequestTiming()
        st = lib().nbx_server_wait(self._h, rid, *out, C.byref(tm))
        if st != 0:
            raise NbxError(st, f"request {rid} failed")
        return np.stack([o[:n] for o in out]), tm.queue_ms, tm.latency_ms

    def stats(self) -> dict:
        s = ServerStats()
        lib().nbx_server_stats(self._h, C.byref(s))
        return {k: getattr(s, k) for k, _ in ServerStats._fields_}
