"""The request server's scheduler on the host, over a stand-in engine.

nbx_server_create_with_ops (include/nbx_serve_ops.h) runs the same
scheduler as nbx_server_create over a table of engine calls.  Here the table
is a small Python engine (ctypes callbacks, no GPU): a "step" adds 1 to every
x, so a request's correct result is its initial x plus its step count, and
failures can be injected into any call.  This checks the scheduler's own
logic: no request ever receives another request's bodies, a failed engine
call fails its batch and leaves the server usable, nbx_server_destroy
releases threads that are waiting, and idle batches are retired.
"""
import ctypes as C
import threading
import time

import numpy as np
import pytest

from next_amd import ic
from next_amd.nbx import MODE_PHYSICS, NbxError, Server, lib

VP = C.c_void_p
DP = C.POINTER(C.c_double)
IP = C.POINTER(C.c_int32)
I64P = C.POINTER(C.c_int64)

CREATE = C.CFUNCTYPE(C.c_int, C.POINTER(VP), VP)
DESTROY = C.CFUNCTYPE(None, VP)
LAST_ERROR = C.CFUNCTYPE(VP, VP)
ADD = C.CFUNCTYPE(C.c_int, VP, C.c_int64, DP, DP, DP, DP, DP, DP, DP, IP, IP)
REMOVE = C.CFUNCTYPE(C.c_int, VP, C.c_int32, DP, DP, DP, DP, DP, DP)
RUN = C.CFUNCTYPE(C.c_int, VP, C.c_double, C.c_int32, DP)
UPLOAD = C.CFUNCTYPE(C.c_int, VP, C.c_int32, I64P, DP, DP, DP, DP, DP, DP, DP, IP)
TRIM = C.CFUNCTYPE(C.c_int, VP)
BYTES = C.CFUNCTYPE(C.c_uint64, VP)


class Ops(C.Structure):
    _fields_ = [("create", CREATE), ("destroy", DESTROY), ("last_error", LAST_ERROR), ("add_system", ADD),
                ("remove_system", REMOVE), ("run", RUN), ("upload_batch", UPLOAD), ("trim", TRIM),
                ("device_bytes", BYTES)]


NBX_ERR_HIP, NBX_ERR_STATE = -5, -7


class FakeEngine:
    """Contexts holding systems as numpy arrays; one step adds 1 to x.
    `fail_remove` / `fail_run`: body counts whose removal / whose batch's
    step fails (after scribbling on the outputs, as a half-done copy would)."""

    def __init__(self, step_s=0.0):
        self.ctx = {}
        self.next = 1
        self.destroyed = []
        self.fail_remove = set()
        self.fail_run = set()
        self.step_s = step_s
        self.lock = threading.Lock()
        self._err = C.create_string_buffer(b"injected failure")
        self.ops = Ops(CREATE(self.create), DESTROY(self.destroy), LAST_ERROR(self.last_error), ADD(self.add),
                       REMOVE(self.remove), RUN(self.run), UPLOAD(self.upload), TRIM(lambda c: 0),
                       BYTES(self.bytes))

    def create(self, out, cfg):
        with self.lock:
            h = self.next
            self.next += 1
            self.ctx[h] = []
        out[0] = h
        return 0

    def destroy(self, h):
        with self.lock:
            self.destroyed.append(h)
            self.ctx.pop(h, None)

    def last_error(self, h):
        return C.addressof(self._err)

    def add(self, h, n, x, y, z, vx, vy, vz, m, t, sys_out):
        arrs = [np.ctypeslib.as_array(a, shape=(n,)).copy() for a in (x, y, z, vx, vy, vz)]
        with self.lock:
            self.ctx[h].append(arrs)
            sys_out[0] = len(self.ctx[h]) - 1
        return 0

    def remove(self, h, k, *outs):
        with self.lock:
            sysl = self.ctx[h]
            arrs = sysl[k]
            n = len(arrs[0])
            dst = [np.ctypeslib.as_array(o, shape=(n,)) for o in outs]
            if n in self.fail_remove:
                for d in dst:
                    d[:] = -12345.0
                return NBX_ERR_HIP
            for d, a in zip(dst, arrs):
                d[:] = a
            del sysl[k]
        return 0

    def run(self, h, dt, nsteps, dts):
        if self.step_s:
            time.sleep(self.step_s)
        with self.lock:
            sysl = self.ctx[h]
            if any(len(a[0]) in self.fail_run for a in sysl):
                return NBX_ERR_HIP
            for a in sysl:
                a[0] += nsteps
        return 0

    def upload(self, h, nsys, counts, *arrs):
        assert nsys == 0
        with self.lock:
            self.ctx[h] = []
        return 0

    def bytes(self, h):
        with self.lock:
            return 100 * sum(len(a[0]) for a in self.ctx.get(h, []))


class HostServer(Server):
    """Server over a FakeEngine (no GPU)."""

    def __init__(self, fake, budget_bytes=1 << 40):
        h = VP()
        st = lib().nbx_server_create_with_ops(C.byref(h), 0, int(budget_bytes), C.byref(fake.ops))
        if st != 0:
            raise NbxError(st, "nbx_server_create_with_ops failed")
        self._h = h
        self._n = {}
        self.fake = fake


@pytest.fixture(autouse=True)
def _bind():
    L = lib()
    L.nbx_server_create_with_ops.argtypes = [C.POINTER(VP), C.c_int, C.c_uint64, C.POINTER(Ops)]


def body(n, seed):
    return ic.uniform_sphere(n, seed=seed)


def stepped(x, nsteps):
    """x after nsteps steps of the stand-in engine (1 added per step)."""
    x = x.copy()
    for _ in range(nsteps):
        x += 1
    return x


def submit(srv, p, nsteps, dt=1e-3):
    return srv.submit(p, nsteps, dt, mode=MODE_PHYSICS)


def test_results_are_each_requests_own():
    fake = FakeEngine()
    ps = [body(n, k) for k, n in enumerate((100, 37, 250, 64, 5))]
    steps = [3, 7, 1, 4, 9]
    with HostServer(fake) as srv:
        ids = [submit(srv, p, s) for p, s in zip(ps, steps)]
        res = [srv.wait(i)[0] for i in ids]
    for p, s, r in zip(ps, steps, res):
        assert np.array_equal(r[0], stepped(p.x, s)) and np.array_equal(r[1], p.y)


def test_failed_remove_fails_its_batch_and_no_one_gets_wrong_bodies():
    """A removal that fails (after writing garbage into the outputs) fails
    every request of that batch; none of them, and no later request,
    receives another request's bodies."""
    fake = FakeEngine(step_s=0.002)
    fake.fail_remove.add(77)
    with HostServer(fake) as srv:
        a = submit(srv, body(100, 1), 6)
        b = submit(srv, body(77, 2), 2)          # its removal fails
        c = submit(srv, body(50, 3), 9)
        d = submit(srv, body(40, 4), 3, dt=2e-3)  # another batch: unaffected
        out = {}
        for name, rid, p0, s in (("a", a, body(100, 1), 6), ("b", b, body(77, 2), 2), ("c", c, body(50, 3), 9),
                                 ("d", d, body(40, 4), 3)):
            try:
                st = srv.wait(rid)[0]
                assert np.array_equal(st[0], stepped(p0.x, s)), name
                out[name] = "ok"
            except NbxError as e:
                out[name] = e.status
        later = submit(srv, body(30, 5), 4)
        st = srv.wait(later)[0]
        assert np.array_equal(st[0], stepped(body(30, 5).x, 4))
    assert out["b"] == NBX_ERR_HIP
    assert out["d"] == "ok"
    assert out["a"] in ("ok", NBX_ERR_HIP) and out["c"] in ("ok", NBX_ERR_HIP)


def test_failed_step_fails_the_batch_and_server_recovers():
    fake = FakeEngine()
    fake.fail_run.add(13)
    with HostServer(fake) as srv:
        bad = submit(srv, body(13, 1), 5)
        with pytest.raises(NbxError) as ei:
            srv.wait(bad)
        assert ei.value.status == NBX_ERR_HIP
        good = submit(srv, body(20, 2), 5)
        assert np.array_equal(srv.wait(good)[0][0], stepped(body(20, 2).x, 5))


def test_destroy_releases_waiting_threads():
    """nbx_server_destroy fails the requests not done and returns only after
    threads already waiting on them have left nbx_server_wait."""
    fake = FakeEngine(step_s=0.01)
    srv = HostServer(fake)
    rid = submit(srv, body(10, 1), 100000)
    got = []

    def waiter():
        try:
            srv.wait(rid)
            got.append(0)
        except NbxError as e:
            got.append(e.status)

    t = threading.Thread(target=waiter)
    t.start()
    time.sleep(0.2)
    srv.close()
    t.join(timeout=10)
    assert not t.is_alive()
    assert got == [NBX_ERR_STATE]


def test_destroy_fails_requests_still_queued():
    fake = FakeEngine(step_s=0.01)
    srv = HostServer(fake, budget_bytes=700 * 100)   # one 100-body request at a time
    r1 = submit(srv, body(100, 1), 1000)
    r2 = submit(srv, body(100, 2), 3)                # waits for budget
    got = []
    t = threading.Thread(target=lambda: got.append(lib().nbx_server_poll(srv._h, r2)))
    t.start()
    t.join()
    assert got == [0]                                # queued
    statuses = []

    def waiter(rid):
        try:
            srv.wait(rid)
            statuses.append(0)
        except NbxError as e:
            statuses.append(e.status)

    ts = [threading.Thread(target=waiter, args=(r,)) for r in (r1, r2)]
    for th in ts:
        th.start()
    time.sleep(0.1)
    srv.close()
    for th in ts:
        th.join(timeout=10)
    assert sorted(statuses) == [NBX_ERR_STATE, NBX_ERR_STATE]


def test_idle_batches_are_retired():
    """Each distinct setting gets a batch; a batch idle for the retirement
    period ends (its context is destroyed), and a later request with the same
    setting gets a fresh batch."""
    fake = FakeEngine()
    with HostServer(fake) as srv:
        for k in range(5):                            # five settings (dt values)
            rid = submit(srv, body(20, k), 2, dt=1e-3 * (k + 1))
            srv.wait(rid)
        assert srv.stats()["batches"] <= 5
        deadline = time.time() + 10
        while srv.stats()["batches"] > 0 and time.time() < deadline:
            time.sleep(0.05)
        assert srv.stats()["batches"] == 0
        assert len(fake.destroyed) == 5 and not fake.ctx
        rid = submit(srv, body(20, 9), 3, dt=1e-3)
        assert np.array_equal(srv.wait(rid)[0][0], stepped(body(20, 9).x, 3))
        assert srv.stats()["batches"] == 1
