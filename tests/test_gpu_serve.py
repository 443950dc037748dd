"""The continuous-batching request loop: each request's result equals the
same simulation run alone; admission respects the memory budget."""
import ctypes as C
import threading
import time

import numpy as np
import pytest

from next_amd import ic
from next_amd.nbx import (MODE_PHYSICS, MODE_REFERENCE, PREC_FP64, PREC_MIXED, Engine, NbxError, RequestTiming, Server,
                          lib)

pytestmark = pytest.mark.gpu

BYTES_PER_BODY = 700   # the server's admission estimate


def requests():
    out = []
    for k in range(8):
        n = [900, 2500, 4000, 1300][k % 4]
        p = ic.plummer(n, seed=200 + k) if k % 2 == 0 else ic.uniform_sphere(n, seed=300 + k)
        mode = MODE_PHYSICS if k % 3 else MODE_REFERENCE
        prec = PREC_MIXED if k % 4 < 2 else PREC_FP64
        theta = 0.5 if mode == MODE_REFERENCE else [0.4, 0.5, 0.7][k % 3]
        out.append(dict(p=p, nsteps=3 + k, dt=1e-3, mode=mode, precision=prec, theta=theta))
    return out


def alone(r):
    with Engine(theta=r["theta"], mode=r["mode"], precision=r["precision"], eps=0.01, leaf_size=16,
                adaptive_dt=False) as e:
        e.upload(r["p"])
        e.run(r["dt"], r["nsteps"])
        return e.download()


def test_results_equal_alone_and_overlap():
    rs = requests()
    with Server(budget_bytes=200 << 30) as srv:
        ids = []
        for r in rs:
            ids.append(srv.submit(r["p"], r["nsteps"], r["dt"], mode=r["mode"], precision=r["precision"],
                                  theta=r["theta"], eps=0.01, leaf_size=16))
            time.sleep(0.002)                      # arrivals while others run
        results = [srv.wait(i) for i in ids]
        st = srv.stats()
    assert st["batches"] >= 3
    for r, (state, q_ms, lat_ms) in zip(rs, results):
        assert lat_ms >= q_ms >= 0
        assert np.array_equal(state, alone(r))


def test_budget_is_respected():
    rs = [dict(p=ic.plummer(3000, seed=k), nsteps=4, dt=1e-3) for k in range(5)]
    budget = int(3000 * BYTES_PER_BODY * 1.5)     # one request at a time
    peak = [0, 0]
    stop = threading.Event()
    with Server(budget_bytes=budget) as srv:
        def watch():
            while not stop.is_set():
                st = srv.stats()
                peak[0] = max(peak[0], st["bytes_in_use"])
                peak[1] = max(peak[1], st["device_bytes"])
                time.sleep(0.0005)
        t = threading.Thread(target=watch)
        t.start()
        ids = [srv.submit(r["p"], r["nsteps"], r["dt"]) for r in rs]
        res = [srv.wait(i) for i in ids]
        stop.set()
        t.join()
        with pytest.raises(NbxError):
            srv.submit(ic.plummer(100000, seed=1), 1, 1e-3)      # can never fit
    assert peak[0] <= budget and peak[1] <= budget
    for r, (state, _, _) in zip(rs, res):
        assert np.array_equal(state, alone(dict(r, mode=MODE_PHYSICS, precision=PREC_MIXED, theta=0.5)))


def test_idle_batches_release_memory_and_waits_are_safe():
    """A batch that empties frees its device buffers (the server's
    device_bytes falls to the few bytes an empty context keeps), and two
    threads waiting on one request get the result once and NBX_ERR_ARG once."""
    p = ic.plummer(20000, seed=77)
    with Server(budget_bytes=200 << 30) as srv:
        rid = srv.submit(p, 3, 1e-3)
        n = len(p)
        got = []

        def waiter():
            out = [np.zeros(n) for _ in range(6)]
            tm = RequestTiming()
            got.append(lib().nbx_server_wait(srv._h, rid, *out, C.byref(tm)))

        ts = [threading.Thread(target=waiter) for _ in range(2)]
        for t in ts:
            t.start()
        for t in ts:
            t.join()
        assert sorted(got) == [-1, 0]                # NBX_ERR_ARG, NBX_OK
        deadline = time.time() + 10
        while srv.stats()["device_bytes"] > 1 << 20 and time.time() < deadline:
            time.sleep(0.01)
        st = srv.stats()
        assert st["batches"] <= 1 and st["running"] == 0 and st["bytes_in_use"] == 0   # idle batches retire
        assert st["device_bytes"] <= 1 << 20         # was ~20,000 bodies' worth (> 10 MB)


@pytest.mark.parametrize("mode", [MODE_REFERENCE, MODE_PHYSICS])
def test_admission_estimate_covers_a_context(mode):
    """The server admits requests by an estimate of BYTES_PER_BODY device
    bytes per body; a context that has stepped a system and computed its
    diagnostics (every buffer a request can make it allocate) holds less."""
    p = ic.plummer(100000, seed=78)
    prec = PREC_MIXED if mode == MODE_PHYSICS else PREC_FP64
    with Engine(theta=0.5, mode=mode, precision=prec, eps=0.01, leaf_size=16, adaptive_dt=True) as e:
        if mode == MODE_PHYSICS:
            e.set_alpha(1e-3)
        e.upload(p)
        e.run(1e-3, 3)
        e.energy(direct=True)
        if mode == MODE_PHYSICS:
            e.energy(direct=False)
        per_body = e.device_bytes() / len(p)
    print(f"device bytes per body: {per_body:.0f}")
    assert per_body <= BYTES_PER_BODY
