"""nbx_accel_noalloc, the entry point that allocates no device memory,
against nbx_accel (the context path: same kernels, so the same bits).

Two calls in a row share the library's static device array, the second
with a smaller input; the array keeps the first call's contents, so a
stale value the second call read would show.  A two-word-key case (bodies
sharing 21 octree digits) and the error paths follow."""
import numpy as np
import pytest

from next_amd import ic, nbx
from next_amd.nbx import Engine

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

GOLDEN_DEEP = "hand_deep"


def on_device(p):
    dev = torch.device("cuda:0")
    cols = [torch.from_numpy(np.ascontiguousarray(a, dtype=np.float64)).to(dev) for a in (p.x, p.y, p.z, p.m)]
    t = torch.from_numpy(np.ascontiguousarray(p.type, dtype=np.int32)).to(dev)
    out = [torch.full((max(len(p), 1),), float("nan"), dtype=torch.float64, device=dev) for _ in range(3)]
    return cols, t, out


def noalloc_accel(p, theta):
    cols, t, out = on_device(p)
    stream = torch.cuda.current_stream().cuda_stream
    print(f"nbx_accel_noalloc n={len(p)} theta={theta}", flush=True)
    nbx.accel_noalloc(len(p), *[c.data_ptr() for c in cols], t.data_ptr(), theta, *[o.data_ptr() for o in out],
                      stream=stream)
    torch.cuda.synchronize()
    return np.stack([o.cpu().numpy()[: len(p)] for o in out], axis=1)


def context_accel(p, theta):
    with Engine(theta=theta) as e:
        e.upload(p)
        return e.accel()


def test_two_calls_in_a_row_equal_nbx_accel():
    first = ic.plummer(nbx_max := 16384, seed=11, type_fraction_dm=0.1)
    second = ic.plummer(3000, seed=12, type_fraction_dm=0.1)
    for p, theta in ((first, 0.5), (second, 0.3)):
        a = noalloc_accel(p, theta)
        b = context_accel(p, theta)
        assert np.all(np.isfinite(a))
        assert np.array_equal(a, b), (len(p), theta, float(np.max(np.abs(a - b))))
    assert nbx_max == 16384


def test_two_word_keys_equal_nbx_accel(gold):
    g = gold(GOLDEN_DEEP)
    p = ic.Particles(*(np.array(g[k]) for k in ("x", "y", "z", "vx", "vy", "vz", "m", "type")))
    for theta in (0.5, 0.0):
        a = noalloc_accel(p, theta)
        assert np.array_equal(a, context_accel(p, theta)), theta


def test_limits_are_errors():
    # more bodies than the array holds: refused on the host
    big = 16385
    with pytest.raises(nbx.NbxError) as e:
        nbx.accel_noalloc(big, 1, 1, 1, 1, 1, 0.5, 1, 1, 1)
    assert e.value.status == -1
    with pytest.raises(nbx.NbxError) as e:
        nbx.accel_noalloc(10, 0, 0, 0, 0, 0, 0.5, 0, 0, 0)
    assert e.value.status == -1
    # 8,192 close pairs: every pair needs a chain of about 30 single-child
    # nodes, more than the array's 65,536 node records -> NBX_ERR_NOMEM,
    # decided from the node counts before any node is written
    rng = np.random.Generator(np.random.PCG64(5))
    c = rng.uniform(-1, 1, size=(8192, 3))
    pos = np.concatenate([c, c + 1e-9])
    z = np.zeros(len(pos))
    p = ic.Particles(pos[:, 0].copy(), pos[:, 1].copy(), pos[:, 2].copy(), z, z, z,
                     np.full(len(pos), 1.0 / len(pos)), np.zeros(len(pos), np.int32))
    with pytest.raises(nbx.NbxError) as e:
        noalloc_accel(p, 0.5)
    assert e.value.status == -4
    # the array is usable again after the refusal
    q = ic.plummer(2000, seed=13)
    assert np.array_equal(noalloc_accel(q, 0.5), context_accel(q, 0.5))
