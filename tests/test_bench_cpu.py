"""bench.py on the CPU: the one-GPU contract (a request for more GPUs exits
with a message and prints no JSON line), the flop convention of both walks,
and the committed rocprofv3 summaries the JSON line reads, each derived
figure recomputed from the raw counters it came from."""
import glob
import os
import subprocess
import sys

import pytest

import bench

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_more_than_one_gpu_exits_with_a_message():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=300)
    assert r.returncode == 2
    assert "one GPU" in r.stderr
    assert r.stdout.strip() == ""


def test_flop_convention():
    # 19 per particle-particle and 55 per particle-cell interaction, for the
    # FP64 reference walk and the physics walk alike
    assert (bench.FLOPS_PP, bench.FLOPS_PC) == (19, 55)
    assert (bench.PHYS_FLOPS_PP, bench.PHYS_FLOPS_PC) == (19, 55)


@pytest.mark.parametrize("name", ["walk_traffic.json", "phys_walk_traffic.json"])
def test_traffic_summary_recomputes(name):
    t = bench.load_profile(name)
    assert t is not None
    ff = t["copy_bytes"] / 1024 / t["copy_fetch_kib"]      # bytes moved per KiB the counter reads
    wf = t["copy_bytes"] / 1024 / t["copy_write_kib"]
    assert t["fetch_factor"] == pytest.approx(ff, abs=1e-3)
    assert t["write_factor"] == pytest.approx(wf, abs=1e-3)
    assert t["fetch_bytes"] == pytest.approx(t["fetch_kib"] * 1024 * ff, rel=1e-5)
    assert t["write_bytes"] == pytest.approx(t["write_kib"] * 1024 * wf, rel=1e-5)
    assert t["bytes_per_launch"] == t["fetch_bytes"] + t["write_bytes"]
    counters = [c.split("--pmc ")[1].split(" -d")[0] for c in t["commands"]]
    assert counters == ["FETCH_SIZE", "WRITE_SIZE", "FETCH_SIZE", "WRITE_SIZE"]
    assert all(" -- python3 bench.py" in c for c in t["commands"])


def test_valu_summary_recomputes():
    v = bench.load_profile("walk_valu.json")
    c, d = v["counters"], v["derived"]
    cycles = c["GRBM_GUI_ACTIVE"] / 8                       # summed over the 8 XCDs
    slots = 1024 * cycles / 4                                # 1,024 SIMDs, 4 cycles per wave64 instruction
    assert d["valu_issue_slots"] == pytest.approx(slots, rel=1e-6)
    assert d["valu_instructions_per_slot"] == pytest.approx(c["SQ_INSTS_VALU"] / slots, abs=1e-4)
    assert d["valu_busy_frac"] == pytest.approx(c["SQ_ACTIVE_INST_VALU"] / slots, abs=1e-4)


def test_kernel_traces_name_their_commands():
    files = glob.glob(os.path.join(ROOT, "profiles", "*.md"))
    assert files
    for f in files:
        text = open(f).read()
        assert "rocprofv3 --kernel-trace --stats" in text and " -- python3 bench.py" in text, f


def test_phase_rates():
    # 124 B per body for kick + drift, 2^20 bodies, 0.03 ms per step
    out = bench.phase_rooflines(1 << 20, {"kick_ms": 0.3, "keys_ms": 0.3, "sort_ms": 2.0}, 10)
    gbs = 124 * (1 << 20) / 0.03e-3 / 1e9
    assert out["kick"]["GB/s"] == pytest.approx(gbs, abs=0.05)
    assert out["kick"]["frac"] == pytest.approx(gbs / 8000.0, abs=1e-4)


def test_phase_bytes_follow_the_tree():
    """Sort bytes follow the passes the depth needs; layout and upward bytes
    the internal nodes per body (1.55 M nodes over 2^20 bodies)."""
    n = 1 << 20
    b16 = bench.phase_bytes(n, 1549025, 16)
    b20 = bench.phase_bytes(n, 1549025, 20)
    assert b16["sort_ms"] == 8 + 24 * 6 + 140 and b20["sort_ms"] == 8 + 24 * 8 + 140
    # builds sorted by the nearly-sorted path; one of ten redone by the radix sort
    assert bench.phase_bytes(n, 1549025, 16, 10, 10, 0)["sort_ms"] == 8 + 8 + 12 + 140
    assert bench.phase_bytes(n, 1549025, 16, 10, 10, 1)["sort_ms"] == pytest.approx(28 + 15.2 + 140, abs=0.05)
    assert bench.phase_bytes(n, 1549025, 16, 10, 4, 0)["sort_ms"] == pytest.approx(0.6 * 152 + 0.4 * 28 + 140, abs=0.05)
    per_int = (1549025 - n) / n
    assert b16["build_ms"] == pytest.approx(188 + 8 * per_int, abs=0.1)
    assert b16["upward_ms"] == pytest.approx(per_int * 156 + (1549025 - 1) / n * 64, abs=0.1)
    out = bench.phase_rooflines(n, {"build_ms": 1.0, "upward_ms": 1.7, "nodes": 1549025, "depth": 16}, 10)
    assert set(out) == {"build", "upward"}
    assert out["build"]["GB/s"] == pytest.approx(b16["build_ms"] * n / 0.1e-3 / 1e9, abs=0.1)
