"""The C-ABI library loads and exports every entry point the C headers
under include/ declare (no device needed)."""
import ctypes
import os
import re

import pytest

from next_amd import nbx

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    inc = os.path.join(ROOT, "include")
    text = "".join(open(os.path.join(inc, f)).read() for f in sorted(os.listdir(inc)) if f.endswith(".h"))
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(nbx_[a-z_0-9]+)\s*\((?!\*)", text)))


def test_header_declares_the_binding_set():
    assert set(declared_functions()) == set(nbx.EXPORTS)


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(nbx.LIB_PATH)
    missing = [f for f in declared_functions() if not hasattr(lib, f)]
    assert not missing, missing


def test_config_defaults_are_nexts():
    cfg = nbx.Config()
    nbx.lib().nbx_config_default(ctypes.byref(cfg))
    assert cfg.theta == 0.5          # step.h:32
    assert cfg.mode == nbx.MODE_REFERENCE and cfg.precision == nbx.PREC_FP64 and cfg.device == 0


def test_null_and_bad_arguments_are_rejected_without_a_device():
    L = nbx.lib()
    assert L.nbx_create(None, None) == -1
    assert L.nbx_get_timing(None, None) == -1
    assert L.nbx_set_theta(None, 0.5) == -1
    assert L.nbx_size(None) == -1
    assert L.nbx_last_error(None) == b"null context"
    cfg = nbx.Config()
    L.nbx_config_default(ctypes.byref(cfg))
    cfg.theta = -1.0
    h = ctypes.c_void_p()
    assert L.nbx_create(ctypes.byref(h), ctypes.byref(cfg)) == -1


def test_create_reports_the_device_state():
    """On a machine without a GPU nbx_create fails with NBX_ERR_NODEVICE; on
    a machine with one it succeeds."""
    try:
        e = nbx.Engine()
    except nbx.NbxError as err:
        assert err.status == -6
    else:
        e.close()


def test_missing_library_fails_loudly(monkeypatch, tmp_path):
    monkeypatch.setattr(nbx, "_lib", None)
    monkeypatch.setattr(nbx, "LIB_PATH", str(tmp_path / "libnbx.so"))
    with pytest.raises(RuntimeError, match="not built"):
        nbx.lib()


def test_next_interface_mirror_compiles_and_runs(tmp_path):
    """include/nbx_next.hpp offers NEXT's computeAdaptiveDt/Step signatures
    for a struct with NEXT's Particle fields; built and run here it reports
    the missing GPU as nbx_next::Error (NBX_ERR_NODEVICE); with a GPU it
    steps."""
    import subprocess
    src = tmp_path / "use.cpp"
    src.write_text(r'''
#include "nbx_next.hpp"
#include <cstdio>
#include <vector>
struct Particle { std::vector<double> x, y, z, vx, vy, vz, ax, ay, az, m; std::vector<int> type; };
using ParticleSystem = Particle;
int main() {
  ParticleSystem p;
  for (int i = 0; i < 3; ++i) {
    p.x.push_back(0.1 * i); p.y.push_back(0.2 * i * i); p.z.push_back(-0.1 * i);
    p.vx.push_back(0); p.vy.push_back(0); p.vz.push_back(0); p.m.push_back(1.0 / 3); p.type.push_back(0);
  }
  try {
    double dt = nbx_next::computeAdaptiveDt(p, 1e-3);
    nbx_next::Step(p, dt);
    std::printf("stepped %.17g\n", p.x[1]);
  } catch (const nbx_next::Error& e) {
    std::printf("error %d\n", (int)e.status);
  }
}
''')
    exe = tmp_path / "use"
    subprocess.run(["g++", "-std=c++17", "-I", os.path.join(ROOT, "include"), str(src), "-L",
                    os.path.join(ROOT, "next_amd"), "-lnbx", f"-Wl,-rpath,{os.path.join(ROOT, 'next_amd')}",
                    "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120).stdout
    assert out.startswith("stepped") or out.strip() == "error -6"
