"""Build the engine in-tree with hipcc for gfx950.

Products:
  next_amd/libnbx.so   the C-ABI library (include/nbx.h)
  next_amd/bin/next    the NEXT-compatible command line driver

    python -m next_amd.build          # or: python next_amd/build.py
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
HOST = os.path.join(PKG, "host")
OBJ = os.path.join(ROOT, "_build")
LIB = os.path.join(PKG, "libnbx.so")
BIN = os.path.join(PKG, "bin", "next")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

COMMON = ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}", "-I", os.path.join(ROOT, "include"),
          "-I", CSRC, "-Wall", "-Wno-unused-function", "-Wno-unused-command-line-argument"]
# Packed FP32 instructions are kept out of every kernel.
NO_PK = ["-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]
# Reference-mode kernels round every operation as written.
EXACT = ["-ffp-contract=off", "-fno-fast-math"]

SOURCES = [
    ("prims.hip", []),
    ("ref_kernels.hip", EXACT),
    ("phys_kernels.hip", []),
    ("nbx_api.cpp", []),
    ("serve.cpp", []),
    ("noalloc.hip", []),
]


def _compile(src: str, extra: list[str]) -> str:
    out = os.path.join(OBJ, os.path.basename(src) + ".o")
    lang = ["-x", "hip"] if src.endswith(".hip") else []
    cmd = [HIPCC, *COMMON, *NO_PK, *extra, *lang, "-c", os.path.join(CSRC, src), "-o", out]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    return out


def build_library() -> str:
    os.makedirs(OBJ, exist_ok=True)
    with ThreadPoolExecutor(max_workers=min(8, len(SOURCES))) as ex:
        objs = list(ex.map(lambda s: _compile(*s), SOURCES))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs, "-lpthread"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    return LIB


HDF5_ROOT = os.environ.get("NBX_HDF5_ROOT", "/opt/conda")


def hdf5_flags() -> list[str]:
    """HDF5 snapshots in the CLI when the image has libhdf5.

    The library's directory also holds an older libstdc++, so the binary
    does not search it: bin/hdf5/ gets links to libhdf5 and its libz only,
    and that is the binary's run path."""
    inc, lib = os.path.join(HDF5_ROOT, "include"), os.path.join(HDF5_ROOT, "lib")
    so = os.path.join(lib, "libhdf5.so")
    if not (os.path.exists(os.path.join(inc, "hdf5.h")) and os.path.exists(so)):
        return []
    side = os.path.join(os.path.dirname(BIN), "hdf5")
    os.makedirs(side, exist_ok=True)
    for name in os.listdir(lib):
        if name.startswith("libhdf5.so.") or name.startswith("libz.so."):
            link = os.path.join(side, name)
            if os.path.lexists(link):
                os.remove(link)
            os.symlink(os.path.join(lib, name), link)
    return ["-DNBX_HDF5", "-I", inc, so, "-Wl,-rpath,$ORIGIN/hdf5", f"-Wl,-rpath-link,{side}"]


def build_cli() -> str | None:
    srcs = sorted(os.path.join(HOST, f) for f in os.listdir(HOST) if f.endswith(".cpp")) if os.path.isdir(HOST) else []
    if not srcs:
        return None
    os.makedirs(os.path.dirname(BIN), exist_ok=True)
    cmd = ["g++", "-O2", "-std=c++17", "-Wall", "-I", os.path.join(ROOT, "include"), "-I", HOST, *srcs,
           "-o", BIN + ".tmp", "-L", PKG, "-lnbx", "-Wl,-rpath,$ORIGIN/..", *hdf5_flags()]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"CLI build failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(BIN + ".tmp", BIN)
    return BIN


def build_oracle() -> None:
    """The CPU oracle (test infrastructure) and, where the reference tree is
    present, the reference harness."""
    if shutil.which("make"):
        subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle")], check=True)


def main() -> None:
    print(build_library())
    print(build_cli())
    build_oracle()


if __name__ == "__main__":
    sys.exit(main())
