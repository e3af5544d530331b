"""Build recipe for the native libraries of the triangle-counting engine.

Produces, in-tree under ``hetegeneous_parallel_programming_amd/lib/``:

* ``libtc_host.so``  -- host graph substrate (g++, OpenMP): CSR build,
  generators, orientation, partitioning, GraphMiner I/O.
* ``libtc_hip.so``   -- HIP kernels for gfx950 plus the device C ABI
  (hipcc), linked against ``libtc_host.so``.
* ``tc_run``          -- command-line driver mirroring the reference's
  ``main`` (successful_push_pull/main.cc:11-36).

Run ``python -m hetegeneous_parallel_programming_amd.build``.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
INCLUDE = os.path.join(ROOT, "include")
CSRC = os.path.join(PKG, "csrc")
LIB = os.path.join(PKG, "lib")
OBJ = os.path.join(PKG, "lib", "obj")

HOST_SOURCES = ["host/csr.cpp", "host/gen.cpp", "host/partition.cpp", "host/metis.cpp", "host/arena.cpp"]
HIP_SOURCES = ["hip/tc_kernels.hip", "hip/tc_runtime.cpp", "hip/tc_batch.cpp",
               "hip/tc_server.cpp", "hip/tc_tune.cpp"]

HOST_FLAGS = ["-O3", "-std=c++17", "-fPIC", "-fopenmp", "-march=x86-64-v3",
              "-Wall", "-Wno-unused-function", "-I" + INCLUDE]
# Packed-FP32 instructions are disabled for all device code.
HIP_FLAGS = ["-O3", "-std=c++17", "-fPIC", "--offload-arch=gfx950",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-Wall", "-Wno-unused-function", "-Wno-unused-command-line-argument",
             "-I" + INCLUDE, "-I" + os.path.join(CSRC, "hip")]


def _run(cmd: list[str]) -> None:
    proc = subprocess.run(cmd, capture_output=True, text=True)
    if proc.returncode != 0:
        sys.stderr.write(" ".join(cmd) + "\n" + proc.stdout + proc.stderr)
        raise RuntimeError(f"build step failed: {cmd[0]} {cmd[-1]}")


def _hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found (ROCm install expected under /opt/rocm)")


def build_host() -> str:
    os.makedirs(OBJ, exist_ok=True)
    out = os.path.join(LIB, "libtc_host.so")
    srcs = [os.path.join(CSRC, s) for s in HOST_SOURCES]
    _run(["g++", *HOST_FLAGS, "-shared", *srcs, "-o", out + ".tmp"])
    os.replace(out + ".tmp", out)
    return out


def build_hip(jobs: int = 8, out_dir: str | None = None, defines: tuple = ()) -> str:
    """Compile the HIP library (into out_dir, default lib/) with extra -D defines."""
    lib_dir = out_dir or LIB
    obj_dir = os.path.join(lib_dir, "obj")
    os.makedirs(obj_dir, exist_ok=True)
    if lib_dir != LIB:
        shutil.copy2(os.path.join(LIB, "libtc_host.so"), os.path.join(lib_dir, "libtc_host.so"))
    hipcc = _hipcc()
    dflags = [f"-D{d}" for d in defines]
    srcs = [s for s in HIP_SOURCES if os.path.exists(os.path.join(CSRC, s))]
    objs = []

    def compile_one(rel: str) -> str:
        src = os.path.join(CSRC, rel)
        obj = os.path.join(obj_dir, os.path.basename(rel) + ".o")
        lang = ["-x", "hip"] if rel.endswith(".hip") else []
        _run([hipcc, *HIP_FLAGS, *dflags, *lang, "-c", src, "-o", obj])
        return obj

    with cf.ThreadPoolExecutor(max_workers=max(1, min(jobs, len(srcs)))) as ex:
        objs = list(ex.map(compile_one, srcs))
    out = os.path.join(lib_dir, "libtc_hip.so")
    _run([hipcc, "--offload-arch=gfx950", "-shared", "-fPIC", *objs, "-L" + lib_dir, "-ltc_host",
          "-Wl,-rpath,$ORIGIN", "-o", out + ".tmp"])
    os.replace(out + ".tmp", out)
    cli = os.path.join(CSRC, "hip", "tc_run.cpp")
    if os.path.exists(cli):
        exe = os.path.join(lib_dir, "tc_run")
        _run([hipcc, *HIP_FLAGS, *dflags, cli, "-L" + lib_dir, "-ltc_hip", "-ltc_host",
              "-Wl,-rpath,$ORIGIN", "-o", exe + ".tmp"])
        os.replace(exe + ".tmp", exe)
    return out


def build_all() -> None:
    build_host()
    build_hip()


if __name__ == "__main__":
    what = sys.argv[1] if len(sys.argv) > 1 else "all"
    {"host": build_host, "hip": build_hip, "all": build_all}[what]()
    print("built", what)
