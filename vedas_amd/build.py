"""Build the native libraries in-tree.

  vedas_amd/libvedas_hip.so  -- HIP kernels + C ABI (hipcc, gfx950)
  vedas_amd/libvedas.so      -- C++ operator API over the C ABI (g++)
  oracle/liboracle.so        -- CPU restatement used by tests only (gcc)

Run as ``python -m vedas_amd.build`` or through ``__graft_entry__.build()``.
Every call recompiles.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "vedas_amd")
CSRC = os.path.join(PKG, "csrc")
HOST = os.path.join(PKG, "host")
INCLUDE = os.path.join(ROOT, "include")
ORACLE = os.path.join(ROOT, "oracle")
OBJ = os.path.join(ROOT, "_build")

ARCH = os.environ.get("VEDAS_OFFLOAD_ARCH", "gfx950")
HIPCC = os.environ.get("HIPCC", shutil.which("hipcc") or "/opt/rocm/bin/hipcc")

HIP_FLAGS = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
    # Packed FP32 instructions off for all device code.
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
    "-I", INCLUDE, "-I", CSRC,
]
HIP_SOURCES = ["vd_ctx.hip", "vd_prim.hip", "vd_radix.hip", "vd_index.hip",
               "vd_join.hip", "vd_relops.hip", "vd_batch.hip"]
HOST_SOURCES_HIPLIB = ["vd_gen.cpp"]


def _run(cmd: list[str]) -> None:
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(" ".join(cmd) + "\n" + r.stdout + r.stderr)
        raise RuntimeError(f"build step failed: {cmd[0]} {cmd[-1]}")


def build_hip() -> str:
    os.makedirs(OBJ, exist_ok=True)
    jobs = []
    for src in HIP_SOURCES:
        obj = os.path.join(OBJ, src.replace(".hip", ".o"))
        jobs.append([HIPCC, *HIP_FLAGS, "-x", "hip", "-c",
                     os.path.join(CSRC, src), "-o", obj])
    for src in HOST_SOURCES_HIPLIB:
        obj = os.path.join(OBJ, src.replace(".cpp", ".o"))
        jobs.append(["g++", "-O3", "-std=c++17", "-fPIC", "-pthread",
                     "-I", INCLUDE, "-c", os.path.join(CSRC, src), "-o", obj])
    with ThreadPoolExecutor(max_workers=min(8, os.cpu_count() or 1)) as ex:
        list(ex.map(_run, jobs))
    objs = [j[-1] for j in jobs]
    out = os.path.join(PKG, "libvedas_hip.so")
    tmp = out + ".tmp"
    _run([HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", *objs,
          "-pthread", "-Wl,-rpath,/opt/rocm/lib", "-o", tmp])
    os.replace(tmp, out)
    return out


def build_host() -> str:
    """The C++ operator API (QueryJob & co.) links only the C ABI."""
    srcs = sorted(f for f in os.listdir(HOST) if f.endswith(".cpp"))
    out = os.path.join(PKG, "libvedas.so")
    tmp = out + ".tmp"
    _run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-pthread",
          "-I", INCLUDE, "-I", HOST,
          *[os.path.join(HOST, s) for s in srcs],
          "-L", PKG, "-lvedas_hip", "-Wl,-rpath,$ORIGIN", "-o", tmp])
    os.replace(tmp, out)
    return out


def build_oracle() -> str:
    """The oracle (test checker) and the CPU port (baseline): test and
    benchmark infrastructure, never linked by the package."""
    out = os.path.join(ORACLE, "liboracle.so")
    tmp = out + ".tmp"
    _run(["gcc", "-O2", "-std=c11", "-fPIC", "-shared",
          os.path.join(ORACLE, "vedas_oracle.c"), "-o", tmp])
    os.replace(tmp, out)
    port = os.path.join(ORACLE, "libcpuport.so")
    _run(["g++", "-O3", "-std=c++17", "-fPIC", "-shared", "-pthread",
          os.path.join(ORACLE, "cpu_port.cpp"), "-o", port + ".tmp"])
    os.replace(port + ".tmp", port)
    return out


def build_all() -> None:
    build_hip()
    build_host()
    build_oracle()


if __name__ == "__main__":
    what = sys.argv[1:] or ["all"]
    for w in what:
        {"all": build_all, "hip": build_hip, "host": build_host,
         "oracle": build_oracle}[w]()
    print("built:", ", ".join(what))
