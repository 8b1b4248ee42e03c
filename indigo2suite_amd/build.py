"""Build recipes for the native libraries (compiled in place, next to the sources).

* ``libindigo_graph.so`` -- host graph substrate (g++ -fopenmp).
* ``libindigo_hip.so``   -- the HIP kernels and the C ABI (hipcc, gfx950).
* ``oracle/liboracle.so`` and, when the reference tree is present,
  ``oracle/_ref/*`` -- test infrastructure, built with ``make`` in ``oracle/``.

Every function compiles unconditionally; callers decide whether to build.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
REPO_DIR = os.path.dirname(PKG_DIR)
CSRC = os.path.join(PKG_DIR, "csrc")
INCLUDE = os.path.join(REPO_DIR, "include")
ORACLE_DIR = os.path.join(REPO_DIR, "oracle")
REFERENCE_DIR = "/root/reference"

GRAPH_LIB = os.path.join(PKG_DIR, "libindigo_graph.so")
HIP_LIB = os.path.join(PKG_DIR, "libindigo_hip.so")
ORACLE_LIB = os.path.join(ORACLE_DIR, "liboracle.so")

HIP_SOURCES = ["runtime.hip", "bfs.hip", "sssp.hip", "cc.hip", "mis.hip", "pr.hip", "tc.hip",
               "batched.hip", "mlp.hip", "features.hip"]
OFFLOAD_ARCH = "gfx950"
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")


def _run(cmd, cwd=None):
    res = subprocess.run(cmd, cwd=cwd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout)
        raise RuntimeError("build step failed: " + " ".join(cmd))
    return res.stdout


def build_graph_lib() -> str:
    src = os.path.join(CSRC, "host", "graph_host.cpp")
    tmp = GRAPH_LIB + ".tmp"
    _run(["g++", "-O3", "-fopenmp", "-fPIC", "-shared", "-std=c++17", "-Wl,-soname,libindigo_graph.so",
          "-I" + INCLUDE, src, "-o", tmp])
    os.replace(tmp, GRAPH_LIB)
    return GRAPH_LIB


def hip_sources():
    return [os.path.join(CSRC, s) for s in HIP_SOURCES if os.path.exists(os.path.join(CSRC, s))]


def build_hip_lib(jobs: int = 8) -> str:
    """Compile each .hip file to an object in parallel, then link the .so."""
    objdir = os.path.join(PKG_DIR, "_build")
    os.makedirs(objdir, exist_ok=True)
    common = [HIPCC, "-O3", "-std=c++17", "-fPIC", "--offload-arch=" + OFFLOAD_ARCH,
              "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
              "-I" + INCLUDE, "-I" + CSRC, "-Wno-unused-result"]
    procs = []
    objs = []
    for src in hip_sources():
        obj = os.path.join(objdir, os.path.basename(src) + ".o")
        objs.append(obj)
        procs.append((src, subprocess.Popen(common + ["-c", src, "-o", obj], stdout=subprocess.PIPE,
                                            stderr=subprocess.STDOUT, text=True)))
        while sum(p.poll() is None for _, p in procs) >= jobs:
            procs[0][1].wait()
    failed = []
    for src, p in procs:
        out, _ = p.communicate()
        if p.returncode != 0:
            failed.append(src)
            sys.stderr.write(out)
    if failed:
        raise RuntimeError("hipcc failed for: " + ", ".join(failed))
    tmp = HIP_LIB + ".tmp"
    _run([HIPCC, "-shared", "-fPIC", "--offload-arch=" + OFFLOAD_ARCH, *objs,
          "-L" + PKG_DIR, "-lindigo_graph", "-o", tmp,
          "-Wl,-rpath,$ORIGIN"])
    os.replace(tmp, HIP_LIB)
    return HIP_LIB


def build_oracle(with_reference: bool | None = None) -> str:
    if with_reference is None:
        with_reference = os.path.isdir(REFERENCE_DIR)
    targets = ["all"] + (["ref"] if with_reference else [])
    _run(["make", "-s", "-j8", "-C", ORACLE_DIR, *targets])
    return ORACLE_LIB


def build_all() -> None:
    build_graph_lib()
    build_hip_lib()
    build_oracle()


if __name__ == "__main__":
    build_all()
    print("built:", GRAPH_LIB, HIP_LIB, ORACLE_LIB)
