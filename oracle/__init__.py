"""ORACLE -- test infrastructure, not product code.

Python access to the C++/OpenMP and C++-threads restatements of the
reference's CPU path (oracle/oracle.cpp and oracle/oracle_cpp.cpp ->
liboracle.so) and to the reference's own OpenMP and C++-threads variants
compiled in place into oracle/_ref/ (present only where the
reference tree was available at build time).  Only tests/, the smoke check
and bench.py's CPU-baseline leg use this package, and only as the checker or
as the timed CPU baseline.
"""
from __future__ import annotations

import ctypes
import os
import subprocess
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
REF_DIR = os.path.join(HERE, "_ref")

_lib = None


class _G(ctypes.Structure):
    _fields_ = [("nodes", ctypes.c_int), ("edges", ctypes.c_int), ("nindex", ctypes.POINTER(ctypes.c_int)),
                ("nlist", ctypes.POINTER(ctypes.c_int)), ("eweight", ctypes.POINTER(ctypes.c_int))]


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} missing: run `make -C oracle` (or __graft_entry__.build())")
        L = ctypes.CDLL(LIB_PATH)
        G = ctypes.POINTER(_G)
        ip = ctypes.POINTER(ctypes.c_int)
        up = ctypes.POINTER(ctypes.c_uint)
        bp = ctypes.POINTER(ctypes.c_ubyte)
        fp = ctypes.POINTER(ctypes.c_float)
        dp = ctypes.POINTER(ctypes.c_double)
        ullp = ctypes.POINTER(ctypes.c_ulonglong)

        def sig(name, res, args):
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args

        sig("orc_bfs_serial", None, [ctypes.c_int, G, ip])
        sig("orc_bfs_omp_topo_push", ctypes.c_double, [ctypes.c_int, G, ip, ip])
        sig("orc_bfs_omp_topo_pull", ctypes.c_double, [ctypes.c_int, G, ip, ip])
        sig("orc_sssp_serial", None, [ctypes.c_int, G, ip])
        sig("orc_sssp_omp_edge_data", ctypes.c_double, [ctypes.c_int, G, ip, ip, ip])
        sig("orc_sssp_omp_topo_pull", ctypes.c_double, [ctypes.c_int, G, ip, ip])
        sig("orc_cc_omp_topo_push", ctypes.c_double, [G, ip, ip])
        sig("orc_cc_omp_topo_pull", ctypes.c_double, [G, ip, ip])
        sig("orc_cc_verify", ctypes.c_int, [G, ip])
        sig("orc_mis_hash", ctypes.c_uint, [ctypes.c_uint])
        sig("orc_mis_omp_topo_push", ctypes.c_double, [G, up, bp, ip])
        sig("orc_mis_omp_edge", ctypes.c_double, [G, ip, bp, ip])
        sig("orc_mis_serial_greedy", None, [G, bp])
        sig("orc_mis_verify", ctypes.c_int, [G, bp])
        sig("orc_pr_omp_pull", ctypes.c_double, [G, fp, ctypes.c_double, ctypes.c_int, ip])
        sig("orc_pr_fp64", ctypes.c_int, [G, dp, ctypes.c_double, ctypes.c_int, ctypes.c_int, dp])
        sig("orc_pr_check", ctypes.c_double, [G, fp])
        sig("orc_tc_serial", ctypes.c_ulonglong, [G])
        sig("orc_tc_forward", ctypes.c_ulonglong, [G])
        sig("orc_tc_omp_vertex", ctypes.c_double, [G, ullp])
        sig("orc_tc_omp_edge", ctypes.c_double, [G, ip, ullp])
        sig("orc_num_threads", ctypes.c_int, [])
        ci = ctypes.c_int
        sig("orc_cpp_bfs_topo_push", ctypes.c_double, [ci, G, ip, ci, ip])
        sig("orc_cpp_sssp_topo_pull", ctypes.c_double, [ci, G, ip, ci, ip])
        sig("orc_cpp_cc_topo_push", ctypes.c_double, [G, ip, ci, ip])
        sig("orc_cpp_mis_topo_push", ctypes.c_double, [G, up, bp, ci, ip])
        sig("orc_cpp_pr_pull", ctypes.c_double, [G, fp, ctypes.c_double, ci, ci, ip])
        sig("orc_cpp_tc_vertex", ctypes.c_double, [G, ullp, ci])
        _lib = L
    return _lib


def _ip(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_int)) if a is not None else ctypes.POINTER(ctypes.c_int)()


def _gs(g):
    return _G(g.nodes, g.edges, _ip(g.nindex), _ip(g.nlist), _ip(g.eweight))


def bfs_serial(g, src):
    out = np.empty(g.nodes, np.int32)
    lib().orc_bfs_serial(src, ctypes.byref(_gs(g)), _ip(out))
    return out


def bfs_omp_push(g, src):
    out = np.empty(g.nodes, np.int32)
    it = ctypes.c_int()
    t = lib().orc_bfs_omp_topo_push(src, ctypes.byref(_gs(g)), _ip(out), ctypes.byref(it))
    return out, t, it.value


def bfs_omp_pull(g, src):
    out = np.empty(g.nodes, np.int32)
    it = ctypes.c_int()
    t = lib().orc_bfs_omp_topo_pull(src, ctypes.byref(_gs(g)), _ip(out), ctypes.byref(it))
    return out, t, it.value


def sssp_serial(g, src):
    out = np.empty(g.nodes, np.int32)
    lib().orc_sssp_serial(src, ctypes.byref(_gs(g)), _ip(out))
    return out


def sssp_omp_edge_data(g, src):
    out = np.empty(g.nodes, np.int32)
    sp = np.ascontiguousarray(g.sp(), np.int32)
    it = ctypes.c_int()
    t = lib().orc_sssp_omp_edge_data(src, ctypes.byref(_gs(g)), _ip(sp), _ip(out), ctypes.byref(it))
    return out, t, it.value


def sssp_omp_pull(g, src):
    out = np.empty(g.nodes, np.int32)
    it = ctypes.c_int()
    t = lib().orc_sssp_omp_topo_pull(src, ctypes.byref(_gs(g)), _ip(out), ctypes.byref(it))
    return out, t, it.value


def cc_omp_push(g):
    out = np.empty(g.nodes, np.int32)
    it = ctypes.c_int()
    t = lib().orc_cc_omp_topo_push(ctypes.byref(_gs(g)), _ip(out), ctypes.byref(it))
    return out, t, it.value


def cc_omp_pull(g):
    out = np.empty(g.nodes, np.int32)
    it = ctypes.c_int()
    t = lib().orc_cc_omp_topo_pull(ctypes.byref(_gs(g)), _ip(out), ctypes.byref(it))
    return out, t, it.value


def cc_verify(g, label):
    lab = np.ascontiguousarray(label, np.int32)
    return lib().orc_cc_verify(ctypes.byref(_gs(g)), _ip(lab))


def mis_hash(v):
    return lib().orc_mis_hash(v)


def mis_omp_push(g):
    prio = np.empty(g.nodes, np.uint32)
    st = np.empty(g.nodes, np.uint8)
    it = ctypes.c_int()
    t = lib().orc_mis_omp_topo_push(ctypes.byref(_gs(g)), prio.ctypes.data_as(ctypes.POINTER(ctypes.c_uint)),
                                    st.ctypes.data_as(ctypes.POINTER(ctypes.c_ubyte)), ctypes.byref(it))
    return st, t, it.value


def mis_omp_edge(g):
    st = np.empty(g.nodes, np.uint8)
    sp = np.ascontiguousarray(g.sp(), np.int32)
    it = ctypes.c_int()
    t = lib().orc_mis_omp_edge(ctypes.byref(_gs(g)), _ip(sp), st.ctypes.data_as(ctypes.POINTER(ctypes.c_ubyte)),
                               ctypes.byref(it))
    return st, t, it.value


def mis_greedy(g):
    st = np.empty(g.nodes, np.uint8)
    lib().orc_mis_serial_greedy(ctypes.byref(_gs(g)), st.ctypes.data_as(ctypes.POINTER(ctypes.c_ubyte)))
    return st


def mis_verify(g, status):
    st = np.ascontiguousarray(status, np.uint8)
    return lib().orc_mis_verify(ctypes.byref(_gs(g)), st.ctypes.data_as(ctypes.POINTER(ctypes.c_ubyte)))


def pr_omp(g, eps=1e-4, max_iter=100):
    sc = np.empty(g.nodes, np.float32)
    it = ctypes.c_int()
    t = lib().orc_pr_omp_pull(ctypes.byref(_gs(g)), sc.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), eps,
                              max_iter, ctypes.byref(it))
    return sc, t, it.value


def pr_fp64(g, eps=1e-6, max_iter=100, fixed_iters=0):
    sc = np.empty(g.nodes, np.float64)
    res = ctypes.c_double()
    it = lib().orc_pr_fp64(ctypes.byref(_gs(g)), sc.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), eps, max_iter,
                           fixed_iters, ctypes.byref(res))
    return sc, it, res.value


def pr_check(g, scores):
    sc = np.ascontiguousarray(scores, np.float32)
    return lib().orc_pr_check(ctypes.byref(_gs(g)), sc.ctypes.data_as(ctypes.POINTER(ctypes.c_float)))


def tc_serial(g):
    return int(lib().orc_tc_serial(ctypes.byref(_gs(g))))


def tc_forward(g):
    """Degree-ordered forward count (checker for large graphs)."""
    return int(lib().orc_tc_forward(ctypes.byref(_gs(g))))


def tc_omp_vertex(g):
    c = ctypes.c_ulonglong()
    t = lib().orc_tc_omp_vertex(ctypes.byref(_gs(g)), ctypes.byref(c))
    return int(c.value), t


def tc_omp_edge(g):
    c = ctypes.c_ulonglong()
    sp = np.ascontiguousarray(g.sp(), np.int32)
    t = lib().orc_tc_omp_edge(ctypes.byref(_gs(g)), _ip(sp), ctypes.byref(c))
    return int(c.value), t


def num_threads():
    return lib().orc_num_threads()


# ------------------------------------------- C++-threads restatement ----
# oracle_cpp.cpp: the reference's codes/cpp variants ([P-BFS], [P-SSSP],
# [P-CC], [P-MIS], [P-PR], [P-TC] in that file's header); each returns
# (result, seconds[, iterations]).
def cpp_bfs(g, src, threads=4):
    out = np.empty(g.nodes, np.int32)
    it = ctypes.c_int()
    t = lib().orc_cpp_bfs_topo_push(src, ctypes.byref(_gs(g)), _ip(out), threads, ctypes.byref(it))
    return out, t, it.value


def cpp_sssp(g, src, threads=4):
    out = np.empty(g.nodes, np.int32)
    it = ctypes.c_int()
    t = lib().orc_cpp_sssp_topo_pull(src, ctypes.byref(_gs(g)), _ip(out), threads, ctypes.byref(it))
    return out, t, it.value


def cpp_cc(g, threads=4):
    out = np.empty(g.nodes, np.int32)
    it = ctypes.c_int()
    t = lib().orc_cpp_cc_topo_push(ctypes.byref(_gs(g)), _ip(out), threads, ctypes.byref(it))
    return out, t, it.value


def cpp_mis(g, threads=4):
    prio = np.empty(g.nodes, np.uint32)
    st = np.empty(g.nodes, np.uint8)
    it = ctypes.c_int()
    t = lib().orc_cpp_mis_topo_push(ctypes.byref(_gs(g)), prio.ctypes.data_as(ctypes.POINTER(ctypes.c_uint)),
                                    st.ctypes.data_as(ctypes.POINTER(ctypes.c_ubyte)), threads, ctypes.byref(it))
    return st, t, it.value


def cpp_pr(g, eps=1e-4, max_iter=100, threads=4):
    sc = np.full(g.nodes, np.float32(1.0) / np.float32(g.nodes), np.float32)
    it = ctypes.c_int()
    t = lib().orc_cpp_pr_pull(ctypes.byref(_gs(g)), sc.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), eps,
                              max_iter, threads, ctypes.byref(it))
    return sc, t, it.value


def cpp_tc(g, threads=4):
    c = ctypes.c_ulonglong()
    t = lib().orc_cpp_tc_vertex(ctypes.byref(_gs(g)), ctypes.byref(c), threads)
    return int(c.value), t


# ---------------------------------------------------------------- _ref ----
REF_OUT = {  # name -> (dtype, is_scalar)
    "bfs_push": (np.int32, False), "bfs_pull": (np.int32, False),
    "sssp_edge_data": (np.int32, False), "sssp_vtopo_pull": (np.int32, False),
    "cc_push": (np.int32, False), "cc_pull": (np.int32, False),
    "mis_vertex": (np.uint8, False), "mis_edge": (np.uint8, False),
    "mis_edata_copy": (np.uint8, False), "mis_edata_merge": (np.uint8, False),
    "bfs_vtopo_pull_determ": (np.int32, False), "bfs_vdata_push_determ": (np.int32, False),
    "sssp_edata_determ": (np.int32, False), "sssp_vtopo_pull_determ": (np.int32, False),
    "sssp_vdata_push_determ": (np.int32, False),
    "cc_vtopo_pull_determ": (np.int32, False), "cc_vdata_push_determ": (np.int32, False),
    "bfs_edata": (np.int32, False), "bfs_edata_rw": (np.int32, False), "bfs_edata_determ": (np.int32, False),
    "bfs_etopo_determ": (np.int32, False), "cc_edata": (np.int32, False), "cc_edata_rw": (np.int32, False),
    "cc_edata_determ": (np.int32, False), "cc_etopo_determ": (np.int32, False),
    "pr_push": (np.float32, False), "pr_push_clausered": (np.float32, False), "tc_vertex_atomic": (np.int64, True), "tc_edge_critical": (np.int64, True),
    "pr_pull": (np.float32, False), "tc_vertex": (np.int64, True), "tc_edge": (np.int64, True),
    "cpp_bfs": (np.int32, False), "cpp_sssp": (np.int32, False), "cpp_cc": (np.int32, False),
    "cpp_mis": (np.uint8, False), "cpp_pr": (np.float32, False), "cpp_tc": (np.int64, True),
    "cpp_bfs_vdata_push_determ": (np.int32, False), "cpp_bfs_vtopo_pull_determ": (np.int32, False),
    "cpp_sssp_vdata_push_determ": (np.int32, False), "cpp_cc_vtopo_pull_determ": (np.int32, False),
    "cpp_cc_vdata_push_determ": (np.int32, False), "cpp_mis_vtopo_pull_copydeterm": (np.uint8, False),
    "cpp_mis_vdata_pull_copydeterm": (np.uint8, False), "cpp_pr_push_atomicadd": (np.float32, False),
    "cpp_tc_vertex_critical": (np.int64, True),
}


def ref_available(name: str) -> bool:
    return os.path.exists(os.path.join(REF_DIR, "ref_" + name))


def ref_run(name: str, g, src: int = 0, threads: int = 4):
    """Run the reference's own OpenMP or C++-threads variant `name` on g
    (OMP_NUM_THREADS, or the thread-count argument, = threads); returns its
    output array."""
    exe = os.path.join(REF_DIR, "ref_" + name)
    dtype, scalar = REF_OUT[name]
    with tempfile.TemporaryDirectory() as td:
        gin = os.path.join(td, "g.egr")
        out = os.path.join(td, "out.bin")
        g.write_egr(gin)
        env = dict(os.environ, OMP_NUM_THREADS=str(threads))
        subprocess.run([exe, gin, out, str(src), str(threads)], check=True, env=env, stdout=subprocess.DEVNULL)
        data = np.fromfile(out, dtype=dtype)
    return int(data[0]) if scalar else data
