"""Python host API over libindigo_hip.so (include/indigo_hip.h).

Mirrors the reference's per-algorithm entry points (include/*_cuda.h) and the
general styled entry ``igx_run``.  Every call goes through the C ABI; there is
no CPU fallback -- without the HIP library or a GPU the calls raise.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass
from typing import Optional, Union

import numpy as np

from . import _native
from ._native import IgxGraph, IgxRunInfo, check
from .graph import Graph

ALGOS = {"bfs": 0, "sssp": 1, "cc": 2, "mis": 3, "pr": 4, "tc": 5}
INT_MAX = 2147483647


def lib():
    return _native.hip_lib()


def init(device: int = 0) -> None:
    check(lib().igx_init(device), "igx_init")


def device_info() -> dict:
    init()
    name = ctypes.create_string_buffer(256)
    cus = ctypes.c_int()
    mt = ctypes.c_int()
    check(lib().igx_device_info(name, 256, ctypes.byref(cus), ctypes.byref(mt)), "igx_device_info")
    return {"name": name.value.decode(), "cus": cus.value, "max_threads_per_cu": mt.value}


def styles(algo: str) -> list:
    a = ALGOS[algo]
    L = lib()
    return [L.igx_style_name(a, i).decode() for i in range(L.igx_num_styles(a))]


def style_id(algo: str, style: Union[int, str, None]) -> int:
    if style is None:
        return lib().igx_get_entry_style(ALGOS[algo], 0)
    if isinstance(style, int):
        return style
    names = styles(algo)
    if style not in names:
        raise KeyError(f"unknown {algo} style {style!r}; known: {names}")
    return names.index(style)


def default_entry_style(algo: str, edge_kind: int = 0) -> int:
    """Style index used by the vertex (0) or edge (1) reference-signature entry."""
    return lib().igx_get_entry_style(ALGOS[algo], 1 if edge_kind else 0)


def set_time_limit(seconds: float) -> None:
    """Wall-clock budget of each following run (0 = none); a round-based
    style past it stops with IGX_ERR_TIMEOUT (igx_set_time_limit)."""
    check(lib().igx_set_time_limit(seconds), "igx_set_time_limit")


def set_pr_params(eps: float, max_iter: int) -> None:
    check(lib().igx_set_pr_params(eps, max_iter), "igx_set_pr_params")


class DeviceBuffer:
    def __init__(self, nbytes: int):
        init()
        self.ptr = ctypes.c_void_p()
        self.nbytes = max(int(nbytes), 4)
        check(lib().igx_alloc(ctypes.byref(self.ptr), self.nbytes), "igx_alloc")

    def to_host(self, dtype, count) -> np.ndarray:
        out = np.empty(count, dtype=dtype)
        if count:
            check(lib().igx_memcpy_d2h(out.ctypes.data_as(ctypes.c_void_p), self.ptr, out.nbytes), "igx_memcpy_d2h")
        return out

    def from_host(self, a: np.ndarray) -> None:
        a = np.ascontiguousarray(a)
        check(lib().igx_memcpy_h2d(self.ptr, a.ctypes.data_as(ctypes.c_void_p), a.nbytes), "igx_memcpy_h2d")

    def free(self):
        if self.ptr:
            lib().igx_free(self.ptr)
            self.ptr = ctypes.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class DeviceGraph:
    """A device copy of a host Graph (igx_graph_upload)."""

    def __init__(self, g: Graph):
        init()
        self.host = g
        self.c = IgxGraph()
        hc = g.c_struct()
        check(lib().igx_graph_upload(ctypes.byref(hc), ctypes.byref(self.c)), "igx_graph_upload")
        self.nodes = g.nodes
        self.edges = g.edges

    def free(self):
        if self.c.nindex:
            lib().igx_graph_free_device(ctypes.byref(self.c))

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.free()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def features_device(dg: DeviceGraph, seed: int = 1):
    """The selector's 16 graph features computed on the device from the
    uploaded CSR (igx_graph_features_device); returns (features, seconds)."""
    from .graph import NUM_FEATURES
    out = (ctypes.c_double * NUM_FEATURES)()
    sec = ctypes.c_double(0.0)
    check(lib().igx_graph_features_device(ctypes.byref(dg.c), seed, out, ctypes.byref(sec)),
          "igx_graph_features_device")
    return np.array(out[:], dtype=np.float64), sec.value


@dataclass
class RunResult:
    out: Union[np.ndarray, int]
    seconds: float
    iterations: int
    launches: int
    residual: float
    style: str


def _out_spec(algo: str, n: int):
    if algo in ("bfs", "sssp", "cc", "mis"):
        return np.int32, n
    if algo == "pr":
        return np.float32, n
    return np.uint64, 1


def run_device(algo: str, dg: DeviceGraph, d_out: DeviceBuffer, style: Union[int, str, None] = None, src: int = 0,
               d_sp: Optional[ctypes.c_void_p] = None) -> IgxRunInfo:
    """Run one style with device-resident output; returns the run info."""
    info = IgxRunInfo()
    sid = style_id(algo, style)
    sp = ctypes.cast(d_sp, _native.c_int_p) if d_sp is not None else ctypes.cast(None, _native.c_int_p)
    check(lib().igx_run(ALGOS[algo], sid, ctypes.byref(dg.c), src, d_out.ptr, sp, ctypes.byref(info)),
          f"igx_run({algo}, {styles(algo)[sid]})")
    return info


def run(algo: str, dg: DeviceGraph, style: Union[int, str, None] = None, src: int = 0) -> RunResult:
    dtype, count = _out_spec(algo, dg.nodes)
    buf = DeviceBuffer(np.dtype(dtype).itemsize * count)
    try:
        info = run_device(algo, dg, buf, style, src)
        out = buf.to_host(dtype, count)
    finally:
        buf.free()
    sid = style_id(algo, style)
    res = out if algo != "tc" else int(out[0])
    return RunResult(res, info.seconds, info.iterations, info.launches, info.residual, styles(algo)[sid])


# ---------------------------------------------- reference-signature entries
def _ip(a):
    return a.ctypes.data_as(_native.c_int_p)


def bfs_vertex(src: int, dg: DeviceGraph):
    """GPUbfs_vertex (bfs_vertex_cuda.h:61): returns (dist, seconds)."""
    dist = np.empty(dg.nodes, np.int32)
    t = ctypes.c_double()
    check(lib().igx_bfs_vertex(src, ctypes.byref(dg.c), _ip(dist), ctypes.byref(t)), "igx_bfs_vertex")
    return dist, t.value


def bfs_edge(src: int, dg: DeviceGraph, sp: np.ndarray):
    dist = np.empty(dg.nodes, np.int32)
    t = ctypes.c_double()
    sp = np.ascontiguousarray(sp, np.int32)
    check(lib().igx_bfs_edge(src, ctypes.byref(dg.c), _ip(dist), _ip(sp), ctypes.byref(t)), "igx_bfs_edge")
    return dist, t.value


def sssp_vertex(src: int, dg: DeviceGraph):
    dist = np.empty(dg.nodes, np.int32)
    t = ctypes.c_double()
    check(lib().igx_sssp_vertex(src, ctypes.byref(dg.c), _ip(dist), ctypes.byref(t)), "igx_sssp_vertex")
    return dist, t.value


def sssp_edge(src: int, dg: DeviceGraph, sp: np.ndarray):
    dist = np.empty(dg.nodes, np.int32)
    t = ctypes.c_double()
    sp = np.ascontiguousarray(sp, np.int32)
    check(lib().igx_sssp_edge(src, ctypes.byref(dg.c), _ip(dist), _ip(sp), ctypes.byref(t)), "igx_sssp_edge")
    return dist, t.value


def cc_vertex(dg: DeviceGraph):
    lab = np.empty(dg.nodes, np.int32)
    t = ctypes.c_double()
    check(lib().igx_cc_vertex(ctypes.byref(dg.c), _ip(lab), ctypes.byref(t)), "igx_cc_vertex")
    return lab, t.value


def cc_edge(dg: DeviceGraph, sp: np.ndarray):
    lab = np.empty(dg.nodes, np.int32)
    t = ctypes.c_double()
    sp = np.ascontiguousarray(sp, np.int32)
    check(lib().igx_cc_edge(ctypes.byref(dg.c), _ip(lab), _ip(sp), ctypes.byref(t)), "igx_cc_edge")
    return lab, t.value


def mis_vertex(dg: DeviceGraph):
    st = np.empty(dg.nodes, np.int32)
    pr = np.empty(dg.nodes, np.uint32)
    t = ctypes.c_double()
    check(lib().igx_mis_vertex(ctypes.byref(dg.c), pr.ctypes.data_as(ctypes.POINTER(ctypes.c_uint)), _ip(st),
                               ctypes.byref(t)), "igx_mis_vertex")
    return st, pr, t.value


def mis_edge(dg: DeviceGraph, d_sp: DeviceBuffer):
    st = np.empty(dg.nodes, np.int32)
    pr = np.empty(dg.nodes, np.uint32)
    t = ctypes.c_double()
    check(lib().igx_mis_edge(ctypes.byref(dg.c), ctypes.cast(d_sp.ptr, _native.c_int_p),
                             pr.ctypes.data_as(ctypes.POINTER(ctypes.c_uint)), _ip(st), ctypes.byref(t)),
          "igx_mis_edge")
    return st, pr, t.value


def pr(g: Graph):
    """PR_GPU (pr_cuda.h:48): host graph in, scores out; returns (scores, seconds, iterations)."""
    scores = np.full(g.nodes, np.float32(1.0) / np.float32(g.nodes), np.float32)
    deg = g.degrees()
    hc = g.c_struct()
    t = ctypes.c_double()
    it = ctypes.c_int()
    check(lib().igx_pr(ctypes.byref(hc), scores.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), _ip(deg),
                       ctypes.byref(t), ctypes.byref(it)), "igx_pr")
    return scores, t.value, it.value


def tc_vertex(dg: DeviceGraph):
    c = ctypes.c_ulonglong()
    t = ctypes.c_double()
    check(lib().igx_tc_vertex(ctypes.byref(c), dg.nodes, dg.c.nindex, dg.c.nlist, ctypes.byref(t)), "igx_tc_vertex")
    return int(c.value), t.value


def tc_edge(dg: DeviceGraph, d_sp: DeviceBuffer):
    c = ctypes.c_ulonglong()
    t = ctypes.c_double()
    check(lib().igx_tc_edge(ctypes.byref(c), dg.edges, dg.c.nindex, dg.c.nlist, ctypes.cast(d_sp.ptr, _native.c_int_p),
                            ctypes.byref(t)), "igx_tc_edge")
    return int(c.value), t.value


def build_sp(dg: DeviceGraph) -> DeviceBuffer:
    """Device sp[] built on the GPU (igx_build_sp), wrapped as a DeviceBuffer."""
    p = ctypes.POINTER(ctypes.c_int)()
    check(lib().igx_build_sp(ctypes.byref(dg.c), ctypes.byref(p)), "igx_build_sp")
    buf = DeviceBuffer.__new__(DeviceBuffer)
    buf.ptr = ctypes.c_void_p(ctypes.cast(p, ctypes.c_void_p).value)
    buf.nbytes = 4 * max(dg.edges, 1)
    return buf
