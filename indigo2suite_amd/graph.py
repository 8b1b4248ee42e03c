"""Host CSR graphs in the ECLgraph layout, backed by numpy arrays.

``Graph`` mirrors ``struct ECLgraph`` (include/ECLgraph.h:44-50): int32
``nindex`` (nodes + 1), int32 ``nlist`` (edges, both directions of every
undirected edge) and optional int32 ``eweight``.  Generators and ``.egr`` I/O
run in libindigo_graph.so (C++/OpenMP); the results are copied into numpy.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass
from typing import Optional

import numpy as np

from . import _native
from ._native import IgxGraph, check

NUM_FEATURES = 16
RMAT_ABC = (0.57, 0.19, 0.19)


def _ptr(a: Optional[np.ndarray]):
    if a is None:
        return ctypes.cast(None, _native.c_int_p)
    return a.ctypes.data_as(_native.c_int_p)


@dataclass
class Graph:
    nindex: np.ndarray
    nlist: np.ndarray
    eweight: Optional[np.ndarray] = None
    name: str = ""

    def __post_init__(self):
        self.nindex = np.ascontiguousarray(self.nindex, dtype=np.int32)
        self.nlist = np.ascontiguousarray(self.nlist, dtype=np.int32)
        if self.eweight is not None:
            self.eweight = np.ascontiguousarray(self.eweight, dtype=np.int32)

    @property
    def nodes(self) -> int:
        return int(self.nindex.shape[0] - 1)

    @property
    def edges(self) -> int:
        return int(self.nlist.shape[0])

    def degrees(self) -> np.ndarray:
        return np.diff(self.nindex).astype(np.int32)

    def c_struct(self) -> IgxGraph:
        """An igx_graph view of this graph's arrays (valid while self lives)."""
        return IgxGraph(self.nodes, self.edges, _ptr(self.nindex), _ptr(self.nlist), _ptr(self.eweight))

    def sp(self) -> np.ndarray:
        """Starting-point array: sp[e] = source of edge e (sssp_edge_cuda.h:149-155)."""
        return np.repeat(np.arange(self.nodes, dtype=np.int32), self.degrees())

    def max_degree_vertex(self) -> int:
        d = self.degrees()
        return int(np.argmax(d)) if d.size else 0

    def with_weights(self, lo: int = 1, hi: int = 255, seed: int = 1) -> "Graph":
        """Copy with symmetric weights uniform in [lo, hi] (hash of the unordered pair)."""
        g = self.c_struct()
        lib = _native.graph_lib()
        check(lib.igx_gen_weights(ctypes.byref(g), lo, hi, seed), "igx_gen_weights")
        w = np.ctypeslib.as_array(g.eweight, shape=(self.edges,)).copy() if self.edges else np.zeros(0, np.int32)
        _free_native_weights(g)
        return Graph(self.nindex, self.nlist, w, self.name)

    def with_reference_weights(self) -> "Graph":
        """Copy with the reference harnesses' fallback weights (sssp_edge_cuda.h:127-137)."""
        g = self.c_struct()
        lib = _native.graph_lib()
        check(lib.igx_ref_weights(ctypes.byref(g)), "igx_ref_weights")
        w = np.ctypeslib.as_array(g.eweight, shape=(self.edges,)).copy() if self.edges else np.zeros(0, np.int32)
        _free_native_weights(g)
        return Graph(self.nindex, self.nlist, w, self.name)

    def is_sorted(self) -> bool:
        g = self.c_struct()
        return _native.graph_lib().igx_check_sorted(ctypes.byref(g)) == _native.IGX_OK

    def features(self, seed: int = 1) -> np.ndarray:
        g = self.c_struct()
        out = (ctypes.c_double * NUM_FEATURES)()
        check(_native.graph_lib().igx_graph_features(ctypes.byref(g), seed, out), "igx_graph_features")
        return np.array(out[:], dtype=np.float64)

    def write_egr(self, path: str) -> None:
        g = self.c_struct()
        check(_native.graph_lib().igx_egr_write(ctypes.byref(g), path.encode()), "igx_egr_write")


def _free_native_weights(g: IgxGraph) -> None:
    # only eweight was allocated by the library; free it alone
    libc = ctypes.CDLL(None)
    libc.free.argtypes = [ctypes.c_void_p]
    libc.free(ctypes.cast(g.eweight, ctypes.c_void_p))


def _take(g: IgxGraph, name: str = "") -> Graph:
    """Copy a library-owned igx_graph into numpy and free the native arrays."""
    n, m = g.nodes, g.edges
    nindex = np.ctypeslib.as_array(g.nindex, shape=(n + 1,)).copy()
    nlist = np.ctypeslib.as_array(g.nlist, shape=(m,)).copy() if m > 0 else np.zeros(0, np.int32)
    ew = None
    if g.eweight:
        ew = np.ctypeslib.as_array(g.eweight, shape=(m,)).copy() if m > 0 else np.zeros(0, np.int32)
    _native.graph_lib().igx_host_graph_free(ctypes.byref(g))
    return Graph(nindex, nlist, ew, name)


def read_egr(path: str) -> Graph:
    g = IgxGraph()
    check(_native.graph_lib().igx_egr_read(path.encode(), ctypes.byref(g)), f"igx_egr_read({path})")
    return _take(g, path)


def erdos_renyi(n: int, avg_degree: float, seed: int = 1) -> Graph:
    g = IgxGraph()
    check(_native.graph_lib().igx_gen_er(n, avg_degree, seed, ctypes.byref(g)), "igx_gen_er")
    return _take(g, f"er_n{n}_d{avg_degree:g}_s{seed}")


def rmat(scale: int, edge_factor: float = 16, abc=RMAT_ABC, seed: int = 1, n_limit: int = 0) -> Graph:
    a, b, c = abc
    g = IgxGraph()
    check(_native.graph_lib().igx_gen_rmat(scale, edge_factor, a, b, c, n_limit, seed, ctypes.byref(g)),
          "igx_gen_rmat")
    return _take(g, f"rmat{scale}_ef{edge_factor:g}_s{seed}" + (f"_n{n_limit}" if n_limit else ""))


def grid2d(side: int, long_frac: float = 0.05, seed: int = 1) -> Graph:
    g = IgxGraph()
    check(_native.graph_lib().igx_gen_grid(side, long_frac, seed, ctypes.byref(g)), "igx_gen_grid")
    return _take(g, f"grid{side}_lr{long_frac:g}_s{seed}")


def chung_lu(n: int, avg_degree: float, gamma: float, seed: int = 1) -> Graph:
    g = IgxGraph()
    check(_native.graph_lib().igx_gen_chunglu(n, avg_degree, gamma, seed, ctypes.byref(g)), "igx_gen_chunglu")
    return _take(g, f"chunglu_n{n}_d{avg_degree:g}_g{gamma:g}_s{seed}")


def from_edges(n: int, src, dst, name: str = "") -> Graph:
    s = np.ascontiguousarray(src, dtype=np.int32)
    d = np.ascontiguousarray(dst, dtype=np.int32)
    g = IgxGraph()
    check(_native.graph_lib().igx_graph_from_edges(n, s.shape[0], _ptr(s), _ptr(d), ctypes.byref(g)),
          "igx_graph_from_edges")
    return _take(g, name)
