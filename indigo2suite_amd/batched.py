"""Batched mode: many uneven graphs in one arena, BFS/CC/MIS for all of them in
one persistent launch per algorithm (optionally one captured hipGraph).

The workload generator follows the batched configuration: graphs 0..G/2-1 are
Erdos-Renyi, the rest R-MAT (a, b, c, d) = (0.57, 0.19, 0.19, 0.05); graph i
draws its vertex count log-uniformly in [2^lo, 2^hi] and its average degree
uniformly in [dmin, dmax] from a generator seeded with 1000 + i, and is
itself generated with seed 1000 + i.  R-MAT graphs use scale ceil(log2 n_i)
and drop samples with an endpoint >= n_i.
"""
from __future__ import annotations

import ctypes
import math
import time
from typing import List, Optional

import numpy as np

from . import _native, device
from ._native import IgxGraph, IgxRunInfo, check
from .graph import Graph, erdos_renyi, rmat

MASK_BFS, MASK_CC, MASK_MIS = 1, 2, 4


class IgxBatch(ctypes.Structure):
    _fields_ = [
        ("graphs", ctypes.c_int),
        ("total_nodes", ctypes.c_longlong),
        ("total_edges", ctypes.c_longlong),
        ("node_off", ctypes.POINTER(ctypes.c_longlong)),
        ("edge_off", ctypes.POINTER(ctypes.c_longlong)),
        ("nindex", ctypes.POINTER(ctypes.c_int)),
        ("nlist", ctypes.POINTER(ctypes.c_int)),
        ("order", ctypes.POINTER(ctypes.c_int)),
    ]


def _lib():
    lib = _native.hip_lib()
    if not hasattr(lib, "_batch_sigs"):
        B = ctypes.POINTER(IgxBatch)
        ip = ctypes.POINTER(ctypes.c_int)
        lib.igx_batch_upload.restype = ctypes.c_int
        lib.igx_batch_upload.argtypes = [ctypes.c_int, ctypes.POINTER(IgxGraph), B]
        lib.igx_batch_free.restype = ctypes.c_int
        lib.igx_batch_free.argtypes = [B]
        lib.igx_batch_graph_view.restype = ctypes.c_int
        lib.igx_batch_graph_view.argtypes = [B, ctypes.c_int, ctypes.POINTER(IgxGraph)]
        lib.igx_batch_run.restype = ctypes.c_int
        lib.igx_batch_run.argtypes = [B, ctypes.c_int, ip, ip, ip, ctypes.c_int, ctypes.c_int,
                                      ctypes.POINTER(IgxRunInfo)]
        lib.igx_batch_run_noalloc.restype = ctypes.c_int
        lib.igx_batch_run_noalloc.argtypes = [B, ctypes.c_int, ip, ip, ip, ctypes.c_int, ctypes.POINTER(IgxRunInfo)]
        lib._batch_sigs = True
    return lib


def make_batch_graphs(count: int = 512, lo: int = 10, hi: int = 18, dmin: float = 4.0, dmax: float = 32.0,
                      seed_base: int = 1000) -> List[Graph]:
    graphs = []
    for i in range(count):
        rng = np.random.default_rng(seed_base + i)
        n = int(round(2.0 ** rng.uniform(lo, hi)))
        d = float(rng.uniform(dmin, dmax))
        if i < count // 2:
            g = erdos_renyi(n, d, seed=seed_base + i)
        else:
            scale = max(1, math.ceil(math.log2(n)))
            g = rmat(scale, d / 2.0, seed=seed_base + i, n_limit=n)
        g.name = f"batch{i}"
        graphs.append(g)
    return graphs


class DeviceBatch:
    def __init__(self, graphs: List[Graph]):
        device.init()
        self.graphs = graphs
        arr = (IgxGraph * len(graphs))(*[g.c_struct() for g in graphs])
        self.b = IgxBatch()
        check(_lib().igx_batch_upload(len(graphs), arr, ctypes.byref(self.b)), "igx_batch_upload")
        self.node_off = np.concatenate([[0], np.cumsum([g.nodes for g in graphs])]).astype(np.int64)
        self.total_nodes = int(self.node_off[-1])
        self.outs = {k: device.DeviceBuffer(4 * self.total_nodes) for k in ("bfs", "cc", "mis")}

    def run(self, mask: int = 7, block: int = 256, use_graph: bool = True) -> IgxRunInfo:
        info = IgxRunInfo()
        p = {k: ctypes.cast(v.ptr, ctypes.POINTER(ctypes.c_int)) for k, v in self.outs.items()}
        check(_lib().igx_batch_run(ctypes.byref(self.b), mask, p["bfs"], p["cc"], p["mis"], block, int(use_graph),
                                   ctypes.byref(info)), "igx_batch_run")
        return info

    def run_noalloc(self, mask: int = 7, block: int = 256) -> IgxRunInfo:
        """igx_batch_run_noalloc: eager run with the library's fixed scratch
        array (no device allocation; at most 2^20 vertices in the batch)."""
        info = IgxRunInfo()
        p = {k: ctypes.cast(v.ptr, ctypes.POINTER(ctypes.c_int)) for k, v in self.outs.items()}
        check(_lib().igx_batch_run_noalloc(ctypes.byref(self.b), mask, p["bfs"], p["cc"], p["mis"], block,
                                           ctypes.byref(info)), "igx_batch_run_noalloc")
        return info

    def results(self, algo: str) -> List[np.ndarray]:
        flat = self.outs[algo].to_host(np.int32, self.total_nodes)
        return [flat[self.node_off[i]:self.node_off[i + 1]] for i in range(len(self.graphs))]

    def view(self, i: int) -> IgxGraph:
        v = IgxGraph()
        check(_lib().igx_batch_graph_view(ctypes.byref(self.b), i, ctypes.byref(v)), "igx_batch_graph_view")
        return v

    def per_graph_baseline(self, styles: Optional[dict] = None) -> float:
        """One launch sequence per graph and algorithm through igx_run (the
        single-graph path), over the same arena; returns total seconds."""
        lib = device.lib()
        styles = styles or {}
        total = 0.0
        for i in range(len(self.graphs)):
            v = self.view(i)
            off = int(self.node_off[i]) * 4
            for algo in ("bfs", "cc", "mis"):
                info = IgxRunInfo()
                sid = device.style_id(algo, styles.get(algo))
                out = ctypes.c_void_p(self.outs[algo].ptr.value + off)
                check(lib.igx_run(device.ALGOS[algo], sid, ctypes.byref(v), 0, out,
                                  ctypes.cast(None, _native.c_int_p), ctypes.byref(info)), "igx_run")
                total += info.seconds
        return total

    def free(self):
        for b in self.outs.values():
            b.free()
        if self.b.graphs:
            _lib().igx_batch_free(ctypes.byref(self.b))

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.free()


def autotune_block(db: DeviceBatch, blocks=(256, 512, 1024), repeats: int = 3) -> dict:
    """Exhaustive sweep of the workgroup size (the batched tile); returns the
    best block and the median time of each."""
    times = {}
    for b in blocks:
        db.run(7, b, use_graph=True)
        ts = [db.run(7, b, use_graph=True).seconds for _ in range(repeats)]
        times[b] = float(np.median(ts))
    best = min(times, key=times.get)
    return {"best_block": best, "seconds": times}
