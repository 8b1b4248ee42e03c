"""Exhaustive style autotuner and the synthetic graph families it trains on.

`sweep` times every registered style of an algorithm on one device graph
(median of `repeats` runs, each run's own hipEvent time); a style that is
slower than `cap_factor` x the best seen so far after its first run is not
repeated.  `collect` builds the selector's training data: seeded graphs from
four families (Erdos-Renyi; R-MAT with a drawn in [0.45, 0.65] and b:c:d kept
at 0.19:0.19:0.05; 2-D grids with long-range edges; Chung-Lu power laws with
exponent in [2.1, 3.0]), their 16 features, and the sweep of every style.
"""
from __future__ import annotations

import math
import time
from typing import Dict, List, Optional

import numpy as np

from . import device
from .graph import Graph, chung_lu, erdos_renyi, grid2d, rmat

FAMILIES = ("er", "rmat", "grid", "chunglu")


def make_graph(family: str, log_n: float, avg_degree: float, seed: int, rng: np.random.Generator) -> Graph:
    n = max(16, int(round(2.0 ** log_n)))
    if family == "er":
        return erdos_renyi(n, avg_degree, seed=seed)
    if family == "rmat":
        a = float(rng.uniform(0.45, 0.65))
        rest = 1.0 - a
        b = c = rest * 0.19 / 0.43
        scale = max(1, math.ceil(math.log2(n)))
        return rmat(scale, avg_degree / 2.0, (a, b, c), seed=seed, n_limit=n if n < (1 << scale) else 0)
    if family == "grid":
        side = max(2, int(round(math.sqrt(n))))
        return grid2d(side, float(rng.uniform(0.0, 0.1)), seed=seed)
    if family == "chunglu":
        return chung_lu(n, avg_degree, float(rng.uniform(2.1, 3.0)), seed=seed)
    raise ValueError(family)


def graph_spec(index: int, seed_base: int, log_n_range=(12, 21), deg_range=(4, 32)) -> dict:
    rng = np.random.default_rng(seed_base + index)
    family = FAMILIES[index % len(FAMILIES)]
    return {"family": family, "log_n": float(rng.uniform(*log_n_range)), "avg_degree": float(rng.uniform(*deg_range)),
            "seed": seed_base + index, "rng_seed": seed_base + index}


def build_from_spec(spec: dict) -> Graph:
    rng = np.random.default_rng(spec["rng_seed"] + 7919)
    g = make_graph(spec["family"], spec["log_n"], spec["avg_degree"], spec["seed"], rng)
    return g.with_weights(1, 255, seed=spec["seed"])


def sweep(dg: device.DeviceGraph, algo: str, src: int = 0, repeats: int = 3, cap_factor: float = 20.0,
          styles: Optional[List[str]] = None) -> Dict[str, float]:
    styles = styles or device.styles(algo)
    times: Dict[str, float] = {}
    best = math.inf
    for st in styles:
        first = device.run(algo, dg, st, src).seconds
        if first > cap_factor * best:
            times[st] = first
            continue
        ts = [first] + [device.run(algo, dg, st, src).seconds for _ in range(repeats - 1)]
        times[st] = float(np.median(ts))
        best = min(best, times[st])
    return times


def collect(count: int, algos=("bfs", "sssp", "cc", "mis", "pr", "tc"), seed_base: int = 100000,
            log_n_range=(12, 21), deg_range=(4, 32), repeats: int = 3, log=print) -> dict:
    """Features and per-style times for `count` seeded graphs."""
    device.init()
    feats = []
    times = {a: [] for a in algos}
    specs = []
    t0 = time.time()
    for i in range(count):
        spec = graph_spec(i, seed_base, log_n_range, deg_range)
        g = build_from_spec(spec)
        src = g.max_degree_vertex()
        feats.append(g.features(seed=spec["seed"]))
        with device.DeviceGraph(g) as dg:
            for a in algos:
                if a == "pr":
                    device.set_pr_params(1e-6, 100)
                sw = sweep(dg, a, src, repeats)
                times[a].append([sw[s] for s in device.styles(a)])
        spec.update({"n": g.nodes, "m": g.edges})
        specs.append(spec)
        if log and (i + 1) % max(1, count // 20) == 0:
            log(f"collected {i + 1}/{count} graphs in {time.time() - t0:.0f}s")
    device.set_pr_params(1e-4, 100)
    return {"features": np.array(feats), "times": {a: np.array(v) for a, v in times.items()},
            "styles": {a: device.styles(a) for a in algos}, "specs": specs}


def save_dataset(path: str, data: dict) -> None:
    out = {"features": data["features"]}
    for a, t in data["times"].items():
        out[f"times_{a}"] = t
        out[f"styles_{a}"] = np.array(data["styles"][a])
    out["spec_n"] = np.array([s["n"] for s in data["specs"]])
    out["spec_m"] = np.array([s["m"] for s in data["specs"]])
    out["spec_family"] = np.array([s["family"] for s in data["specs"]])
    np.savez_compressed(path, **out)


def load_dataset(path: str) -> dict:
    d = np.load(path)
    algos = [k[len("times_"):] for k in d.files if k.startswith("times_")]
    return {"features": d["features"], "times": {a: d[f"times_{a}"] for a in algos},
            "styles": {a: [str(s) for s in d[f"styles_{a}"]] for a in algos}}
