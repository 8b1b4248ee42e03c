"""Exhaustive style autotuner and the synthetic graph families it trains on.

`sweep` times every registered style of an algorithm on one device graph
(median of `repeats` runs, each run's own hipEvent time), the library's
default style first; a style whose first run is slower than `cap_factor` x
the best seen so far, or than `slow_seconds`, is not repeated.  `collect` builds the selector's training data: seeded graphs from
four families (Erdos-Renyi; R-MAT with a drawn in [0.45, 0.65] and b:c:d kept
at 0.19:0.19:0.05; 2-D grids with long-range edges; Chung-Lu power laws with
exponent in [2.1, 3.0]), their 16 features (computed on the device, as a
caller of `selector.pick_style_device` gets them), and the sweep of every style.
"""
from __future__ import annotations

import math
import time
from concurrent.futures import ThreadPoolExecutor
from typing import Dict, List, Optional

import numpy as np

from . import device
from ._native import IGX_ERR_TIMEOUT, IgxError
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


# Triangle-counting styles without orientation do O(d(v)^2) work per vertex
# in one thread, wave or workgroup: on a graph whose largest degree is in the
# tens of thousands a single run takes minutes, and in the measured sweeps
# they never beat the oriented style once the sum of squared degrees passes
# 1e9; the thread-per-vertex merge never won with a vertex above degree 256,
# and there cost 100x the best.  They are not run there; their time is
# recorded as +inf ("too slow to measure"), which the selector learns as the
# largest slowdown and never picks.
UNORIENTED_TC = ("v_merge_thread", "v_merge_wave", "e_bsearch_thread", "v_merge_block", "e_bsearch_wave",
                 "e_bsearch_block")
TC_MAX_DEGREE = 8192
TC_MAX_SUM_D2 = 1e9
TC_THREAD_MERGE_MAX_DEGREE = 256


def too_slow(algo: str, style: str, max_degree: int, sum_d2: float = 0.0) -> bool:
    if algo != "tc" or style not in UNORIENTED_TC:
        return False
    if style == "v_merge_thread" and max_degree > TC_THREAD_MERGE_MAX_DEGREE:
        return True
    return max_degree > TC_MAX_DEGREE or sum_d2 > TC_MAX_SUM_D2


def sum_sq_degrees(features: np.ndarray) -> float:
    """n (variance + mean^2) from the features log2(n), mean and standard
    deviation of the degrees (features 0, 2 and 3)."""
    return float(2.0 ** features[0] * (features[3] ** 2 + features[2] ** 2))


def sweep(dg: device.DeviceGraph, algo: str, src: int = 0, repeats: int = 3, cap_factor: float = 10.0,
          styles: Optional[List[str]] = None, slow_seconds: float = 1.0, max_degree: int = 0,
          repeat_factor: float = 3.0, limit_floor: float = 0.005, sum_d2: float = 0.0) -> Dict[str, float]:
    """Seconds of every style (median of `repeats` runs).  The default style
    runs first; every later run gets a time limit of cap_factor x the best
    median so far plus `limit_floor` (igx_set_time_limit), and a style that
    hits it is recorded as +inf (too slow: the selector learns the largest
    slowdown).  Styles slower than repeat_factor x the best, or than
    slow_seconds, are run once.  Output stays on the device (one buffer for
    all runs)."""
    styles = list(styles or device.styles(algo))
    # the library's default style first: its time caps the repeats of the slow ones
    default = device.styles(algo)[device.default_entry_style(algo, 0)]
    if default in styles:
        styles.remove(default)
        styles.insert(0, default)
    times: Dict[str, float] = {}
    best = math.inf
    out = device.DeviceBuffer(max(8 * dg.nodes, 64))

    def timed(st):
        limit = cap_factor * best + limit_floor if math.isfinite(best) else 0.0
        device.set_time_limit(limit)
        try:
            return device.run_device(algo, dg, out, st, src).seconds
        except IgxError as e:
            if e.status == IGX_ERR_TIMEOUT:
                return math.inf
            raise
        finally:
            device.set_time_limit(0.0)

    try:
        for st in styles:
            if too_slow(algo, st, max_degree, sum_d2):
                times[st] = math.inf
                continue
            first = timed(st)
            if not math.isfinite(first) or first > repeat_factor * best or first > slow_seconds:
                times[st] = first
                continue
            ts = [first] + [timed(st) for _ in range(repeats - 1)]
            times[st] = float(np.median(ts))
            best = min(best, times[st])
    finally:
        out.free()
    return times


def collect(count: int, algos=("bfs", "sssp", "cc", "mis", "pr", "tc"), seed_base: int = 100000,
            log_n_range=(12, 21), deg_range=(4, 32), repeats: int = 3, log=print, checkpoint=None,
            prefetch: int = 4, checkpoint_every: int = 1, max_seconds: float = 0.0) -> dict:
    """Features (computed on the device) and per-style times for `count`
    seeded graphs.  `log` gets a line per graph; checkpoint(partial_dataset)
    is called every `checkpoint_every` graphs.  With max_seconds > 0 the
    collection ends after the first graph that finishes past that many
    seconds (the dataset then holds fewer than `count` graphs)."""
    device.init()
    feats = []
    times = {a: [] for a in algos}
    specs = []
    t0 = time.time()

    def dataset():
        return {"features": np.array(feats), "times": {a: np.array(v) for a, v in times.items()},
                "styles": {a: device.styles(a) for a in algos}, "specs": specs}

    # graphs are generated on host threads ahead of the sweeps (the native
    # generators release the GIL), so the GPU does not wait for them
    pool = ThreadPoolExecutor(max_workers=prefetch)
    specs_all = [graph_spec(i, seed_base, log_n_range, deg_range) for i in range(count)]
    pending = {i: pool.submit(build_from_spec, specs_all[i]) for i in range(min(prefetch, count))}
    try:
        for i in range(count):
            spec = specs_all[i]
            g = pending.pop(i).result()
            if i + prefetch < count:
                pending[i + prefetch] = pool.submit(build_from_spec, specs_all[i + prefetch])
            src = g.max_degree_vertex()
            row = {}
            deg = np.diff(g.nindex).astype(np.float64)
            dmax = int(deg.max()) if g.nodes > 0 else 0
            sum_d2 = float(np.sum(deg * deg))
            with device.DeviceGraph(g) as dg:
                feat, feat_s = device.features_device(dg, spec["seed"])
                for a in algos:
                    device.set_pr_params(1e-6, 100) if a == "pr" else device.set_pr_params(1e-4, 100)
                    ta = time.time()
                    sw = sweep(dg, a, src, repeats, max_degree=dmax, sum_d2=sum_d2)
                    row[a] = [sw[s] for s in device.styles(a)]
                    if log and g.nodes >= (1 << 20):
                        log(f"  graph {i + 1} {a}: {len(sw)} styles in {time.time() - ta:.1f}s")
            feats.append(feat)
            for a in algos:
                times[a].append(row[a])
            spec.update({"n": g.nodes, "m": g.edges, "features_s": feat_s})
            specs.append(spec)
            if log:
                log(f"graph {i + 1}/{count} {spec['family']} n={g.nodes} m={g.edges} done at {time.time() - t0:.0f}s")
            out_of_time = max_seconds > 0 and time.time() - t0 > max_seconds
            if checkpoint and ((i + 1) % checkpoint_every == 0 or i + 1 == count or out_of_time):
                checkpoint(dataset())
            if out_of_time:
                if log:
                    log(f"time budget of {max_seconds:.0f}s spent after {i + 1} graphs")
                break
    finally:
        device.set_pr_params(1e-4, 100)
        pool.shutdown(wait=True, cancel_futures=True)
    return dataset()


def save_dataset(path: str, data: dict) -> None:
    out = {"features": data["features"]}
    for a, t in data["times"].items():
        out[f"times_{a}"] = t
        out[f"styles_{a}"] = np.array(data["styles"][a])
    out["spec_n"] = np.array([s["n"] for s in data["specs"]])
    out["spec_m"] = np.array([s["m"] for s in data["specs"]])
    out["spec_family"] = np.array([s["family"] for s in data["specs"]])
    out["spec_features_s"] = np.array([s.get("features_s", np.nan) for s in data["specs"]])
    out["spec_seed"] = np.array([s["seed"] for s in data["specs"]])
    np.savez_compressed(path, **out)


def retime(path: str, out: str, algo: str, styles: List[str], seed_base: int, log_n_range=(12, 21),
           deg_range=(4, 32), repeats: int = 3, log=print, prefetch: int = 4) -> int:
    """Time `styles` of `algo` again on the graphs of the dataset at `path`
    (regenerated from the seeds that collect drew them from, checked against
    the stored vertex and edge counts) and write the dataset with those
    columns replaced, or added for a style registered since, to `out`.  For
    refreshing the labels of a style whose kernels changed without re-timing
    every style.  Returns the graph count."""
    d = dict(np.load(path))
    names = [str(x) for x in d[f"styles_{algo}"]]
    reg = device.styles(algo)
    for st in styles:  # a style registered after the dataset was made gets a new column
        if st not in names:
            names.append(st)
            d[f"times_{algo}"] = np.concatenate([d[f"times_{algo}"], np.full((d["features"].shape[0], 1), np.nan)],
                                                axis=1)
    if names != [x for x in reg if x in names]:
        raise ValueError("dataset styles are not in registry order")
    d[f"styles_{algo}"] = np.array(names)
    cols = [names.index(st) for st in styles]
    count = d["features"].shape[0]
    device.init()
    specs = [graph_spec(i, seed_base, log_n_range, deg_range) for i in range(count)]
    pool = ThreadPoolExecutor(max_workers=prefetch)
    pending = {i: pool.submit(build_from_spec, specs[i]) for i in range(min(prefetch, count))}
    t = d[f"times_{algo}"].copy()
    t0 = time.time()
    try:
        for i in range(count):
            g = pending.pop(i).result()
            if i + prefetch < count:
                pending[i + prefetch] = pool.submit(build_from_spec, specs[i + prefetch])
            if g.nodes != int(d["spec_n"][i]) or g.edges != int(d["spec_m"][i]):
                raise ValueError(f"graph {i} does not match the dataset (seed base or ranges differ)")
            if algo == "pr":
                device.set_pr_params(1e-6, 100)
            with device.DeviceGraph(g) as dg:
                sw = sweep(dg, algo, g.max_degree_vertex(), repeats, styles=list(styles))
            for st, c in zip(styles, cols):
                t[i, c] = sw[st]
            if log and (i + 1) % 100 == 0:
                log(f"retimed {i + 1}/{count} at {time.time() - t0:.0f}s")
    finally:
        device.set_pr_params(1e-4, 100)
        pool.shutdown(wait=True, cancel_futures=True)
    d[f"times_{algo}"] = t
    np.savez_compressed(out, **d)
    return count


def load_dataset(path: str) -> dict:
    d = np.load(path)
    algos = [k[len("times_"):] for k in d.files if k.startswith("times_")]
    return {"features": d["features"], "times": {a: d[f"times_{a}"] for a in algos},
            "styles": {a: [str(s) for s in d[f"styles_{a}"]] for a in algos}}
