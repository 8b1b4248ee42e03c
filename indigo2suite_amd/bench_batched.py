"""Batched-mode benchmark: BFS (source 0), CC and MIS over a batch of uneven
graphs, one persistent launch per algorithm replayed as one hipGraph, against
one launch sequence per graph and algorithm through the single-graph entry.

  python -m indigo2suite_amd.bench_batched [--count 512] [--lo 10] [--hi 18]
        [--repeats 3] [--verify 16]

Default workload: 512 graphs, half Erdos-Renyi and half R-MAT
(0.57, 0.19, 0.19, 0.05), vertex counts log-uniform in [2^10, 2^18], average
degree uniform in [4, 32], seeds 1000 + i.  Prints one JSON line.
"""
from __future__ import annotations

import argparse
import json
import sys
import time

import numpy as np

from . import batched as BT
from . import device as D


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--count", type=int, default=512)
    ap.add_argument("--lo", type=int, default=10)
    ap.add_argument("--hi", type=int, default=18)
    ap.add_argument("--repeats", type=int, default=3)
    ap.add_argument("--verify", type=int, default=16, help="graphs checked against the oracle (0: none)")
    a = ap.parse_args(argv)
    print(json.dumps(measure(a.count, a.lo, a.hi, a.repeats, a.verify)), flush=True)


def measure(count=512, lo=10, hi=18, repeats=3, verify=16) -> dict:
    """Build the batch, pick the workgroup size by an exhaustive sweep, time
    the captured batch against one launch sequence per graph; returns the
    record main() prints."""
    D.init()
    t = time.time()
    graphs = BT.make_batch_graphs(count, lo, hi)
    n_tot = sum(g.nodes for g in graphs)
    m_tot = sum(g.edges for g in graphs)
    print(f"generated {count} graphs, n={n_tot} m={m_tot} in {time.time() - t:.1f}s", file=sys.stderr, flush=True)
    db = BT.DeviceBatch(graphs)
    tune = BT.autotune_block(db, repeats=repeats)
    block = tune["best_block"]
    print("block sweep", tune, file=sys.stderr, flush=True)
    ts = []
    for _ in range(repeats):
        D.lib().igx_sync()
        t0 = time.perf_counter()
        info = db.run(7, block, use_graph=True)
        D.lib().igx_sync()
        ts.append((time.perf_counter() - t0, info.seconds))
    batched_wall = float(np.median([x[0] for x in ts]))
    batched_gpu = float(np.median([x[1] for x in ts]))
    checked = 0
    if verify:
        import oracle as O
        outs = {k: db.results(k) for k in ("bfs", "cc", "mis")}
        idx = np.linspace(0, count - 1, min(verify, count)).astype(int)
        for i in idx:
            g = graphs[i]
            assert np.array_equal(outs["bfs"][i], O.bfs_serial(g, 0)), ("bfs", i)
            assert np.array_equal(outs["cc"][i], O.cc_omp_pull(g)[0]), ("cc", i)
            assert np.array_equal(outs["mis"][i], O.mis_greedy(g).astype(np.int32)), ("mis", i)
            checked += 1
    # one launch sequence per graph and algorithm (single-graph default styles)
    D.lib().igx_sync()
    t0 = time.perf_counter()
    per_gpu = db.per_graph_baseline()
    D.lib().igx_sync()
    per_wall = time.perf_counter() - t0
    line = {
        "metric": "batched BFS+CC+MIS throughput vs one launch per graph",
        "graphs": count, "total_nodes": n_tot, "total_edges": m_tot, "block": block,
        "block_sweep_seconds": tune["seconds"],
        "batched_seconds_wall": round(batched_wall, 6), "batched_seconds_events": round(batched_gpu, 6),
        "per_graph_seconds_wall": round(per_wall, 6), "per_graph_seconds_events": round(per_gpu, 6),
        "speedup_wall": round(per_wall / batched_wall, 2), "speedup_events": round(per_gpu / batched_gpu, 2),
        "batched_gteps": round(3 * m_tot / batched_wall * 1e-9, 3),
        "graphs_verified": checked,
    }
    db.free()
    return line


if __name__ == "__main__":
    main(sys.argv[1:])
