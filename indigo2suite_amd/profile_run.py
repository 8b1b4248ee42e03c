"""Workload driver for profiling one algorithm with rocprofv3.

Runs `--runs` timed runs of one algorithm's default style (or `--style`) on
one of the measured graphs, with nothing else in the process but the graph
upload, so that every kernel dispatch rocprofv3 records (other than the
runtime's own copy and fill kernels) belongs to those runs.  With
`--copy-probe MIB` it ends with one device-to-device copy of MIB MiB, the
known-size kernel that calibrates the FETCH_SIZE / WRITE_SIZE counters.

  rocprofv3 --kernel-trace --stats -d OUT -o run --output-format csv -- \\
      python3 -m indigo2suite_amd.profile_run --algo sssp --graph rmat22 --runs 5
  rocprofv3 --pmc FETCH_SIZE -d OUT -o run --output-format csv -- \\
      python3 -m indigo2suite_amd.profile_run --algo sssp --graph rmat22 --runs 5 --copy-probe 256

Graphs: rmat22 / rmat20 (R-MAT scale 22 / 20, edge factor 16, (a, b, c) =
(0.57, 0.19, 0.19), seed 1, weights U[1, 255]), grid2048 (2048 x 2048 grid,
5% of the vertices with one long-range edge, seed 1), er16384 (Erdos-Renyi,
n = 16384, average degree 8, seed 1).  BFS and SSSP start at the
highest-degree vertex; PageRank stops at L1 < 1e-6 or 100 iterations.
Prints one JSON line with the per-run seconds, iterations and the sizes.
"""
from __future__ import annotations

import argparse
import json
import sys

import numpy as np

from . import device as D
from . import graph as G

GRAPHS = {
    "rmat22": lambda: G.rmat(22, 16, seed=1).with_weights(1, 255, seed=1),
    "rmat20": lambda: G.rmat(20, 16, seed=1).with_weights(1, 255, seed=1),
    "grid2048": lambda: G.grid2d(2048, 0.05, seed=1).with_weights(1, 255, seed=1),
    "er16384": lambda: G.erdos_renyi(16384, 8, seed=1).with_weights(1, 255, seed=1),
}


def main(argv=None) -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--algo", required=True, choices=sorted(D.ALGOS))
    ap.add_argument("--graph", required=True, choices=sorted(GRAPHS))
    ap.add_argument("--style", default=None)
    ap.add_argument("--runs", type=int, default=5)
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB")
    a = ap.parse_args(argv)
    D.init(0)
    g = GRAPHS[a.graph]()
    src = g.max_degree_vertex()
    dg = D.DeviceGraph(g)
    out = D.DeviceBuffer(max(4 * g.nodes, 8))
    if a.algo == "pr":
        D.set_pr_params(1e-6, 100)
    style = D.styles(a.algo)[D.style_id(a.algo, a.style)]
    secs, iters, work = [], [], []
    for _ in range(a.runs):
        info = D.run_device(a.algo, dg, out, style, src)
        secs.append(info.seconds)
        iters.append(info.iterations)
        work.append((int(info.edges_processed), int(info.edges_relaxed)))
    D.lib().igx_sync()
    if a.copy_probe > 0:
        nb = a.copy_probe << 20
        x, y = D.DeviceBuffer(nb), D.DeviceBuffer(nb)
        D.check(D.lib().igx_memcpy_d2d(y.ptr, x.ptr, nb), "igx_memcpy_d2d")
        x.free()
        y.free()
    print(json.dumps({"algo": a.algo, "graph": a.graph, "style": style, "n": g.nodes, "m": g.edges,
                      "runs": a.runs, "seconds": secs, "median_s": float(np.median(secs)), "iterations": iters,
                      "edges_scanned_per_m": [round(w[0] / max(g.edges, 1), 3) for w in work],
                      "edges_relaxed_per_m": [round(w[1] / max(g.edges, 1), 3) for w in work],
                      "copy_probe_mib": a.copy_probe}), flush=True)
    out.free()
    dg.free()
    return 0


if __name__ == "__main__":
    sys.exit(main())
