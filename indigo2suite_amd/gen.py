"""Input tooling: seeded synthetic graphs and conversions, written as .egr.

The .egr format is the reference's ECLgraph file (include/ECLgraph.h:82-96,
written by igx_egr_write): int32 nodes, int32 edges, nindex[nodes + 1],
nlist[edges] and, when present, eweight[edges].  Every graph written here is
undirected in the reference's sense: both directions of every edge stored,
adjacency sorted, no self loops and no duplicates.

  python -m indigo2suite_amd.gen er      --n 16384 --degree 8 --seed 1 -o er.egr
  python -m indigo2suite_amd.gen rmat    --scale 22 --edge-factor 16 --seed 1 [--abc .57 .19 .19] -o rmat22.egr
  python -m indigo2suite_amd.gen grid    --side 2048 --long 0.05 --seed 1 -o grid.egr
  python -m indigo2suite_amd.gen chunglu --n 1000000 --degree 16 --gamma 2.5 --seed 1 -o cl.egr
  python -m indigo2suite_amd.gen convert edges.txt -o g.egr [--nodes N] [--one-based]
  python -m indigo2suite_amd.gen info g.egr
  python -m indigo2suite_amd.gen large-inputs [--dir large_inputs]

Generators take --weights LO HI to store symmetric weights uniform in
[LO, HI] (both directions of an edge get the same weight), or
--reference-weights for the reference harnesses' fallback 1 + (u * v mod n)
(sssp_vertex_cuda.h:127-137); without either, no weights are stored.
`convert` reads a text edge list (one "u v" pair per line; '#' and '%'
comment lines, as in SNAP and Matrix Market files, are skipped; a Matrix
Market size line is recognised) and symmetrizes it.  `large-inputs` looks for
the five graphs that download_large_inputs.sh:39-43 fetches, under --dir or
$IGX_LARGE_INPUTS, and reports which are there (nothing is downloaded).
"""
from __future__ import annotations

import argparse
import os
import sys
from typing import List, Optional

import numpy as np

from . import graph as G

# download_large_inputs.sh:39-43
LARGE_INPUTS = ("2d-2e20.sym.egr", "coPapersDBLP.egr", "rmat22.sym.egr", "soc-LiveJournal1.egr",
                "USA-road-d.NY.egr")


def large_input_dir(d: Optional[str] = None) -> str:
    return d or os.environ.get("IGX_LARGE_INPUTS", "large_inputs")


def large_input_path(name: str, d: Optional[str] = None) -> Optional[str]:
    """Path of one of the reference's large inputs if it is present, else None."""
    if name not in LARGE_INPUTS:
        raise KeyError(f"{name!r} is not one of {LARGE_INPUTS}")
    p = os.path.join(large_input_dir(d), name)
    return p if os.path.isfile(p) else None


def load_large_input(name: str, d: Optional[str] = None) -> G.Graph:
    p = large_input_path(name, d)
    if p is None:
        raise FileNotFoundError(f"{name} not found under {large_input_dir(d)} (download_large_inputs.sh fetches it)")
    return G.read_egr(p)


def read_edge_list(path: str, one_based: bool = False):
    """Text edge list -> (src, dst, max id + 1).  Skips '#' / '%' comments and
    a Matrix Market size line (three integers after a '%%MatrixMarket'
    header); extra columns (weights) are ignored."""
    src: List[int] = []
    dst: List[int] = []
    mm = False
    size_seen = False
    with open(path) as f:
        for line in f:
            s = line.strip()
            if not s:
                continue
            if s.startswith("%%MatrixMarket"):
                mm = True
                one_based = True
                continue
            if s[0] in "#%":
                continue
            parts = s.split()
            if mm and not size_seen:
                size_seen = True
                continue
            if len(parts) < 2:
                raise ValueError(f"{path}: bad edge line {line!r}")
            src.append(int(parts[0]))
            dst.append(int(parts[1]))
    a = np.array(src, dtype=np.int64)
    b = np.array(dst, dtype=np.int64)
    if one_based:
        a -= 1
        b -= 1
    if a.size and (a.min() < 0 or b.min() < 0):
        raise ValueError(f"{path}: negative vertex id (one-based ids need --one-based)")
    n = int(max(a.max(), b.max()) + 1) if a.size else 0
    return a, b, n


def convert(path: str, nodes: Optional[int] = None, one_based: bool = False) -> G.Graph:
    a, b, n = read_edge_list(path, one_based)
    n = max(n, nodes or 0, 1)
    if n >= 2 ** 31:
        raise ValueError("more than 2^31 - 1 vertices do not fit int32 ids")
    return G.from_edges(n, a.astype(np.int32), b.astype(np.int32), name=os.path.basename(path))


def info(g: G.Graph) -> dict:
    deg = g.degrees()
    out = {"nodes": g.nodes, "edges": g.edges, "weights": g.eweight is not None,
           "avg_degree": round(g.edges / max(g.nodes, 1), 3), "max_degree": int(deg.max()) if deg.size else 0,
           "isolated": int(np.sum(deg == 0)), "sorted_no_self_loops": g.is_sorted()}
    if g.eweight is not None and g.edges:
        out["weight_range"] = [int(g.eweight.min()), int(g.eweight.max())]
    return out


def _weights(g: G.Graph, args) -> G.Graph:
    if args.reference_weights:
        return g.with_reference_weights()
    if args.weights:
        lo, hi = args.weights
        return g.with_weights(lo, hi, seed=args.seed)
    return g


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(prog="python -m indigo2suite_amd.gen", description=__doc__.split("\n")[0])
    sub = ap.add_subparsers(dest="cmd", required=True)

    def gen_parser(name):
        p = sub.add_parser(name)
        p.add_argument("--seed", type=int, default=1)
        p.add_argument("--weights", type=int, nargs=2, metavar=("LO", "HI"))
        p.add_argument("--reference-weights", action="store_true")
        p.add_argument("-o", "--output", required=True)
        return p

    p = gen_parser("er")
    p.add_argument("--n", type=int, required=True)
    p.add_argument("--degree", type=float, default=8.0)
    p = gen_parser("rmat")
    p.add_argument("--scale", type=int, required=True)
    p.add_argument("--edge-factor", type=float, default=16.0)
    p.add_argument("--abc", type=float, nargs=3, default=list(G.RMAT_ABC))
    p.add_argument("--n-limit", type=int, default=0, help="drop samples with an endpoint >= this (0: none)")
    p = gen_parser("grid")
    p.add_argument("--side", type=int, required=True)
    p.add_argument("--long", type=float, default=0.05, help="fraction of vertices with one long-range edge")
    p = gen_parser("chunglu")
    p.add_argument("--n", type=int, required=True)
    p.add_argument("--degree", type=float, default=16.0)
    p.add_argument("--gamma", type=float, default=2.5)
    p = gen_parser("convert")
    p.add_argument("input")
    p.add_argument("--nodes", type=int, default=None)
    p.add_argument("--one-based", action="store_true")
    p = sub.add_parser("info")
    p.add_argument("input")
    p = sub.add_parser("large-inputs")
    p.add_argument("--dir", default=None)
    args = ap.parse_args(argv)

    if args.cmd == "info":
        import json
        print(json.dumps(info(G.read_egr(args.input))))
        return 0
    if args.cmd == "large-inputs":
        d = large_input_dir(args.dir)
        for name in LARGE_INPUTS:
            p = large_input_path(name, args.dir)
            print(f"{name}: {'present' if p else 'missing'} ({os.path.join(d, name)})")
        return 0
    if args.cmd == "er":
        g = G.erdos_renyi(args.n, args.degree, seed=args.seed)
    elif args.cmd == "rmat":
        g = G.rmat(args.scale, args.edge_factor, tuple(args.abc), seed=args.seed, n_limit=args.n_limit)
    elif args.cmd == "grid":
        g = G.grid2d(args.side, args.long, seed=args.seed)
    elif args.cmd == "chunglu":
        g = G.chung_lu(args.n, args.degree, args.gamma, seed=args.seed)
    else:
        g = convert(args.input, args.nodes, args.one_based)
    g = _weights(g, args)
    g.write_egr(args.output)
    print(f"wrote {args.output}: {g.nodes} nodes, {g.edges} edges{', weights' if g.eweight is not None else ''}",
          file=sys.stderr)
    return 0


if __name__ == "__main__":
    sys.exit(main())
