"""Regenerate tests/golden/seeded_*.npz from the reference's own OpenMP code.

Run in a checkout where the reference tree is present and `make -C oracle ref`
has built oracle/_ref/.  For each seeded graph below this script stores the
CSR (plus weights) and the outputs of the reference variants compiled from
codes/omp and codes/cpp (BFS, SSSP, CC, MIS, PR with the reference's
EPSILON = 1e-4, TC).
The fixtures are data only; the GPU box uses them without the reference.

usage: python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from indigo2suite_amd import graph as G  # noqa: E402
import oracle as O  # noqa: E402

GRAPHS = {
    "er512": lambda: G.erdos_renyi(512, 8, seed=7),
    "rmat10": lambda: G.rmat(10, 16, seed=3),
    "grid24": lambda: G.grid2d(24, 0.05, seed=5),
    "chunglu1000": lambda: G.chung_lu(1000, 10, 2.3, seed=11),
}


def main():
    for name, make in GRAPHS.items():
        g = make().with_weights(1, 255, seed=1)
        src = g.max_degree_vertex()
        out = {
            "nindex": g.nindex, "nlist": g.nlist, "eweight": g.eweight, "src": np.int32(src),
            "bfs": O.ref_run("bfs_push", g, src),
            "bfs_pull": O.ref_run("bfs_pull", g, src),
            "sssp": O.ref_run("sssp_vtopo_pull", g, src),
            "cc": O.ref_run("cc_pull", g),
            "mis": O.ref_run("mis_vertex", g),
            "mis_edge": O.ref_run("mis_edge", g),
            "pr_eps1e-4": O.ref_run("pr_pull", g),
            "tc": np.int64(O.ref_run("tc_vertex", g)),
            "tc_edge": np.int64(O.ref_run("tc_edge", g)),
            # C++-threads variants (codes/cpp), 4 threads
            "cpp_bfs": O.ref_run("cpp_bfs", g, src),
            "cpp_sssp": O.ref_run("cpp_sssp", g, src),
            "cpp_cc": O.ref_run("cpp_cc", g),
            "cpp_mis": O.ref_run("cpp_mis", g),
            "cpp_pr_eps1e-4": O.ref_run("cpp_pr", g),
            "cpp_tc": np.int64(O.ref_run("cpp_tc", g)),
        }
        path = os.path.join(HERE, f"seeded_{name}.npz")
        np.savez_compressed(path, **out)
        print("wrote", path, g.nodes, g.edges)


if __name__ == "__main__":
    main()
