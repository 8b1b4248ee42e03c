"""Long-row chunk list at its size limits: graphs of about 2^23 vertices just
below and just above the row count whose range offsets the fill kernel scans
in LDS (4096 ranges of 2048 rows), with rows longer than one 256-edge chunk
on both sides of range boundaries and at the last vertex.  The BFS, SSSP and
CC styles that consume the list are checked against the oracle."""
import numpy as np
import pytest

import oracle as O
from indigo2suite_amd import graph as G

pytestmark = pytest.mark.gpu

LDS_ROWS = 4096 * 2048


def _graph(n, seed):
    rng = np.random.default_rng(seed)
    hubs = {0: 3000, 2047: 600, 2048: 300, LDS_ROWS - 1: 257, n - 1: 1000, n // 2: 5000}
    src, dst = [], []
    for h, d in hubs.items():
        if h < n:
            src.append(np.full(d, h, np.int64))
            dst.append(rng.choice(n, d, replace=False))
    k = 200000
    src.append(rng.integers(0, n, k))
    dst.append(rng.integers(0, n, k))
    s = np.concatenate(src)
    t = np.concatenate(dst)
    keep = s != t
    return G.from_edges(n, s[keep], t[keep], name=f"long_rows_{n}").with_weights(1, 255, seed=seed)


@pytest.mark.parametrize("n", [LDS_ROWS - 7, LDS_ROWS + 2100])
def test_long_row_list_consumers(gpu, n):
    g = _graph(n, 11)
    src = g.max_degree_vertex()
    print("long rows", g.name, "edges", g.edges, flush=True)
    dg = gpu.DeviceGraph(g)
    try:
        ref = O.bfs_serial(g, src)
        for st in ("do_hybrid", "v_topo_push_thread"):
            print("bfs", st, flush=True)
            assert np.array_equal(gpu.run("bfs", dg, st, src).out, ref), st
        ref = O.sssp_serial(g, src)
        print("sssp v_do_balanced", flush=True)
        assert np.array_equal(gpu.run("sssp", dg, "v_do_balanced", src).out, ref)
        ref = O.cc_omp_pull(g)[0]
        for st in ("uf_afforest", "uf_hook_compress"):
            print("cc", st, flush=True)
            assert np.array_equal(gpu.run("cc", dg, st).out, ref), st
    finally:
        dg.free()
