"""Regression tests for two device-memory hazards of the runtime.

* A captured batched hipGraph keeps the buffer addresses it was recorded with.
  The batch therefore owns its counter and scratch buffers; growing or
  releasing the library's shared workspaces between two replays (a
  reference-signature entry on a larger graph, PageRank on a larger graph,
  igx_release_workspace) must not leave the graph writing freed memory.
* The device starting-point array sp[] is rebuilt for every edge-based run: a
  graph freed element by element with igx_free and another graph with the same
  edge count uploaded in its place (the allocator hands back the same
  addresses) must get its own sp[].
"""
import ctypes

import numpy as np
import pytest

import oracle as O
from indigo2suite_amd import _native
from indigo2suite_amd import batched as BT
from indigo2suite_amd import device as D
from indigo2suite_amd import graph as G

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)


def _check_batch(db, refs, tag):
    for algo, k in (("bfs", 0), ("cc", 1), ("mis", 2)):
        outs = db.results(algo)
        for i, (got, ref) in enumerate(zip(outs, refs)):
            assert np.array_equal(got, ref[k]), (tag, algo, i, db.graphs[i].name)


def _poison(db):
    for buf in db.outs.values():
        buf.from_host(np.full(db.total_nodes, -7, np.int32))


def test_batched_replay_survives_workspace_growth_and_release(gpu):
    graphs = BT.make_batch_graphs(count=16, lo=6, hi=11, dmin=2, dmax=20, seed_base=7100)
    refs = [(O.bfs_serial(g, 0), O.cc_omp_pull(g)[0], O.mis_greedy(g).astype(np.int32)) for g in graphs]
    db = BT.DeviceBatch(graphs)
    mid = G.erdos_renyi(10000, 8, seed=11)                 # grows the entries' output workspace
    big = G.rmat(16, 16, seed=12)                           # grows PageRank's workspaces
    dmid = D.DeviceGraph(mid)
    try:
        _poison(db)
        _say("hazard: first replay")
        assert db.run(7, 256, use_graph=True).launches == 1
        _check_batch(db, refs, "replay0")
        for rep in range(1, 4):
            _say("hazard: igx_bfs_vertex on ER n=10000, PR on R-MAT 16, release", rep)
            dist, _ = D.bfs_vertex(0, dmid)
            assert np.array_equal(dist, O.bfs_serial(mid, 0))
            D.set_pr_params(1e-6, 100)
            try:
                scores, _, it = D.pr(big)
            finally:
                D.set_pr_params(1e-4, 100)
            ref, _, _ = O.pr_fp64(big, 1e-6, 100, fixed_iters=it)
            assert np.max(np.abs(scores - ref) / ref) <= 1e-5
            if rep >= 2:
                _native.check(D.lib().igx_release_workspace(), "igx_release_workspace")
            _poison(db)
            _say("hazard: replay", rep)
            assert db.run(7, 256, use_graph=True).launches == 1
            _check_batch(db, refs, f"replay{rep}")
        # a larger batch in eager mode regrows the batch scratch; the first
        # batch's graph must be recaptured, not replayed with freed buffers
        graphs2 = BT.make_batch_graphs(count=8, lo=12, hi=13, dmin=4, dmax=8, seed_base=7300)
        refs2 = [(O.bfs_serial(g, 0), O.cc_omp_pull(g)[0], O.mis_greedy(g).astype(np.int32)) for g in graphs2]
        db2 = BT.DeviceBatch(graphs2)
        try:
            _poison(db2)
            _say("hazard: larger batch eager")
            db2.run(7, 256, use_graph=False)
            _check_batch(db2, refs2, "eager2")
        finally:
            db2.free()
        _poison(db)
        _say("hazard: replay after larger batch")
        db.run(7, 256, use_graph=True)
        _check_batch(db, refs, "replay-after-regrow")
    finally:
        dmid.free()
        db.free()


class _PieceGraph:
    """A device graph whose arrays are allocated one by one with igx_alloc and
    freed one by one with igx_free (not igx_graph_free_device)."""

    def __init__(self, g):
        self.nodes, self.edges = g.nodes, g.edges
        self.bufs = [D.DeviceBuffer(4 * (g.nodes + 1)), D.DeviceBuffer(4 * max(g.edges, 1)),
                     D.DeviceBuffer(4 * max(g.edges, 1))]
        self.bufs[0].from_host(g.nindex)
        self.bufs[1].from_host(g.nlist)
        self.bufs[2].from_host(g.eweight)
        self.c = _native.IgxGraph()
        self.c.nodes, self.c.edges = g.nodes, g.edges
        self.c.nindex, self.c.nlist, self.c.eweight = [ctypes.cast(b.ptr, _native.c_int_p) for b in self.bufs]

    def addresses(self):
        return tuple(b.ptr.value for b in self.bufs)

    def free(self):
        for b in self.bufs:
            b.free()


def _same_edges_pair():
    """Two graphs on the same vertex count with the same number of directed
    edges but different rows: a path and a star."""
    n = 4097
    a = G.from_edges(n, np.arange(n - 1, dtype=np.int32), np.arange(1, n, dtype=np.int32), name="path")
    b = G.from_edges(n, np.zeros(n - 1, np.int32), np.arange(1, n, dtype=np.int32), name="star")
    assert a.edges == b.edges
    return a.with_weights(1, 9, seed=2), b.with_weights(1, 9, seed=3)


def test_edge_styles_get_fresh_sp_after_free_and_reupload(gpu):
    ga, gb = _same_edges_pair()
    edge_styles = [s for i, s in enumerate(D.styles("bfs")) if D.lib().igx_style_edge_based(0, i) == 1]
    assert edge_styles
    pa = _PieceGraph(ga)
    addr_a = pa.addresses()
    for st in edge_styles:
        _say("sp: bfs", st, "on", ga.name)
        assert np.array_equal(D.run("bfs", pa, st, 0).out, O.bfs_serial(ga, 0)), st
    pa.free()
    pb = _PieceGraph(gb)
    try:
        _say("sp: reuploaded, same addresses:", pb.addresses() == addr_a)
        for algo, ref in (("bfs", O.bfs_serial(gb, 0)), ("sssp", O.sssp_serial(gb, 0)), ("cc", O.cc_omp_pull(gb)[0])):
            a = D.ALGOS[algo]
            for i, st in enumerate(D.styles(algo)):
                if D.lib().igx_style_edge_based(a, i) != 1:
                    continue
                _say("sp:", algo, st, "on", gb.name)
                assert np.array_equal(D.run(algo, pb, st, 0).out, ref), (algo, st)
    finally:
        pb.free()


def test_time_limit_stops_a_long_run_and_the_next_run_is_whole(gpu):
    """igx_set_time_limit: a run needing ~100k rounds (deterministic push
    along a path) stops at the limit with IGX_ERR_TIMEOUT; with the limit
    cleared, the next runs on the same graph are complete and correct."""
    from indigo2suite_amd._native import IGX_ERR_TIMEOUT, IgxError
    n = 100000
    g = G.from_edges(n, np.arange(n - 1, dtype=np.int32), np.arange(1, n, dtype=np.int32), name="path")
    g = g.with_weights(1, 9, seed=4)
    dg = D.DeviceGraph(g)
    try:
        _say("time limit: sssp v_topo_push_thread_determ on a 100k path, limit 0.05 s")
        D.set_time_limit(0.05)
        try:
            with pytest.raises(IgxError) as ei:
                D.run("sssp", dg, "v_topo_push_thread_determ", 0)
            assert ei.value.status == IGX_ERR_TIMEOUT
        finally:
            D.set_time_limit(0.0)
        _say("time limit cleared: sssp and bfs defaults")
        assert np.array_equal(D.run("sssp", dg, None, 0).out, O.sssp_serial(g, 0))
        assert np.array_equal(D.run("bfs", dg, None, 0).out, O.bfs_serial(g, 0))
    finally:
        dg.free()


_SHUTDOWN_SCRIPT = r"""
import numpy as np, oracle as O
from indigo2suite_amd import device as D, graph as G
g = G.rmat(12, 8, seed=9).with_weights(1, 9, seed=9)
for k in range(2):
    D.init(0)
    dg = D.DeviceGraph(g)
    assert np.array_equal(D.run("sssp", dg, None, 0).out, O.sssp_serial(g, 0))
    assert np.array_equal(D.run("bfs", dg, "e_data_push_thread", 0).out, O.bfs_serial(g, 0))
    dg.free()
    D.check(D.lib().igx_shutdown(), "igx_shutdown")
    D.check(D.lib().igx_shutdown(), "igx_shutdown (again)")
print("shutdown ok", flush=True)
"""


def test_shutdown_then_init_again_in_a_fresh_process(gpu):
    """igx_shutdown releases the stream, events, control block and
    workspaces; a later call initialises the library again.  Run in a child
    process so this session's library state is untouched."""
    import os
    import subprocess
    import sys
    repo = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    _say("shutdown / re-init in a child process")
    out = subprocess.run([sys.executable, "-c", _SHUTDOWN_SCRIPT], cwd=repo, stdout=subprocess.PIPE,
                         stderr=subprocess.STDOUT, text=True, timeout=300,
                         env=dict(os.environ, PYTHONPATH=repo))
    assert out.returncode == 0 and "shutdown ok" in out.stdout, out.stdout[-3000:]
