"""The allocation-free batched entry (igx_batch_run_noalloc) through the C ABI:
its BFS levels, CC labels and MIS sets equal those of igx_batch_run (eager)
and the oracle's, on two calls in a row, the second with a smaller batch
(the shared scratch array keeps the first call's contents); a batch above
its vertex limit is rejected with IGX_ERR_LIMIT before anything runs."""
import numpy as np
import pytest

import oracle as O
from indigo2suite_amd import _native
from indigo2suite_amd import batched as BT
from indigo2suite_amd import graph as G

pytestmark = pytest.mark.gpu

MAX_NODES = 1 << 20


def _say(*a):
    print(*a, flush=True)


def _poison(db):
    for buf in db.outs.values():
        buf.from_host(np.full(db.total_nodes, -7, np.int32))


def _outputs(db):
    return {a: db.outs[a].to_host(np.int32, db.total_nodes) for a in ("bfs", "cc", "mis")}


@pytest.mark.parametrize("block", [256, 1024])
def test_noalloc_matches_batch_run_twice_then_smaller(gpu, block):
    big = BT.make_batch_graphs(count=48, lo=10, hi=15, seed_base=9100)
    small = BT.make_batch_graphs(count=12, lo=6, hi=11, seed_base=9300)
    for tag, graphs in (("first", big), ("second, smaller", small)):
        db = BT.DeviceBatch(graphs)
        try:
            assert db.total_nodes <= MAX_NODES
            _poison(db)
            _say("noalloc:", tag, "igx_batch_run eager, block", block, "graphs", len(graphs), "n", db.total_nodes)
            db.run(7, block, use_graph=False)
            want = _outputs(db)
            _poison(db)
            _say("noalloc:", tag, "igx_batch_run_noalloc")
            info = db.run_noalloc(7, block)
            assert info.launches == 2 and info.iterations == len(graphs)
            got = _outputs(db)
            for a in ("bfs", "cc", "mis"):
                assert np.array_equal(got[a], want[a]), (tag, a)
            for i, g in enumerate(graphs):
                lo, hi = int(db.node_off[i]), int(db.node_off[i + 1])
                assert np.array_equal(got["bfs"][lo:hi], O.bfs_serial(g, 0)), (tag, i)
                assert np.array_equal(got["cc"][lo:hi], O.cc_omp_pull(g)[0]), (tag, i)
                assert np.array_equal(got["mis"][lo:hi], O.mis_greedy(g).astype(np.int32)), (tag, i)
        finally:
            db.free()


def test_noalloc_rejects_batches_above_its_limit(gpu):
    # two edgeless graphs of 2^19 + 1 vertices: one vertex over the limit
    half = MAX_NODES // 2 + 1
    graphs = [G.from_edges(half, np.zeros(0, np.int32), np.zeros(0, np.int32), name=f"edgeless{i}") for i in (0, 1)]
    db = BT.DeviceBatch(graphs)
    try:
        assert db.total_nodes == MAX_NODES + 2
        _say("noalloc: oversize batch is rejected")
        with pytest.raises(_native.IgxError) as e:
            db.run_noalloc(7, 256)
        assert "LIMIT" in str(e.value) and "at most" in str(e.value)
    finally:
        db.free()
