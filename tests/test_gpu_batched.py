"""Batched mode: every graph of a packed batch gets the same BFS levels, CC
labels and MIS set as the oracle, in eager mode, for each workgroup size, and
on repeated replays of the captured hipGraph with other GPU work between."""
import numpy as np
import pytest

import oracle as O
from indigo2suite_amd import batched as BT
from indigo2suite_amd import device as D

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def batch(gpu):
    graphs = BT.make_batch_graphs(count=24, lo=6, hi=12, dmin=2, dmax=24, seed_base=5000)
    db = BT.DeviceBatch(graphs)
    refs = [(O.bfs_serial(g, 0), O.cc_omp_pull(g)[0], O.mis_greedy(g).astype(np.int32)) for g in graphs]
    yield db, refs
    db.free()


def _check(db, refs, tag):
    for algo, k in (("bfs", 0), ("cc", 1), ("mis", 2)):
        outs = db.results(algo)
        for i, (got, ref) in enumerate(zip(outs, refs)):
            assert np.array_equal(got, ref[k]), (tag, algo, i, db.graphs[i].name)


@pytest.mark.parametrize("block", [256, 512, 1024])
def test_batched_eager_matches_oracle(batch, block):
    db, refs = batch
    for k in db.outs.values():
        k.from_host(np.full(db.total_nodes, -7, np.int32))
    print("batched eager block", block, flush=True)
    db.run(7, block, use_graph=False)
    _check(db, refs, f"eager{block}")


def test_batched_graph_replays(batch):
    db, refs = batch
    g = BT.make_batch_graphs(count=1, lo=10, hi=10, seed_base=77)[0]
    dg = D.DeviceGraph(g)
    try:
        for rep in range(4):
            for k in db.outs.values():
                k.from_host(np.full(db.total_nodes, -7, np.int32))
            print("batched graph replay", rep, flush=True)
            info = db.run(7, 256, use_graph=True)
            assert info.launches == 1
            _check(db, refs, f"replay{rep}")
            # unrelated eager work between replays
            r = D.run("bfs", dg, "do_hybrid", 0)
            assert np.array_equal(r.out, O.bfs_serial(g, 0))
    finally:
        dg.free()


def test_per_graph_baseline_matches(batch):
    db, refs = batch
    for k in db.outs.values():
        k.from_host(np.full(db.total_nodes, -7, np.int32))
    db.per_graph_baseline()
    _check(db, refs, "per-graph")
