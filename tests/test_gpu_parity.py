"""GPU parity: every registered style of every algorithm, through the C ABI,
against the oracle on the same inputs.  Bit-exact for BFS/SSSP/CC/MIS/TC;
PageRank within 1e-5 relative of the FP64 oracle at the same iteration count."""
import json
import os
import sys

import numpy as np
import pytest

import oracle as O
from indigo2suite_amd import graph as G

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = os.path.join(HERE, "golden")
PR_RTOL = 1e-5  # PageRank tolerance vs the FP64 oracle (max relative error)


def _fixtures():
    with open(os.path.join(GOLDEN, "small_inputs.json")) as f:
        fx = json.load(f)["fixtures"]
    out = []
    for f in fx:
        e = np.array(f["edges"], dtype=np.int32).reshape(-1, 2)
        out.append(G.from_edges(3, e[:, 0], e[:, 1], name=f"fixture{f['id']}"))
    return out


def _seeded():
    out = []
    for name in ["er512", "rmat10", "grid24", "chunglu1000"]:
        d = np.load(os.path.join(GOLDEN, f"seeded_{name}.npz"))
        out.append(G.Graph(d["nindex"], d["nlist"], d["eweight"], name))
    return out


def _graphs():
    gs = _fixtures() + _seeded()
    gs.append(G.erdos_renyi(16384, 8, seed=1))                 # the BFS oracle config
    gs.append(G.rmat(14, 16, seed=1))
    gs.append(G.grid2d(64, 0.05, seed=1))
    gs.append(G.chung_lu(20000, 12, 2.1, seed=3))
    k = 5000                                                   # hub: multi-chunk rows
    star = G.from_edges(k + 10, np.zeros(k, np.int32), np.arange(1, k + 1), name="star5000")
    gs.append(star)
    gs.append(G.from_edges(5, np.zeros(0, np.int32), np.zeros(0, np.int32), name="edgeless5"))
    gs.append(G.from_edges(1, np.zeros(0, np.int32), np.zeros(0, np.int32), name="single"))
    return [g if g.eweight is not None else g.with_weights(1, 255, seed=7) for g in gs]


GRAPHS = _graphs()
IDS = [g.name for g in GRAPHS]


@pytest.fixture(scope="module")
def dev(gpu):
    return gpu


@pytest.fixture(scope="module")
def dgraphs(dev):
    dgs = {}
    for g in GRAPHS:
        dgs[g.name] = dev.DeviceGraph(g)
    yield dgs
    for d in dgs.values():
        d.free()


def _styles(dev, algo):
    return dev.styles(algo)


def _say(*a):
    print(*a, flush=True)


@pytest.mark.parametrize("gi", range(len(GRAPHS)), ids=IDS)
def test_bfs_all_styles(dev, dgraphs, gi):
    g = GRAPHS[gi]
    src = g.max_degree_vertex()
    ref = O.bfs_serial(g, src)
    for st in _styles(dev, "bfs"):
        _say("bfs", g.name, st)
        r = dev.run("bfs", dgraphs[g.name], st, src)
        assert np.array_equal(r.out, ref), (g.name, st)


@pytest.mark.parametrize("gi", range(len(GRAPHS)), ids=IDS)
def test_sssp_all_styles(dev, dgraphs, gi):
    g = GRAPHS[gi]
    src = g.max_degree_vertex()
    ref = O.sssp_serial(g, src)
    for st in _styles(dev, "sssp"):
        _say("sssp", g.name, st)
        r = dev.run("sssp", dgraphs[g.name], st, src)
        assert np.array_equal(r.out, ref), (g.name, st)


@pytest.mark.parametrize("gi", range(len(GRAPHS)), ids=IDS)
def test_cc_all_styles(dev, dgraphs, gi):
    g = GRAPHS[gi]
    ref = O.cc_omp_pull(g)[0]
    for st in _styles(dev, "cc"):
        _say("cc", g.name, st)
        r = dev.run("cc", dgraphs[g.name], st)
        assert np.array_equal(r.out, ref), (g.name, st)


@pytest.mark.parametrize("gi", range(len(GRAPHS)), ids=IDS)
def test_mis_all_styles(dev, dgraphs, gi):
    g = GRAPHS[gi]
    ref = O.mis_greedy(g).astype(np.int32)
    for st in _styles(dev, "mis"):
        _say("mis", g.name, st)
        r = dev.run("mis", dgraphs[g.name], st)
        assert np.array_equal(r.out, ref), (g.name, st)


@pytest.mark.parametrize("gi", range(len(GRAPHS)), ids=IDS)
def test_tc_all_styles(dev, dgraphs, gi):
    g = GRAPHS[gi]
    ref = O.tc_serial(g)
    for st in _styles(dev, "tc"):
        _say("tc", g.name, st)
        r = dev.run("tc", dgraphs[g.name], st)
        assert r.out == ref, (g.name, st)


@pytest.mark.parametrize("gi", range(len(GRAPHS)), ids=IDS)
def test_pr_all_styles(dev, dgraphs, gi):
    g = GRAPHS[gi]
    dev.set_pr_params(1e-6, 100)
    try:
        _, it_ref, _ = O.pr_fp64(g, 1e-6, 100)
        for st in _styles(dev, "pr"):
            _say("pr", g.name, st)
            r = dev.run("pr", dgraphs[g.name], st)
            ref, _, _ = O.pr_fp64(g, 1e-6, 100, fixed_iters=r.iterations)
            deg = g.degrees()
            # isolated vertices keep (1-d)/n; compare every vertex
            rel = np.abs(r.out.astype(np.float64) - ref) / np.abs(ref)
            assert np.max(rel) <= PR_RTOL, (g.name, st, float(np.max(rel)))
            assert abs(r.iterations - it_ref) <= 1, (g.name, st, r.iterations, it_ref)
            assert deg.shape[0] == g.nodes
    finally:
        dev.set_pr_params(1e-4, 100)


def test_tc_rejects_unsorted(dev):
    g = G.Graph(np.array([0, 2, 3, 4]), np.array([2, 1, 0, 0]), None, "unsorted")
    dg = dev.DeviceGraph(g)
    try:
        with pytest.raises(Exception) as e:
            dev.run("tc", dg, "v_merge_thread")
        assert "UNSORTED" in str(e.value)
    finally:
        dg.free()


def test_sssp_needs_weights_and_source_in_range(dev):
    g = G.rmat(8, 8, seed=1)
    dg = dev.DeviceGraph(g)
    try:
        with pytest.raises(Exception) as e:
            dev.run("sssp", dg, None, 0)
        assert "NO_WEIGHTS" in str(e.value)
        with pytest.raises(Exception) as e:
            dev.run("bfs", dg, None, g.nodes)
        assert "ERR_ARG" in str(e.value)
    finally:
        dg.free()


@pytest.mark.parametrize("k", [600, 1100])
def test_tc_complete_graph_closed_form(dev, k):
    """K_k has C(k, 3) triangles; its oriented out-lists reach k - 1 > 512
    entries, so the oriented style hashes them in slices."""
    i, j = np.triu_indices(k, 1)
    g = G.from_edges(k, i.astype(np.int32), j.astype(np.int32), name=f"complete{k}")
    dg = dev.DeviceGraph(g)
    try:
        want = k * (k - 1) * (k - 2) // 6
        for st in ("oriented_lds", "v_merge_wave"):
            _say("tc", g.name, st)
            assert dev.run("tc", dg, st).out == want, st
    finally:
        dg.free()
