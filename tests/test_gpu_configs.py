"""GPU parity at the full sizes of the measured workloads, against the oracle.

* BFS on Erdos-Renyi n = 16384, average degree 8, from source 0: every style.
* SSSP, CC and MIS on R-MAT scale 22 (n = 4,194,304, edge factor 16, seed 1,
  weights U[1,255] equal in both directions), SSSP from the highest-degree
  vertex: every style, plus the reference-signature entries with host sp[].
  BFS (default style) on the same graph.
* PageRank (d = 0.85, L1 < 1e-6 or 100 iterations) and triangle counting on
  R-MAT scale 20 and on the 2048 x 2048 grid with 5% long-range edges:
  PageRank within 1e-5 relative of the FP64 oracle at the same iteration
  count, triangle counts exact.
* The 512-graph batch (Erdos-Renyi and R-MAT halves, n log-uniform in
  [2^10, 2^18], degree U[4,32]): BFS, CC and MIS of every graph, eager and
  as a captured hipGraph.

Oracle: sssp_serial (Dijkstra, sssp_edge_omp.h:105-130), bfs_serial,
cc_omp_pull ([O-CC] pull form, component-minimum labels), mis_greedy
(unsigned hash order, mis_vertex_omp.h:127-132), pr_fp64 (pr_omp.h), and
tc_forward / tc_serial.
"""
import numpy as np
import pytest

import oracle as O
from indigo2suite_amd import autotune
from indigo2suite_amd import batched as BT
from indigo2suite_amd import graph as G

pytestmark = pytest.mark.gpu

PR_RTOL = 1e-5  # PageRank: max relative error vs the FP64 oracle at the same iteration count
PR_EPS = 1e-6
PR_MAX_ITER = 100


def _say(*a):
    print(*a, flush=True)


# ------------------------------------------------------------- ER, BFS src 0
@pytest.fixture(scope="module")
def er(gpu):
    g = G.erdos_renyi(16384, 8, seed=1)
    dg = gpu.DeviceGraph(g)
    yield g, dg, O.bfs_serial(g, 0), O.bfs_omp_push(g, 0)[0]
    dg.free()


def test_er16384_bfs_source0_all_styles(gpu, er):
    g, dg, ref, ref_push = er
    assert np.array_equal(ref, ref_push)          # the plain OpenMP method agrees with the serial BFS
    assert g.nodes == 16384 and abs(g.edges - 131072) < 4096
    for st in gpu.styles("bfs"):
        _say("er16384 bfs src 0", st)
        assert np.array_equal(gpu.run("bfs", dg, st, 0).out, ref), st
    dist, _ = gpu.bfs_vertex(0, dg)
    assert np.array_equal(dist, ref)
    dist, _ = gpu.bfs_edge(0, dg, g.sp())
    assert np.array_equal(dist, ref)


# --------------------------------------------------------- R-MAT 22: SSSP, CC
@pytest.fixture(scope="module")
def rmat22(gpu):
    _say("generating R-MAT 22")
    g = G.rmat(22, 16, seed=1).with_weights(1, 255, seed=1)
    src = g.max_degree_vertex()
    dg = gpu.DeviceGraph(g)
    refs = {}
    yield g, dg, src, refs
    dg.free()
    gpu.lib().igx_release_workspace()


def _ref(refs, key, fn):
    if key not in refs:
        refs[key] = fn()
    return refs[key]


@pytest.mark.parametrize("si", range(16))
def test_rmat22_sssp_style(gpu, rmat22, si):
    g, dg, src, refs = rmat22
    names = gpu.styles("sssp")
    if si >= len(names):
        pytest.skip("fewer styles")
    ref = _ref(refs, "sssp", lambda: O.sssp_serial(g, src))
    _say("rmat22 sssp", names[si], "src", src)
    assert np.array_equal(gpu.run("sssp", dg, names[si], src).out, ref), names[si]


def test_rmat22_sssp_reference_entries(gpu, rmat22):
    g, dg, src, refs = rmat22
    ref = _ref(refs, "sssp", lambda: O.sssp_serial(g, src))
    _say("rmat22 igx_sssp_edge (host sp) and igx_sssp_vertex")
    dist, t = gpu.sssp_edge(src, dg, g.sp())
    assert np.array_equal(dist, ref) and t > 0
    dist, t = gpu.sssp_vertex(src, dg)
    assert np.array_equal(dist, ref) and t > 0


@pytest.mark.parametrize("si", range(16))
def test_rmat22_cc_style(gpu, rmat22, si):
    g, dg, _, refs = rmat22
    names = gpu.styles("cc")
    if si >= len(names):
        pytest.skip("fewer styles")
    ref = _ref(refs, "cc", lambda: O.cc_omp_pull(g)[0])
    _say("rmat22 cc", names[si])
    assert np.array_equal(gpu.run("cc", dg, names[si]).out, ref), names[si]


def test_rmat22_cc_reference_entries(gpu, rmat22):
    g, dg, _, refs = rmat22
    ref = _ref(refs, "cc", lambda: O.cc_omp_pull(g)[0])
    _say("rmat22 igx_cc_vertex and igx_cc_edge (host sp)")
    lab, _ = gpu.cc_vertex(dg)
    assert np.array_equal(lab, ref)
    lab, _ = gpu.cc_edge(dg, g.sp())
    assert np.array_equal(lab, ref)
    assert O.cc_verify(g, lab) == 0


def test_rmat22_bfs_and_mis_defaults(gpu, rmat22):
    g, dg, src, _ = rmat22
    _say("rmat22 bfs default src", src)
    assert np.array_equal(gpu.run("bfs", dg, None, src).out, O.bfs_serial(g, src))
    _say("rmat22 mis default")
    st = gpu.run("mis", dg, None).out
    assert np.array_equal(st, O.mis_greedy(g).astype(np.int32))
    assert O.mis_verify(g, st) == 0


def test_rmat22_mis_all_styles(gpu, rmat22):
    g, dg, _, refs = rmat22
    ref = _ref(refs, "mis", lambda: O.mis_greedy(g).astype(np.int32))
    for st in gpu.styles("mis"):
        _say("rmat22 mis", st)
        assert np.array_equal(gpu.run("mis", dg, st).out, ref), st


# ------------------------------------------------ R-MAT 20 and grid: PR, TC
@pytest.fixture(scope="module", params=["rmat20", "grid2048"])
def prtc(gpu, request):
    _say("generating", request.param)
    if request.param == "rmat20":
        g = G.rmat(20, 16, seed=1)
    else:
        g = G.grid2d(2048, 0.05, seed=1)
    dg = gpu.DeviceGraph(g)
    yield g, dg
    dg.free()
    gpu.lib().igx_release_workspace()


def test_pr_full_size_all_styles(gpu, prtc):
    g, dg = prtc
    _, it_ref, _ = O.pr_fp64(g, PR_EPS, PR_MAX_ITER)
    gpu.set_pr_params(PR_EPS, PR_MAX_ITER)
    try:
        for st in gpu.styles("pr"):
            _say("pr", g.name, st)
            r = gpu.run("pr", dg, st)
            ref, _, _ = O.pr_fp64(g, PR_EPS, PR_MAX_ITER, fixed_iters=r.iterations)
            rel = float(np.max(np.abs(r.out.astype(np.float64) - ref) / ref))
            assert rel <= PR_RTOL, (g.name, st, rel)
            assert abs(r.iterations - it_ref) <= 1, (g.name, st, r.iterations, it_ref)
    finally:
        gpu.set_pr_params(1e-4, 100)


def test_tc_full_size(gpu, prtc):
    g, dg = prtc
    want = O.tc_forward(g)
    deg = g.degrees().astype(np.float64)
    maxdeg = int(deg.max())
    for st in gpu.styles("tc"):
        if autotune.too_slow("tc", st, maxdeg, float(np.sum(deg * deg))):
            continue                          # unoriented merges on a 64k-degree hub: quadratic
        _say("tc", g.name, st)
        assert gpu.run("tc", dg, st).out == want, (g.name, st)
    c, _ = gpu.tc_vertex(dg)
    assert c == want


# -------------------------------------------------------- 512-graph batch
@pytest.fixture(scope="module")
def batch512(gpu):
    _say("generating the 512-graph batch")
    graphs = BT.make_batch_graphs(512)
    db = BT.DeviceBatch(graphs)
    _say("oracle for 512 graphs")
    refs = [(O.bfs_serial(g, 0), O.cc_omp_pull(g)[0], O.mis_greedy(g).astype(np.int32)) for g in graphs]
    yield db, refs
    db.free()


def _check_all(db, refs, tag):
    for algo, k in (("bfs", 0), ("cc", 1), ("mis", 2)):
        outs = db.results(algo)
        bad = [i for i, (got, ref) in enumerate(zip(outs, refs)) if not np.array_equal(got, ref[k])]
        assert not bad, (tag, algo, bad[:10])


@pytest.mark.parametrize("use_graph,block", [(False, 256), (True, 1024)], ids=["eager256", "hipgraph1024"])
def test_batch512_every_graph(gpu, batch512, use_graph, block):
    db, refs = batch512
    assert len(db.graphs) == 512
    for rep in range(2 if use_graph else 1):
        for buf in db.outs.values():
            buf.from_host(np.full(db.total_nodes, -7, np.int32))
        _say("batch512", "hipgraph" if use_graph else "eager", "rep", rep)
        db.run(7, block, use_graph=use_graph)
        _check_all(db, refs, f"{'graph' if use_graph else 'eager'}{rep}")
