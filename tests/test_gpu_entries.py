"""The reference-signature entry points (include/indigo_hip.h, one per
include/*_cuda.h harness entry) through the C ABI, against the oracle."""
import numpy as np
import pytest

import oracle as O
from indigo2suite_amd import graph as G

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def setup(gpu):
    g = G.rmat(13, 16, seed=5).with_weights(1, 255, seed=5)
    dg = gpu.DeviceGraph(g)
    yield gpu, g, dg
    dg.free()


def test_bfs_entries(setup):
    dev, g, dg = setup
    ref = O.bfs_serial(g, 3)
    d, t = dev.bfs_vertex(3, dg)
    assert np.array_equal(d, ref) and t > 0
    d, t = dev.bfs_edge(3, dg, g.sp())
    assert np.array_equal(d, ref)


def test_sssp_entries(setup):
    dev, g, dg = setup
    src = g.max_degree_vertex()
    ref = O.sssp_serial(g, src)
    assert np.array_equal(dev.sssp_vertex(src, dg)[0], ref)
    assert np.array_equal(dev.sssp_edge(src, dg, g.sp())[0], ref)


def test_cc_entries(setup):
    dev, g, dg = setup
    ref = O.cc_omp_pull(g)[0]
    assert np.array_equal(dev.cc_vertex(dg)[0], ref)
    assert np.array_equal(dev.cc_edge(dg, g.sp())[0], ref)


def test_mis_entries(setup):
    dev, g, dg = setup
    ref = O.mis_greedy(g).astype(np.int32)
    st, pr, _ = dev.mis_vertex(dg)
    assert np.array_equal(st, ref)
    assert pr[:3].tolist() == [2241873249, 1358751764, 3624587312]
    sp = dev.build_sp(dg)
    try:
        st, _, _ = dev.mis_edge(dg, sp)
        assert np.array_equal(st, ref)
        assert np.array_equal(sp.to_host(np.int32, g.edges), g.sp())
        c, _ = dev.tc_edge(dg, sp)
        assert c == O.tc_serial(g)
    finally:
        sp.free()


def test_pr_entry(setup):
    dev, g, _ = setup
    s, t, it = dev.pr(g)  # reference defaults: eps 1e-4, 100 iterations
    ref, _, _ = O.pr_fp64(g, 1e-4, 100, fixed_iters=it)
    assert np.max(np.abs(s - ref) / ref) <= 1e-5
    assert O.pr_check(g, s) < 1e-4


def test_tc_entry(setup):
    dev, g, dg = setup
    c, _ = dev.tc_vertex(dg)
    assert c == O.tc_serial(g)
