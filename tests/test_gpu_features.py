"""Selector features computed on the device (igx_graph_features_device)
against the host computation (igx_graph_features, itself pinned by closed
forms in test_graph.py): every feature built from integer statistics, the
BFS sweeps and the sampled wedges and edges is identical; the Gini
coefficient and the top-1% share (sums taken in another order) agree to
1e-12 relative."""
import sys

import numpy as np
import pytest

from indigo2suite_amd import graph as G

pytestmark = pytest.mark.gpu

EXACT = [0, 1, 2, 3, 4, 6, 7, 8, 9, 10, 12, 13, 14, 15]
ROUNDED = [5, 11]


def _graphs():
    n = 3000
    star = G.from_edges(n, np.zeros(n - 1, np.int32), np.arange(1, n, dtype=np.int32), name="star")
    path = G.from_edges(n, np.arange(n - 1, dtype=np.int32), np.arange(1, n, dtype=np.int32), name="path")
    two = G.from_edges(12, np.array([0, 1, 5], np.int32), np.array([1, 2, 6], np.int32), name="pieces")
    lonely = G.from_edges(7, np.zeros(0, np.int32), np.zeros(0, np.int32), name="no-edges")
    return [star, path, two, lonely,
            G.rmat(14, 8, seed=3), G.rmat(17, 16, seed=4), G.grid2d(200, 0.05, seed=2),
            G.erdos_renyi(50000, 6, seed=5), G.chung_lu(40000, 12, 2.2, seed=6)]


@pytest.mark.parametrize("seed", [1, 77])
def test_device_features_match_host(gpu, seed):
    for g in _graphs():
        print("features:", g.name, g.nodes, g.edges, "seed", seed, flush=True)
        host = g.features(seed=seed)
        with gpu.DeviceGraph(g) as dg:
            dev, secs = gpu.features_device(dg, seed)
        assert secs > 0.0
        np.testing.assert_array_equal(dev[EXACT], host[EXACT], err_msg=g.name)
        np.testing.assert_allclose(dev[ROUNDED], host[ROUNDED], rtol=1e-12, atol=1e-15, err_msg=g.name)


def test_device_features_repeatable_and_fast(gpu):
    g = G.rmat(20, 16, seed=1)
    with gpu.DeviceGraph(g) as dg:
        a, _ = gpu.features_device(dg, 1)
        b, secs = gpu.features_device(dg, 1)
    np.testing.assert_array_equal(a, b)
    print("device features, R-MAT 20:", secs * 1e3, "ms", file=sys.stderr)
    assert secs < 0.5
