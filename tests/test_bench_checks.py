"""The output checks bench.py applies to the timed step at full size
(verify_bfs / verify_sssp / verify_cc, size-independent characterisations)
accept the oracle's results and reject planted mistakes (CPU)."""
import os
import sys

import numpy as np
import pytest

import oracle as O
from indigo2suite_amd import graph as G

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
import bench  # noqa: E402

INF = 2147483647


@pytest.fixture(scope="module")
def graphs():
    out = [G.rmat(10, 8, seed=3).with_weights(1, 255, seed=3),
           G.grid2d(16, 0.05, seed=2).with_weights(1, 9, seed=2)]
    # two components plus isolated vertices
    e = np.array([[0, 1], [1, 2], [4, 5]], dtype=np.int32)
    out.append(G.from_edges(8, e[:, 0], e[:, 1], name="pieces").with_weights(1, 5, seed=1))
    return out


def test_bfs_check(graphs):
    for g in graphs:
        src = g.max_degree_vertex()
        d = O.bfs_serial(g, src).astype(np.int32)
        bench.verify_bfs(g, src, d)
        reached = np.nonzero((d != INF) & (np.arange(g.nodes) != src))[0]
        if reached.size == 0:
            continue
        far = reached[np.argmax(d[reached])]
        bad = d.copy()
        bad[far] += 1                      # a level one too deep: no parent one level up
        with pytest.raises(AssertionError):
            bench.verify_bfs(g, src, bad)
        bad = d.copy()
        bad[reached[0]] = INF              # a reached vertex reported unreached
        with pytest.raises(AssertionError):
            bench.verify_bfs(g, src, bad)


def test_sssp_check(graphs):
    for g in graphs:
        src = g.max_degree_vertex()
        d = O.sssp_serial(g, src).astype(np.int32)
        bench.verify_sssp(g, src, d)
        reached = np.nonzero((d != INF) & (np.arange(g.nodes) != src))[0]
        if reached.size == 0:
            continue
        v = reached[np.argmax(d[reached])]
        bad = d.copy()
        bad[v] += 1                        # too long: no tight incoming edge
        with pytest.raises(AssertionError):
            bench.verify_sssp(g, src, bad)
        bad = d.copy()
        bad[v] -= 1                        # too short: no tight edge either
        with pytest.raises(AssertionError):
            bench.verify_sssp(g, src, bad)


def test_cc_check(graphs):
    for g in graphs:
        lab = O.cc_omp_pull(g)[0].astype(np.int32)
        bench.verify_cc(g, lab)
        comps = np.unique(lab)
        if comps.size < 2:
            continue
        bad = lab.copy()
        bad[lab == comps[1]] = comps[0]    # two components merged into one label
        with pytest.raises(AssertionError):
            bench.verify_cc(g, bad)
        v = int(np.nonzero(lab == comps[-1])[0][-1])
        bad = lab.copy()
        bad[v] = v if v != lab[v] else (v + 1) % g.nodes   # one vertex split off its component
        with pytest.raises(AssertionError):
            bench.verify_cc(g, bad)
