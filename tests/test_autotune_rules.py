"""Autotuner rules that need no GPU: the sum of squared degrees and the
largest degree recovered from the host features (as the selector's pick
does), and which triangle-counting styles are skipped as too slow."""
import numpy as np
import pytest

from indigo2suite_amd import autotune as A
from indigo2suite_amd import graph as G


@pytest.mark.parametrize("make", [lambda: G.rmat(12, 8, seed=2), lambda: G.grid2d(32, 0.05, seed=1),
                                  lambda: G.chung_lu(3000, 10, 2.3, seed=4),
                                  lambda: G.from_edges(6, np.zeros(5, np.int32), np.arange(1, 6, dtype=np.int32))])
def test_degree_statistics_from_features(make):
    g = make()
    f = g.features(seed=1)
    deg = np.diff(g.nindex).astype(np.float64)
    assert A.sum_sq_degrees(f) == pytest.approx(float(np.sum(deg * deg)), rel=1e-9)
    assert int(round(2.0 ** f[4] - 1.0)) == int(deg.max())


def test_too_slow_rules():
    lim_d, lim_s = A.TC_MAX_DEGREE, A.TC_MAX_SUM_D2
    # only the unoriented TC styles are ever skipped
    for algo in ("bfs", "sssp", "cc", "mis", "pr"):
        assert not A.too_slow(algo, "v_merge_thread", 10 ** 6, 1e12)
    assert not A.too_slow("tc", "oriented_lds", 10 ** 6, 1e12)
    for st in A.UNORIENTED_TC:
        assert not A.too_slow("tc", st, 200, 1e6)
        assert A.too_slow("tc", st, lim_d + 1, 1e6)
        assert A.too_slow("tc", st, 200, lim_s * 1.01)
        assert not A.too_slow("tc", st, lim_d, lim_s) or st == "v_merge_thread"
    # the thread-per-vertex merge has its own, lower degree bound
    assert A.too_slow("tc", "v_merge_thread", A.TC_THREAD_MERGE_MAX_DEGREE + 1, 1e6)
    assert not A.too_slow("tc", "v_merge_thread", A.TC_THREAD_MERGE_MAX_DEGREE, 1e6)
    assert not A.too_slow("tc", "v_merge_wave", A.TC_THREAD_MERGE_MAX_DEGREE + 1, 1e6)
