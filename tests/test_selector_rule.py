"""The selector's decision rule on the CPU (network scores stubbed)."""
import numpy as np
import pytest

from indigo2suite_amd import selector as S


def test_margin_rule_without_the_network():
    """pick(): the default stays unless another style's predicted slowdown is
    lower by more than the margin (scores stubbed, no GPU needed)."""
    sel = S.Selector("sssp", ["a", "b", "c"])
    scores = np.array([[0.10, 0.05, 0.30], [0.50, 0.05, 0.40], [0.2, 0.2, 0.2]])
    sel.predict_scores = lambda f: scores
    feats = np.zeros((3, 16))
    assert sel.pick(feats) == [1, 1, 0]
    sel.default, sel.margin = "a", 0.1
    assert sel.pick(feats) == [0, 1, 0]
    sel.default = "c"
    assert sel.pick(feats) == [1, 1, 2]


def test_targets_from_times():
    """Targets are log(t / best) per graph, clipped at log(kClip); a style
    too slow to run (+inf) gets the clipped maximum; NaN (not measured) is
    masked out."""
    t = np.array([[1.0, 2.0, np.inf, np.nan],
                  [4.0, 1.0, 1e9, 2.0]])
    y, m = S.targets_from_times(t)
    clip = np.log(S.kClip)
    np.testing.assert_allclose(y[0], [0.0, np.log(2.0), clip, 0.0], rtol=1e-6)
    np.testing.assert_allclose(y[1], [np.log(4.0), 0.0, clip, np.log(2.0)], rtol=1e-6)
    assert m.tolist() == [[1, 1, 1, 0], [1, 1, 1, 1]]


def test_evaluate_with_stubbed_scores():
    """evaluate(): regret against the per-graph best, speedup over the default
    as a geometric mean, and a pick that could not be timed counted as the
    clipped slowdown."""
    sel = S.Selector("bfs", ["a", "b", "c"])
    t = np.array([[1.0, 2.0, 4.0],      # picks b: regret 1, speedup over a 0.5
                  [3.0, 1.0, 2.0],      # picks b: best, speedup 3
                  [2.0, np.inf, 1.0]])  # picks b: untimed -> clipped slowdown
    sel.predict_scores = lambda f: np.tile([0.5, 0.0, 0.9], (np.atleast_2d(f).shape[0], 1))
    ev = sel.evaluate(np.zeros((3, 16)), t, default_style=0)
    assert ev["picks"] == ["b", "b", "b"]
    clip = S.kClip
    np.testing.assert_allclose(ev["mean_regret"], (1.0 + 0.0 + (clip - 1.0)) / 3)
    np.testing.assert_allclose(ev["geomean_speedup_vs_default"], (0.5 * 3.0 * (2.0 / clip)) ** (1 / 3))
    assert ev["frac_within_10pct"] == pytest.approx(1 / 3)
    assert ev["frac_faster_than_default"] == pytest.approx(1 / 3)
