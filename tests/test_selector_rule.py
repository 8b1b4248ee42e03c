"""The selector's decision rule on the CPU (network scores stubbed)."""
import numpy as np

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
