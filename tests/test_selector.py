"""Selector network: the HIP forward pass and gradients against an FP64 numpy
implementation of the same MLP, training reduces the loss, and a trained
selector beats the fixed default style on a synthetic timing problem."""
import numpy as np
import pytest

from indigo2suite_amd import selector as S

pytestmark = pytest.mark.gpu


def np_forward(p, x, n_in, n_out):
    W1, b1, W2, b2, W3, b3 = S.unpack(p, n_in, n_out)
    z1 = x @ W1 + b1
    h1 = np.maximum(z1, 0)
    z2 = h1 @ W2 + b2
    h2 = np.maximum(z2, 0)
    return z1, h1, z2, h2, h2 @ W3 + b3


def np_grads(p, x, t, m, norm, n_in, n_out):
    W1, b1, W2, b2, W3, b3 = S.unpack(p, n_in, n_out)
    z1, h1, z2, h2, y = np_forward(p, x, n_in, n_out)
    d3 = 2.0 * (y - t) * m / norm
    loss = float(np.sum(((y - t) * m) ** 2) / norm)
    gW3, gb3 = h2.T @ d3, d3.sum(0)
    d2 = (d3 @ W3.T) * (z2 > 0)
    gW2, gb2 = h1.T @ d2, d2.sum(0)
    d1 = (d2 @ W2.T) * (z1 > 0)
    gW1, gb1 = x.T @ d1, d1.sum(0)
    return np.concatenate([gW1.ravel(), gb1, gW2.ravel(), gb2, gW3.ravel(), gb3]), loss


@pytest.fixture(scope="module")
def dev(gpu):
    return gpu


def test_forward_and_gradients_match_fp64(dev):
    rng = np.random.default_rng(0)
    n_in, n_out, B = 16, 11, 37
    net = S.Mlp(n_in, n_out, seed=3)
    try:
        p = net.params().astype(np.float64)
        x = rng.normal(size=(B, n_in)).astype(np.float32)
        t = rng.normal(size=(B, n_out)).astype(np.float32)
        m = (rng.random((B, n_out)) < 0.8).astype(np.float32)
        y = net.forward(x)
        ref = np_forward(p, x.astype(np.float64), n_in, n_out)[4]
        assert np.max(np.abs(y - ref)) < 1e-4 * (1 + np.max(np.abs(ref)))
        g, loss = net.gradients(x, t, m)
        gref, lref = np_grads(p, x.astype(np.float64), t, m, float(m.sum()), n_in, n_out)
        assert abs(loss - lref) <= 1e-4 * abs(lref)
        scale = np.max(np.abs(gref))
        assert np.max(np.abs(g - gref)) <= 1e-4 * scale
        # a planted mistake in the reference (missing ReLU mask) must be detectable
        bad = gref.copy()
        bad[: n_in * 64] *= 1.5
        assert np.max(np.abs(g - bad)) > 1e-3 * scale
    finally:
        net.free()


def test_training_reduces_loss_and_selector_beats_default(dev):
    rng = np.random.default_rng(1)
    n, K = 600, 6
    f = rng.normal(size=(n, 16))
    # synthetic timings: style k is fastest when feature k is the largest of f[:, :K]
    base = np.exp(rng.normal(0, 0.05, size=(n, K)))
    times = base * (1.0 + 3.0 * (f[:, :K].max(axis=1, keepdims=True) - f[:, :K]))
    sel = S.Selector("bfs", [f"s{k}" for k in range(K)], seed=2)
    y, m = S.targets_from_times(times)
    first = None
    try:
        first = sel.fit(f[:400], times[:400], epochs=1, batch=32, lr=2e-3)
        last = sel.fit(f[:400], times[:400], epochs=150, batch=32, lr=2e-3)
        assert last < 0.5 * first
        ev = sel.evaluate(f[400:], times[400:], default_style=0)
        assert ev["geomean_speedup_vs_default"] > 1.2
        assert ev["mean_regret"] < 0.25
    finally:
        if sel.net:
            sel.net.free()


def test_margin_keeps_the_default_unless_the_gain_is_clear(dev):
    """Default style 0 is the best on most graphs by a hair (noise); styles 1-2
    win by 3x where feature 0 is large.  After fit_with_margin the picks are
    faster than the default overall and lose no more than the noise where the
    default is the best."""
    rng = np.random.default_rng(7)
    n, K = 800, 3
    f = rng.normal(size=(n, 16))
    times = np.exp(rng.normal(0, 0.03, size=(n, K)))
    times[:, 0] *= 0.99
    big = f[:, 0] > 1.0
    times[big, 0] *= 3.0
    sel = S.Selector("bfs", ["s0", "s1", "s2"], seed=3)
    try:
        res = sel.fit_with_margin(f[:600], times[:600], "s0", epochs=120, batch=32, lr=2e-3)
        print("margin fit:", res)
        ev = sel.evaluate(f[600:], times[600:], default_style=0)
        assert ev["geomean_speedup_vs_default"] > 1.05, (res, ev)
        # where the default is the best in expectation, the picks cost at most
        # the timing noise
        noise = ~big[600:]
        ev_noise = sel.evaluate(f[600:][noise], times[600:][noise], default_style=0)
        assert ev_noise["geomean_speedup_vs_default"] > 0.97, (res, ev_noise)
    finally:
        if sel.net:
            sel.net.free()

