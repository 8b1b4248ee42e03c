"""Shipped selector models and timing datasets: their style lists match the
library's registry (CPU), and on the two held-out sets (graphs of 2^14 to
2^22 vertices; graphs of 2^23 vertices with average degree 16) the trained
picks stay within a few percent of the exhaustive best and, on average, no
slower than the default style (GPU: the MLP runs on the device)."""
import os

import numpy as np
import pytest

from indigo2suite_amd import device as D
from indigo2suite_amd import selector as S

ALGOS = ("bfs", "sssp", "cc", "mis", "pr", "tc")
MD = S.MODEL_DIR


def test_models_and_datasets_match_registry():
    for algo in ALGOS:
        m = np.load(os.path.join(MD, f"selector_{algo}.npz"))
        # the models name their styles; a pick is looked up by name, so a
        # model may know fewer styles than the library registers
        names = [str(x) for x in m["styles"]]
        reg = D.styles(algo)
        assert set(names) <= set(reg) and names == sorted(names, key=reg.index)
        assert m["mean"].shape == (16,) and np.all(m["std"] > 0)
    for name in sorted(os.listdir(MD)):
        if not name.startswith(("train_", "test_")):
            continue
        d = np.load(os.path.join(MD, name))
        n = d["features"].shape[0]
        assert d["features"].shape == (n, 16) and n > 0
        for algo in ALGOS:
            t = d[f"times_{algo}"]
            names = [str(x) for x in d[f"styles_{algo}"]]
            assert t.shape == (n, len(names)) and set(names) <= set(D.styles(algo))
            # every graph has at least one measured style
            assert np.all(np.any(np.isfinite(t) & (t > 0), axis=1))


# held-out sets: (file, minimum geometric-mean speedup over the default style)
HELD_OUT = (("test_2100000.npz", 1.0),   # 152 graphs, 2^14 to 2^22 vertices, degree U[4, 32]
            ("test_2000000.npz", 0.98))  # 50 graphs at 2^23 vertices, degree 16: the defaults are the tuned best


@pytest.mark.gpu
@pytest.mark.parametrize("name,min_speedup", HELD_OUT)
def test_shipped_models_on_held_out_graphs(gpu, name, min_speedup):
    d = np.load(os.path.join(MD, name))
    for algo in ALGOS:
        sel = S.Selector.load(os.path.join(MD, f"selector_{algo}.npz"))
        try:
            dflt = D.default_entry_style(algo)
            ev = sel.evaluate(d["features"], d[f"times_{algo}"], dflt)
            print(name, algo, {k: v for k, v in ev.items() if k != "picks"}, flush=True)
            assert ev["geomean_speedup_vs_default"] >= min_speedup, (algo, ev)
            assert ev["frac_within_10pct"] >= 0.85, (algo, ev)
            assert ev["mean_regret"] <= 0.05, (algo, ev)
        finally:
            sel.net.free()


@pytest.mark.gpu
def test_pick_style_on_a_graph(gpu):
    from indigo2suite_amd.autotune import make_graph
    g = make_graph("rmat", 12, 8, 5, np.random.default_rng(5))
    for algo in ALGOS:
        st = S.pick_style(g, algo)
        assert st in D.styles(algo)
    for sel in S._loaded.values():
        sel.net.free()
    S._loaded.clear()
