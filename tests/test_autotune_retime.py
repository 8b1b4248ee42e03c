"""autotune.retime on the CPU, with the device calls stubbed: it regenerates
the dataset's graphs from their seeds, checks their sizes, replaces the
columns of the styles it re-times, appends a column for a style registered
after the dataset was made, and refuses a dataset from other seeds."""
import numpy as np
import pytest

from indigo2suite_amd import autotune as A
from indigo2suite_amd import device as D


class _FakeDeviceGraph:
    def __init__(self, g):
        self.nodes = g.nodes

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        return False


def _dataset(path, count, seed_base, styles):
    specs = [A.graph_spec(i, seed_base, (10, 11), (4, 6)) for i in range(count)]
    graphs = [A.build_from_spec(s) for s in specs]
    np.savez_compressed(path, features=np.zeros((count, 16)), times_sssp=np.ones((count, len(styles))),
                        styles_sssp=np.array(styles), spec_n=np.array([g.nodes for g in graphs]),
                        spec_m=np.array([g.edges for g in graphs]),
                        spec_family=np.array([s["family"] for s in specs]))


@pytest.fixture
def stubbed(monkeypatch):
    reg = ["a", "b", "c"]
    monkeypatch.setattr(D, "styles", lambda algo: list(reg))
    monkeypatch.setattr(D, "init", lambda *a: None)
    monkeypatch.setattr(D, "set_pr_params", lambda *a: None)
    monkeypatch.setattr(D, "DeviceGraph", _FakeDeviceGraph)
    calls = []

    def fake_sweep(dg, algo, src, repeats, styles=None, **kw):
        calls.append(list(styles))
        return {st: 0.001 * dg.nodes + i for i, st in enumerate(styles)}
    monkeypatch.setattr(A, "sweep", fake_sweep)
    return reg, calls


def test_retime_replaces_and_appends_columns(tmp_path, stubbed):
    reg, calls = stubbed
    src, out = str(tmp_path / "d.npz"), str(tmp_path / "e.npz")
    _dataset(src, 6, 4242, ["a", "b"])            # made before style "c" was registered
    assert A.retime(src, out, "sssp", ["b", "c"], 4242, (10, 11), (4, 6), log=None, prefetch=2) == 6
    d = np.load(out)
    assert [str(x) for x in d["styles_sssp"]] == ["a", "b", "c"]
    t = d["times_sssp"]
    n = d["spec_n"]
    assert np.array_equal(t[:, 0], np.ones(6))                 # untouched
    assert np.allclose(t[:, 1], 0.001 * n) and np.allclose(t[:, 2], 0.001 * n + 1)
    assert calls == [["b", "c"]] * 6


def test_retime_refuses_other_seeds(tmp_path, stubbed):
    src, out = str(tmp_path / "d.npz"), str(tmp_path / "e.npz")
    _dataset(src, 4, 4242, ["a", "b", "c"])
    with pytest.raises(ValueError):
        A.retime(src, out, "sssp", ["a"], 9999, (10, 11), (4, 6), log=None, prefetch=2)
