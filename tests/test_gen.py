"""Input tooling (indigo2suite_amd.gen): the generator and convert commands
write .egr files that read back as the library's own generators' graphs
(byte layout of ECLgraph.h:82-96: int32 nodes, edges, nindex, nlist
[, eweight]); conversions symmetrize, drop self loops and duplicates."""
import json
import os
import struct
import subprocess
import sys

import numpy as np
import pytest

from indigo2suite_amd import gen
from indigo2suite_amd import graph as G

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _cli(*args):
    return subprocess.run([sys.executable, "-m", "indigo2suite_amd.gen", *args], cwd=REPO, check=True,
                          stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True)


def _same(a, b):
    assert a.nodes == b.nodes and a.edges == b.edges
    assert np.array_equal(a.nindex, b.nindex) and np.array_equal(a.nlist, b.nlist)
    if a.eweight is None or b.eweight is None:
        assert a.eweight is None and b.eweight is None
    else:
        assert np.array_equal(a.eweight, b.eweight)


@pytest.mark.parametrize("cmd,args,ref", [
    ("er", ["--n", "3000", "--degree", "6", "--seed", "4"], lambda: G.erdos_renyi(3000, 6, seed=4)),
    ("rmat", ["--scale", "12", "--edge-factor", "8", "--seed", "2"], lambda: G.rmat(12, 8, seed=2)),
    ("grid", ["--side", "40", "--long", "0.05", "--seed", "3"], lambda: G.grid2d(40, 0.05, seed=3)),
    ("chunglu", ["--n", "2000", "--degree", "10", "--gamma", "2.3", "--seed", "5"],
     lambda: G.chung_lu(2000, 10, 2.3, seed=5)),
])
def test_generators_write_what_the_library_generates(tmp_path, cmd, args, ref):
    out = str(tmp_path / f"{cmd}.egr")
    _cli(cmd, *args, "-o", out)
    got = G.read_egr(out)
    _same(got, ref())
    assert got.eweight is None
    # file layout: header, offsets, neighbours and nothing else
    assert os.path.getsize(out) == 8 + 4 * (got.nodes + 1) + 4 * got.edges


def test_weights_options(tmp_path):
    out = str(tmp_path / "w.egr")
    _cli("rmat", "--scale", "10", "--seed", "7", "--weights", "1", "255", "-o", out)
    g = G.read_egr(out)
    _same(g, G.rmat(10, 16, seed=7).with_weights(1, 255, seed=7))
    assert g.eweight.min() >= 1 and g.eweight.max() <= 255
    # symmetric: w(u, v) == w(v, u)
    src = g.sp()
    key = {(int(u), int(v)): int(w) for u, v, w in zip(src, g.nlist, g.eweight)}
    assert all(key[(v, u)] == w for (u, v), w in key.items())
    assert os.path.getsize(out) == 8 + 4 * (g.nodes + 1) + 8 * g.edges
    out2 = str(tmp_path / "r.egr")
    _cli("er", "--n", "500", "--degree", "4", "--reference-weights", "-o", out2)
    g2 = G.read_egr(out2)
    assert np.array_equal(g2.eweight, G.erdos_renyi(500, 4, seed=1).with_reference_weights().eweight)


def test_convert_edge_list_and_matrix_market(tmp_path):
    txt = tmp_path / "edges.txt"
    txt.write_text("# a SNAP-style comment\n0 1\n1 0\n1 2\n2 2\n3 1 17\n\n0 1\n")
    out = str(tmp_path / "e.egr")
    _cli("convert", str(txt), "--nodes", "6", "-o", out)
    g = G.read_egr(out)
    assert g.nodes == 6
    assert g.nindex.tolist() == [0, 1, 4, 5, 6, 6, 6]
    assert g.nlist.tolist() == [1, 0, 2, 3, 1, 1]
    mtx = tmp_path / "m.mtx"
    mtx.write_text("%%MatrixMarket matrix coordinate pattern symmetric\n% comment\n4 4 3\n1 2\n2 3\n4 1\n")
    out2 = str(tmp_path / "m.egr")
    _cli("convert", str(mtx), "-o", out2)
    g2 = G.read_egr(out2)
    assert g2.nodes == 4 and g2.nlist.tolist() == [1, 3, 0, 2, 1, 0]
    with open(out2, "rb") as f:
        assert struct.unpack("<2i", f.read(8)) == (4, 6)


def test_convert_rejects_negative_ids(tmp_path):
    txt = tmp_path / "bad.txt"
    txt.write_text("0 -1\n")
    with pytest.raises(ValueError):
        gen.convert(str(txt))


def test_info_and_large_inputs(tmp_path, monkeypatch):
    out = str(tmp_path / "g.egr")
    _cli("grid", "--side", "8", "--long", "0", "-o", out)
    rec = json.loads(_cli("info", out).stdout)
    assert rec["nodes"] == 64 and rec["edges"] == 2 * 2 * 8 * 7 and rec["max_degree"] == 4
    assert rec["sorted_no_self_loops"] and not rec["weights"]
    d = tmp_path / "large"
    d.mkdir()
    G.rmat(8, 4, seed=1).write_egr(str(d / "rmat22.sym.egr"))
    monkeypatch.setenv("IGX_LARGE_INPUTS", str(d))
    assert gen.large_input_path("rmat22.sym.egr") == str(d / "rmat22.sym.egr")
    assert gen.large_input_path("coPapersDBLP.egr") is None
    assert gen.load_large_input("rmat22.sym.egr").nodes == 256
    with pytest.raises(FileNotFoundError):
        gen.load_large_input("USA-road-d.NY.egr")
    with pytest.raises(KeyError):
        gen.large_input_path("other.egr")
    lines = _cli("large-inputs", "--dir", str(d)).stdout.splitlines()
    assert len(lines) == 5 and sum("present" in ln for ln in lines) == 1
