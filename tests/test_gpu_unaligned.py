"""Device graphs whose nlist / eweight are views at arbitrary element offsets
into larger allocations (as a batch arena or a caller's own buffers give, not
aligned to more than 4 bytes): BFS, SSSP and CC, whose kernels read ids and
weights through chunk lists and streamed slices, must give the oracle's
results there as on freshly allocated arrays."""
import ctypes

import numpy as np
import pytest

import oracle as O
from indigo2suite_amd import _native
from indigo2suite_amd import device as D
from indigo2suite_amd import graph as G

pytestmark = pytest.mark.gpu


class OffsetDeviceGraph:
    """nindex at offset 0, nlist at element offset `off_l`, eweight at `off_w`."""

    def __init__(self, g, off_l, off_w):
        self.nodes, self.edges = g.nodes, g.edges
        self.bufs = []

        def put(a, off):
            a = np.ascontiguousarray(a, np.int32)
            buf = D.DeviceBuffer(4 * (a.size + off + 1))
            self.bufs.append(buf)
            ptr = buf.ptr.value + 4 * off
            if a.size:
                _native.check(D.lib().igx_memcpy_h2d(ctypes.c_void_p(ptr), a.ctypes.data_as(ctypes.c_void_p),
                                                     a.nbytes), "igx_memcpy_h2d")
            return ctypes.cast(ctypes.c_void_p(ptr), _native.c_int_p)

        self.c = _native.IgxGraph()
        self.c.nodes, self.c.edges = g.nodes, g.edges
        self.c.nindex = put(g.nindex, 0)
        self.c.nlist = put(g.nlist, off_l)
        self.c.eweight = put(g.eweight, off_w)

    def free(self):
        for b in self.bufs:
            b.free()


GRAPHS = [
    G.rmat(14, 16, seed=3).with_weights(1, 255, seed=3),
    G.from_edges(3000, np.zeros(2900, np.int32), np.arange(1, 2901), name="star2900").with_weights(1, 9, seed=4),
    G.chung_lu(20000, 12, 2.1, seed=5).with_weights(1, 255, seed=5),
]


@pytest.mark.parametrize("offs", [(1, 2), (3, 0), (0, 1), (0, 0), (4, 8)])
@pytest.mark.parametrize("gi", range(len(GRAPHS)))
def test_offset_arrays(gpu, gi, offs):
    g = GRAPHS[gi]
    src = g.max_degree_vertex()
    dg = OffsetDeviceGraph(g, *offs)
    try:
        print("offsets", offs, g.name, flush=True)
        assert np.array_equal(gpu.run("sssp", dg, "v_do_balanced", src).out, O.sssp_serial(g, src))
        assert np.array_equal(gpu.run("bfs", dg, "do_hybrid", src).out, O.bfs_serial(g, src))
        assert np.array_equal(gpu.run("cc", dg, "uf_afforest").out, O.cc_omp_pull(g)[0])
    finally:
        dg.free()
