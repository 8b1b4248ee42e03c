"""ctypes bindings of the two native libraries.

``libindigo_graph.so`` (host graph substrate, include/igx_graph.h) and
``libindigo_hip.so`` (HIP kernels and C ABI, include/indigo_hip.h) are loaded
from this package directory.  Loading fails loudly when a library is missing:
there is no Python or CPU fallback for the GPU path.
"""
from __future__ import annotations

import ctypes
import os

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
GRAPH_LIB_PATH = os.path.join(PKG_DIR, "libindigo_graph.so")
HIP_LIB_PATH = os.path.join(PKG_DIR, "libindigo_hip.so")

c_int_p = ctypes.POINTER(ctypes.c_int)


class IgxGraph(ctypes.Structure):
    """Layout of ECLgraph (include/ECLgraph.h:44-50) / igx_graph."""
    _fields_ = [
        ("nodes", ctypes.c_int),
        ("edges", ctypes.c_int),
        ("nindex", c_int_p),
        ("nlist", c_int_p),
        ("eweight", c_int_p),
    ]


class IgxRunInfo(ctypes.Structure):
    _fields_ = [
        ("seconds", ctypes.c_double),
        ("iterations", ctypes.c_int),
        ("launches", ctypes.c_int),
        ("count", ctypes.c_ulonglong),
        ("residual", ctypes.c_double),
        ("edges_processed", ctypes.c_ulonglong),
        ("vertices_processed", ctypes.c_ulonglong),
        ("edges_relaxed", ctypes.c_ulonglong),
    ]


# status codes (include/igx_graph.h)
IGX_OK = 0
STATUS_NAMES = {
    0: "IGX_OK", 1: "IGX_ERR_ARG", 2: "IGX_ERR_IO", 3: "IGX_ERR_FORMAT", 4: "IGX_ERR_NOMEM",
    5: "IGX_ERR_HIP", 6: "IGX_ERR_UNSORTED", 7: "IGX_ERR_NO_WEIGHTS", 8: "IGX_ERR_STYLE",
    9: "IGX_ERR_NODEV", 10: "IGX_ERR_LIMIT", 11: "IGX_ERR_TIMEOUT",
}
IGX_ERR_TIMEOUT = 11


class IgxError(RuntimeError):
    def __init__(self, status: int, what: str, detail: str = ""):
        self.status = status
        msg = f"{what} failed with {STATUS_NAMES.get(status, status)}"
        if detail:
            msg += f" ({detail})"
        super().__init__(msg)


_graph_lib = None
_hip_lib = None


def _sig(lib, name, restype, argtypes):
    fn = getattr(lib, name)
    fn.restype = restype
    fn.argtypes = argtypes
    return fn


def graph_lib():
    global _graph_lib
    if _graph_lib is None:
        if not os.path.exists(GRAPH_LIB_PATH):
            raise ImportError(f"{GRAPH_LIB_PATH} is missing: run __graft_entry__.build() first")
        lib = ctypes.CDLL(GRAPH_LIB_PATH)
        G = ctypes.POINTER(IgxGraph)
        u64 = ctypes.c_uint64
        _sig(lib, "igx_egr_read", ctypes.c_int, [ctypes.c_char_p, G])
        _sig(lib, "igx_egr_write", ctypes.c_int, [G, ctypes.c_char_p])
        _sig(lib, "igx_host_graph_free", None, [G])
        _sig(lib, "igx_check_csr", ctypes.c_int, [G])
        _sig(lib, "igx_check_sorted", ctypes.c_int, [G])
        _sig(lib, "igx_gen_er", ctypes.c_int, [ctypes.c_int, ctypes.c_double, u64, G])
        _sig(lib, "igx_gen_rmat", ctypes.c_int, [ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                                 ctypes.c_double, ctypes.c_int, u64, G])
        _sig(lib, "igx_gen_grid", ctypes.c_int, [ctypes.c_int, ctypes.c_double, u64, G])
        _sig(lib, "igx_gen_chunglu", ctypes.c_int, [ctypes.c_int, ctypes.c_double, ctypes.c_double, u64, G])
        _sig(lib, "igx_graph_from_edges", ctypes.c_int, [ctypes.c_int, ctypes.c_int64, c_int_p, c_int_p, G])
        _sig(lib, "igx_gen_weights", ctypes.c_int, [G, ctypes.c_int, ctypes.c_int, u64])
        _sig(lib, "igx_ref_weights", ctypes.c_int, [G])
        _sig(lib, "igx_build_sp_host", ctypes.c_int, [G, c_int_p])
        _sig(lib, "igx_max_degree_vertex", ctypes.c_int, [G])
        _sig(lib, "igx_graph_features", ctypes.c_int, [G, u64, ctypes.POINTER(ctypes.c_double)])
        _sig(lib, "igx_feature_draws", ctypes.c_int, [u64, u64, ctypes.c_int, ctypes.POINTER(u64)])
        _graph_lib = lib
    return _graph_lib


def hip_lib():
    """Load libindigo_hip.so (the product path).  Raises if it is missing."""
    global _hip_lib
    if _hip_lib is None:
        if not os.path.exists(HIP_LIB_PATH):
            raise ImportError(f"{HIP_LIB_PATH} is missing: run __graft_entry__.build() first "
                              "(the HIP path has no fallback)")
        graph_lib()  # dependency, resolved through the rpath as well
        lib = ctypes.CDLL(HIP_LIB_PATH)
        G = ctypes.POINTER(IgxGraph)
        vp = ctypes.c_void_p
        dp = ctypes.POINTER(ctypes.c_double)
        ip = c_int_p
        up = ctypes.POINTER(ctypes.c_uint)
        ullp = ctypes.POINTER(ctypes.c_ulonglong)
        _sig(lib, "igx_init", ctypes.c_int, [ctypes.c_int])
        _sig(lib, "igx_device_info", ctypes.c_int, [ctypes.c_char_p, ctypes.c_int, ip, ip])
        _sig(lib, "igx_sync", ctypes.c_int, [])
        _sig(lib, "igx_release_workspace", ctypes.c_int, [])
        _sig(lib, "igx_shutdown", ctypes.c_int, [])
        _sig(lib, "igx_last_error", ctypes.c_char_p, [])
        _sig(lib, "igx_status_string", ctypes.c_char_p, [ctypes.c_int])
        _sig(lib, "igx_alloc", ctypes.c_int, [ctypes.POINTER(vp), ctypes.c_size_t])
        _sig(lib, "igx_free", ctypes.c_int, [vp])
        _sig(lib, "igx_memcpy_h2d", ctypes.c_int, [vp, vp, ctypes.c_size_t])
        _sig(lib, "igx_memcpy_d2h", ctypes.c_int, [vp, vp, ctypes.c_size_t])
        _sig(lib, "igx_memset", ctypes.c_int, [vp, ctypes.c_int, ctypes.c_size_t])
        _sig(lib, "igx_memcpy_d2d", ctypes.c_int, [vp, vp, ctypes.c_size_t])
        _sig(lib, "igx_graph_upload", ctypes.c_int, [G, G])
        _sig(lib, "igx_graph_free_device", ctypes.c_int, [G])
        _sig(lib, "igx_build_sp", ctypes.c_int, [G, ctypes.POINTER(ip)])
        _sig(lib, "igx_num_styles", ctypes.c_int, [ctypes.c_int])
        _sig(lib, "igx_style_name", ctypes.c_char_p, [ctypes.c_int, ctypes.c_int])
        _sig(lib, "igx_style_edge_based", ctypes.c_int, [ctypes.c_int, ctypes.c_int])
        _sig(lib, "igx_get_entry_style", ctypes.c_int, [ctypes.c_int, ctypes.c_int])
        _sig(lib, "igx_set_entry_style", ctypes.c_int, [ctypes.c_int, ctypes.c_int, ctypes.c_int])
        _sig(lib, "igx_run", ctypes.c_int, [ctypes.c_int, ctypes.c_int, G, ctypes.c_int, vp, ip,
                                            ctypes.POINTER(IgxRunInfo)])
        _sig(lib, "igx_set_pr_params", ctypes.c_int, [ctypes.c_double, ctypes.c_int])
        _sig(lib, "igx_set_time_limit", ctypes.c_int, [ctypes.c_double])
        _sig(lib, "igx_profile_kernel", ctypes.c_int, [ctypes.c_char_p])
        _sig(lib, "igx_profile_read", ctypes.c_int, [dp, ip])
        _sig(lib, "igx_bfs_vertex", ctypes.c_int, [ctypes.c_int, G, ip, dp])
        _sig(lib, "igx_bfs_edge", ctypes.c_int, [ctypes.c_int, G, ip, ip, dp])
        _sig(lib, "igx_sssp_vertex", ctypes.c_int, [ctypes.c_int, G, ip, dp])
        _sig(lib, "igx_sssp_edge", ctypes.c_int, [ctypes.c_int, G, ip, ip, dp])
        _sig(lib, "igx_cc_vertex", ctypes.c_int, [G, ip, dp])
        _sig(lib, "igx_cc_edge", ctypes.c_int, [G, ip, ip, dp])
        _sig(lib, "igx_mis_vertex", ctypes.c_int, [G, up, ip, dp])
        _sig(lib, "igx_mis_edge", ctypes.c_int, [G, ip, up, ip, dp])
        _sig(lib, "igx_pr", ctypes.c_int, [G, ctypes.POINTER(ctypes.c_float), ip, dp, ip])
        _sig(lib, "igx_tc_vertex", ctypes.c_int, [ullp, ctypes.c_int, ip, ip, dp])
        _sig(lib, "igx_tc_edge", ctypes.c_int, [ullp, ctypes.c_int, ip, ip, ip, dp])
        _sig(lib, "igx_graph_features_device", ctypes.c_int, [G, ctypes.c_uint64, dp, dp])
        _hip_lib = lib
    return _hip_lib


def check(status: int, what: str) -> None:
    if status != IGX_OK:
        detail = ""
        if _hip_lib is not None:
            err = _hip_lib.igx_last_error()
            detail = err.decode() if err else ""
        raise IgxError(status, what, detail)
