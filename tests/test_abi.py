"""The C-ABI libraries load and export every function their headers declare;
without a GPU the product path fails loudly (no CPU fallback)."""
import ctypes
import os
import re

import pytest

from indigo2suite_amd import _native

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared(header):
    text = open(os.path.join(REPO, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(igx_\w+)\s*\(", text)))


def test_graph_lib_exports_header():
    lib = ctypes.CDLL(_native.GRAPH_LIB_PATH)
    names = declared("igx_graph.h")
    assert len(names) >= 15
    for name in names:
        assert hasattr(lib, name), name


def test_hip_lib_exports_header():
    if not os.path.exists(_native.HIP_LIB_PATH):
        pytest.fail("libindigo_hip.so not built")
    lib = ctypes.CDLL(_native.HIP_LIB_PATH)
    names = declared("indigo_hip.h")
    assert len(names) >= 30
    for name in names:
        assert hasattr(lib, name), name


def test_style_registry_without_gpu():
    lib = _native.hip_lib()
    counts = [lib.igx_num_styles(a) for a in range(6)]
    assert all(c >= 4 for c in counts)
    for a, c in enumerate(counts):
        names = [lib.igx_style_name(a, i).decode() for i in range(c)]
        assert len(set(names)) == c
    assert lib.igx_style_name(0, 999) is None
    assert lib.igx_num_styles(17) == 0


def test_no_gpu_means_error_not_fallback():
    try:
        import torch  # noqa: F401  (only to see whether a GPU exists here)
        has_gpu = torch.cuda.is_available()
    except Exception:
        has_gpu = False
    if has_gpu:
        pytest.skip("a GPU is present")
    lib = _native.hip_lib()
    assert lib.igx_init(0) == 9  # IGX_ERR_NODEV
