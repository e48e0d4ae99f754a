"""The C ABI libraries load and export every symbol their headers declare;
failure paths that need no GPU."""
import ctypes
import os
import re

import pytest

import itsca

ROOT = os.path.abspath(os.path.join(os.path.dirname(__file__), ".."))


def declared(header):
    text = open(os.path.join(ROOT, "include", header)).read()
    return sorted(set(re.findall(r"\b(itsca_[a-z_0-9]+)\s*\(", text)))


@pytest.mark.parametrize("header,lib", [("itsca.h", "libitsca_hip.so"), ("itsca_host.h", "libitsca.so")])
def test_library_exports_every_declared_symbol(header, lib):
    so = ctypes.CDLL(os.path.join(itsca.LIB_DIR, lib))
    names = declared(header)
    assert names
    missing = [n for n in names if not hasattr(so, n)]
    assert not missing


def test_binding_lists_match_headers():
    assert sorted(itsca.HIP_SYMBOLS) == declared("itsca.h")
    assert sorted(itsca.HOST_SYMBOLS) == declared("itsca_host.h")


def test_create_fails_loudly_without_a_device():
    try:
        import torch
        if torch.cuda.device_count() > 0:
            pytest.skip("a GPU is present")
    except ImportError:
        pass
    with pytest.raises(itsca.ItscaError):
        itsca.GpuTracker(device=0)


def test_bad_device_is_an_argument_error():
    with pytest.raises(itsca.ItscaError):
        itsca.GpuTracker(device=4096)


def test_native_weights_loader_matches_python():
    """The C++ API reads the learned filter's weights file (itsca_filter_load)
    to exactly the values the Python loader gives."""
    import ctypes

    import numpy as np

    from itsca import filter_training as ft

    params, thr, _ = ft.load_filter()
    out = np.zeros(itsca.MLP_PARAMS, dtype=np.float32)
    t = ctypes.c_float(0)
    lib = itsca.host_lib()
    rc = lib.itsca_filter_load(ft.DEFAULT_WEIGHTS.encode(), out.ctypes.data_as(ctypes.POINTER(ctypes.c_float)),
                               itsca.MLP_PARAMS, ctypes.byref(t))
    assert rc == 0
    assert np.array_equal(out.view(np.uint32), params.view(np.uint32))
    assert t.value == np.float32(thr)
    # wrong parameter count and missing file are errors
    assert lib.itsca_filter_load(ft.DEFAULT_WEIGHTS.encode(), out.ctypes.data_as(ctypes.POINTER(ctypes.c_float)),
                                 100, ctypes.byref(t)) == -1
    assert lib.itsca_filter_load(b"/nonexistent/w.json", out.ctypes.data_as(ctypes.POINTER(ctypes.c_float)),
                                 itsca.MLP_PARAMS, ctypes.byref(t)) == -1
    # the per-layer phi gates: the same float32 values; a file without them gives 0.3 (no gate)
    gates = np.zeros(itsca.TRACKLET_LAYERS, dtype=np.float32)
    gp = gates.ctypes.data_as(ctypes.POINTER(ctypes.c_float))
    assert lib.itsca_filter_load_gates(ft.DEFAULT_WEIGHTS.encode(), gp, itsca.TRACKLET_LAYERS) == 0
    want = ft.load_gates()
    assert want is not None and np.array_equal(gates, np.array(want, dtype=np.float32))
    assert all(0.0 < g <= 0.3 for g in want)
    import json
    import tempfile
    with tempfile.NamedTemporaryFile("w", suffix=".json", delete=False) as f:
        doc = json.load(open(ft.DEFAULT_WEIGHTS))
        doc.pop("phi_gates")
        json.dump(doc, f)
    assert lib.itsca_filter_load_gates(f.name.encode(), gp, itsca.TRACKLET_LAYERS) == 0
    assert np.array_equal(gates, np.full(itsca.TRACKLET_LAYERS, 0.3, dtype=np.float32))
    assert ft.load_gates(f.name) is None
    os.unlink(f.name)
