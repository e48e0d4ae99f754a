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
