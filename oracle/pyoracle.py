"""ctypes bindings of the CPU oracle (oracle/libitsca_oracle.so).

TEST INFRASTRUCTURE ONLY: imported by the tests, the smoke check and the
CPU-baseline leg of the benchmark, never by the product.
"""
from __future__ import annotations

import ctypes
import os
import tempfile
from typing import Sequence

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
_lib = None


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        path = os.path.join(HERE, "libitsca_oracle.so")
        if not os.path.exists(path):
            raise RuntimeError(f"{path} is missing: run make -C oracle")
        L = ctypes.CDLL(path)
        P, I = ctypes.c_void_p, ctypes.c_int
        F, I32 = ctypes.POINTER(ctypes.c_float), ctypes.POINTER(ctypes.c_int32)
        L.itsca_oracle_new.restype = P
        L.itsca_oracle_free.argtypes = [P]
        L.itsca_oracle_run.argtypes = [P, F, F, F, I32, I32, I32, F, I]
        L.itsca_oracle_dump.argtypes = [P, ctypes.c_char_p]
        L.itsca_oracle_dump.restype = ctypes.c_long
        L.itsca_oracle_count.argtypes = [P, I, I]
        L.itsca_oracle_count.restype = ctypes.c_long
        L.itsca_oracle_copy.argtypes = [P, I, I, P]
        L.itsca_oracle_copy.restype = ctypes.c_long
        L.itsca_oracle_time.argtypes = [F, F, F, I32, I32, I32, F, I, I, I]
        L.itsca_oracle_time.restype = ctypes.c_double
        L.itsca_check_atan2f.argtypes = [ctypes.c_long, ctypes.c_uint64]
        L.itsca_check_atan2f.restype = ctypes.c_long
        _lib = L
    return _lib


def _fp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))


def _ip(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))


class OracleVertex:
    """Runs the CPU restatement for one vertex of one event."""

    def __init__(self, event, vertex_index: int = 0, stable: bool = True, exclusive_bound: bool = False):
        """exclusive_bound (not the reference): no extra candidate past each
        window row, so every tracklet the search finds appears once."""
        L = lib()
        self.h = L.itsca_oracle_new()
        self._keep = [np.ascontiguousarray(a, dtype=np.float32) for a in (event.x, event.y, event.z)]
        self._ids = np.ascontiguousarray(event.cluster_id, dtype=np.int32)
        self._mc = np.ascontiguousarray(event.mc, dtype=np.int32)
        self._offs = np.ascontiguousarray(event.layer_offsets, dtype=np.int32)
        v = np.ascontiguousarray(event.vertices[vertex_index], dtype=np.float32)
        st = L.itsca_oracle_run(self.h, _fp(self._keep[0]), _fp(self._keep[1]), _fp(self._keep[2]),
                                _ip(self._ids), _ip(self._mc), _ip(self._offs), _fp(v),
                                (1 if stable else 0) | (2 if exclusive_bound else 0))
        if st != 0:
            raise RuntimeError("oracle run failed")

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.itsca_oracle_free(self.h)
            self.h = None

    def dump(self) -> str:
        with tempfile.NamedTemporaryFile("r", suffix=".txt") as f:
            if lib().itsca_oracle_dump(self.h, f.name.encode()) < 0:
                raise RuntimeError("oracle dump failed")
            return open(f.name).read()

    def count(self, what: int, layer: int = 0) -> int:
        return int(lib().itsca_oracle_count(self.h, what, layer))

    def records(self, what: int, layer: int = 0) -> np.ndarray:
        fields = {0: 8, 1: 4, 2: 10, 3: 8}[what]
        n = self.count(what, layer)
        buf = np.zeros((max(n, 1), fields), dtype=np.int32)
        lib().itsca_oracle_copy(self.h, what, layer, buf.ctypes.data)
        return buf[:n]


def time_events(events: Sequence, repeats: int, threads: int) -> float:
    """Wall-clock seconds for `repeats` passes over the events (first vertex of
    each) on `threads` threads, one independent tracker state per thread."""
    xs, ys, zs, ids, mcs, offs, vs = [], [], [], [], [], [], []
    base = 0
    for ev in events:
        xs.append(ev.x), ys.append(ev.y), zs.append(ev.z), ids.append(ev.cluster_id), mcs.append(ev.mc)
        offs.append(np.asarray(ev.layer_offsets, dtype=np.int64) + base)
        vs.append(ev.vertices[0])
        base += len(ev.x)
    x = np.ascontiguousarray(np.concatenate(xs), dtype=np.float32)
    y = np.ascontiguousarray(np.concatenate(ys), dtype=np.float32)
    z = np.ascontiguousarray(np.concatenate(zs), dtype=np.float32)
    i = np.ascontiguousarray(np.concatenate(ids), dtype=np.int32)
    m = np.ascontiguousarray(np.concatenate(mcs), dtype=np.int32)
    o = np.ascontiguousarray(np.concatenate(offs), dtype=np.int32)
    v = np.ascontiguousarray(np.stack(vs), dtype=np.float32)
    return float(lib().itsca_oracle_time(_fp(x), _fp(y), _fp(z), _ip(i), _ip(m), _ip(o), _fp(v), len(events),
                                         int(repeats), int(threads)))


def check_atan2f(n: int, seed: int) -> int:
    """Number of inputs (of n) where the kernels' atan2f differs from glibc's."""
    return int(lib().itsca_check_atan2f(int(n), int(seed)))
