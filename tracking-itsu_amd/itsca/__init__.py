"""Python bindings (ctypes) of the gfx950 ITS cellular-automaton track finder.

The product is the C ABI of ``lib/libitsca_hip.so`` (``include/itsca.h``) and the
event API of ``lib/libitsca.so`` (``include/itsca_host.h``).  This module only
loads them, moves numpy arrays across the boundary and renders results in the
canonical text form used for bit-exact comparisons.  There is no fallback: if
the native libraries are missing, every entry point raises.
"""
from __future__ import annotations

import ctypes
import os
from dataclasses import dataclass
from typing import List, Optional, Sequence

import numpy as np

PKG_DIR = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_DIR = os.path.join(PKG_DIR, "lib")
BIN_DIR = os.path.join(PKG_DIR, "bin")

LAYERS = 7
TRACKLET_LAYERS = 6
CELL_LAYERS = 5
TABLE_SIZE = 401

# Record kinds of itsca_download (include/itsca.h)
CLUSTERS, INDEX_TABLE, TRACKLETS, TRACKLET_LOOKUP, CELLS, CELL_LOOKUP, NEIGHBOUR_OFFSETS, NEIGHBOURS, ROADS = range(9)

_hip = None
_host = None


class ItscaError(RuntimeError):
    pass


class _Config(ctypes.Structure):
    _fields_ = [("device", ctypes.c_int), ("persistent_blocks", ctypes.c_int), ("capacity_scale", ctypes.c_double)]


class Tuning(ctypes.Structure):
    _fields_ = [("block_size", ctypes.c_int), ("persistent_blocks", ctypes.c_int), ("z_subbins", ctypes.c_int),
                ("phi_subrows", ctypes.c_int)]


class Capacities(ctypes.Structure):
    _fields_ = [("tracklets", ctypes.c_long), ("cells", ctypes.c_long), ("neighbours", ctypes.c_long),
                ("roads", ctypes.c_long)]


class DeviceEvent(ctypes.Structure):
    """itsca_device_event: device addresses of one event's cluster arrays,
    host layer offsets and vertex."""
    _fields_ = [("x", ctypes.c_void_p), ("y", ctypes.c_void_p), ("z", ctypes.c_void_p),
                ("cluster_id", ctypes.c_void_p), ("mc_id", ctypes.c_void_p),
                ("layer_offsets", ctypes.c_int32 * 8), ("vertex", ctypes.c_float * 3)]


class Counts(ctypes.Structure):
    _fields_ = [
        ("clusters", ctypes.c_int * 7),
        ("tracklets", ctypes.c_int * 6),
        ("cells", ctypes.c_int * 5),
        ("neighbours", ctypes.c_int * 4),
        ("roads", ctypes.c_int),
    ]


HIP_SYMBOLS = [
    "itsca_create", "itsca_destroy", "itsca_last_error", "itsca_clear_events", "itsca_add_event", "itsca_upload",
    "itsca_run", "itsca_run_async", "itsca_finish", "itsca_num_items", "itsca_get_counts", "itsca_download",
    "itsca_get_capacities", "itsca_run_noalloc", "itsca_download_roads",
    "itsca_set_profiling", "itsca_stage_times", "itsca_kernel_times", "itsca_kernel_name", "itsca_set_graph",
    "itsca_set_tuning", "itsca_get_tuning", "itsca_autotune",
    "itsca_mlp_init", "itsca_mlp_set_params", "itsca_mlp_get_params", "itsca_set_filter", "itsca_set_filter_gates",
    "itsca_train_clear",
    "itsca_train_collect", "itsca_train_set", "itsca_train_size", "itsca_train_get", "itsca_train_step",
    "itsca_train_epoch", "itsca_train_scores",
]
MLP_PARAMS = 1441
TRACKLET_SLOTS = 64  # candidate slots per cluster between the tracklet count and fill (kTrkSlots)
MLP_FEATURES = 16
HOST_SYMBOLS = [
    "itsca_events_generate", "itsca_events_load", "itsca_events_write", "itsca_events_free", "itsca_events_count",
    "itsca_events_num_vertices", "itsca_events_layer_offsets", "itsca_events_copy", "itsca_filter_load",
    "itsca_filter_load_gates",
    "itsca_events_write_binary", "itsca_events_load_labels", "itsca_events_merge_pileup",
    "itsca_events_roads_report",
]


def _path(name: str) -> str:
    p = os.path.join(LIB_DIR, name)
    if not os.path.exists(p):
        raise ItscaError(f"{p} is missing: build the package first (__graft_entry__.build())")
    return p


def hip_lib() -> ctypes.CDLL:
    """Loads libitsca_hip.so (the gfx950 kernels and the C ABI)."""
    global _hip
    if _hip is None:
        _hip = bind_hip(ctypes.CDLL(_path("libitsca_hip.so")))
    return _hip


def bind_hip(lib: ctypes.CDLL) -> ctypes.CDLL:
    """Declares the C ABI of include/itsca.h on a loaded library."""
    if True:
        P, I, F = ctypes.c_void_p, ctypes.c_int, ctypes.POINTER(ctypes.c_float)
        I32 = ctypes.POINTER(ctypes.c_int32)
        lib.itsca_create.argtypes = [ctypes.POINTER(P), ctypes.POINTER(_Config)]
        lib.itsca_destroy.argtypes = [P]
        lib.itsca_destroy.restype = None
        lib.itsca_last_error.argtypes = [P]
        lib.itsca_last_error.restype = ctypes.c_char_p
        lib.itsca_clear_events.argtypes = [P]
        lib.itsca_add_event.argtypes = [P, F, F, F, I32, I32, I32, F, I, ctypes.POINTER(I)]
        for f in ("itsca_upload", "itsca_run", "itsca_run_async", "itsca_finish", "itsca_num_items"):
            getattr(lib, f).argtypes = [P]
        lib.itsca_get_counts.argtypes = [P, I, ctypes.POINTER(Counts)]
        lib.itsca_get_capacities.argtypes = [P, ctypes.POINTER(Capacities)]
        lib.itsca_download_roads.argtypes = [P, P, ctypes.c_size_t, P, ctypes.POINTER(ctypes.c_size_t)]
        lib.itsca_run_noalloc.argtypes = [P, ctypes.POINTER(DeviceEvent), P, I, ctypes.POINTER(I)]
        lib.itsca_download.argtypes = [P, I, I, I, P, ctypes.c_size_t, ctypes.POINTER(ctypes.c_size_t)]
        lib.itsca_set_profiling.argtypes = [P, I]
        lib.itsca_set_graph.argtypes = [P, I]
        lib.itsca_set_tuning.argtypes = [P, ctypes.POINTER(Tuning)]
        lib.itsca_get_tuning.argtypes = [P, ctypes.POINTER(Tuning)]
        lib.itsca_autotune.argtypes = [P, I, ctypes.POINTER(Tuning), F]
        U16, U8 = ctypes.POINTER(ctypes.c_uint16), ctypes.POINTER(ctypes.c_uint8)
        lib.itsca_mlp_init.argtypes = [P, ctypes.c_uint64]
        lib.itsca_mlp_set_params.argtypes = [P, F]
        lib.itsca_mlp_get_params.argtypes = [P, F]
        lib.itsca_set_filter.argtypes = [P, I, ctypes.c_float]
        lib.itsca_set_filter_gates.argtypes = [P, F]
        lib.itsca_train_clear.argtypes = [P, I]
        lib.itsca_train_collect.argtypes = [P, I, I, ctypes.POINTER(I)]
        lib.itsca_train_set.argtypes = [P, U16, U8, I]
        lib.itsca_train_size.argtypes = [P]
        lib.itsca_train_get.argtypes = [P, U16, U8, I]
        lib.itsca_train_step.argtypes = [P, I, I, ctypes.c_long, ctypes.c_long, ctypes.c_float, ctypes.c_float,
                                         ctypes.c_float, F, F]
        lib.itsca_train_epoch.argtypes = [P, I, ctypes.c_float, ctypes.c_float, ctypes.c_float, ctypes.c_uint64, F]
        lib.itsca_train_scores.argtypes = [P, F, I]
        lib.itsca_stage_times.argtypes = [P, F]
        lib.itsca_kernel_times.argtypes = [P, F, I]
        lib.itsca_kernel_name.argtypes = [P, I]
        lib.itsca_kernel_name.restype = ctypes.c_char_p
    return lib


def host_lib() -> ctypes.CDLL:
    """Loads libitsca.so (event generator and reference text formats)."""
    global _host
    if _host is None:
        lib = ctypes.CDLL(_path("libitsca.so"))
        P, I = ctypes.c_void_p, ctypes.c_int
        lib.itsca_events_generate.argtypes = [I, ctypes.c_uint64, I]
        lib.itsca_events_generate.restype = P
        lib.itsca_events_load.argtypes = [ctypes.c_char_p]
        lib.itsca_events_load.restype = P
        lib.itsca_events_write.argtypes = [P, ctypes.c_char_p, ctypes.c_char_p]
        lib.itsca_events_free.argtypes = [P]
        lib.itsca_events_free.restype = None
        lib.itsca_events_count.argtypes = [P]
        lib.itsca_events_num_vertices.argtypes = [P, I]
        lib.itsca_events_layer_offsets.argtypes = [P, I, ctypes.POINTER(ctypes.c_int32)]
        F, I32 = ctypes.POINTER(ctypes.c_float), ctypes.POINTER(ctypes.c_int32)
        lib.itsca_events_copy.argtypes = [P, I, F, F, F, I32, I32, F]
        lib.itsca_filter_load.argtypes = [ctypes.c_char_p, F, I, F]
        lib.itsca_filter_load_gates.argtypes = [ctypes.c_char_p, F, I]
        lib.itsca_events_write_binary.argtypes = [P, ctypes.c_char_p]
        lib.itsca_events_load_labels.argtypes = [P, ctypes.c_char_p]
        lib.itsca_events_merge_pileup.argtypes = [P, I]
        lib.itsca_events_merge_pileup.restype = P
        lib.itsca_events_roads_report.argtypes = [P, I, ctypes.POINTER(ctypes.c_void_p), I32, I,
                                                  ctypes.POINTER(ctypes.c_long)]
        _host = lib
    return _host


@dataclass
class Event:
    """One event: clusters of all layers (layer after layer, file order)."""

    x: np.ndarray
    y: np.ndarray
    z: np.ndarray
    cluster_id: np.ndarray
    mc: np.ndarray
    layer_offsets: np.ndarray  # 8 x int32
    vertices: np.ndarray  # (nv, 3) float32

    def layer(self, l: int) -> slice:
        return slice(int(self.layer_offsets[l]), int(self.layer_offsets[l + 1]))

    def drop_layers(self, layers: Sequence[int]) -> "Event":
        """Copy with the clusters of `layers` removed (ids keep their values)."""
        keep = np.ones(len(self.x), dtype=bool)
        for l in layers:
            keep[self.layer(l)] = False
        counts = [int(self.layer_offsets[l + 1] - self.layer_offsets[l]) for l in range(LAYERS)]
        for l in layers:
            counts[l] = 0
        offs = np.zeros(8, dtype=np.int32)
        offs[1:] = np.cumsum(counts)
        return Event(self.x[keep], self.y[keep], self.z[keep], self.cluster_id[keep], self.mc[keep], offs,
                     self.vertices.copy())


class _EventsHandle:
    def __init__(self, handle: int):
        if not handle:
            raise ItscaError("event batch could not be created")
        self.h = handle

    def __del__(self):
        if getattr(self, "h", None) and _host is not None:
            _host.itsca_events_free(self.h)
            self.h = None


def _read_events(handle: _EventsHandle) -> List[Event]:
    lib = host_lib()
    out = []
    for e in range(lib.itsca_events_count(handle.h)):
        offs = np.zeros(8, dtype=np.int32)
        lib.itsca_events_layer_offsets(handle.h, e, offs.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)))
        n = int(offs[7])
        nv = lib.itsca_events_num_vertices(handle.h, e)
        x, y, z = (np.zeros(n, dtype=np.float32) for _ in range(3))
        cid, mc = np.zeros(n, dtype=np.int32), np.zeros(n, dtype=np.int32)
        v = np.zeros((max(nv, 1), 3), dtype=np.float32)
        fp = lambda a: a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))  # noqa: E731
        ip = lambda a: a.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))  # noqa: E731
        if lib.itsca_events_copy(handle.h, e, fp(x), fp(y), fp(z), ip(cid), ip(mc), fp(v)) != 0:
            raise ItscaError("itsca_events_copy failed")
        out.append(Event(x, y, z, cid, mc, offs, v[:nv].copy()))
    return out


def generate_events(preset: int, seed: int, num_events: int = 1) -> List[Event]:
    """Seeded synthetic events (preset 1 small, 2 central, 3 mixed, 5 pile-up)."""
    h = _EventsHandle(host_lib().itsca_events_generate(int(preset), int(seed), int(num_events)))
    return _read_events(h)


def write_events(preset: int, seed: int, num_events: int, path: str, labels_path: Optional[str] = None) -> None:
    h = _EventsHandle(host_lib().itsca_events_generate(int(preset), int(seed), int(num_events)))
    lp = labels_path.encode() if labels_path else None
    if host_lib().itsca_events_write(h.h, path.encode(), lp) != 0:
        raise ItscaError(f"cannot write {path}")


def load_events(path: str) -> List[Event]:
    """Reads the reference event text format (src/CA/IOUtils.cxx:38-78) or its
    binary side-car (written by convert_events_binary)."""
    return _read_events(_EventsHandle(host_lib().itsca_events_load(path.encode())))


def convert_events_binary(src: str, dst: str) -> None:
    """Writes the binary side-car of an event file (text or binary): the same
    content, read back through mmap by load_events."""
    h = _EventsHandle(host_lib().itsca_events_load(src.encode()))
    if host_lib().itsca_events_write_binary(h.h, dst.encode()) != 0:
        raise ItscaError(f"cannot write {dst}")


def merge_pileup(events_path: str, pile_up: int, out_path: str, labels_path: Optional[str] = None,
                 out_labels_path: Optional[str] = None) -> int:
    """Pile-up merger (benchmarks/scripts/root/generatePileUpData.cxx:44-207):
    consecutive events merged into events of `pile_up` vertices, MC labels of
    the k-th collision offset by (k - 1) * 100000; writes the reference text
    formats and returns the number of merged events."""
    lib = host_lib()
    h = _EventsHandle(lib.itsca_events_load(events_path.encode()))
    if labels_path and lib.itsca_events_load_labels(h.h, labels_path.encode()) != 0:
        raise ItscaError(f"cannot read {labels_path}")
    m = _EventsHandle(lib.itsca_events_merge_pileup(h.h, int(pile_up)))
    lp = out_labels_path.encode() if out_labels_path else None
    if lib.itsca_events_write(m.h, out_path.encode(), lp) != 0:
        raise ItscaError(f"cannot write {out_path}")
    return lib.itsca_events_count(m.h)


class GpuTracker:
    """A context of the gfx950 tracker (one HIP stream, device-resident batch)."""

    def __init__(self, device: int = 0, persistent_blocks: int = 0, capacity_scale: float = 0.0,
                 lib: Optional[ctypes.CDLL] = None):
        lib = lib if lib is not None else hip_lib()
        self._lib = lib
        self.ctx = ctypes.c_void_p()
        cfg = _Config(device, persistent_blocks, capacity_scale)
        st = lib.itsca_create(ctypes.byref(self.ctx), ctypes.byref(cfg))
        if st != 0:
            msg = lib.itsca_last_error(self.ctx).decode() if self.ctx else "itsca_create failed"
            lib.itsca_destroy(self.ctx)
            self.ctx = None
            raise ItscaError(msg)
        self.events: List[Event] = []

    def close(self):
        if self.ctx:
            self._lib.itsca_destroy(self.ctx)
            self.ctx = None

    def __del__(self):
        self.close()

    def _check(self, st: int, what: str):
        if st != 0:
            raise ItscaError(f"{what}: {self._lib.itsca_last_error(self.ctx).decode()} (status {st})")

    def set_events(self, events: Sequence[Event]) -> None:
        lib = self._lib
        self._check(lib.itsca_clear_events(self.ctx), "itsca_clear_events")
        fp = lambda a: np.ascontiguousarray(a, dtype=np.float32).ctypes.data_as(ctypes.POINTER(ctypes.c_float))  # noqa
        for ev in events:
            arrs = [np.ascontiguousarray(a) for a in (ev.x, ev.y, ev.z)]
            ids = np.ascontiguousarray(ev.cluster_id, dtype=np.int32)
            mcs = np.ascontiguousarray(ev.mc, dtype=np.int32)
            offs = np.ascontiguousarray(ev.layer_offsets, dtype=np.int32)
            vtx = np.ascontiguousarray(ev.vertices, dtype=np.float32).reshape(-1)
            ip = lambda a: a.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))  # noqa: E731
            self._check(lib.itsca_add_event(self.ctx, fp(arrs[0]), fp(arrs[1]), fp(arrs[2]), ip(ids), ip(mcs),
                                            ip(offs), fp(vtx), int(len(ev.vertices)), None), "itsca_add_event")
        self._check(lib.itsca_upload(self.ctx), "itsca_upload")
        self.events = list(events)

    def run(self) -> None:
        self._check(self._lib.itsca_run(self.ctx), "itsca_run")

    def run_async(self) -> None:
        self._check(self._lib.itsca_run_async(self.ctx), "itsca_run_async")

    def finish(self) -> None:
        self._check(self._lib.itsca_finish(self.ctx), "itsca_finish")

    def set_tuning(self, block_size: int, persistent_blocks: int = 0, z_subbins: int = 0,
                   phi_subrows: int = 0) -> None:
        t = Tuning(block_size, persistent_blocks, z_subbins, phi_subrows)
        self._check(self._lib.itsca_set_tuning(self.ctx, ctypes.byref(t)), "itsca_set_tuning")

    def tuning(self) -> Tuning:
        t = Tuning()
        self._check(self._lib.itsca_get_tuning(self.ctx, ctypes.byref(t)), "itsca_get_tuning")
        return t

    def autotune(self, repeats: int = 3):
        """Picks the fastest tuning (launch shape, and z sub-bins and phi
        sub-rows of the tracklet search index) for the uploaded batch; returns
        (Tuning, ms per run)."""
        t = Tuning()
        ms = ctypes.c_float(0)
        self._check(self._lib.itsca_autotune(self.ctx, int(repeats), ctypes.byref(t), ctypes.byref(ms)),
                    "itsca_autotune")
        return t, ms.value

    # --- learned tracklet filter -------------------------------------------
    def mlp_init(self, seed: int) -> None:
        self._check(self._lib.itsca_mlp_init(self.ctx, int(seed)), "itsca_mlp_init")

    def mlp_params(self) -> np.ndarray:
        p = np.zeros(MLP_PARAMS, dtype=np.float32)
        self._check(self._lib.itsca_mlp_get_params(self.ctx, _fptr(p)), "itsca_mlp_get_params")
        return p

    def set_mlp_params(self, params) -> None:
        p = np.ascontiguousarray(params, dtype=np.float32)
        assert p.size == MLP_PARAMS
        self._check(self._lib.itsca_mlp_set_params(self.ctx, _fptr(p)), "itsca_mlp_set_params")

    def set_filter(self, on: bool, threshold: float = 0.0, fused: bool = False) -> None:
        """Learned filter on or off; `fused` scores inside the tracklet search
        instead of in a kernel of its own after it (same tracklets)."""
        mode = (2 if fused else 1) if on else 0
        self._check(self._lib.itsca_set_filter(self.ctx, mode, float(threshold)), "itsca_set_filter")

    def set_filter_gates(self, gates=None) -> None:
        """Per tracklet layer, the phi gate of filtered runs (radians, in
        (0, 0.3]); None: no gates."""
        if gates is None:
            self._check(self._lib.itsca_set_filter_gates(self.ctx, None), "itsca_set_filter_gates")
            return
        g = np.ascontiguousarray(gates, dtype=np.float32)
        if g.shape != (TRACKLET_LAYERS,):
            raise ValueError("one phi gate per tracklet layer")
        self._check(self._lib.itsca_set_filter_gates(self.ctx, g.ctypes.data_as(ctypes.POINTER(ctypes.c_float))),
                    "itsca_set_filter_gates")

    def train_clear(self, capacity: int) -> None:
        self._check(self._lib.itsca_train_clear(self.ctx, int(capacity)), "itsca_train_clear")

    def train_collect(self, stride: int = 1, phase: int = 0) -> int:
        added = ctypes.c_int(0)
        self._check(self._lib.itsca_train_collect(self.ctx, int(stride), int(phase), ctypes.byref(added)),
                    "itsca_train_collect")
        return added.value

    def train_set(self, X: np.ndarray, y: np.ndarray) -> None:
        X = np.ascontiguousarray(X, dtype=np.uint16).reshape(-1, MLP_FEATURES)
        y = np.ascontiguousarray(y, dtype=np.uint8)
        self._check(self._lib.itsca_train_set(self.ctx, X.ctypes.data_as(ctypes.POINTER(ctypes.c_uint16)),
                                              y.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)), len(y)),
                    "itsca_train_set")

    def train_size(self) -> int:
        return self._lib.itsca_train_size(self.ctx)

    def train_get(self, n: int):
        X = np.zeros((max(n, 1), MLP_FEATURES), dtype=np.uint16)
        y = np.zeros(max(n, 1), dtype=np.uint8)
        self._check(self._lib.itsca_train_get(self.ctx, X.ctypes.data_as(ctypes.POINTER(ctypes.c_uint16)),
                                              y.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)), int(n)),
                    "itsca_train_get")
        return X[:n], y[:n]

    def train_step(self, first: int, count: int, mult: int = 1, offset: int = 0, pos_weight: float = 1.0,
                   lr: float = 0.0, weight_decay: float = 0.0):
        loss = np.zeros(1, dtype=np.float32)
        grad = np.zeros(MLP_PARAMS, dtype=np.float32)
        self._check(self._lib.itsca_train_step(self.ctx, int(first), int(count), int(mult), int(offset),
                                               float(pos_weight), float(lr), float(weight_decay), _fptr(loss),
                                               _fptr(grad)), "itsca_train_step")
        return float(loss[0]), grad

    def train_epoch(self, batch: int, lr: float, weight_decay: float, pos_weight: float, seed: int) -> float:
        loss = np.zeros(1, dtype=np.float32)
        self._check(self._lib.itsca_train_epoch(self.ctx, int(batch), float(lr), float(weight_decay),
                                                float(pos_weight), int(seed), _fptr(loss)), "itsca_train_epoch")
        return float(loss[0])

    def train_scores(self, n: int) -> np.ndarray:
        z = np.zeros(max(n, 1), dtype=np.float32)
        self._check(self._lib.itsca_train_scores(self.ctx, _fptr(z), int(n)), "itsca_train_scores")
        return z[:n]

    def set_graph(self, on: bool) -> None:
        self._check(self._lib.itsca_set_graph(self.ctx, 1 if on else 0), "itsca_set_graph")

    def set_profiling(self, on: bool) -> None:
        self._check(self._lib.itsca_set_profiling(self.ctx, 1 if on else 0), "itsca_set_profiling")

    def stage_times(self) -> np.ndarray:
        ms = np.zeros(5, dtype=np.float32)
        self._check(self._lib.itsca_stage_times(self.ctx, ms.ctypes.data_as(ctypes.POINTER(ctypes.c_float))),
                    "itsca_stage_times")
        return ms

    def kernel_times(self):
        """[(kernel name, ms)] of every launch of the last profiled run."""
        ms = np.zeros(64, dtype=np.float32)
        n = self._lib.itsca_kernel_times(self.ctx, ms.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), 64)
        if n < 0:
            raise ItscaError("itsca_kernel_times: no profiled run")
        return [(self._lib.itsca_kernel_name(self.ctx, i).decode(), float(ms[i])) for i in range(min(n, 64))]

    def num_items(self) -> int:
        return self._lib.itsca_num_items(self.ctx)

    def capacities(self) -> Capacities:
        c = Capacities()
        self._check(self._lib.itsca_get_capacities(self.ctx, ctypes.byref(c)), "itsca_get_capacities")
        return c

    def all_roads(self):
        """(rows, first): the roads of every work item in one copy, rows of 8
        int; item i's roads are rows[first[i]:first[i + 1]]."""
        first = np.zeros(self.num_items() + 1, dtype=np.int32)
        total = ctypes.c_size_t(0)
        st = self._lib.itsca_download_roads(self.ctx, None, 0, first.ctypes.data, ctypes.byref(total))
        rows = np.empty((max(total.value, 1), 8), dtype=np.int32)  # every returned row is written
        if st != 0 and total.value == 0:
            self._check(st, "itsca_download_roads")
        self._check(self._lib.itsca_download_roads(self.ctx, rows.ctypes.data, total.value, first.ctypes.data,
                                                   ctypes.byref(total)), "itsca_download_roads")
        return rows[:total.value], first

    def run_noalloc(self, arrays, layer_offsets, vertex, roads, roads_cap: int) -> int:
        """itsca_run_noalloc: `arrays` are the device addresses of x, y, z,
        cluster ids and MC ids, `roads` the device address of roads_cap rows
        of 8 int; returns the number of roads."""
        ev = DeviceEvent(*[ctypes.c_void_p(int(a)) for a in arrays],
                         (ctypes.c_int32 * 8)(*[int(o) for o in layer_offsets]),
                         (ctypes.c_float * 3)(*[float(v) for v in vertex]))
        n = ctypes.c_int(0)
        st = self._lib.itsca_run_noalloc(self.ctx, ctypes.byref(ev), ctypes.c_void_p(int(roads)), int(roads_cap),
                                         ctypes.byref(n))
        if st != 0:
            raise ItscaError("itsca_run_noalloc: %s (status %d, roads %d)"
                             % (self._lib.itsca_last_error(self.ctx).decode(), st, n.value))
        return n.value

    def counts(self, item: int) -> Counts:
        c = Counts()
        self._check(self._lib.itsca_get_counts(self.ctx, item, ctypes.byref(c)), "itsca_get_counts")
        return c

    def download(self, item: int, what: int, layer: int = 0, rows: int = 0, fields: int = 1) -> np.ndarray:
        buf = np.zeros(max(rows * fields, 1), dtype=np.int32)
        written = ctypes.c_size_t(0)
        self._check(self._lib.itsca_download(self.ctx, item, what, layer, buf.ctypes.data, buf.nbytes,
                                             ctypes.byref(written)), "itsca_download")
        n = written.value // 4
        out = buf[:n]
        return out.reshape(-1, fields) if fields > 1 else out

    def roads(self, item: int) -> np.ndarray:
        """(R, 8) int32: cell per layer 0..4, size, label, fake."""
        return self.download(item, ROADS, 0, self.counts(item).roads, 8)


def roads_report(events_path: str, labels_path: str, roads) -> List[tuple]:
    """Correct, duplicate and fake roads of every event, classified like
    IOUtils::writeRoadsReport (IOUtils.cxx:125-170) against the labels file;
    roads[e][v] holds the (R, 8) road rows of vertex v of event e."""
    lib = host_lib()
    h = lib.itsca_events_load(events_path.encode())
    if not h:
        raise ItscaError("cannot read events from %s" % events_path)
    try:
        if lib.itsca_events_load_labels(h, labels_path.encode()) != 0:
            raise ItscaError("cannot read labels from %s" % labels_path)
        out = []
        for e, per_vertex in enumerate(roads):
            rows = [np.ascontiguousarray(r, dtype=np.int32).reshape(-1, 8) for r in per_vertex]
            ptrs = (ctypes.c_void_p * max(len(rows), 1))(*[r.ctypes.data for r in rows])
            n = np.array([len(r) for r in rows] or [0], dtype=np.int32)
            counts = (ctypes.c_long * 3)()
            if lib.itsca_events_roads_report(h, e, ptrs, n.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)), len(rows),
                                             counts) != 0:
                raise ItscaError("no labels for event %d" % e)
            out.append(tuple(counts))
        return out
    finally:
        lib.itsca_events_free(h)


def road_tracks(tracker: GpuTracker, item: int) -> List[tuple]:
    """Roads of one work item as tuples of the 7 input cluster ids (layer 0..6),
    in road order: the identity of a track, independent of the cell and
    tracklet indices of a particular run."""
    n = tracker.counts(item)
    if n.roads == 0:
        return []
    ids = [tracker.download(item, CLUSTERS, l, n.clusters[l], 8)[:, 0] for l in range(LAYERS)]
    cells = [tracker.download(item, CELLS, k, n.cells[k], 10) for k in range(CELL_LAYERS)]
    r = tracker.roads(item)
    cols = [cells[0][r[:, 0], 0], cells[0][r[:, 0], 1]] + [cells[k][r[:, k], 2] for k in range(CELL_LAYERS)]
    return list(zip(*[ids[l][cols[l]].tolist() for l in range(LAYERS)]))


def cell_keys(tracker: GpuTracker, item: int) -> set:
    """Cells of one work item as (cell layer, cluster id, cluster id, cluster
    id): identities independent of the indices of a particular run."""
    n = tracker.counts(item)
    out = set()
    ids = [tracker.download(item, CLUSTERS, l, n.clusters[l], 8)[:, 0] for l in range(LAYERS)]
    for k in range(CELL_LAYERS):
        if n.cells[k] == 0:
            continue
        c = tracker.download(item, CELLS, k, n.cells[k], 10)
        out.update(zip([k] * len(c), ids[k][c[:, 0]].tolist(), ids[k + 1][c[:, 1]].tolist(),
                       ids[k + 2][c[:, 2]].tolist()))
    return out


def _fptr(a: np.ndarray):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))


def _hex(a: np.ndarray) -> List[str]:
    return ["%08x" % v for v in np.asarray(a).astype(np.int32).view(np.uint32).tolist()]


def _ints_line(a) -> str:
    return " ".join(str(int(v)) for v in a) + "\n"


def dump_item(tracker: GpuTracker, item: int, vertex: Sequence[float]) -> str:
    """Canonical text form of one work item's full state (same format as the
    CPU oracle's dump: every record, floats as IEEE bit patterns)."""
    n = tracker.counts(item)
    vx = np.asarray(vertex, dtype=np.float32)
    parts = ["vertex %s\n" % " ".join(_hex(vx.view(np.int32)))]
    for l in range(LAYERS):
        c = tracker.download(item, CLUSTERS, l, n.clusters[l], 8)
        parts.append("clusters %d %d\n" % (l, n.clusters[l]))
        if len(c):
            ph, rr = _hex(c[:, 1]), _hex(c[:, 2])
            parts.append("".join("%d %s %s %d\n" % (i, p, r, b) for i, p, r, b in
                                 zip(c[:, 0].tolist(), ph, rr, c[:, 3].tolist())))
    for l in range(1, LAYERS):
        t = tracker.download(item, INDEX_TABLE, l, TABLE_SIZE, 1)
        parts.append("table %d %d\n" % (l, TABLE_SIZE))
        parts.append(_ints_line(t))
    for l in range(TRACKLET_LAYERS):
        t = tracker.download(item, TRACKLETS, l, n.tracklets[l], 4)
        parts.append("tracklets %d %d\n" % (l, n.tracklets[l]))
        if len(t):
            parts.append("".join("%d %d %s %s\n" % q for q in
                                 zip(t[:, 0].tolist(), t[:, 1].tolist(), _hex(t[:, 2]), _hex(t[:, 3]))))
    for l in range(CELL_LAYERS):
        lk = tracker.download(item, TRACKLET_LOOKUP, l, n.clusters[l + 1], 1)
        parts.append("tlookup %d %d\n" % (l, len(lk)))
        parts.append(_ints_line(lk))
    for l in range(CELL_LAYERS):
        c = tracker.download(item, CELLS, l, n.cells[l], 10)
        parts.append("cells %d %d\n" % (l, n.cells[l]))
        if len(c):
            cols = [c[:, k].tolist() for k in range(5)] + [_hex(c[:, k]) for k in range(5, 9)] + [c[:, 9].tolist()]
            parts.append("".join("%d %d %d %d %d %s %s %s %s %d\n" % q for q in zip(*cols)))
    for l in range(CELL_LAYERS - 1):
        lk = tracker.download(item, CELL_LOOKUP, l, n.tracklets[l + 1], 1)
        parts.append("clookup %d %d\n" % (l, len(lk)))
        parts.append(_ints_line(lk))
    for l in range(CELL_LAYERS - 1):
        cn = n.cells[l + 1]
        off = tracker.download(item, NEIGHBOUR_OFFSETS, l, cn + 1, 1)
        lst = tracker.download(item, NEIGHBOURS, l, n.neighbours[l], 1).tolist()
        parts.append("neighbours %d %d\n" % (l, cn))
        lines = []
        for i in range(cn):
            seg = lst[int(off[i]):int(off[i + 1])]
            lines.append(" ".join([str(len(seg))] + [str(v) for v in seg]) + "\n")
        parts.append("".join(lines))
    r = tracker.roads(item)
    parts.append("roads %d\n" % len(r))
    parts.append("".join("%d %d %d %d %d %d %d %d\n" % tuple(row) for row in r.tolist()))
    return "".join(parts)
