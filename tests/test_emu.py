"""The gfx950 kernels run on the CPU: the product's kernel and runtime sources
compiled for the host against the HIP shim of tests/emu/ (one fiber per GPU
thread, the wavefront exchanges and the MFMA instructions emulated), with
AddressSanitizer.  Checks the kernels' logic and bounds without a GPU: golden
fixtures bit for bit under several tunings (launch shapes, z sub-bins and phi
sub-rows of the search index), the learned filter's paths (scored
after the search and inside it), an event
of 1831 clusters per layer and an event whose tracklets overflow the count's
slots, against the oracle; and the same with the workgroups of each launch
running concurrently on several OS threads."""
import glob
import os
import subprocess
import sys

import pytest

ROOT = os.path.abspath(os.path.join(os.path.dirname(__file__), ".."))
EMU = os.path.join(ROOT, "tests", "emu")
CLANG = "/opt/rocm/llvm/bin/clang++"

SCRIPT = r"""
import ctypes, os, sys
from collections import Counter
import numpy as np
sys.path[:0] = [os.path.join(ROOT, "tracking-itsu_amd"), os.path.join(ROOT, "tests"), os.path.join(ROOT, "oracle")]
import itsca
from golden_cases import fixture, small_cases
lib = itsca.bind_hip(ctypes.CDLL(os.path.join(ROOT, "tests", "emu", "libitsca_emu.so")))
cases = {n: p for n, p, _ in small_cases()}
for name in ("small_s1", "sparse_s6", "small_s5_nolayer3"):
    ev = itsca.load_events(cases[name])[0]
    t = itsca.GpuTracker(persistent_blocks=4, lib=lib)
    t.set_events([ev])
    t.run()
    assert itsca.dump_item(t, 0, ev.vertices[0]) == fixture(name, "stable", 0), name
    if name in ("small_s1", "small_s5_nolayer3"):  # the second: a layer whose tracklet count never ran
        plain = Counter(itsca.road_tracks(t, 0))
        n = t.counts(0)
        trk = [t.download(0, itsca.TRACKLETS, l, n.tracklets[l], 4)[:, :2] for l in range(6)]
        total = sum(len(x) for x in trk)
        t.mlp_init(3)
        t.train_clear(total)
        assert t.train_collect(1, 0) == total
        z = t.train_scores(total)
        t.set_filter(True, -3.0e38)
        t.run()
        assert itsca.dump_item(t, 0, ev.vertices[0]) == fixture(name, "stable", 0)
        thr = float(np.quantile(z, 0.5))
        t.set_filter(True, thr, fused=True)  # the search and the filter in one kernel: the same tracklets
        t.run()
        fused = itsca.dump_item(t, 0, ev.vertices[0])
        t.set_filter(True, thr)
        t.run()
        assert itsca.dump_item(t, 0, ev.vertices[0]) == fused, name
        assert not (Counter(itsca.road_tracks(t, 0)) - plain)
        k = 0
        for l in range(6):
            raw = np.bincount(trk[l][:, 0], minlength=n.clusters[l])
            keep = (z[k:k + len(trk[l])] >= thr) | (raw[trk[l][:, 0]] > itsca.TRACKLET_SLOTS)
            k += len(trk[l])
            got = t.download(0, itsca.TRACKLETS, l, t.counts(0).tracklets[l], 4)[:, :2]
            assert np.array_equal(got, trk[l][keep]), (name, l)
    if name == "small_s1":
        # the MFMA forward pass (swizzled LDS tile, emulated MFMA lane layouts)
        # against float64, as in the GPU test: two full tiles and a partial one
        import mlp_ref as M
        rng = np.random.default_rng(5)
        p = (rng.normal(size=M.PARAMS) * 0.3).astype(np.float32)
        X = M.pack_rows((rng.normal(size=(130, 10)) * 1.5).astype(np.float32))
        t.set_mlp_params(p)
        t.train_set(X, np.zeros(130, np.uint8))
        z = t.train_scores(130)
        zr = M.forward(p, X)[3]
        err = np.abs(z - zr)
        assert np.max(err / np.maximum(1.0, np.abs(zr))) < 3e-2
        assert np.quantile(err, 0.99) < 1e-4 * max(1.0, float(np.max(np.abs(zr))))
        print("mfma forward vs float64 ok", flush=True)
    t.close()
    print(name, "ok", flush=True)
# phi gates: a filtered run whose threshold keeps every candidate builds exactly
# the plain tracklets inside the gates, in the plain order -- scored after the
# search and inside it, with and without the search index (phi sub-rows)
from gpu_items import gated_tracklets
for name in ("small_s1", "pileup2_s7"):
    ev = itsca.load_events(cases[name])[0]
    t = itsca.GpuTracker(persistent_blocks=4, lib=lib)
    t.set_events([ev])
    t.run()
    items = t.num_items()
    plain = [[t.download(w, itsca.TRACKLETS, l, t.counts(w).tracklets[l], 4) for l in range(6)] for w in range(items)]
    t.mlp_init(3)
    for gates, tuning, fused in (([0.02, 0.02, 0.15, 0.05, 0.1, 0.05], (256, 4, 1, 1), False),
                                 ([0.004] * 6, (128, 4, 2, 4), False),
                                 ([0.02, 0.3, 0.15, 0.05, 0.1, 0.05], (256, 4, 1, 1), True)):
        t.set_tuning(*tuning)
        t.set_filter_gates(gates)
        t.set_filter(True, -3.0e38, fused=fused)
        t.run()
        for w in range(items):
            want = gated_tracklets(t, w, plain[w], gates)
            n = t.counts(w)
            for l in range(6):
                got = t.download(w, itsca.TRACKLETS, l, n.tracklets[l], 4)[:, :2]
                assert np.array_equal(got, want[l]), (name, gates, tuning, w, l)
    t.set_filter(False)
    t.set_filter_gates(None)
    t.close()
    print(name, "phi gates ok", flush=True)
# other launch shapes and search-index binnings: an odd number of z sub-bins,
# phi sub-rows alone and together with z sub-bins
p = cases["pileup2_s7"]
ev = itsca.load_events(p)[0]
for tuning in ((128, 4, 3, 1), (64, 4, 1, 4), (256, 4, 3, 4)):
    t = itsca.GpuTracker(persistent_blocks=4, lib=lib)
    t.set_tuning(*tuning)
    t.set_events([ev])
    t.run()
    for v in range(len(ev.vertices)):
        assert itsca.dump_item(t, v, ev.vertices[v]) == fixture("pileup2_s7", "stable", v), (tuning, v)
    t.close()
print("pileup2_s7 under three search-index binnings ok", flush=True)
# 1831 clusters per layer: two tiles of the stable sort, dense bins searched by z sub-bin, against the oracle
import pyoracle
ev = itsca.generate_events(3, 3006, 1)[0]
assert int(np.diff(ev.layer_offsets).min()) > 1024
want = pyoracle.OracleVertex(ev, 0, stable=True).dump()
for tuning in ((256, 8, 8, 1), (256, 8, 2, 4)):
    t = itsca.GpuTracker(persistent_blocks=8, lib=lib)
    t.set_tuning(*tuning)
    t.set_events([ev])
    t.run()
    assert itsca.dump_item(t, 0, ev.vertices[0]) == want, tuning
    t.close()
print("mixed event 3006 vs oracle ok (z sub-bins; z sub-bins and phi sub-rows)", flush=True)
# a cluster with more tracklets than the count's slots: the fill's own search, against the oracle
from golden_cases import dense_spot_event
ev = dense_spot_event()
want = pyoracle.OracleVertex(ev, 0, stable=True).dump()
t = itsca.GpuTracker(persistent_blocks=4, lib=lib)
t.set_tuning(256, 4, 4, 2)  # the fill's own search through the search index with phi sub-rows
t.set_events([ev])
t.run()
assert itsca.dump_item(t, 0, ev.vertices[0]) == want
t.set_tuning(256, 4, 8, 1)
t.run()
n = t.counts(0)
assert np.bincount(t.download(0, itsca.TRACKLETS, 0, n.tracklets[0], 4)[:, 0]).max() > itsca.TRACKLET_SLOTS
assert itsca.dump_item(t, 0, ev.vertices[0]) == want
t.mlp_init(4)
t.set_filter(True, 0.0, fused=True)
t.run()
fused = itsca.dump_item(t, 0, ev.vertices[0])
t.set_filter(True, 0.0)
t.run()
assert itsca.dump_item(t, 0, ev.vertices[0]) == fused
t.close()
print("slot-overflow event vs oracle ok", flush=True)
print("ALL OK")
"""


def _asan_runtime():
    found = sorted(glob.glob("/opt/rocm/llvm/lib/clang/*/lib/linux/libclang_rt.asan-x86_64.so"))
    return found[-1] if found else None


def test_kernels_under_cpu_emulation_with_asan():
    rt = _asan_runtime()
    if not os.path.exists(CLANG) or rt is None:
        pytest.skip("ROCm clang++ with its AddressSanitizer runtime is needed")
    build = subprocess.run(["make", "-C", EMU], capture_output=True, text=True, timeout=1200)
    assert build.returncode == 0, build.stdout[-2000:] + build.stderr[-4000:]
    env = dict(os.environ)
    env["LD_PRELOAD"] = rt + (":" + env["LD_PRELOAD"] if env.get("LD_PRELOAD") else "")
    env["ASAN_OPTIONS"] = "detect_leaks=0"
    out = subprocess.run([sys.executable, "-c", "ROOT = %r\n" % ROOT + SCRIPT], env=env, capture_output=True,
                         text=True, timeout=1200)
    assert out.returncode == 0 and "ALL OK" in out.stdout, out.stdout[-2000:] + out.stderr[-6000:]


CONCURRENT_SCRIPT = r"""
import ctypes, os, sys
import numpy as np
sys.path[:0] = [os.path.join(ROOT, "tracking-itsu_amd"), os.path.join(ROOT, "tests"), os.path.join(ROOT, "oracle")]
import itsca, pyoracle
from golden_cases import dense_spot_event, fixture, small_cases
from gpu_items import gated_tracklets
assert os.environ.get("EMU_THREADS") == "4"
lib = itsca.bind_hip(ctypes.CDLL(os.path.join(ROOT, "tests", "emu", "libitsca_emu.so")))
cases = {n: p for n, p, _ in small_cases()}
# fixtures bit for bit, several workgroups per CU so that chunks of one stage interleave
for name, tuning in (("small_s1", (256, 16, 1, 1)), ("sparse_s6", (64, 16, 1, 1)), ("pileup2_s7", (128, 16, 3, 2))):
    ev = itsca.load_events(cases[name])[0]
    t = itsca.GpuTracker(persistent_blocks=16, lib=lib)
    t.set_tuning(*tuning)
    t.set_events([ev])
    t.run()
    for v in range(len(ev.vertices)):
        assert itsca.dump_item(t, v, ev.vertices[v]) == fixture(name, "stable", v), (name, v)
    if name == "small_s1":  # the learned filter: -inf = plain; gated, keeping all = the gated plain tracklets
        plain = [t.download(0, itsca.TRACKLETS, l, t.counts(0).tracklets[l], 4) for l in range(6)]
        t.mlp_init(3)
        t.set_filter(True, -3.0e38)
        t.run()
        assert itsca.dump_item(t, 0, ev.vertices[0]) == fixture(name, "stable", 0)
        gates = [0.02, 0.02, 0.15, 0.05, 0.1, 0.05]
        t.set_filter_gates(gates)
        t.run()
        want = gated_tracklets(t, 0, plain, gates)
        n = t.counts(0)
        for l in range(6):
            assert np.array_equal(t.download(0, itsca.TRACKLETS, l, n.tracklets[l], 4)[:, :2], want[l]), l
    t.close()
    print(name, "ok", flush=True)
# an event of 1831 clusters per layer and the slot-overflow event, against the oracle
for ev in (itsca.generate_events(3, 3006, 1)[0], dense_spot_event()):
    t = itsca.GpuTracker(persistent_blocks=16, lib=lib)
    t.set_tuning(256, 16, 1, 1)
    t.set_events([ev])
    t.run()
    assert itsca.dump_item(t, 0, ev.vertices[0]) == pyoracle.OracleVertex(ev, 0, stable=True).dump()
    t.close()
print("oracle events ok", flush=True)
print("ALL OK")
"""


def test_concurrent_workgroups_under_cpu_emulation_with_asan():
    """The same kernels with the workgroups of every launch run at once on four
    OS threads (each with its own LDS), dealt from the last one down: the
    fixtures and the oracle still match bit for bit, so no kernel depends on
    the order in which workgroups run or on another workgroup's writes within
    one launch."""
    rt = _asan_runtime()
    if not os.path.exists(CLANG) or rt is None:
        pytest.skip("ROCm clang++ with its AddressSanitizer runtime is needed")
    build = subprocess.run(["make", "-C", EMU], capture_output=True, text=True, timeout=1200)
    assert build.returncode == 0, build.stdout[-2000:] + build.stderr[-4000:]
    env = dict(os.environ)
    env["LD_PRELOAD"] = rt + (":" + env["LD_PRELOAD"] if env.get("LD_PRELOAD") else "")
    env["ASAN_OPTIONS"] = "detect_leaks=0"
    env["EMU_THREADS"] = "4"
    env["EMU_ORDER"] = "reverse"
    out = subprocess.run([sys.executable, "-c", "ROOT = %r\n" % ROOT + CONCURRENT_SCRIPT], env=env, capture_output=True,
                         text=True, timeout=1200)
    assert out.returncode == 0 and "ALL OK" in out.stdout, out.stdout[-2000:] + out.stderr[-6000:]


EDGE_SCRIPT = r"""
import ctypes, os, sys
import numpy as np
sys.path[:0] = [os.path.join(ROOT, "tracking-itsu_amd"), os.path.join(ROOT, "tests"), os.path.join(ROOT, "oracle")]
import itsca, pyoracle
from golden_cases import fixture, small_cases
lib = itsca.bind_hip(ctypes.CDLL(os.path.join(ROOT, "tests", "emu", "libitsca_emu.so")))
# Clusters outside the detector in z: below the layer (a negative z bin, so a
# negative bin in phi row 0 and an aliased bin in row 1), and at |z| = 1e10
# (an index outside the int range).  The reference's sort puts negative bins
# first and its index table counts them inside bin 0; every array must still
# equal the oracle's.
base = itsca.generate_events(1, 1, 1)[0]
vx, vy, vz = (float(a) for a in base.vertices[0])
radii = [2.34, 3.15, 3.93, 19.6, 24.55, 34.39, 39.34]
extra = {0: [(-20.0, 0.1), (1.0e10, 0.1), (-1.0e10, 0.12), (-30.0, 0.5), (-20.0, 0.1)],
         1: [(-25.0, 0.05), (-19.0, 0.2), (3.0e9, 0.02), (-21.0, 0.6)]}
cols = {k: [] for k in ("x", "y", "z", "mc")}
offs = [0]
for layer in range(itsca.LAYERS):
    sl = base.layer(layer)
    for k, a in (("x", base.x), ("y", base.y), ("z", base.z), ("mc", base.mc)):
        cols[k] += list(a[sl])
    for z, phi in extra.get(layer, []):
        cols["x"].append(vx + radii[layer] * np.cos(phi))
        cols["y"].append(vy + radii[layer] * np.sin(phi))
        cols["z"].append(z)
        cols["mc"].append(-1)
    offs.append(len(cols["x"]))
n = offs[-1]
ev = itsca.Event(np.array(cols["x"], np.float32), np.array(cols["y"], np.float32), np.array(cols["z"], np.float32),
                 np.arange(n, dtype=np.int32), np.array(cols["mc"], np.int32), np.array(offs, np.int32),
                 base.vertices.copy())
for zsub in (1, 4):
    t = itsca.GpuTracker(persistent_blocks=4, lib=lib)
    t.set_tuning(256, 4, zsub)
    t.set_events([ev])
    t.run()
    cl = t.download(0, itsca.CLUSTERS, 0, n, 8)
    assert int(cl[:, 3].min()) < 0  # negative bins are kept as the reference computes them
    assert itsca.dump_item(t, 0, ev.vertices[0]) == pyoracle.OracleVertex(ev, 0, stable=True).dump(), zsub
    t.close()
print("clusters outside the layers vs oracle ok", flush=True)
# non-finite coordinates are refused with an argument error
bad = itsca.Event(ev.x.copy(), ev.y.copy(), ev.z.copy(), ev.cluster_id, ev.mc, ev.layer_offsets, ev.vertices)
bad.z[3] = np.nan
t = itsca.GpuTracker(persistent_blocks=4, lib=lib)
try:
    t.set_events([bad])
    raise SystemExit("a NaN coordinate was accepted")
except itsca.ItscaError as e:
    assert "non-finite" in str(e), e
# a refused event leaves the events staged before it intact, and the next one is staged after them
L = t._lib
F = lambda a: np.ascontiguousarray(a, np.float32).ctypes.data_as(ctypes.POINTER(ctypes.c_float))  # noqa: E731
I = lambda a: np.ascontiguousarray(a, np.int32).ctypes.data_as(ctypes.POINTER(ctypes.c_int32))  # noqa: E731
def add(e):
    return L.itsca_add_event(t.ctx, F(e.x), F(e.y), F(e.z), I(e.cluster_id), I(e.mc), I(e.layer_offsets),
                             F(np.asarray(e.vertices, np.float32).reshape(-1)), len(e.vertices), None)
assert L.itsca_clear_events(t.ctx) == 0
assert add(ev) == 0 and add(bad) != 0 and add(ev) == 0
assert L.itsca_upload(t.ctx) == 0 and L.itsca_run(t.ctx) == 0
want = pyoracle.OracleVertex(ev, 0, stable=True).dump()
assert t.num_items() == 2
assert itsca.dump_item(t, 0, ev.vertices[0]) == want and itsca.dump_item(t, 1, ev.vertices[0]) == want
t.close()
print("non-finite input refused ok", flush=True)
# search-index binnings outside the 16-byte sub-bin entries are refused
t = itsca.GpuTracker(persistent_blocks=4, lib=lib)
for bad_tuning in ((256, 0, 8, 2), (256, 0, 4, 4), (256, 0, 2, 3)):
    try:
        t.set_tuning(*bad_tuning)
        raise SystemExit("tuning %r was accepted" % (bad_tuning,))
    except itsca.ItscaError as e:
        assert "phi_subrows" in str(e), e
t.set_tuning(256, 0, 3, 4)
assert (t.tuning().z_subbins, t.tuning().phi_subrows) == (3, 4)
t.close()
print("search-index binning bounds ok", flush=True)
# roads per root cell bounded at 3 in this build: events within the bound give
# the fixture, an event with a root of 4 roads fails with a capacity error
cap = itsca.bind_hip(ctypes.CDLL(os.path.join(ROOT, "tests", "emu", "libitsca_emu_cap3.so")))
cases = {nm: p for nm, p, _ in small_cases()}
for name, ok in (("small_s1", True), ("small_s4", True), ("small_s2", False)):
    e = itsca.load_events(cases[name])[0]
    t = itsca.GpuTracker(persistent_blocks=4, lib=cap)
    t.set_events([e])
    try:
        t.run()
        assert ok, name
        assert itsca.dump_item(t, 0, e.vertices[0]) == fixture(name, "stable", 0), name
    except itsca.ItscaError as err:
        assert not ok and "more than 3 roads" in str(err), (name, err)
    t.close()
print("root road bound ok", flush=True)
# itsca_run_noalloc: the same roads as itsca_run, no hipMalloc or hipFree
# during the call, nothing written past the roads; a smaller event second
so = ctypes.CDLL(os.path.join(ROOT, "tests", "emu", "libitsca_emu.so"))
so.emu_alloc_calls.restype = ctypes.c_long
t = itsca.GpuTracker(persistent_blocks=4, lib=lib)
for name in ("sparse_s6", "small_s3"):
    e = itsca.load_events(cases[name])[0]
    t.set_events([e])
    t.run()
    want = t.roads(0)
    arrs = [np.ascontiguousarray(a) for a in (e.x, e.y, e.z, e.cluster_id.astype(np.int32), e.mc.astype(np.int32))]
    rows = np.full((len(want) + 50, 8), -7, np.int32)
    before = so.emu_alloc_calls()
    n = t.run_noalloc([a.ctypes.data for a in arrs], e.layer_offsets, e.vertices[0], rows.ctypes.data, len(rows))
    assert so.emu_alloc_calls() == before, name
    assert n == len(want) and np.array_equal(rows[:n], want) and (rows[n:] == -7).all(), name
t.close()
print("no-alloc entry point ok", flush=True)
print("ALL OK")
"""


def test_edge_inputs_under_cpu_emulation_with_asan():
    rt = _asan_runtime()
    if not os.path.exists(CLANG) or rt is None:
        pytest.skip("ROCm clang++ with its AddressSanitizer runtime is needed")
    build = subprocess.run(["make", "-C", EMU, "all"], capture_output=True, text=True, timeout=1200)
    assert build.returncode == 0, build.stdout[-2000:] + build.stderr[-4000:]
    env = dict(os.environ)
    env["LD_PRELOAD"] = rt + (":" + env["LD_PRELOAD"] if env.get("LD_PRELOAD") else "")
    env["ASAN_OPTIONS"] = "detect_leaks=0"
    out = subprocess.run([sys.executable, "-c", "ROOT = %r\n" % ROOT + EDGE_SCRIPT], env=env, capture_output=True,
                         text=True, timeout=1200)
    assert out.returncode == 0 and "ALL OK" in out.stdout, out.stdout[-2000:] + out.stderr[-6000:]
