#!/usr/bin/env python3
"""Throughput benchmark: plain GPU CA vs batched CA with the learned tracklet filter.

Workload: 256 central Pb-Pb-like events (8000 primaries + 2% secondaries + 2%
noise, about 8.3e3 clusters per layer, one primary vertex), generated from the
held-out seeds 20000..20255 that the filter never saw in training or threshold
tuning, and uploaded to HBM before timing.  One step
= one pass of the pipeline (index tables, tracklets, [filter], cells,
neighbours, levels, roads) over a batch of such events, replayed as one HIP
graph, with the launch shape and the search-index binning that the autotuner
picks for that form.  Both forms run on the same batch; `value` is the learned form, whose
roads are checked against the plain ones in the same run (recall, and tracks
the plain CA does not find).  Prints one JSON line.
"""
import argparse
import json
import os
import subprocess
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "tracking-itsu_amd"))

HBM_PEAK_GBS = 8000.0
STAGES = ["prep+index tables", "tracklets", "cells", "neighbours", "levels+roads"]
METRIC = "events/s on 1 MI355X, Pb-Pb-like 8k tracks: plain GPU vs batched+learned filter"


def bytes_alg(c):
    """Compulsory HBM bytes of one (event, vertex): every stage reads its inputs
    once and writes its outputs once, at the reference record sizes (Cluster 36,
    Tracklet 16, Cell 40, Road 32, index/level/edge 4 bytes)."""
    N, T, C, E, R = list(c.clusters), list(c.tracklets), list(c.cells), list(c.neighbours), c.roads
    b = sum(n * (16 + 36) for n in N) + 6 * 401 * 4
    b += sum((N[l] + N[l + 1]) * 36 + T[l] * 16 + N[l + 1] * 4 for l in range(6))
    b += sum((T[l] + T[l + 1]) * 16 + (N[l] + N[l + 1] + N[l + 2]) * 36 + C[l] * 40 + T[l + 1] * 4 for l in range(5))
    b += sum((C[l] + C[l + 1]) * 40 + E[l] * 4 + C[l + 1] * 4 for l in range(4))
    b += sum(C) * 4 + sum(E) * 4 + R * 32
    return b


def kernel_bytes(name, c, kept=None):
    """Algorithmic HBM bytes of one launch of kernel `name` for one work item,
    from the same record sizes (plain-CA counts; `kept`: the learned form's
    counts, for the filter): each kernel reads its inputs once and writes its
    outputs once."""
    N, T, C, E, R = list(c.clusters), list(c.tracklets), list(c.cells), list(c.neighbours), c.roads
    if name == "kPrep":  # raw x, y, z, id, mc in; sorted cluster + 401-entry table out
        return sum(n * (20 + 36) for n in N) + 7 * 401 * 4
    if name == "kTrackletsCount":  # both layers' clusters in; one count per cluster out
        return sum((N[l] + N[l + 1]) * 36 + N[l] * 4 for l in range(6))
    if name == "kTrackletsScore":  # candidate slots in, both cluster layers in once; kept slots and counts out
        K = list(kept.tracklets) if kept is not None else T
        return sum(T[l] * 4 + (N[l] + N[l + 1]) * 36 + K[l] * 4 + N[l] * 4 for l in range(6))
    if name == "kTrackletsFill":  # tracklet records and per-cluster offsets out
        return sum(T[l] * 16 + N[l] * 4 for l in range(6))
    if name == "kCellsCount":  # two tracklet layers and three cluster layers in; one count per tracklet out
        return sum((T[l] + T[l + 1]) * 16 + (N[l] + N[l + 1] + N[l + 2]) * 36 + T[l] * 4 for l in range(5))
    if name == "kCellsFill":  # cell records and per-tracklet offsets out
        return sum(C[l] * 40 + T[l] * 4 for l in range(5))
    if name in ("kEdgesCount", "kEdgesFill"):  # two cell layers in; edges and counts out
        return sum((C[l] + C[l + 1]) * 40 + E[l] * 4 + C[l + 1] * 4 for l in range(4))
    if name == "kLevels":
        return sum(C) * 4 + sum(E) * 4
    if name == "kRoadsFill":
        return R * 32 + sum(E) * 4
    return 0


STAGE_KERNELS = {"prep+index tables": ("kPrep",), "tracklets": ("kTrackletsCount", "kTrackletsScore", "kTrackletsFill"),
                 "cells": ("kCellsCount", "kCellsFill"), "neighbours": ("kEdgesCount", "kEdgesFill", "kEdgesTranspose"),
                 "levels+roads": ("kLevels", "kRoadsFill")}


def stage_bandwidth(counts, kernel_ms, kept=None):
    """Per stage: kernel time (HIP events), algorithmic bytes (the per-kernel
    figures of kernel_bytes summed over the stage's kernels and the batch) and
    their ratio as GB/s and as a fraction of the 8 TB/s HBM peak."""
    out = {}
    kept = kept or [None] * len(counts)
    for stage, kernels in STAGE_KERNELS.items():
        ms = sum(kernel_ms.get(k, 0.0) for k in kernels)
        b = sum(kernel_bytes(k, c, kk) for k in kernels if k in kernel_ms for c, kk in zip(counts, kept))
        if ms > 0:
            gbs = b / (ms * 1e-3) / 1e9
            out[stage] = {"ms": round(ms, 4), "bytes_alg": b, "GBps": round(gbs, 1),
                          "frac_of_peak": round(gbs / HBM_PEAK_GBS, 4)}
    return out


TRAFFIC_FILE = os.path.join(ROOT, "profiles", "ca_central_batch256_hbm_traffic.md")


HELD_OUT_SEED = 20000  # first seed of the held-out central events
MIXED_SEED = 3000      # first seed of the mixed-multiplicity batch
PILEUP_SEED = 50000    # first seed of the pile-up batch
TRAFFIC_BATCH = 256     # the workload of the counter runs in TRAFFIC_FILE


def measured_traffic(kernel, batch, seed):
    """HBM bytes of one launch of `kernel` (learned form), from the calibrated
    FETCH_SIZE + WRITE_SIZE counter runs summarised in TRAFFIC_FILE; those runs
    used this benchmark's default workload (256 central events, seeds
    20000..20255), so the figure is reported only for that workload."""
    if batch != TRAFFIC_BATCH or seed != HELD_OUT_SEED or not os.path.exists(TRAFFIC_FILE):
        return None
    name = kernel.split("<")[0]
    with open(TRAFFIC_FILE) as f:
        for line in f:
            parts = line.split()
            if len(parts) == 3 and parts[0] == "traffic_bytes" and parts[1] == name:
                return int(parts[2])
    return None


def cpu_baseline(itsca, tmp, n_events, seed, per_core):
    """The reference CPU tracker itself (built from its sources into
    oracle/_ref/) on a bounded sample of the same workload: one independent
    tracker process per host core (at most 16, the host share of one GPU),
    each on its own slice of per_core events, timed together (events over the
    slowest process's time); and the same binary on one core."""
    exe = os.path.join(ROOT, "oracle", "_ref", "ref_harness")
    if os.path.exists(exe):
        events_path = os.path.join(tmp, "cpu_sample.txt")
        itsca.write_events(2, seed, n_events, events_path)
        out = subprocess.run([exe, "time", events_path, "1"], check=True, capture_output=True, text=True).stdout
        single = json.loads(out.strip().splitlines()[-1])
        cores = max(1, min(16, os.cpu_count() or 1))
        paths = []
        for k in range(cores):
            p = os.path.join(tmp, "cpu_core%d.txt" % k)
            itsca.write_events(2, seed + k * per_core, per_core, p)
            paths.append(p)
        t0 = time.perf_counter()
        procs = [subprocess.Popen([exe, "time", p, "1"], stdout=subprocess.PIPE, text=True) for p in paths]
        res = [json.loads(pr.communicate()[0].strip().splitlines()[-1]) for pr in procs]
        wall = time.perf_counter() - t0
        slowest = max(r["seconds"] for r in res)
        total = sum(r["events"] for r in res)
        return {"value": round(total / slowest, 3), "unit": "events/s", "cores": cores, "kind": "reference",
                "sample": "%d central events (seeds %d..%d), %d per process, one process per core, each a single "
                          "pass after one warm-up event; slowest process %.1f s (wall %.1f s)"
                          % (total, seed, seed + cores * per_core - 1, per_core, slowest, wall),
                "single_core": {"value": round(single["events"] / single["seconds"], 3), "cores": 1,
                                "sample": "%d central events (seeds %d..%d), %.1f s"
                                          % (n_events, seed, seed + n_events - 1, single["seconds"])}}
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle  # noqa: E402
    evs = itsca.generate_events(2, seed, n_events)
    cores = max(1, min(16, os.cpu_count() or 1))
    s = pyoracle.time_events(evs, 1, cores)
    return {"value": round(len(evs) / s, 3), "unit": "events/s", "cores": cores, "kind": "port",
            "sample": "%d central events, single pass, one tracker per thread" % len(evs)}


def tuning_dict(t):
    return {"block_size": t.block_size, "persistent_blocks": t.persistent_blocks, "z_subbins": t.z_subbins,
            "phi_subrows": t.phi_subrows}


def time_form(tracker, steps, warmup):
    """Events are already resident; the pipeline is replayed as a HIP graph."""
    tracker.set_graph(False)
    tracker.run()  # sizes the buffers for this form
    tracker.set_graph(True)
    for _ in range(max(warmup, 1)):
        tracker.run_async()
    tracker.finish()
    t0 = time.perf_counter()
    for _ in range(steps):
        tracker.run_async()
    tracker.finish()
    elapsed = time.perf_counter() - t0
    tracker.set_graph(False)
    return elapsed


def kernel_profile(tracker, runs=5):
    """Mean per-kernel HIP-event times (ms) over `runs` profiled runs, recorded
    on the library's stream (the stream every kernel is launched on); kernels
    launched several times per run are summed."""
    tracker.set_profiling(True)
    kernel_ms, launches = {}, {}
    stage_ms = [0.0] * 5
    for _ in range(runs):
        tracker.run()
        stage_ms = [a + b / runs for a, b in zip(stage_ms, tracker.stage_times().tolist())]
        seen = {}
        for name, ms in tracker.kernel_times():
            kernel_ms[name] = kernel_ms.get(name, 0.0) + ms / runs
            seen[name] = seen.get(name, 0) + 1
        launches = seen
    tracker.set_profiling(False)
    return kernel_ms, launches, stage_ms


def road_report(itsca, tracker, events, events_path, labels_path):
    """Correct, duplicate and fake roads against the MC labels map, classified
    like the reference's road report (IOUtils.cxx:80-170; pions with 7
    clusters), summed over the events of the batch."""
    roads, item = [], 0
    for ev in events:
        roads.append([tracker.roads(item + v) for v in range(len(ev.vertices))])
        item += len(ev.vertices)
    counts = itsca.roads_report(events_path, labels_path, roads)
    return {"correct": sum(c[0] for c in counts), "duplicate": sum(c[1] for c in counts),
            "fake": sum(c[2] for c in counts)}


def end_to_end(itsca, events, params, threshold, gates, batches=6):
    """The whole path a caller of the reference entry sees, per batch of
    events held in host memory: events in (itsca_add_event + itsca_upload),
    the pipeline, roads out to host memory (itsca_download_roads), no graph.
    Two contexts on two host threads, each with its own HIP stream, take
    alternate batches, so one batch's copies overlap the other's kernels."""
    import threading
    out = {}
    for form in ("plain", "learned"):
        trackers = [itsca.GpuTracker(device=0) for _ in range(2)]
        roads = [0, 0]
        for t in trackers:
            t.set_mlp_params(params)
            t.set_filter_gates(gates)
            t.set_filter(form == "learned", threshold)
            t.set_events(events)  # sizes the buffers
            t.run()
            t.all_roads()

        def worker(k):
            t = trackers[k]
            for _ in range(k, batches, 2):
                t.set_events(events)
                t.run()
                rows, _ = t.all_roads()
                roads[k] += len(rows)

        t0 = time.perf_counter()
        threads = [threading.Thread(target=worker, args=(k,)) for k in range(2)]
        for th in threads:
            th.start()
        for th in threads:
            th.join()
        elapsed = time.perf_counter() - t0
        for t in trackers:
            t.close()
        out[form] = {"events_per_s": round(batches * len(events) / elapsed, 2),
                     "ms_per_batch": round(1e3 * elapsed / batches, 3), "roads_to_host": sum(roads)}
    out["workload"] = ("%d batches of the %d central events, host memory to host memory: upload, run, roads "
                       "download; two contexts on two threads alternating batches" % (batches, len(events)))
    return out


def cpp_api_latency(itsca, tmp, seed, n=12):
    """Tracker<true>::clustersToTracks(event), the C++ API a reference caller
    uses, one event per call (upload, run, roads back by value), through
    bin/itsca-track: mean ms per central event after its warm-up event."""
    import re
    path = os.path.join(tmp, "api_events.txt")
    itsca.write_events(2, seed, n, path)
    exe = os.path.join(itsca.BIN_DIR, "itsca-track")
    out = {}
    for form, extra in (("plain", []), ("learned", ["--filter", os.path.join(ROOT, "tracking-itsu_amd", "itsca",
                                                                           "filter_weights.json")])):
        r = subprocess.run([exe, path] + extra, capture_output=True, text=True, timeout=300)
        if r.returncode != 0:
            out[form] = {"error": r.stdout[-300:] + r.stderr[-300:]}
            continue
        ms = [float(x) for x in re.findall(r"processed in: ([0-9.eE+-]+)ms", r.stdout)]
        out[form] = {"ms_per_event": round(sum(ms) / len(ms), 4), "events": len(ms),
                     "events_per_s": round(1e3 * len(ms) / sum(ms), 2)}
    out["workload"] = "%d central events (seeds %d..%d), one Tracker<true>::clustersToTracks call each" % (
        n, seed, seed + n - 1)
    return out


def single_events(itsca, steps=50):
    """One event per launch, plain form, graph replay, autotuned: the latency
    of a small event (100 tracks) and of one central event (8000 tracks)."""
    out = {}
    for key, preset, seed in (("small_100_tracks", 1, 1), ("central_8000_tracks", 2, HELD_OUT_SEED)):
        t = itsca.GpuTracker(device=0)
        t.set_events(itsca.generate_events(preset, seed, 1))
        tuning, _ = t.autotune(5)
        s = time_form(t, steps, 3)
        t.close()
        out[key] = {"ms_per_event": round(1e3 * s / steps, 4), "events_per_s": round(steps / s, 1),
                    "form": "plain, one event per launch, hipGraph replay", "seed": seed,
                    "tuning": tuning_dict(tuning)}
    return out


def mixed_batch(itsca, steps, tuning_fixed=None):
    """256 events of log-uniform multiplicity in [50, 8000] tracks in one
    launch, plain form, launch shape chosen by the autotuner, graph replay;
    per-stage kernel times and algorithmic GB/s."""
    evs = itsca.generate_events(3, MIXED_SEED, 256)
    t = itsca.GpuTracker(device=0)
    t.set_events(evs)
    t.run()
    if tuning_fixed:
        t.set_tuning(*[int(x) for x in tuning_fixed.split(",")])
        tuning = t.tuning()
    else:
        tuning, _ = t.autotune(3)
    s = time_form(t, steps, 2)
    counts = [t.counts(i) for i in range(t.num_items())]
    kernel_ms, _, _ = kernel_profile(t)
    out = {"workload": "256 events, multiplicity log-uniform in [50, 8000] tracks (seeds %d..%d), one launch" % (MIXED_SEED, MIXED_SEED + 255),
           "form": "plain, hipGraph replay", "events_per_s": round(256 * steps / s, 2),
           "ms_per_step": round(1e3 * s / steps, 4),
           "tuning": tuning_dict(tuning),
           "kernel_ms": {k: round(v, 4) for k, v in sorted(kernel_ms.items(), key=lambda kv: -kv[1])},
           "stage_bandwidth": stage_bandwidth(counts, kernel_ms),
           "bytes_alg_per_step": sum(bytes_alg(c) for c in counts)}
    t.close()
    return out


BINNINGS = [(1, 1), (2, 1), (4, 1), (8, 1), (1, 2), (2, 2), (4, 2), (1, 4), (2, 4), (3, 4)]


def binning_sweep(itsca, tracker, threshold, steps=10):
    """The search-index binnings (z sub-bins x phi sub-rows) on the uploaded
    central batch, both forms, and on the mixed-multiplicity batch (plain), at
    block 256 and 16 workgroups per CU: ms per step (graph replay) and the
    tracklet count's ms (HIP events)."""
    def sweep(t, forms):
        rows = []
        for form in forms:
            if form != "mixed":
                t.set_filter(form == "learned", threshold)
            for zs, ps in BINNINGS:
                t.set_tuning(256, 0, zs, ps)
                s = time_form(t, steps, 2)
                km, _, _ = kernel_profile(t, 3)
                rows.append({"form": form, "z_subbins": zs, "phi_subrows": ps, "ms_per_step": round(1e3 * s / steps, 4),
                             "kTrackletsCount_ms": round(km.get("kTrackletsCount", 0.0), 4),
                             "kPrep_ms": round(km.get("kPrep", 0.0), 4)})
        return rows
    out = {"central": sweep(tracker, ("plain", "learned"))}
    evs = itsca.generate_events(3, MIXED_SEED, 256)
    t = itsca.GpuTracker(device=0)
    t.set_events(evs)
    out["mixed"] = sweep(t, ("mixed",))
    t.close()
    return out


def pileup_batch(itsca, ft, params, threshold, gates, steps, events=64):
    """Pile-up events (4 collisions, z ~ U(-10, 10) cm, 2000-8000 primaries
    each, 3 % noise), every vertex its own work item, plain and learned."""
    with tempfile.TemporaryDirectory() as tmp:
        ev_path, lab_path = os.path.join(tmp, "pileup.txt"), os.path.join(tmp, "pileup_labels.txt")
        itsca.write_events(5, PILEUP_SEED, events, ev_path, lab_path)
        evs = itsca.load_events(ev_path)
        t = itsca.GpuTracker(device=0)
        t.set_events(evs)
        t.set_mlp_params(params)
        t.set_filter_gates(gates)
        t.set_filter(False)
        s_plain = time_form(t, steps, 1)
        q_plain = road_report(itsca, t, evs, ev_path, lab_path)
        t.set_filter(True, threshold)
        s_learned = time_form(t, steps, 1)
        q_learned = road_report(itsca, t, evs, ev_path, lab_path)
        quality = ft.evaluate(t, evs, threshold)
        t.close()
    return {"workload": "%d pile-up events (4 vertices each, seeds %d..%d), one launch" % (events, PILEUP_SEED, PILEUP_SEED + events - 1),
            "plain_events_per_s": round(events * steps / s_plain, 2),
            "learned_events_per_s": round(events * steps / s_learned, 2),
            "speedup": round(s_plain / s_learned, 4), "track_recall": round(quality["recall"], 6),
            "plain_tracks": quality["plain_tracks"], "tracks_outside_plain": quality["extra"],
            "road_report": {"plain": q_plain, "learned": q_learned,
                            "method": "IOUtils::writeRoadsReport classification against the labels map"},
            "tracklets_kept": round(quality["tracklets_kept"], 4)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--batch", type=int, default=256, help="central events per step")
    ap.add_argument("--seed", type=int, default=HELD_OUT_SEED, help="first seed of the central events")
    ap.add_argument("--cpu-events", type=int, default=200, help="events in the single-core CPU-baseline sample")
    ap.add_argument("--cpu-events-per-core", type=int, default=40,
                    help="events per process in the all-cores CPU baseline")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-extra", action="store_true", help="skip the mixed-multiplicity and pile-up batches")
    ap.add_argument("--weights", default=None, help="learned filter weights (default: the shipped file)")
    ap.add_argument("--tuning-plain", default=None, metavar="B,P,Z[,F]",
                    help="fixed block size, persistent workgroups, z sub-bins and phi sub-rows of the plain form "
                         "instead of the autotuner's choice (for profiler runs)")
    ap.add_argument("--tuning-learned", default=None, metavar="B,P,Z[,F]", help="the same for the learned form")
    ap.add_argument("--only", choices=["central", "mixed"], default=None,
                    help="run one workload alone (for profiler runs): the central batch (both forms) or the "
                         "mixed-multiplicity batch")
    ap.add_argument("--binning-sweep", action="store_true",
                    help="also time every search-index binning (z sub-bins x phi sub-rows) on the central and the "
                         "mixed batches")
    ap.add_argument("--copy-bytes", type=int, default=0,
                    help="first run one device-to-device copy kernel of this many bytes (a reference dispatch of "
                         "known traffic for hardware-counter runs)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU only: every (event, vertex) is independent, so N GPUs would each run "
              "their own share; run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    import itsca
    from itsca import filter_training as ft

    if args.copy_bytes:
        import torch
        src = torch.ones(args.copy_bytes // 4, dtype=torch.float32, device="cuda:0")
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        print("# copy kernel: %d bytes read, %d bytes written" % (args.copy_bytes, args.copy_bytes), file=sys.stderr,
              flush=True)
        del src, dst

    if args.only == "mixed":
        print(json.dumps({"mixed_batch": mixed_batch(itsca, args.steps, args.tuning_plain)}), flush=True)
        return

    t_gen = time.time()
    events = itsca.generate_events(2, args.seed, args.batch)
    tracker = itsca.GpuTracker(device=0)
    tracker.set_events(events)
    params, threshold, meta = ft.load_filter(args.weights)
    gates = ft.load_gates(args.weights)
    tracker.set_mlp_params(params)
    tracker.set_filter_gates(gates)
    print("# generated and uploaded %d events in %.1f s" % (len(events), time.time() - t_gen), file=sys.stderr,
          flush=True)

    def tuned(form):  # the autotuner's choice for this form on this batch (results never depend on it)
        fixed = args.tuning_plain if form == "plain" else args.tuning_learned
        if fixed:
            tracker.set_tuning(*[int(x) for x in fixed.split(",")])
            t = tracker.tuning()
        else:
            t, _ = tracker.autotune(2)
        print("# %s: tuning block %d, %d workgroups, %d z sub-bins, %d phi sub-rows"
              % (form, t.block_size, t.persistent_blocks, t.z_subbins, t.phi_subrows), file=sys.stderr, flush=True)
        return tuning_dict(t)

    tracker.set_filter(False)
    plain_tuning = tuned("plain")
    plain_s = time_form(tracker, args.steps, args.warmup)
    counts = [tracker.counts(i) for i in range(tracker.num_items())]  # plain-CA counts
    plain_kernels, _, plain_stages = kernel_profile(tracker)
    print("# plain: %.1f events/s" % (args.batch * args.steps / plain_s), file=sys.stderr, flush=True)

    tracker.set_filter(True, threshold)
    learned_tuning = tuned("learned")
    learned_s = time_form(tracker, args.steps, args.warmup)
    kernel_ms, launches, stage_ms = kernel_profile(tracker)
    print("# learned: %.1f events/s" % (args.batch * args.steps / learned_s), file=sys.stderr, flush=True)

    kept = [tracker.counts(i) for i in range(tracker.num_items())]  # learned-form counts
    quality = ft.evaluate(tracker, events, threshold)  # plain and filtered roads of the same batch
    quality.pop("plain")
    tracker.set_filter(False)

    ms_per_step = 1e3 * learned_s / args.steps
    events_per_s = args.batch * args.steps / learned_s
    plain_eps = args.batch * args.steps / plain_s
    balg = sum(bytes_alg(c) for c in counts)
    # roofline of the learned form's dominant kernel: algorithmic bytes of one
    # launch (one launch covers the whole batch) over its mean HIP-event time
    top = max(kernel_ms, key=kernel_ms.get)
    top_bytes = sum(kernel_bytes(top, c, k) for c, k in zip(counts, kept))
    top_ms = kernel_ms[top] / launches.get(top, 1)
    achieved = top_bytes / (top_ms * 1e-3) / 1e9

    line = {
        "metric": METRIC,
        "value": round(events_per_s, 2),
        "unit": "events/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic",
        "config": {"workload": "central Pb-Pb-like events (8000 tracks, ~8.3e3 clusters/layer, held-out seeds "
                               "%d..%d), %d events per step, inputs resident in HBM, pipeline replayed as one HIP graph"
                               % (args.seed, args.seed + args.batch - 1, args.batch),
                   "events_per_step": args.batch,
                   "form": "batched CA + learned tracklet filter (10-32-32-1, BF16 MFMA, FP32 accumulate)",
                   "filter_threshold": threshold,
                   "filter_phi_gates": gates,
                   "filter_training": meta.get("training_events"),
                   "tuning": learned_tuning},
        "plain": {"value": round(plain_eps, 2), "unit": "events/s", "ms_per_step": round(1e3 * plain_s / args.steps, 4),
                  "tuning": plain_tuning,
                  "kernel_ms": {k: round(v, 4) for k, v in sorted(plain_kernels.items(), key=lambda kv: -kv[1])},
                  "stage_ms": {STAGES[k]: round(plain_stages[k], 4) for k in range(5)},
                  "stage_bandwidth": stage_bandwidth(counts, plain_kernels)},
        "learned_vs_plain": {"speedup": round(events_per_s / plain_eps, 4),
                             "track_recall": round(quality["recall"], 6), "plain_tracks": quality["plain_tracks"],
                             "tracks_found": quality["found"], "tracks_outside_plain": quality["extra"],
                             "tracklets_kept": round(quality["tracklets_kept"], 4),
                             "cells_kept": round(quality["cells_kept"], 4),
                             "road_cells_kept": round(quality["road_cells_kept"], 6)},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 5),
                     "traffic": measured_traffic(top, args.batch, args.seed),
                     "kernel": top, "kernel_ms": round(top_ms, 4), "kernel_bytes_alg": top_bytes,
                     "pipeline_bytes_alg_per_step": balg,
                     "pipeline_GBps": round(balg / (ms_per_step * 1e-3) / 1e9, 2)},
        "kernel_ms": {k: round(v, 4) for k, v in sorted(kernel_ms.items(), key=lambda kv: -kv[1])},
        "stage_ms": {STAGES[k]: round(stage_ms[k], 4) for k in range(5)},
        "stage_bandwidth": stage_bandwidth(counts, kernel_ms, kept),
        "counts_event0": {"tracklets": list(counts[0].tracklets), "cells": list(counts[0].cells),
                          "roads": counts[0].roads},
    }
    if args.binning_sweep:
        line["binning_sweep"] = binning_sweep(itsca, tracker, threshold)
        print("# binning_sweep done", file=sys.stderr, flush=True)
    tracker.close()
    if args.only == "central":
        print(json.dumps(line), flush=True)
        return
    if not args.no_extra:
        tmpdir = tempfile.TemporaryDirectory()
        for key, fn in (("single_event", lambda: single_events(itsca)),
                        ("mixed_batch", lambda: mixed_batch(itsca, 5)),
                        ("pileup_batch", lambda: pileup_batch(itsca, ft, params, threshold, gates, 3)),
                        ("end_to_end", lambda: end_to_end(itsca, events, params, threshold, gates)),
                        ("cpp_api_per_event", lambda: cpp_api_latency(itsca, tmpdir.name, args.seed))):
            try:
                line[key] = fn()
            except Exception as e:  # reported in the line; the main measurement stands
                line[key] = {"error": "%s: %s" % (type(e).__name__, e)}
            print("# %s done" % key, file=sys.stderr, flush=True)
        tmpdir.cleanup()
    if not args.no_cpu:
        with tempfile.TemporaryDirectory() as tmp:
            line["cpu_baseline"] = cpu_baseline(itsca, tmp, args.cpu_events, args.seed, args.cpu_events_per_core)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
