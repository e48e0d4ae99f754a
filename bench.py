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


TRAFFIC_FILE = os.path.join(ROOT, "profiles", "ca_central_batch256_hbm_traffic.md")


HELD_OUT_SEED = 20000  # first seed of the held-out central events
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


def cpu_baseline(events_path, n_events, seed):
    """The reference CPU tracker itself (built from its sources into
    oracle/_ref/), single thread, on a bounded sample of the same workload."""
    exe = os.path.join(ROOT, "oracle", "_ref", "ref_harness")
    if os.path.exists(exe):
        out = subprocess.run([exe, "time", events_path, "1"], check=True, capture_output=True, text=True).stdout
        r = json.loads(out.strip().splitlines()[-1])
        return {"value": round(r["events"] / r["seconds"], 3), "unit": "events/s", "cores": 1, "kind": "reference",
                "sample": "%d central events (seeds %d..%d), single pass after one warm-up event, %.1f s"
                          % (n_events, seed, seed + n_events - 1, r["seconds"])}
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle  # noqa: E402
    import itsca  # noqa: E402
    evs = itsca.load_events(events_path)
    s = pyoracle.time_events(evs, 1, 1)
    return {"value": round(len(evs) / s, 3), "unit": "events/s", "cores": 1, "kind": "port",
            "sample": "%d central events, single pass" % len(evs)}


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


def road_quality(tracker):
    """Fake roads (Moore vote of the MC labels, Tracker.cxx:573-652), and
    distinct MC labels among the other roads, over all work items."""
    fake = correct = 0
    for i in range(tracker.num_items()):
        r = tracker.roads(i)
        if len(r) == 0:
            continue
        fake += int((r[:, 7] != 0).sum())
        good = r[r[:, 7] == 0][:, 6]
        correct += len(set(int(x) for x in good.tolist() if x >= 0))
    return fake, correct


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
                    "tuning": {"block_size": tuning.block_size, "persistent_blocks": tuning.persistent_blocks,
                               "z_subbins": tuning.z_subbins}}
    return out


def mixed_batch(itsca, steps):
    """256 events of log-uniform multiplicity in [50, 8000] tracks in one
    launch, plain form, launch shape chosen by the autotuner, graph replay."""
    evs = itsca.generate_events(3, 1000, 256)
    t = itsca.GpuTracker(device=0)
    t.set_events(evs)
    t.run()
    tuning, _ = t.autotune(3)
    s = time_form(t, steps, 2)
    out = {"workload": "256 events, multiplicity log-uniform in [50, 8000] tracks (seeds 1000..1255), one launch",
           "form": "plain, hipGraph replay", "events_per_s": round(256 * steps / s, 2),
           "ms_per_step": round(1e3 * s / steps, 4),
           "tuning": {"block_size": tuning.block_size, "persistent_blocks": tuning.persistent_blocks,
                      "z_subbins": tuning.z_subbins}}
    t.close()
    return out


def pileup_batch(itsca, ft, params, threshold, steps, events=64):
    """Pile-up events (4 collisions, z ~ U(-10, 10) cm, 2000-8000 primaries
    each, 3 % noise), every vertex its own work item, plain and learned."""
    evs = itsca.generate_events(5, 40000, events)
    t = itsca.GpuTracker(device=0)
    t.set_events(evs)
    t.set_mlp_params(params)
    t.set_filter(False)
    s_plain = time_form(t, steps, 1)
    q_plain = road_quality(t)
    t.set_filter(True, threshold)
    s_learned = time_form(t, steps, 1)
    q_learned = road_quality(t)
    quality = ft.evaluate(t, evs, threshold)
    t.close()
    return {"workload": "%d pile-up events (4 vertices each, seeds 40000..%d), one launch" % (events, 40000 + events - 1),
            "plain_events_per_s": round(events * steps / s_plain, 2),
            "learned_events_per_s": round(events * steps / s_learned, 2),
            "speedup": round(s_plain / s_learned, 4), "track_recall": round(quality["recall"], 6),
            "plain_tracks": quality["plain_tracks"], "tracks_outside_plain": quality["extra"],
            "fake_roads": {"plain": q_plain[0], "learned": q_learned[0]},
            "correct_labels": {"plain": q_plain[1], "learned": q_learned[1]},
            "tracklets_kept": round(quality["tracklets_kept"], 4)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--batch", type=int, default=256, help="central events per step")
    ap.add_argument("--seed", type=int, default=HELD_OUT_SEED, help="first seed of the central events")
    ap.add_argument("--cpu-events", type=int, default=200, help="events in the CPU-baseline sample")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-extra", action="store_true", help="skip the mixed-multiplicity and pile-up batches")
    ap.add_argument("--weights", default=None, help="learned filter weights (default: the shipped file)")
    ap.add_argument("--tuning-plain", default=None, metavar="B,P,Z",
                    help="fixed block size, persistent workgroups and z sub-bins of the plain form instead of the "
                         "autotuner's choice (for profiler runs)")
    ap.add_argument("--tuning-learned", default=None, metavar="B,P,Z", help="the same for the learned form")
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

    t_gen = time.time()
    events = itsca.generate_events(2, args.seed, args.batch)
    tracker = itsca.GpuTracker(device=0)
    tracker.set_events(events)
    params, threshold, meta = ft.load_filter(args.weights)
    tracker.set_mlp_params(params)
    print("# generated and uploaded %d events in %.1f s" % (len(events), time.time() - t_gen), file=sys.stderr,
          flush=True)

    def tuned(form):  # the autotuner's choice for this form on this batch (results never depend on it)
        fixed = args.tuning_plain if form == "plain" else args.tuning_learned
        if fixed:
            tracker.set_tuning(*[int(x) for x in fixed.split(",")])
            t = tracker.tuning()
        else:
            t, _ = tracker.autotune(2)
        print("# %s: tuning block %d, %d workgroups, %d z sub-bins" % (form, t.block_size, t.persistent_blocks,
                                                                       t.z_subbins), file=sys.stderr, flush=True)
        return {"block_size": t.block_size, "persistent_blocks": t.persistent_blocks, "z_subbins": t.z_subbins}

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
                   "filter_training": meta.get("training_events"),
                   "tuning": learned_tuning},
        "plain": {"value": round(plain_eps, 2), "unit": "events/s", "ms_per_step": round(1e3 * plain_s / args.steps, 4),
                  "tuning": plain_tuning,
                  "kernel_ms": {k: round(v, 4) for k, v in sorted(plain_kernels.items(), key=lambda kv: -kv[1])},
                  "stage_ms": {STAGES[k]: round(plain_stages[k], 4) for k in range(5)}},
        "learned_vs_plain": {"speedup": round(events_per_s / plain_eps, 4),
                             "track_recall": round(quality["recall"], 6), "plain_tracks": quality["plain_tracks"],
                             "tracks_found": quality["found"], "tracks_outside_plain": quality["extra"],
                             "tracklets_kept": round(quality["tracklets_kept"], 4),
                             "cells_kept": round(quality["cells_kept"], 4)},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 5),
                     "traffic": measured_traffic(top, args.batch, args.seed),
                     "kernel": top, "kernel_ms": round(top_ms, 4), "kernel_bytes_alg": top_bytes,
                     "pipeline_bytes_alg_per_step": balg,
                     "pipeline_GBps": round(balg / (ms_per_step * 1e-3) / 1e9, 2)},
        "kernel_ms": {k: round(v, 4) for k, v in sorted(kernel_ms.items(), key=lambda kv: -kv[1])},
        "stage_ms": {STAGES[k]: round(stage_ms[k], 4) for k in range(5)},
        "counts_event0": {"tracklets": list(counts[0].tracklets), "cells": list(counts[0].cells),
                          "roads": counts[0].roads},
    }
    tracker.close()
    if not args.no_extra:
        for key, fn in (("single_event", lambda: single_events(itsca)),
                        ("mixed_batch", lambda: mixed_batch(itsca, 5)),
                        ("pileup_batch", lambda: pileup_batch(itsca, ft, params, threshold, 3))):
            try:
                line[key] = fn()
            except Exception as e:  # reported in the line; the main measurement stands
                line[key] = {"error": "%s: %s" % (type(e).__name__, e)}
            print("# %s done" % key, file=sys.stderr, flush=True)
    if not args.no_cpu:
        with tempfile.TemporaryDirectory() as tmp:
            p = os.path.join(tmp, "cpu_sample.txt")
            itsca.write_events(2, args.seed, args.cpu_events, p)
            line["cpu_baseline"] = cpu_baseline(p, args.cpu_events, args.seed)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
