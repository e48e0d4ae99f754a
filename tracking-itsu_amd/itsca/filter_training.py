"""Training of the learned tracklet filter on the GPU, and its weights file.

The filter has two parts, both fitted here:
- per tracklet layer, a phi gate of the tracklet search: 1.1 times the
  largest |dphi| of the positive tracklets of the first gate-sample events
  (capped at the reference's 0.3 rad cut).  Filtered runs search and keep only the
  candidates inside it;
- the network, which scores the gated candidates (its count features are the
  gated counts).

The labels come from CA runs: after a run every tracklet is labelled 1 when it
is one of the two tracklets of a cell on a final road, else 0
(itsca_train_collect).  The gates are fitted on plain runs; the training rows
then come from gated runs whose threshold keeps every candidate.  Training runs
the device's own forward/backward and AdamW kernels (itsca_train_epoch).  The
logit threshold is then tuned on separate validation events by running the
filtered pipeline and measuring the fraction of plain roads it still finds.

    python -m itsca.filter_training --out weights.json
"""
import argparse
import json
import math
import os
import sys
import time
from collections import Counter
from typing import Dict, List, Optional, Sequence

import numpy as np

from . import MLP_PARAMS, GpuTracker, cell_keys, generate_events, road_tracks

DEFAULT_WEIGHTS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "filter_weights.json")


def save_filter(path: str, params: np.ndarray, threshold: float, meta: Dict, gates=None) -> None:
    p = np.ascontiguousarray(params, dtype=np.float32)
    meta = {k: v for k, v in meta.items() if k != "phi_gates"}  # the gates are kept once, at the top level
    doc = {"format": "itsca tracklet filter 10-32-32-1", "params_f32_bits": ["%08x" % v for v in p.view(np.uint32)],
           "threshold": float(threshold), "meta": meta}
    if gates is not None:
        doc["phi_gates"] = [float(np.float32(g)) for g in gates]
    with open(path, "w") as f:
        json.dump(doc, f, indent=1)


def load_filter(path: Optional[str] = None):
    """(params float32[1441], threshold, meta) of a weights file (default: the
    file shipped with the package)."""
    with open(path or DEFAULT_WEIGHTS) as f:
        doc = json.load(f)
    p = np.array([int(h, 16) for h in doc["params_f32_bits"]], dtype=np.uint32).view(np.float32)
    if p.size != MLP_PARAMS:
        raise ValueError("weights file holds %d parameters, expected %d" % (p.size, MLP_PARAMS))
    return p, float(doc["threshold"]), doc.get("meta", {})


def load_gates(path: Optional[str] = None):
    """The per-layer phi gates of a weights file (float32 values), or None
    when it has none."""
    with open(path or DEFAULT_WEIGHTS) as f:
        doc = json.load(f)
    g = doc.get("phi_gates")
    return None if g is None else [float(np.float32(v)) for v in g]


def prepare(tracker: GpuTracker, path: Optional[str] = None) -> float:
    """Loads a weights file's parameters and phi gates into the tracker (the
    filter stays as it is); returns its threshold."""
    p, thr, _ = load_filter(path)
    tracker.set_mlp_params(p)
    tracker.set_filter_gates(load_gates(path))
    return thr


def use_filter(tracker: GpuTracker, path: Optional[str] = None) -> float:
    """Loads a weights file into the tracker and turns the filter on."""
    thr = prepare(tracker, path)
    tracker.set_filter(True, thr)
    return thr


def _log(msg):
    print(msg, file=sys.stderr, flush=True)


KEEP_ALL = -3.0e38  # a threshold that keeps every candidate the gates let through


def collect(tracker: GpuTracker, preset: int, seeds: Sequence[int], stride: int, capacity: int, batch: int = 25,
            gates=None) -> int:
    """Runs over the events of `seeds` (in batches): plain, or with the phi
    gates and a threshold that keeps every gated candidate; every stride-th
    tracklet becomes a training row (the phase rotates between batches).
    Gated runs score with the tracker's current parameters, which must be
    finite."""
    if gates is None:
        tracker.set_filter(False)
    else:
        tracker.set_filter_gates(gates)
        tracker.set_filter(True, KEEP_ALL)
    tracker.train_clear(capacity)
    seeds = list(seeds)
    for k in range(0, len(seeds), batch):
        evs = [e for s in seeds[k:k + batch] for e in generate_events(preset, s, 1)]
        tracker.set_events(evs)
        tracker.run()
        tracker.train_collect(stride, (k // batch) % stride)
        _log("  collected events %d..%d: %d rows" % (seeds[k], seeds[min(k + batch, len(seeds)) - 1],
                                                   tracker.train_size()))
    tracker.set_filter(False)
    return tracker.train_size()


GATE_MARGIN = 1.1  # gate = this x the largest |dphi| of a positive tracklet of the gate sample


def fit_gates(tracker: GpuTracker, preset: int, seeds: Sequence[int], margin: float = GATE_MARGIN):
    """Per tracklet layer: margin x the largest |dphi| of the positive rows of
    plain runs over `seeds`, capped at the reference's 0.3 rad cut.  Read from
    the rows' features: layer = f0 / 0.2, dphi = f1 * 0.3 (BF16, 2^-9
    relative)."""
    tracker.set_filter_gates(None)
    tracker.set_filter(False)
    tracker.set_events(generate_events(preset, seeds[0], 1))
    tracker.run()
    per_event = sum(sum(tracker.counts(i).tracklets) for i in range(tracker.num_items()))
    n = collect(tracker, preset, seeds, 1, int(per_event * 1.2 * len(seeds)) + 1024)
    X, Y = tracker.train_get(n)
    f = (X[:, :2].astype(np.uint32) << 16).view(np.float32)
    layer = np.rint(f[:, 0] / np.float32(0.2)).astype(int)
    dphi = np.abs(f[:, 1]) * np.float32(0.3)
    gates = []
    for l in range(6):
        pos = dphi[(layer == l) & (Y > 0)]
        g = margin * float(pos.max()) if pos.size else 0.3
        gates.append(float(np.float32(min(max(g, 1e-3), 0.3))))
    return gates


def train(tracker: GpuTracker, epochs: int, batch: int = 4096, lr: float = 3e-3, lr_min: float = 1e-4,
          weight_decay: float = 1e-4, pos_weight: float = 4.0, seed: int = 1) -> List[float]:
    """AdamW epochs over the tracker's training set with a cosine learning rate."""
    losses = []
    for e in range(epochs):
        lr_e = lr_min + 0.5 * (lr - lr_min) * (1 + math.cos(math.pi * e / max(epochs - 1, 1)))
        losses.append(tracker.train_epoch(batch, lr_e, weight_decay, pos_weight, seed * 1000003 + e))
        _log("  epoch %d: lr %.2e, mean loss %.5f" % (e, lr_e, losses[-1]))
    return losses


def evaluate(tracker: GpuTracker, events, threshold: Optional[float], plain=None) -> Dict:
    """Runs `events` plain (unless their plain tracks are given) and with the
    filter at `threshold`; returns recall of the plain tracks (with
    multiplicity), tracks outside the plain set, the tracklets and cells kept,
    and the road cells kept: of the distinct cells of the plain roads (the
    cells the training labels mark), the share the filtered run still builds."""
    tracker.set_events(events)
    items = None
    if plain is None:
        tracker.set_filter(False)
        tracker.run()
        items = tracker.num_items()
        tracks = [Counter(road_tracks(tracker, i)) for i in range(items)]
        plain = {"tracks": tracks,
                 "road_cells": [{(k,) + t[k:k + 3] for t in tr for k in range(5)} for tr in tracks],
                 "tracklets": sum(sum(tracker.counts(i).tracklets) for i in range(items)),
                 "cells": sum(sum(tracker.counts(i).cells) for i in range(items))}
    tracker.set_filter(True, threshold)
    tracker.run()
    items = tracker.num_items()
    found = extra = total = 0
    for i in range(items):
        learned = Counter(road_tracks(tracker, i))
        found += sum((learned & plain["tracks"][i]).values())
        extra += sum((learned - plain["tracks"][i]).values())
        total += sum(plain["tracks"][i].values())
    tl = sum(sum(tracker.counts(i).tracklets) for i in range(items))
    ce = sum(sum(tracker.counts(i).cells) for i in range(items))
    rc = rc_kept = 0
    for i in range(items):
        want = plain["road_cells"][i]
        rc += len(want)
        rc_kept += len(want & cell_keys(tracker, i)) if want else 0
    return {"plain": plain, "recall": found / max(total, 1), "plain_tracks": total, "found": found, "extra": extra,
            "tracklets_kept": tl / max(plain["tracklets"], 1), "cells_kept": ce / max(plain["cells"], 1),
            "road_cells_kept": rc_kept / max(rc, 1)}


def tune_threshold(tracker: GpuTracker, events, target: float, scores_pos: np.ndarray) -> Dict:
    """The largest threshold whose road recall on `events` reaches `target`.
    Candidates are the logits at which 99.9 %, 99.95 %, ... of the positive
    training rows are kept (from strict to loose)."""
    qs = [1e-3, 5e-4, 2e-4, 1e-4, 5e-5, 2e-5, 1e-5, 0.0]
    cands = sorted({float(np.quantile(scores_pos, q)) for q in qs}, reverse=True)
    plain = None
    best = None
    for thr in cands:
        r = evaluate(tracker, events, thr, plain)
        plain = r["plain"]
        _log("  threshold %.4f: recall %.5f (%d/%d), extra %d, tracklets kept %.3f, cells kept %.3f"
             % (thr, r["recall"], r["found"], r["plain_tracks"], r["extra"], r["tracklets_kept"], r["cells_kept"]))
        if r["recall"] >= target:
            best = dict(r, threshold=thr)
            break
    if best is None:
        best = dict(r, threshold=-3.0e38)
    best.pop("plain")
    tracker.set_filter(False)
    return best


def _replay_ms(tracker: GpuTracker, steps: int = 10) -> float:
    """Mean ms per run of the uploaded batch, replayed as a HIP graph."""
    tracker.set_graph(False)
    tracker.run()
    tracker.set_graph(True)
    tracker.run_async()
    tracker.finish()
    t0 = time.perf_counter()
    for _ in range(steps):
        tracker.run_async()
    tracker.finish()
    ms = 1e3 * (time.perf_counter() - t0) / steps
    tracker.set_graph(False)
    return ms


def sweep(events, params: np.ndarray, thresholds: Sequence[float], gates=None) -> List[Dict]:
    """The trade-off of the threshold on `events`: per threshold, the road
    recall of the plain tracks, tracks outside them, the tracklets and cells
    kept, and the time per run (graph replay, block 256, 8 workgroups per CU,
    one z sub-bin), with the plain form's time for reference."""
    tracker = GpuTracker(device=0)
    tracker.set_mlp_params(params)
    tracker.set_filter_gates(gates)
    tracker.set_tuning(256, 0, 1)
    tracker.set_events(events)
    tracker.set_filter(False)
    plain_ms = _replay_ms(tracker)
    plain = None
    out = []
    for thr in sorted(thresholds):
        r = evaluate(tracker, events, thr, plain)
        plain = r.pop("plain")
        tracker.set_filter(True, thr)
        ms = _replay_ms(tracker)
        row = dict(r, threshold=thr, ms=ms, plain_ms=plain_ms, speedup=plain_ms / ms)
        _log("  threshold %8.4f: recall %.5f, outside %d, tracklets kept %.4f, cells kept %.4f, road cells kept "
             "%.5f, %.3f ms (plain %.3f ms, %.3fx)" % (thr, r["recall"], r["extra"], r["tracklets_kept"],
                                                      r["cells_kept"], r["road_cells_kept"], ms, plain_ms,
                                                      plain_ms / ms))
        out.append(row)
    tracker.close()
    return out


def train_filter(tracker: GpuTracker, preset: int, seeds: Sequence[int], samples: float, val_seeds: Sequence[int],
                 epochs: int = 12, batch: int = 4096, lr: float = 3e-3, weight_decay: float = 1e-4,
                 pos_weight: float = 4.0, init_seed: int = 7, target_recall: float = 0.999, gate_events: int = 50,
                 gate_margin: float = GATE_MARGIN):
    """The whole training: the phi gates from plain runs over the first
    `gate_events` events of `seeds` (gate_margin <= 0: no gates), labels from
    gated runs over the events of `seeds` (about `samples` rows), AdamW epochs
    from a fresh initialisation, and the threshold tuned to `target_recall` of
    the plain roads of the `val_seeds` events.  Returns (params, threshold,
    meta); meta["phi_gates"] holds the gates.  The tracker keeps the trained
    weights and the gates, filter off."""
    seeds, val_seeds = list(seeds), list(val_seeds)
    t0 = time.time()
    gates = fit_gates(tracker, preset, seeds[:max(gate_events, 1)], gate_margin) if gate_margin > 0 else None
    _log("phi gates: %s (%.1f s)" % (gates, time.time() - t0))
    tracker.mlp_init(init_seed)  # finite parameters for the gated runs, which keep every candidate
    tracker.set_filter_gates(gates)
    tracker.set_filter(gates is not None, KEEP_ALL)
    tracker.set_events(generate_events(preset, seeds[0], 1))
    tracker.run()
    per_event = sum(sum(tracker.counts(i).tracklets) for i in range(tracker.num_items()))
    stride = max(1, int(round(per_event * len(seeds) / samples)))
    cap = int(per_event * 1.1 * len(seeds) / stride) + 1024
    n = collect(tracker, preset, seeds, stride, cap, gates=gates)
    X, Y = tracker.train_get(n)
    _log("collected %d rows (stride %d, %d tracklets in the first event), %.4f positive, %.1f s"
         % (n, stride, per_event, Y.mean(), time.time() - t0))

    tracker.mlp_init(init_seed)
    losses = train(tracker, epochs, batch, lr, weight_decay=weight_decay, pos_weight=pos_weight, seed=init_seed)
    _log("epoch losses " + " ".join("%.5f" % v for v in losses) + "  (%.1f s)" % (time.time() - t0))
    z = tracker.train_scores(n)
    params = tracker.mlp_params()

    val = [e for s in val_seeds for e in generate_events(preset, s, 1)]
    best = tune_threshold(tracker, val, target_recall, z[Y > 0])
    meta = {"training_events": "preset %d seeds %d..%d" % (preset, seeds[0], seeds[-1]),
            "rows": int(n), "stride": stride, "positive_fraction": float(Y.mean()), "epochs": epochs,
            "batch": batch, "lr": lr, "weight_decay": weight_decay, "pos_weight": pos_weight,
            "final_loss": losses[-1], "validation_events": "preset %d seeds %d..%d" % (
                preset, val_seeds[0], val_seeds[-1]),
            "validation_road_recall": best["recall"], "validation_tracklets_kept": best["tracklets_kept"],
            "validation_cells_kept": best["cells_kept"],
            "training_row_recall": float(np.mean(z[Y > 0] >= best["threshold"])),
            "training_rows_kept": float(np.mean(z >= best["threshold"])), "phi_gates": gates,
            "gate_events": "preset %d seeds %d..%d" % (preset, seeds[0], seeds[min(gate_events, len(seeds)) - 1]),
            "gate_margin": gate_margin}
    return params, best["threshold"], meta


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--sweep", action="store_true",
                    help="instead of training: the threshold trade-off of the weights file (--out) on held-out "
                         "events (--sweep-first-seed, --sweep-events)")
    ap.add_argument("--sweep-first-seed", type=int, default=20000)
    ap.add_argument("--sweep-events", type=int, default=256)
    ap.add_argument("--sweep-offsets", type=float, nargs="+",
                    default=[-40.0, -30.0, -24.0, -18.0, -12.0, -8.0, -4.0, -2.0, -1.0, 0.0, 0.5, 1.0, 2.0],
                    help="thresholds of the sweep, relative to the file's")
    ap.add_argument("--out", default=DEFAULT_WEIGHTS)
    ap.add_argument("--preset", type=int, default=2, help="event preset of the training events (2: central)")
    ap.add_argument("--first-seed", type=int, default=10000)
    ap.add_argument("--events", type=int, default=500)
    ap.add_argument("--samples", type=float, default=2e7, help="training rows to keep (sets the stride)")
    ap.add_argument("--val-first-seed", type=int, default=10500)
    ap.add_argument("--val-events", type=int, default=32)
    ap.add_argument("--epochs", type=int, default=12)
    ap.add_argument("--batch", type=int, default=4096)
    ap.add_argument("--lr", type=float, default=3e-3)
    ap.add_argument("--weight-decay", type=float, default=1e-4)
    ap.add_argument("--pos-weight", type=float, default=4.0)
    ap.add_argument("--init-seed", type=int, default=7)
    ap.add_argument("--target-recall", type=float, default=0.999, help="road recall on the validation events")
    ap.add_argument("--gate-events", type=int, default=50, help="events whose positive tracklets set the phi gates")
    ap.add_argument("--gate-margin", type=float, default=GATE_MARGIN,
                    help="phi gate = this x the largest |dphi| of a positive tracklet (<= 0: no gates)")
    args = ap.parse_args(argv)

    if args.sweep:
        params, thr, _ = load_filter(args.out)
        events = generate_events(args.preset, args.sweep_first_seed, args.sweep_events)
        thresholds = [thr + d for d in args.sweep_offsets]
        rows = sweep(events, params, thresholds, load_gates(args.out))
        print(json.dumps({"events": "preset %d seeds %d..%d" % (args.preset, args.sweep_first_seed,
                                                                  args.sweep_first_seed + args.sweep_events - 1),
                          "shipped_threshold": thr, "rows": rows}))
        return

    t0 = time.time()
    tracker = GpuTracker(device=0)
    params, threshold, meta = train_filter(
        tracker, args.preset, range(args.first_seed, args.first_seed + args.events), args.samples,
        range(args.val_first_seed, args.val_first_seed + args.val_events), epochs=args.epochs, batch=args.batch,
        lr=args.lr, weight_decay=args.weight_decay, pos_weight=args.pos_weight, init_seed=args.init_seed,
        target_recall=args.target_recall, gate_events=args.gate_events, gate_margin=args.gate_margin)
    save_filter(args.out, params, threshold, meta, meta.get("phi_gates"))
    _log(json.dumps(dict(meta, threshold=threshold, seconds=round(time.time() - t0, 1))))
    tracker.close()


if __name__ == "__main__":
    main()
