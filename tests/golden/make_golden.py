#!/usr/bin/env python3
"""Writes the golden fixtures of the parity tests by running the REFERENCE CPU
tracker (built from its own sources by `make -C oracle ref` into oracle/_ref/)
on seeded events.

Run in the build container only (it needs the reference checkout to build the
harness).  Output, all under tests/golden/:
  <case>.events.txt            the input event (reference text format)
  <case>.ref.v<k>.txt          reference dump of vertex k, as-is build (std::sort)
  <case>.stable.v<k>.txt       reference dump of vertex k, stable-order build
  report_*                     road reports of events with MC labels (see
                               write_report_fixtures)
  central_s2.json              for the 8000-track central event (not committed
                               as text): SHA-256 of the generated event file and
                               of both dumps, and the output counts

Dump format: oracle/ca_oracle.cpp (dump); the harness is oracle/ref_harness.cpp.
"""
import hashlib
import json
import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.abspath(os.path.join(HERE, "..", ".."))
sys.path.insert(0, os.path.join(ROOT, "tracking-itsu_amd"))
import itsca  # noqa: E402

sys.path.insert(0, os.path.join(ROOT, "tests"))
import dumps  # noqa: E402

HARNESS = os.path.join(ROOT, "oracle", "_ref", "ref_harness")
HARNESS_STABLE = os.path.join(ROOT, "oracle", "_ref", "ref_harness_stable")


def write_event_file(events, path):
    """Reference event format (src/CA/IOUtils.cxx:38-78); ids follow file order."""
    with open(path, "w") as f:
        for ev in events:
            for v in ev.vertices:
                f.write("-1\t%.9g\t%.9g\t%.9g\n" % tuple(float(c) for c in v))
            for l in range(itsca.LAYERS):
                sl = ev.layer(l)
                for x, y, z, mc in zip(ev.x[sl].tolist(), ev.y[sl].tolist(), ev.z[sl].tolist(), ev.mc[sl].tolist()):
                    f.write("%d\t%.9g\t%.9g\t%.9g\t0\t0\t0\t0\t%d\n" % (l, x, y, z, mc))


def subset(ev, keep_mask):
    counts = []
    for l in range(itsca.LAYERS):
        counts.append(int(keep_mask[ev.layer(l)].sum()))
    offs = np.zeros(8, dtype=np.int32)
    offs[1:] = np.cumsum(counts)
    return itsca.Event(ev.x[keep_mask], ev.y[keep_mask], ev.z[keep_mask], ev.cluster_id[keep_mask],
                       ev.mc[keep_mask], offs, ev.vertices.copy())


def merge(a, b):
    """Two collisions in one event (pile-up): per layer a's clusters then b's;
    b's MC ids offset by 100000 (benchmarks/scripts/root/generatePileUpData.cxx:31,94)."""
    xs, ys, zs, ids, mcs, counts = [], [], [], [], [], []
    for l in range(itsca.LAYERS):
        for ev, off in ((a, 0), (b, 100000)):
            sl = ev.layer(l)
            xs.append(ev.x[sl]), ys.append(ev.y[sl]), zs.append(ev.z[sl]), ids.append(ev.cluster_id[sl])
            mcs.append(np.where(ev.mc[sl] >= 0, ev.mc[sl] + off, ev.mc[sl]))
        counts.append(int(a.layer_offsets[l + 1] - a.layer_offsets[l] + b.layer_offsets[l + 1] - b.layer_offsets[l]))
    offs = np.zeros(8, dtype=np.int32)
    offs[1:] = np.cumsum(counts)
    return itsca.Event(np.concatenate(xs), np.concatenate(ys), np.concatenate(zs), np.concatenate(ids),
                       np.concatenate(mcs), offs, np.concatenate([a.vertices, b.vertices]))


def small_cases():
    cases = {}
    for seed in (1, 2, 3, 4):
        cases["small_s%d" % seed] = [itsca.generate_events(1, seed)[0]]
    # an empty layer: the reference returns early from the tracklet loop
    # (Tracker.cxx:49-52), so no tracklets from layer 2 on
    cases["small_s5_nolayer3"] = [itsca.generate_events(1, 5)[0].drop_layers([3])]
    # a sparse event (12 tracks + noise): many empty bins, so the inclusive
    # candidate bound (Tracker.cxx:88) reaches into neighbouring bins and rows
    ev = itsca.generate_events(1, 6)[0]
    cases["sparse_s6"] = [subset(ev, (ev.mc < 12))]
    # no clusters at all
    empty = itsca.Event(np.zeros(0, np.float32), np.zeros(0, np.float32), np.zeros(0, np.float32),
                        np.zeros(0, np.int32), np.zeros(0, np.int32), np.zeros(8, np.int32),
                        np.array([[0.0, 0.0, 1.5]], dtype=np.float32))
    cases["empty"] = [empty]
    # two collisions in one event, two primary vertices
    a, b = itsca.generate_events(1, 7)[0], itsca.generate_events(1, 8)[0]
    cases["pileup2_s7"] = [merge(a, b)]
    # clusters outside the layers in z: negative z bins (negative bins in phi
    # row 0, aliased bins in row 1) and |z| = 1e10 (an index outside the int
    # range, INT_MIN on x86)
    cases["outside_s1"] = [outside_event()]
    return cases


def outside_event():
    base = itsca.generate_events(1, 1)[0]
    vx, vy, vz = (float(a) for a in base.vertices[0])
    radii = [2.34, 3.15, 3.93, 19.6, 24.55, 34.39, 39.34]
    extra = {0: [(-20.0, 0.1), (1.0e10, 0.1), (-1.0e10, 0.12), (-30.0, 0.5), (-20.0, 0.1)],
             1: [(-25.0, 0.05), (-19.0, 0.2), (3.0e9, 0.02), (-21.0, 0.6)]}
    cols = {k: [] for k in ("x", "y", "z", "mc")}
    offs = [0]
    for layer in range(itsca.LAYERS):
        sl = base.layer(layer)
        for k, arr in (("x", base.x), ("y", base.y), ("z", base.z), ("mc", base.mc)):
            cols[k] += list(arr[sl])
        for z, phi in extra.get(layer, []):
            cols["x"].append(vx + radii[layer] * np.cos(phi))
            cols["y"].append(vy + radii[layer] * np.sin(phi))
            cols["z"].append(z)
            cols["mc"].append(-1)
        offs.append(len(cols["x"]))
    n = offs[-1]
    return itsca.Event(np.array(cols["x"], np.float32), np.array(cols["y"], np.float32),
                       np.array(cols["z"], np.float32), np.arange(n, dtype=np.int32), np.array(cols["mc"], np.int32),
                       np.array(offs, np.int32), base.vertices.copy())


REPORT_FILES = ("CorrectRoads.txt", "DuplicateRoads.txt", "FakeRoads.txt")


def write_report_fixtures():
    """The reference's road report (IOUtils::loadLabels + writeRoadsReport run
    by the harness) for events with MC labels:
      report_s21       4 small events (preset 1, seeds 21..24)
      report_pileup3   6 small events (seeds 31..36) merged into 2 pile-up
                       events of 3 vertices (labels of the k-th collision
                       offset by (k - 1) * 100000)
    committed as <case>.ev.txt, <case>.labels.txt and
    <case>.<ref|stable>.<report file>; and for one central event (preset 2,
    seed 20000) the SHA-256 of each report file in report_central.json."""
    with tempfile.TemporaryDirectory() as tmp:
        itsca.write_events(1, 21, 4, os.path.join(HERE, "report_s21.ev.txt"),
                           os.path.join(HERE, "report_s21.labels.txt"))
        src, lab = os.path.join(tmp, "src.txt"), os.path.join(tmp, "src_lab.txt")
        itsca.write_events(1, 31, 6, src, lab)
        itsca.merge_pileup(src, 3, os.path.join(HERE, "report_pileup3.ev.txt"), lab,
                           os.path.join(HERE, "report_pileup3.labels.txt"))
        for case in ("report_s21", "report_pileup3"):
            for kind, exe in (("ref", HARNESS), ("stable", HARNESS_STABLE)):
                out = os.path.join(tmp, case + kind)
                os.makedirs(out)
                subprocess.run([exe, "report", os.path.join(HERE, case + ".ev.txt"),
                                os.path.join(HERE, case + ".labels.txt"), out], check=True, capture_output=True)
                for f in REPORT_FILES:
                    with open(os.path.join(out, f)) as a, open(os.path.join(HERE, "%s.%s.%s" % (case, kind, f)), "w") as b:
                        b.write(a.read())
            print(case, "report written")
        ev, lab = os.path.join(tmp, "central.txt"), os.path.join(tmp, "central_lab.txt")
        itsca.write_events(2, 20000, 1, ev, lab)
        info = {"preset": 2, "seed": 20000, "events_sha256": hashlib.sha256(open(ev, "rb").read()).hexdigest(),
                "labels_sha256": hashlib.sha256(open(lab, "rb").read()).hexdigest()}
        for kind, exe in (("ref", HARNESS), ("stable", HARNESS_STABLE)):
            out = os.path.join(tmp, "central" + kind)
            os.makedirs(out)
            subprocess.run([exe, "report", ev, lab, out], check=True, capture_output=True)
            for f in REPORT_FILES:
                text = open(os.path.join(out, f)).read()
                info["%s_%s_sha256" % (kind, f)] = hashlib.sha256(text.encode()).hexdigest()
                info["%s_%s_lines" % (kind, f)] = text.count("\n")
        with open(os.path.join(HERE, "report_central.json"), "w") as f:
            json.dump(info, f, indent=1)
        print("report_central", info)


def run_harness(exe, events_path, prefix):
    out = subprocess.run([exe, "dump", events_path, prefix], check=True, capture_output=True, text=True)
    return out.stdout


def main():
    for exe in (HARNESS, HARNESS_STABLE):
        if not os.path.exists(exe):
            sys.exit(f"{exe} missing: run make -C oracle ref")
    write_report_fixtures()
    for name, events in small_cases().items():
        ev_path = os.path.join(HERE, name + ".events.txt")
        write_event_file(events, ev_path)
        with tempfile.TemporaryDirectory() as tmp:
            log = run_harness(HARNESS, ev_path, os.path.join(tmp, "ref"))
            run_harness(HARNESS_STABLE, ev_path, os.path.join(tmp, "stable"))
            nv = sum(len(e.vertices) for e in events)
            for v in range(nv):
                for kind in ("ref", "stable"):
                    src = os.path.join(tmp, "%s_e0_v%d.txt" % (kind, v))
                    dst = os.path.join(HERE, "%s.%s.v%d.txt" % (name, kind, v))
                    with open(src) as f, open(dst, "w") as g:
                        g.write(f.read())
        print(name, log.strip().replace("\n", " | "))
    # the central event: digests only
    with tempfile.TemporaryDirectory() as tmp:
        ev_path = os.path.join(tmp, "central_s2.events.txt")
        itsca.write_events(2, 2, 1, ev_path)
        log = run_harness(HARNESS, ev_path, os.path.join(tmp, "ref"))
        run_harness(HARNESS_STABLE, ev_path, os.path.join(tmp, "stable"))
        digest = lambda p: hashlib.sha256(open(p, "rb").read()).hexdigest()  # noqa: E731
        ref_dump = open(os.path.join(tmp, "ref_e0_v0.txt")).read()
        counts = {}
        for line in ref_dump.splitlines():
            parts = line.split()
            if parts and parts[0] in ("clusters", "tracklets", "cells") and len(parts) == 3:
                counts.setdefault(parts[0], []).append(int(parts[2]))
            elif parts and parts[0] == "roads":
                counts["roads"] = int(parts[1])
        stable_dump = open(os.path.join(tmp, "stable_e0_v0.txt")).read()
        info = {
            "preset": 2, "seed": 2,
            "events_sha256": digest(ev_path),
            "ref_sha256": digest(os.path.join(tmp, "ref_e0_v0.txt")),
            "stable_sha256": digest(os.path.join(tmp, "stable_e0_v0.txt")),
            # cluster-id-keyed sets, multiplicities dropped: the same for both
            # sort orders (they differ only in duplicates of the inclusive
            # candidate bound); duplicates counted per stage
            "ref_distinct_sha256": dumps.distinct_digest(ref_dump),
            "stable_distinct_sha256": dumps.distinct_digest(stable_dump),
            "ref_duplicates": dumps.duplicates(ref_dump),
            "stable_duplicates": dumps.duplicates(stable_dump),
            "counts": counts,
            "harness_log": log.strip(),
        }
        with open(os.path.join(HERE, "central_s2.json"), "w") as f:
            json.dump(info, f, indent=1)
        print("central_s2", log.strip())


if __name__ == "__main__":
    main()
