"""gfx950 pipeline against the reference: bit-exact, position for position,
with the stable-order reference build (the reference with std::stable_sort in
place of std::sort for the per-layer bin order; see golden/make_golden.py)."""
import hashlib
import os
import sys
import tempfile

import numpy as np
import pytest

import itsca
import pyoracle
from golden_cases import central, dense_spot_event, fixture, small_cases

pytestmark = pytest.mark.gpu
CASES = small_cases()


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def tracker():
    t = itsca.GpuTracker(device=0)
    yield t
    t.close()


@pytest.mark.parametrize("name,path,nv", CASES, ids=[c[0] for c in CASES])
def test_single_event_bit_exact(tracker, name, path, nv):
    say(f"gpu case {name}")
    ev = itsca.load_events(path)[0]
    tracker.set_events([ev])
    tracker.run()
    assert tracker.num_items() == nv
    for v in range(nv):
        assert itsca.dump_item(tracker, v, ev.vertices[v]) == fixture(name, "stable", v)


def test_batch_of_all_cases_bit_exact(tracker):
    say("gpu batch of all golden cases")
    evs = [itsca.load_events(p)[0] for _, p, _ in CASES]
    tracker.set_events(evs)
    tracker.run()
    item = 0
    for (name, _, nv), ev in zip(CASES, evs):
        for v in range(nv):
            assert itsca.dump_item(tracker, item, ev.vertices[v]) == fixture(name, "stable", v), name
            item += 1
    assert item == tracker.num_items()


def test_central_event_matches_reference_digest(tracker):
    say("gpu central event (8000 tracks)")
    info = central()
    with tempfile.TemporaryDirectory() as tmp:
        p = os.path.join(tmp, "c.txt")
        itsca.write_events(info["preset"], info["seed"], 1, p)
        assert hashlib.sha256(open(p, "rb").read()).hexdigest() == info["events_sha256"]
        ev = itsca.load_events(p)[0]
    tracker.set_events([ev])
    tracker.run()
    c = tracker.counts(0)
    assert list(c.tracklets) == info["counts"]["tracklets"]
    assert list(c.cells) == info["counts"]["cells"]
    text = itsca.dump_item(tracker, 0, ev.vertices[0])
    assert hashlib.sha256(text.encode()).hexdigest() == info["stable_sha256"]


@pytest.mark.parametrize("name,path,nv", CASES, ids=[c[0] for c in CASES])
def test_matches_unmodified_reference_as_keyed_sets(tracker, name, path, nv):
    """Against the reference as-is (std::sort, the committed ref fixtures):
    the cluster-id-keyed tracklets, cells and roads are the same sets, and
    every multiplicity difference is a duplicate of the inclusive candidate
    bound (Tracker.cxx:84-89): the oracle without that extra candidate gives
    the same set with every tracklet and cell once."""
    from dumps import keyed

    say(f"gpu case {name} vs unmodified reference (keyed)")
    ev = itsca.load_events(path)[0]
    tracker.set_events([ev])
    tracker.run()
    for v in range(nv):
        gpu = keyed(itsca.dump_item(tracker, v, ev.vertices[v]))
        ref = keyed(fixture(name, "ref", v))
        once = keyed(pyoracle.OracleVertex(ev, v, stable=True, exclusive_bound=True).dump())
        for stage in ref:
            assert set(gpu[stage]) == set(ref[stage]) == set(once[stage]), (name, v, stage)
            if stage != "roads":
                assert max(once[stage].values(), default=1) == 1
            for k in set(gpu[stage]) | set(ref[stage]):
                if gpu[stage][k] != ref[stage][k]:  # a duplicated record in at least one of the two
                    assert max(gpu[stage][k], ref[stage][k]) >= 2, (name, v, stage, k)


def test_central_event_matches_unmodified_reference_keyed(tracker):
    from dumps import distinct_digest, duplicates

    say("gpu central event vs unmodified reference (keyed)")
    info = central()
    with tempfile.TemporaryDirectory() as tmp:
        p = os.path.join(tmp, "c.txt")
        itsca.write_events(info["preset"], info["seed"], 1, p)
        ev = itsca.load_events(p)[0]
    tracker.set_events([ev])
    tracker.run()
    text = itsca.dump_item(tracker, 0, ev.vertices[0])
    assert distinct_digest(text) == info["ref_distinct_sha256"]
    assert duplicates(text) == info["ref_duplicates"]


def test_mixed_multiplicity_batch_matches_oracle(tracker):
    say("gpu mixed batch vs oracle")
    evs = itsca.generate_events(3, 3000, 6)
    tracker.set_events(evs)
    tracker.run()
    for i, ev in enumerate(evs):
        assert itsca.dump_item(tracker, i, ev.vertices[0]) == pyoracle.OracleVertex(ev, 0, stable=True).dump()


def test_pileup_event_with_large_layers_matches_oracle(tracker):
    """A pile-up event (4 vertices, up to 27 088 clusters per layer): layers
    above the LDS sort size take the global-scratch sort, and dense bins exceed
    the z sub-bin index (they are searched whole)."""
    say("gpu pile-up event (large layers) vs oracle")
    ev = itsca.generate_events(5, 40003, 1)[0]
    assert int(np.diff(ev.layer_offsets).max()) > 16384
    tracker.set_tuning(256, 0, 8)
    tracker.set_events([ev])
    tracker.run()
    tracker.set_tuning(256, 0, 0)
    for v in (0, 3):
        assert itsca.dump_item(tracker, v, ev.vertices[v]) == pyoracle.OracleVertex(ev, v, stable=True).dump(), v


def test_slot_overflow_event_matches_oracle(tracker):
    """A dense spot gives one cluster of layer 0 and one of layer 2 more
    tracklets than the 64 slots the tracklet count keeps per cluster: the fill
    searches those clusters again, and every array still equals the oracle's."""
    say("gpu slot-overflow event vs oracle")
    ev = dense_spot_event()
    tracker.set_tuning(256, 0, 0)
    tracker.set_events([ev])
    tracker.run()
    n = tracker.counts(0)
    for layer in (0, 2):
        t = tracker.download(0, itsca.TRACKLETS, layer, n.tracklets[layer], 4)
        assert np.bincount(t[:, 0]).max() > itsca.TRACKLET_SLOTS, layer
    assert itsca.dump_item(tracker, 0, ev.vertices[0]) == pyoracle.OracleVertex(ev, 0, stable=True).dump()


def test_cli_on_merged_pileup_in_binary_format(tracker):
    """bin/itsca-events merges small events into pile-up events of 3 vertices
    (vertices sharing x and y share their prepared clusters) and writes the
    binary side-car; bin/itsca-track reads it with the merged labels, writes
    the road reports, and finds per vertex the road counts of the same events
    run through the C ABI."""
    import re
    import subprocess
    say("gpu cli on merged pile-up, binary events")
    with tempfile.TemporaryDirectory() as tmp:
        src, lab = os.path.join(tmp, "e.txt"), os.path.join(tmp, "l.txt")
        merged, mlab, binp = os.path.join(tmp, "m.txt"), os.path.join(tmp, "ml.txt"), os.path.join(tmp, "m.itsb")
        itsca.write_events(1, 21, 6, src, lab)
        tool = os.path.join(itsca.BIN_DIR, "itsca-events")
        subprocess.run([tool, "pileup", "3", src, merged, lab, mlab], check=True, capture_output=True)
        subprocess.run([tool, "binary", merged, binp], check=True, capture_output=True)
        out = subprocess.run([os.path.join(itsca.BIN_DIR, "itsca-track"), binp, mlab], capture_output=True, text=True,
                             timeout=120)
        assert out.returncode == 0, out.stdout + out.stderr
        got = [int(n) for n in re.findall(r"Found (\d+) roads for vertex", out.stdout)]
        evs = itsca.load_events(binp)
        assert [len(e.vertices) for e in evs] == [3, 3]
        tracker.set_events(evs)
        tracker.run()
        assert got == [tracker.counts(i).roads for i in range(tracker.num_items())] and sum(got) > 0
        for name in ("CorrectRoads.txt", "DuplicateRoads.txt", "FakeRoads.txt"):
            assert os.path.exists(os.path.join(tmp, name)), name


def test_runs_are_deterministic(tracker):
    say("gpu determinism")
    evs = itsca.generate_events(2, 5, 2)
    tracker.set_events(evs)
    tracker.run()
    a = [tracker.roads(i).copy() for i in range(2)]
    tracker.run()
    b = [tracker.roads(i) for i in range(2)]
    for x, y in zip(a, b):
        assert np.array_equal(x, y)


def _item_arrays(t, w):
    n = t.counts(w)
    out = [np.array(list(n.clusters) + list(n.tracklets) + list(n.cells) + list(n.neighbours) + [n.roads])]
    out += [t.download(w, itsca.TRACKLETS, l, n.tracklets[l], 4) for l in range(itsca.TRACKLET_LAYERS)]
    out += [t.download(w, itsca.CELLS, l, n.cells[l], 10) for l in range(itsca.CELL_LAYERS)]
    for l in range(itsca.CELL_LAYERS - 1):
        out.append(t.download(w, itsca.NEIGHBOUR_OFFSETS, l, n.cells[l + 1] + 1, 1))
        out.append(t.download(w, itsca.NEIGHBOURS, l, n.neighbours[l], 1))
    out.append(t.roads(w))
    return out


def test_large_batch_matches_single_event_runs(tracker):
    """12 central events in one launch: the cell stage has about 21 000 chunks
    of 256 tracklets, more than one tile (16 384) of the chunk scan.  Every
    event's tracklets, cells, neighbours and roads equal those of the event run
    on its own."""
    say("gpu large batch vs single events")
    evs = itsca.generate_events(2, 60, 12)
    tracker.set_tuning(256, 0, 0)
    tracker.set_events(evs)
    tracker.run()
    batch = [_item_arrays(tracker, w) for w in range(len(evs))]
    # cell-stage chunks: tracklets of layers 0..4 in chunks of 256 per (event, layer)
    assert sum(-(-int(x) // 256) for a in batch for x in a[0][7:12]) > 16384
    for w, ev in enumerate(evs):
        tracker.set_events([ev])
        tracker.run()
        single = _item_arrays(tracker, 0)
        for k, (x, y) in enumerate(zip(batch[w], single)):
            assert np.array_equal(x, y), (w, k)


def test_capacity_growth_keeps_results(golden_dir):
    say("gpu capacity growth")
    ev = itsca.load_events(os.path.join(golden_dir, "small_s1.events.txt"))[0]
    t = itsca.GpuTracker(device=0, capacity_scale=0.001)  # forces the overflow/grow path
    t.set_events([ev])
    t.run()
    assert itsca.dump_item(t, 0, ev.vertices[0]) == fixture("small_s1", "stable", 0)
    t.close()


@pytest.mark.parametrize("block,zsub,phisub", [(64, 1, 1), (128, 2, 1), (256, 8, 1), (256, 3, 1), (64, 1, 4),
                                               (128, 3, 4), (256, 4, 2)])
def test_launch_shapes_do_not_change_results(tracker, golden_dir, block, zsub, phisub):
    say(f"gpu tuning block={block} z_subbins={zsub} phi_subrows={phisub}")
    evs = [itsca.load_events(os.path.join(golden_dir, n + ".events.txt"))[0] for n in ("small_s3", "sparse_s6")]
    t = itsca.GpuTracker(device=0)
    t.set_tuning(block, 0, zsub, phisub)
    t.set_events(evs)
    t.run()
    for i, name in enumerate(("small_s3", "sparse_s6")):
        assert itsca.dump_item(t, i, evs[i].vertices[0]) == fixture(name, "stable", 0)
    t.close()


@pytest.mark.parametrize("zsub,phisub", [(1, 1), (2, 1), (4, 1), (8, 1), (1, 2), (1, 4), (2, 4), (4, 2), (3, 4)])
def test_search_index_binning_keeps_central_event(zsub, phisub):
    """The z sub-bins and phi sub-rows of the search index only decide which
    candidates are skipped untested: the central event's full dump stays
    bit-exact."""
    say(f"gpu central event z_subbins={zsub} phi_subrows={phisub}")
    info = central()
    with tempfile.TemporaryDirectory() as tmp:
        p = os.path.join(tmp, "c.txt")
        itsca.write_events(info["preset"], info["seed"], 1, p)
        ev = itsca.load_events(p)[0]
    t = itsca.GpuTracker(device=0)
    t.set_tuning(256, 0, zsub, phisub)
    assert (t.tuning().z_subbins, t.tuning().phi_subrows) == (zsub, phisub)
    t.set_events([ev])
    t.run()
    text = itsca.dump_item(t, 0, ev.vertices[0])
    t.close()
    assert hashlib.sha256(text.encode()).hexdigest() == info["stable_sha256"]


def test_graph_replays_match_eager():
    say("gpu graph replays")
    evs = itsca.generate_events(3, 3100, 4)
    t = itsca.GpuTracker(device=0)
    t.set_events(evs)
    t.run()
    ref = [itsca.dump_item(t, i, ev.vertices[0]) for i, ev in enumerate(evs)]
    other = itsca.GpuTracker(device=0)  # eager work of the same program between replays
    other.set_events(itsca.generate_events(1, 77, 2))
    t.set_graph(True)
    for rep in range(4):
        t.run_async()
        t.run_async()
        t.finish()
        for i, ev in enumerate(evs):
            assert itsca.dump_item(t, i, ev.vertices[0]) == ref[i], (rep, i)
        other.run()
    t.close()
    other.close()


def test_autotune_keeps_results():
    say("gpu autotune")
    evs = itsca.generate_events(3, 3200, 3)
    t = itsca.GpuTracker(device=0)
    t.set_events(evs)
    t.run()
    ref = [t.roads(i).copy() for i in range(3)]
    tuning, ms = t.autotune(repeats=2)
    assert tuning.block_size in (64, 128, 256) and ms > 0
    for i in range(3):
        assert np.array_equal(t.roads(i), ref[i])
    t.close()


REPORT_FILES = ("CorrectRoads.txt", "DuplicateRoads.txt", "FakeRoads.txt")


def _track_report(tmp, events_path, labels_path):
    """bin/itsca-track <events> <labels> in tmp: the report files it writes."""
    import shutil
    import subprocess
    ev, lab = os.path.join(tmp, "ev.txt"), os.path.join(tmp, "labels.txt")
    shutil.copy(events_path, ev)
    shutil.copy(labels_path, lab)
    out = subprocess.run([os.path.join(itsca.BIN_DIR, "itsca-track"), ev, lab], capture_output=True, text=True,
                         timeout=120)
    assert out.returncode == 0, out.stdout[-2000:] + out.stderr[-2000:]
    return {f: open(os.path.join(tmp, f)).read() for f in REPORT_FILES}


def _blocks(text):
    """Rows per event of a report file with -1 separators, as sets."""
    out = []
    for line in text.splitlines():
        if line.strip() == "-1":
            out.append(set())
        elif out:
            out[-1].add(line)
    return out


@pytest.mark.parametrize("case", ["report_s21", "report_pileup3"])
def test_cli_road_report_matches_reference(golden_dir, case):
    """itsca-track writes the reference's road report files: byte for byte
    those of the stable-order reference build, and per event the same sets of
    correct and fake labels as the reference as-is.  The as-is build's sort
    order moves roads and changes which roads the inclusive candidate bound
    duplicates (test_matches_unmodified_reference_as_keyed_sets), so its rows
    come in another order and its DuplicateRoads.txt lists other repeats."""
    say(f"gpu cli road report {case}")
    with tempfile.TemporaryDirectory() as tmp:
        got = _track_report(tmp, os.path.join(golden_dir, case + ".ev.txt"),
                            os.path.join(golden_dir, case + ".labels.txt"))
    for f in REPORT_FILES:
        want = open(os.path.join(golden_dir, "%s.stable.%s" % (case, f))).read()
        assert got[f] == want, f
    for f in ("CorrectRoads.txt", "FakeRoads.txt"):
        ref = open(os.path.join(golden_dir, "%s.ref.%s" % (case, f))).read()
        assert _blocks(got[f]) == _blocks(ref), f


def test_cli_road_report_central_event(golden_dir):
    import json
    say("gpu cli road report, central event")
    info = json.load(open(os.path.join(golden_dir, "report_central.json")))
    with tempfile.TemporaryDirectory() as tmp:
        ev, lab = os.path.join(tmp, "c.txt"), os.path.join(tmp, "c_lab.txt")
        itsca.write_events(info["preset"], info["seed"], 1, ev, lab)
        assert hashlib.sha256(open(ev, "rb").read()).hexdigest() == info["events_sha256"]
        assert hashlib.sha256(open(lab, "rb").read()).hexdigest() == info["labels_sha256"]
        os.makedirs(os.path.join(tmp, "run"))
        got = _track_report(os.path.join(tmp, "run"), ev, lab)
    for f in REPORT_FILES:
        assert hashlib.sha256(got[f].encode()).hexdigest() == info["stable_%s_sha256" % f], f
