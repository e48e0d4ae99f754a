"""Generator determinism and the reference text formats."""
import os
import subprocess
import tempfile

import numpy as np
import pytest

import itsca


def test_generator_is_deterministic_and_seeded():
    a = itsca.generate_events(1, 3)[0]
    b = itsca.generate_events(1, 3)[0]
    c = itsca.generate_events(1, 4)[0]
    assert np.array_equal(a.x, b.x) and np.array_equal(a.z, b.z) and np.array_equal(a.mc, b.mc)
    assert not np.array_equal(a.x, c.x)


def test_small_preset_shape():
    ev = itsca.generate_events(1, 9)[0]
    assert len(ev.vertices) == 1 and ev.vertices[0][0] == 0 and ev.vertices[0][1] == 0
    n = np.diff(ev.layer_offsets)
    assert all(95 <= k <= 110 for k in n)
    r = np.hypot(ev.x, ev.y)
    radii = [2.34, 3.15, 3.93, 19.6, 24.55, 34.39, 39.34]
    for l in range(7):
        assert np.allclose(r[ev.layer(l)], radii[l], atol=1e-3)
    assert np.array_equal(ev.cluster_id, np.arange(len(ev.x)))


def test_pileup_preset_has_four_vertices_with_offset_labels():
    ev = itsca.generate_events(5, 50000)[0]
    assert len(ev.vertices) == 4
    assert np.all(np.abs(ev.vertices[:, 2]) <= 10)
    assert set(np.unique(ev.mc[ev.mc >= 0] // 100000)) == {0, 1, 2, 3}


def test_text_round_trip_and_cli_agree():
    with tempfile.TemporaryDirectory() as tmp:
        p1, p2 = os.path.join(tmp, "a.txt"), os.path.join(tmp, "b.txt")
        itsca.write_events(3, 77, 2, p1, os.path.join(tmp, "labels.txt"))
        subprocess.run([os.path.join(itsca.BIN_DIR, "itsca-gen"), "3", "77", "2", p2], check=True)
        assert open(p1).read() == open(p2).read()
        back = itsca.load_events(p1)
        gen = itsca.generate_events(3, 77, 2)
        for a, b in zip(back, gen):
            assert np.array_equal(a.x, b.x) and np.array_equal(a.y, b.y) and np.array_equal(a.z, b.z)
            assert np.array_equal(a.mc, b.mc) and np.array_equal(a.layer_offsets, b.layer_offsets)
        labels = open(os.path.join(tmp, "labels.txt")).read().splitlines()
        assert labels[0].split()[0] == "mc" and labels.count("-1") == 1


def test_loader_skips_malformed_lines_and_splits_events():
    with tempfile.TemporaryDirectory() as tmp:
        p = os.path.join(tmp, "e.txt")
        with open(p, "w") as f:
            f.write("-1 0 0 1.5\n0 1 2 3 0 0 0 0 7\ngarbage line\n1 4 5 6\n-1 0 0 2.5\n-1 0 0 3.5\n2 7 8 9 0 0 0 0 8\n")
        evs = itsca.load_events(p)
        assert len(evs) == 2
        assert len(evs[0].x) == 1 and evs[0].cluster_id.tolist() == [0]
        assert len(evs[1].vertices) == 2 and evs[1].cluster_id.tolist() == [0]


def test_binary_side_car_round_trip_matches_text():
    with tempfile.TemporaryDirectory() as tmp:
        txt, binp = os.path.join(tmp, "e.txt"), os.path.join(tmp, "e.itsb")
        itsca.write_events(5, 50001, 2, txt)
        itsca.convert_events_binary(txt, binp)
        a, b = itsca.load_events(txt), itsca.load_events(binp)
        assert len(a) == len(b) == 2
        for u, v in zip(a, b):
            for f in ("x", "y", "z", "cluster_id", "mc", "layer_offsets", "vertices"):
                assert np.array_equal(getattr(u, f), getattr(v, f)), f
        # the CLI writes the same bytes
        cli = os.path.join(tmp, "cli.itsb")
        subprocess.run([os.path.join(itsca.BIN_DIR, "itsca-events"), "binary", txt, cli], check=True)
        assert open(cli, "rb").read() == open(binp, "rb").read()
        # a truncated file is an error, not a short batch
        bad = os.path.join(tmp, "bad.itsb")
        with open(bad, "wb") as f:
            f.write(open(binp, "rb").read()[:-100])
        with pytest.raises(itsca.ItscaError):
            itsca.load_events(bad)


def _reference_merge(lines, pile_up):
    """generatePileUpData.cxx:44-119 (mergeEvents) restated on text lines:
    vertex lines are written at once, cluster lines are held until the merged
    event holds pile_up vertices; clusters after the k-th vertex get their MC
    label (column 8) offset by (k - 1) * 100000."""
    out, held, vertices = [], [], 0
    for line in lines:
        parts = line.split()
        if not parts:
            continue
        if int(parts[0]) == -1:
            if vertices == pile_up:
                out += held
                vertices, held = 0, []
            out.append(line)
            vertices += 1
        else:
            if vertices > 1:
                parts[8] = str(int(parts[8]) + (vertices - 1) * 100000)
                line = "\t".join(parts)
            held.append(line)
    return out + held


def test_pileup_merger_follows_the_reference_algorithm():
    with tempfile.TemporaryDirectory() as tmp:
        src, lab = os.path.join(tmp, "e.txt"), os.path.join(tmp, "l.txt")
        itsca.write_events(1, 11, 7, src, lab)
        ref = os.path.join(tmp, "ref.txt")
        with open(ref, "w") as f:
            f.write("\n".join(_reference_merge(open(src).read().splitlines(), 3)) + "\n")
        out, out_lab = os.path.join(tmp, "m.txt"), os.path.join(tmp, "ml.txt")
        assert itsca.merge_pileup(src, 3, out, lab, out_lab) == 3  # 7 events -> 3 + 3 + 1
        got, want, orig = itsca.load_events(out), itsca.load_events(ref), itsca.load_events(src)
        assert [len(e.vertices) for e in got] == [3, 3, 1]
        for u, v in zip(got, want):
            for f in ("x", "y", "z", "cluster_id", "mc", "layer_offsets", "vertices"):
                assert np.array_equal(getattr(u, f), getattr(v, f)), f
        assert sum(len(e.x) for e in got) == sum(len(e.x) for e in orig)
        # the CLI gives the same file
        cli = os.path.join(tmp, "cli.txt")
        subprocess.run([os.path.join(itsca.BIN_DIR, "itsca-events"), "pileup", "3", src, cli], check=True)
        assert open(cli).read() == open(out).read()
        # label rows carry the offsets of their collision's clusters; like the
        # reference, noise labels (-1) are offset too (k * 100000 - 1)
        rows = [l.split() for l in open(out_lab).read().splitlines()[1:]]
        events, cur = [], set()
        for r in rows:
            if r[0] == "-1":
                events.append(cur)
                cur = set()
            else:
                cur.add(int(r[0]))
        events.append(cur)
        for e, labels in zip(got, events):
            tracks = set(int(m) for m in e.mc if m >= 0 and (m + 1) % 100000 != 0)
            assert tracks <= labels and {m // 100000 for m in labels} == set(range(len(e.vertices)))
