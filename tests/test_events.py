"""Generator determinism and the reference text formats."""
import os
import subprocess
import tempfile

import numpy as np

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
