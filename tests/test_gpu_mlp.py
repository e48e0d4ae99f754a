"""The learned tracklet filter on the GPU: the BF16 MFMA forward pass, the
backward pass and AdamW against the float64 reference (tests/mlp_ref.py,
itself pinned by test_mlp_ref.py); the training rows (features bit-exact with
their float32 restatement, labels from the plain roads); and the pipeline
invariants (a filter that keeps everything reproduces the plain CA bit for
bit; any filter finds a sub-multiset of the plain roads)."""
import sys
from collections import Counter

import numpy as np
import pytest

import itsca
import mlp_ref as M
from golden_cases import dense_spot_event, fixture, small_cases
from gpu_items import gated_tracklets

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


def _random_params(seed):
    rng = np.random.default_rng(seed)
    p = np.zeros(M.PARAMS, np.float32)
    p[M.OFF_W1:M.OFF_B1] = rng.normal(size=320) * np.sqrt(2 / 10)
    p[M.OFF_B1:M.OFF_W2] = rng.normal(size=32) * 0.2
    p[M.OFF_W2:M.OFF_B2] = rng.normal(size=1024) * np.sqrt(2 / 32)
    p[M.OFF_B2:M.OFF_W3] = rng.normal(size=32) * 0.2
    p[M.OFF_W3:M.OFF_B3] = rng.normal(size=32) * np.sqrt(1 / 32)
    p[M.OFF_B3] = 0.3
    return p


def _random_rows(seed, n):
    rng = np.random.default_rng(seed)
    f = rng.normal(size=(n, 10)).astype(np.float32) * np.float32(1.5)
    f[: n // 10] *= 20  # some large inputs
    return M.pack_rows(f), (rng.random(n) < 0.25).astype(np.uint8)


@pytest.mark.parametrize("n", [1, 63, 64, 1000, 4099])
def test_mfma_forward_matches_float64(tracker, n):
    say(f"gpu mlp forward n={n}")
    p = _random_params(n)
    X, y = _random_rows(n + 1, n)
    tracker.set_mlp_params(p)
    tracker.train_set(X, y)
    z = tracker.train_scores(n)
    zr = M.forward(p, X)[3]
    err = np.abs(z - zr)
    # FP32 accumulation order differs from float64; a BF16 rounding of a hidden
    # activation may then flip by one unit, hence the bound on the worst row
    assert np.max(err / np.maximum(1.0, np.abs(zr))) < 3e-2
    assert np.quantile(err, 0.99) < 1e-4 * max(1.0, float(np.max(np.abs(zr))))


@pytest.mark.parametrize("mult,offset,first,count", [(1, 0, 0, 3000), (7, 3, 10, 513), (1, 0, 0, 1)])
def test_gradient_matches_float64(tracker, mult, offset, first, count):
    say(f"gpu mlp gradient {mult} {offset} {first} {count}")
    n = 3000
    p = _random_params(11)
    X, y = _random_rows(12, n)
    tracker.set_mlp_params(p)
    tracker.train_set(X, y)
    loss, g = tracker.train_step(first, count, mult, offset, pos_weight=4.0)
    rows = (mult * np.arange(first, first + count) + offset) % n
    lr_, gr = M.loss_and_grad(p, X[rows], y[rows], 4.0)
    assert abs(loss - lr_) < 1e-5 * max(1.0, abs(lr_))
    assert np.linalg.norm(g - gr) < 2e-3 * np.linalg.norm(gr)
    # every block of the parameter vector, not only the dominant one
    for a, b in [(M.OFF_W1, M.OFF_B1), (M.OFF_B1, M.OFF_W2), (M.OFF_W2, M.OFF_B2), (M.OFF_B2, M.OFF_W3),
                 (M.OFF_W3, M.OFF_B3), (M.OFF_B3, M.PARAMS)]:
        assert np.linalg.norm(g[a:b] - gr[a:b]) < 5e-3 * np.linalg.norm(gr[a:b]) + 1e-7, (a, b)
    assert np.array_equal(tracker.mlp_params(), p)  # lr 0: no update


def test_adamw_matches_closed_form(tracker):
    say("gpu adamw")
    p0 = _random_params(21)
    X, y = _random_rows(22, 2000)
    tracker.set_mlp_params(p0)
    tracker.train_set(X, y)
    lr, wd = 2e-3, 0.1
    grads = []
    for first in (0, 1000, 500):
        _, g = tracker.train_step(first, 1000, 1, 0, pos_weight=2.0, lr=lr, weight_decay=wd)
        grads.append(g.astype(np.float64))
    ref = M.adamw(p0, grads, lr, wd)
    np.testing.assert_allclose(tracker.mlp_params(), ref, rtol=1e-5, atol=2e-7)
    # the gradient of each step is taken at the weights of the step before
    _, g_ref = M.loss_and_grad(M.adamw(p0, grads[:1], lr, wd), X[1000:2000], y[1000:2000], 2.0)
    assert np.linalg.norm(grads[1] - g_ref) < 2e-3 * np.linalg.norm(g_ref)


def _reference_rows(tracker, item, vertex):
    """Features and labels of every tracklet of one item, restated on the host
    from the downloaded clusters, tracklets, cells and roads."""
    n = tracker.counts(item)
    cl = [tracker.download(item, itsca.CLUSTERS, l, n.clusters[l], 8) for l in range(itsca.LAYERS)]
    cols = []
    for l in range(itsca.LAYERS):
        c = cl[l]
        f = c[:, [4, 5, 6, 2, 1]].copy().view(np.float32)
        cols.append({"x": f[:, 0], "y": f[:, 1], "z": f[:, 2], "r": f[:, 3], "phi": f[:, 4]})
    trk = [tracker.download(item, itsca.TRACKLETS, l, n.tracklets[l], 4) for l in range(itsca.TRACKLET_LAYERS)]
    cnt = [np.bincount(trk[l][:, 0], minlength=n.clusters[l]) if len(trk[l]) else np.zeros(n.clusters[l], int)
           for l in range(itsca.TRACKLET_LAYERS)]
    labels = [np.zeros(n.tracklets[l], np.uint8) for l in range(itsca.TRACKLET_LAYERS)]
    if n.roads:
        cells = [tracker.download(item, itsca.CELLS, k, n.cells[k], 10) for k in range(itsca.CELL_LAYERS)]
        r = tracker.roads(item)
        for k in range(itsca.CELL_LAYERS):
            labels[k][cells[k][r[:, k], 3]] = 1
            labels[k + 1][cells[k][r[:, k], 4]] = 1
    X, Y = [], []
    for l in range(itsca.TRACKLET_LAYERS):
        if not n.tracklets[l]:
            continue
        a, b = trk[l][:, 0], trk[l][:, 1]
        ci = {k: v[a] for k, v in cols[l].items()}
        cj = {k: v[b] for k, v in cols[l + 1].items()}
        cj_cnt = cnt[l + 1][b] if l + 1 < itsca.TRACKLET_LAYERS else np.zeros(len(b), int)
        X.append(M.pack_rows(M.tracklet_features(ci, cj, vertex, l, cnt[l][a], cj_cnt)))
        Y.append(labels[l])
    if not X:
        return np.zeros((0, 16), np.uint16), np.zeros(0, np.uint8)
    return np.concatenate(X), np.concatenate(Y)


def test_training_rows_match_restatement(tracker):
    say("gpu training rows: features and labels")
    evs = [itsca.load_events(p)[0] for _, p, _ in CASES] + itsca.generate_events(5, 41, 1)
    tracker.set_filter(False)
    tracker.set_events(evs)
    tracker.run()
    vertices = [v for ev in evs for v in ev.vertices]
    assert len(vertices) == tracker.num_items()
    Xr = np.concatenate([_reference_rows(tracker, i, vertices[i])[0] for i in range(len(vertices))])
    Yr = np.concatenate([_reference_rows(tracker, i, vertices[i])[1] for i in range(len(vertices))])
    total = sum(sum(tracker.counts(i).tracklets) for i in range(len(vertices)))
    assert len(Yr) == total and 0 < Yr.sum() < total
    tracker.train_clear(total + 10)
    assert tracker.train_collect(1, 0) == total
    X, Y = tracker.train_get(total)
    assert np.array_equal(Y, Yr)
    assert np.array_equal(X, Xr)
    # a strided sample continues the set; capacity bounds what is kept
    tracker.train_clear(total)
    k = tracker.train_collect(5, 2)
    assert k == len(range(2, total, 5))
    X5, Y5 = tracker.train_get(k)
    assert np.array_equal(X5, Xr[2::5]) and np.array_equal(Y5, Yr[2::5])
    k2 = tracker.train_collect(1, 0)
    assert k2 == total - k and tracker.train_size() == total


def test_filter_keeping_everything_is_the_plain_ca(tracker):
    say("gpu filter with threshold -inf")
    tracker.mlp_init(5)
    tracker.set_filter(True, -3.0e38)
    evs = [itsca.load_events(p)[0] for _, p, _ in CASES]
    tracker.set_events(evs)
    tracker.run()
    item = 0
    for (name, _, nv), ev in zip(CASES, evs):
        for v in range(nv):
            assert itsca.dump_item(tracker, item, ev.vertices[v]) == fixture(name, "stable", v), name
            item += 1
    tracker.set_filter(True, 3.0e38)
    tracker.run()
    assert all(sum(tracker.counts(i).tracklets) == 0 for i in range(item))
    tracker.set_filter(False)


def _cell_keys(tracker, item):
    n = tracker.counts(item)
    ids = [tracker.download(item, itsca.CLUSTERS, l, n.clusters[l], 8)[:, 0] for l in range(itsca.LAYERS)]
    keys = []
    for k in range(itsca.CELL_LAYERS):
        c = tracker.download(item, itsca.CELLS, k, n.cells[k], 10)
        keys += [(k, a, b, d) for a, b, d in zip(ids[k][c[:, 0]].tolist(), ids[k + 1][c[:, 1]].tolist(),
                                                 ids[k + 2][c[:, 2]].tolist())]
    return Counter(keys)


@pytest.mark.parametrize("threshold", [-2.0, 0.0, 1.0])
def test_filtered_roads_are_plain_roads(tracker, threshold):
    say(f"gpu filter subset threshold={threshold}")
    evs = itsca.generate_events(2, 7, 1) + itsca.generate_events(5, 8, 1)
    tracker.set_filter(False)
    tracker.set_events(evs)
    tracker.run()
    items = tracker.num_items()
    plain = [Counter(itsca.road_tracks(tracker, i)) for i in range(items)]
    plain_cells = [_cell_keys(tracker, i) for i in range(items)]
    tracker.set_mlp_params(_random_params(99))
    tracker.set_filter(True, threshold)
    tracker.run()
    kept = 0
    for i in range(items):
        learned = Counter(itsca.road_tracks(tracker, i))
        assert not (learned - plain[i]), i
        assert not (_cell_keys(tracker, i) - plain_cells[i]), i
        kept += sum(learned.values())
    eager = [itsca.road_tracks(tracker, i) for i in range(items)]
    tracker.set_graph(True)  # the filter inside the captured pipeline
    for _ in range(3):
        tracker.run()
        assert [itsca.road_tracks(tracker, i) for i in range(items)] == eager
    tracker.set_graph(False)
    tracker.set_filter(False)
    say(f"  roads kept {kept} of {sum(sum(p.values()) for p in plain)}")


@pytest.mark.parametrize("quantile", [0.3, 0.9])
def test_filter_keeps_exactly_the_tracklets_scored_above_threshold(tracker, quantile):
    """The in-pipeline scoring sees the same features and runs the same MFMA
    network as the training rows, so the filtered tracklet list is the plain
    list restricted to rows with logit >= threshold, in the same order
    (clusters whose candidates overflowed their slots are kept whole)."""
    say(f"gpu filter keeps exactly the rows above the {quantile} quantile")
    evs = itsca.generate_events(2, 31, 1) + itsca.generate_events(5, 32, 1) + [dense_spot_event()]
    tracker.set_filter(False)
    tracker.set_events(evs)
    tracker.run()
    items = tracker.num_items()
    plain = [[tracker.download(i, itsca.TRACKLETS, l, tracker.counts(i).tracklets[l], 4)[:, :2]
              for l in range(itsca.TRACKLET_LAYERS)] for i in range(items)]
    total = sum(len(t) for it in plain for t in it)
    tracker.set_mlp_params(_random_params(5))
    tracker.train_clear(total)
    assert tracker.train_collect(1, 0) == total
    z = tracker.train_scores(total)
    threshold = float(np.quantile(z, quantile))
    tracker.set_filter(True, threshold)
    tracker.run()
    k = 0
    dropped = overflowed = 0
    for i in range(items):
        n = tracker.counts(i)
        for l in range(itsca.TRACKLET_LAYERS):
            t = plain[i][l]
            raw = np.bincount(t[:, 0], minlength=n.clusters[l]) if len(t) else np.zeros(n.clusters[l], int)
            keep = (z[k:k + len(t)] >= threshold) | (raw[t[:, 0]] > itsca.TRACKLET_SLOTS)
            k += len(t)
            got = tracker.download(i, itsca.TRACKLETS, l, n.tracklets[l], 4)[:, :2]
            assert np.array_equal(got, t[keep]), (i, l)
            dropped += int((~keep).sum())
            overflowed += int((raw > itsca.TRACKLET_SLOTS).sum())
    assert k == total and 0 < dropped < total
    assert overflowed > 0  # the dense spot's clusters are kept whole
    tracker.set_filter(False)


@pytest.mark.parametrize("quantile", [0.3, 0.9])
def test_gated_filter_keeps_exactly_the_gated_tracklets_scored_above_threshold(tracker, quantile):
    """With phi gates, a filtered run that keeps every candidate builds exactly
    the plain tracklets inside the gates, in the plain order (restated on the
    host from the downloaded cluster phis), and a threshold then keeps exactly
    those of them whose training row (gated counts) scores at or above it;
    clusters whose gated candidates overflowed their slots are kept whole."""
    from itsca import filter_training as ft
    say(f"gpu gated filter keeps exactly the gated rows above the {quantile} quantile")
    gates = [0.02, 0.02, 0.15, 0.05, 0.1, 0.05]
    evs = itsca.generate_events(2, 33, 1) + itsca.generate_events(5, 34, 1) + [dense_spot_event()]
    tracker.set_filter_gates(None)
    tracker.set_filter(False)
    tracker.set_events(evs)
    tracker.run()
    items = tracker.num_items()
    plain = [[tracker.download(i, itsca.TRACKLETS, l, tracker.counts(i).tracklets[l], 4)
              for l in range(itsca.TRACKLET_LAYERS)] for i in range(items)]
    tracker.set_mlp_params(_random_params(6))
    tracker.set_filter_gates(gates)
    tracker.set_filter(True, ft.KEEP_ALL)
    tracker.run()
    gated = [gated_tracklets(tracker, i, plain[i], gates) for i in range(items)]
    for i in range(items):
        n = tracker.counts(i)
        for l in range(itsca.TRACKLET_LAYERS):
            assert np.array_equal(tracker.download(i, itsca.TRACKLETS, l, n.tracklets[l], 4)[:, :2], gated[i][l]), (i, l)
    total = sum(len(t) for it in gated for t in it)
    n_plain = sum(len(t) for it in plain for t in it)
    say(f"  gated candidates {total} of {n_plain} plain tracklets")
    assert 0 < total < 0.7 * n_plain
    tracker.train_clear(total)
    assert tracker.train_collect(1, 0) == total
    z = tracker.train_scores(total)
    threshold = float(np.quantile(z, quantile))
    tracker.set_filter(True, threshold)
    tracker.run()
    k = dropped = 0
    for i in range(items):
        n = tracker.counts(i)
        for l in range(itsca.TRACKLET_LAYERS):
            t = gated[i][l]
            raw = np.bincount(t[:, 0], minlength=n.clusters[l]) if len(t) else np.zeros(n.clusters[l], int)
            keep = (z[k:k + len(t)] >= threshold) | (raw[t[:, 0]] > itsca.TRACKLET_SLOTS)
            k += len(t)
            got = tracker.download(i, itsca.TRACKLETS, l, n.tracklets[l], 4)[:, :2]
            assert np.array_equal(got, t[keep]), (i, l)
            dropped += int((~keep).sum())
    assert k == total and 0 < dropped < total
    tracker.set_filter(False)
    tracker.set_filter_gates(None)


def _device_ms(tracker, runs=3):
    """(all kernels, cell-building kernels) in ms per run, from HIP events."""
    tracker.set_profiling(True)
    total = cells = 0.0
    for _ in range(runs):
        tracker.run()
        kt = tracker.kernel_times()
        total += sum(ms for _, ms in kt) / runs
        cells += sum(ms for name, ms in kt if name in ("kCellsCount", "kCellsFill")) / runs
    tracker.set_profiling(False)
    return total, cells


def test_fused_filter_equals_separate_filter(tracker):
    """The filter scored inside the tracklet search (one launch per layer)
    keeps exactly the tracklets of the filter scored by its own kernel: every
    array of every item, on central events and the fixtures, at the shipped
    threshold and at the median score of a random network."""
    from itsca import filter_training as ft
    say("gpu fused filter vs separate filter")
    evs = itsca.generate_events(2, 30100, 6) + [itsca.load_events(p)[0] for _, p, _ in CASES]
    tracker.set_events(evs)
    items = tracker.num_items()
    params, thr, _ = ft.load_filter()
    tracker.set_filter_gates(ft.load_gates())  # both forms search inside the shipped phi gates
    for p, t in ((params, thr), (_random_params(41), None)):
        tracker.set_mlp_params(p)
        if t is None:
            tracker.set_filter(False)
            tracker.run()
            total = sum(sum(tracker.counts(i).tracklets) for i in range(items))
            tracker.train_clear(total)
            tracker.train_collect(1, 0)
            t = float(np.median(tracker.train_scores(total)))
        outs = []
        for fused in (False, True):
            tracker.set_filter(True, t, fused=fused)
            tracker.run()
            outs.append([itsca.dump_item(tracker, i, (0.0, 0.0, 0.0)) for i in range(items)])
        kept = sum(sum(tracker.counts(i).tracklets) for i in range(items))
        say(f"  threshold {t:.4f}: {kept} tracklets kept, fused == separate: {outs[0] == outs[1]}")
        assert outs[0] == outs[1]
    tracker.set_filter(False)
    tracker.set_filter_gates(None)


def test_trained_filter_gain_over_plain_and_untrained(tracker):
    """Ablation on held-out central events (seeds the training never saw): the
    shipped network keeps >= 99.5 % of the plain tracks, finds none outside
    them, and cuts the device time; held to the same recall of positive rows,
    it drops far more of the negative rows than an untrained network (the
    gates have already dropped the easy ones, so the rows compared are the
    gated candidates and the measure is the share of negatives dropped)."""
    from itsca import filter_training as ft
    say("gpu ablation: trained filter vs plain and vs untrained")
    evs = itsca.generate_events(2, 30000, 24)
    thr = ft.prepare(tracker)  # the shipped weights and phi gates
    gates = ft.load_gates()
    r = ft.evaluate(tracker, evs, thr)
    say(f"  trained: recall {r['recall']:.5f}, outside {r['extra']}, tracklets kept {r['tracklets_kept']:.3f}, "
        f"road cells kept {r['road_cells_kept']:.5f}")
    assert r["extra"] == 0
    assert r["recall"] >= 0.995
    assert r["road_cells_kept"] >= 0.995
    assert r["tracklets_kept"] < 0.3
    tracker.set_filter_gates(None)
    everything = ft.evaluate(tracker, evs, -3.0e38, plain=r["plain"])  # no gates, a cut no score fails
    assert (everything["recall"], everything["road_cells_kept"], everything["cells_kept"]) == (1.0, 1.0, 1.0)
    tracker.set_filter_gates(gates)
    tracker.set_filter(True, thr)
    learned_ms, learned_cells = _device_ms(tracker)
    tracker.set_filter(False)
    plain_ms, plain_cells = _device_ms(tracker)
    say(f"  device time per run: plain {plain_ms:.3f} ms (cells {plain_cells:.3f}), "
        f"learned {learned_ms:.3f} ms (cells {learned_cells:.3f})")
    assert learned_cells < 0.5 * plain_cells
    assert learned_ms < 0.9 * plain_ms

    # rows as the network sees them: the gated candidates, from a run that keeps all of them
    tracker.set_filter(True, ft.KEEP_ALL)
    tracker.run()
    total = sum(sum(tracker.counts(i).tracklets) for i in range(tracker.num_items()))
    tracker.train_clear(total)
    tracker.train_collect(1, 0)
    tracker.set_filter(False)
    tracker.set_filter_gates(None)
    _, y = tracker.train_get(total)
    z_trained = tracker.train_scores(total)
    row_recall = float(np.mean(z_trained[y > 0] >= thr))
    tracker.mlp_init(123)
    z_untrained = tracker.train_scores(total)
    thr_u = float(np.quantile(z_untrained[y > 0], 1.0 - row_recall))
    kept_u = float(np.mean(z_untrained[y == 0] >= thr_u))
    kept_t = float(np.mean(z_trained[y == 0] >= thr))
    say(f"  positive-row recall {row_recall:.5f}: negative rows kept: trained {kept_t:.3f}, untrained {kept_u:.3f}")
    assert 1.0 - kept_t > 0.2
    assert 1.0 - kept_t > 5.0 * (1.0 - kept_u)


def test_training_end_to_end(tracker):
    """The whole training at a small size: labels from plain runs over 32
    central events, four AdamW epochs from a fresh initialisation, the
    threshold tuned on 8 validation events; then, on 8 events neither set
    holds, the trained filter keeps the plain roads (recall >= 99 %, none
    outside) while cutting most tracklets, and its training rows separate far
    better than the initial network's (negative rows dropped at equal recall
    of the positive ones)."""
    from itsca import filter_training as ft
    say("gpu training end to end: collect, train, tune, evaluate on held-out events")
    params, thr, meta = ft.train_filter(tracker, 2, range(11000, 11032), 2e6, range(11500, 11508), epochs=4,
                                        target_recall=0.995)
    say(f"  {meta['rows']} rows, final loss {meta['final_loss']:.5f}, threshold {thr:.4f}, "
        f"validation recall {meta['validation_road_recall']:.5f}")
    assert meta["validation_road_recall"] >= 0.995 and thr > -3.0e38
    n = meta["rows"]
    _, y = tracker.train_get(n)
    z_trained = tracker.train_scores(n)
    tracker.mlp_init(7)
    z_initial = tracker.train_scores(n)
    recall = float(np.mean(z_trained[y > 0] >= thr))
    kept_t = float(np.mean(z_trained[y == 0] >= thr))
    kept_i = float(np.mean(z_initial[y == 0] >= np.quantile(z_initial[y > 0], 1.0 - recall)))
    say(f"  training rows at positive recall {recall:.5f}: negative rows kept: trained {kept_t:.3f}, "
        f"initial {kept_i:.3f}")
    assert 1.0 - kept_t > 0.1
    assert 1.0 - kept_t > 4.0 * (1.0 - kept_i)
    assert len(meta["phi_gates"]) == itsca.TRACKLET_LAYERS and max(meta["phi_gates"]) <= 0.3

    tracker.set_mlp_params(params)
    r = ft.evaluate(tracker, itsca.generate_events(2, 31000, 8), thr)
    say(f"  held out: recall {r['recall']:.5f}, outside {r['extra']}, tracklets kept {r['tracklets_kept']:.3f}, "
        f"road cells kept {r['road_cells_kept']:.5f}")
    assert r["extra"] == 0
    assert r["recall"] >= 0.99
    assert r["tracklets_kept"] < 0.4
    tracker.set_filter(False)
    tracker.set_filter_gates(None)


def test_cli_plain_and_with_learned_filter(golden_dir):
    """The reference-style CLI (bin/itsca-track) finds the fixture's road count
    plain, and a subset of it with --filter and the shipped weights."""
    import os
    import re
    import subprocess

    from itsca import filter_training as ft
    say("gpu cli plain and --filter")
    exe = os.path.join(os.path.dirname(itsca.__file__), "..", "bin", "itsca-track")
    events = os.path.join(golden_dir, "small_s1.events.txt")

    def roads(extra):
        out = subprocess.run([exe, events] + extra, capture_output=True, text=True, timeout=120)
        assert out.returncode == 0, out.stdout + out.stderr
        return [int(n) for n in re.findall(r"Found (\d+) roads for vertex", out.stdout)]

    want = int(re.search(r"^roads (\d+)$", fixture("small_s1", "stable", 0), re.M).group(1))
    plain = roads([])
    learned = roads(["--filter", ft.DEFAULT_WEIGHTS])
    assert plain == [want]
    assert len(learned) == 1 and 0 < learned[0] <= want
