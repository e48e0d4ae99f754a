"""The CPU oracle against the reference CPU tracker's own output (golden
fixtures written from the reference built from its sources), and against
closed forms and brute force for what the fixtures do not cover."""
import hashlib
import os
import tempfile

import numpy as np
import pytest

import itsca
import pyoracle
from golden_cases import central, fixture, small_cases

CASES = small_cases()


@pytest.mark.parametrize("name,path,nv", CASES, ids=[c[0] for c in CASES])
def test_oracle_matches_reference_dump(name, path, nv):
    ev = itsca.load_events(path)[0]
    assert len(ev.vertices) == nv
    for v in range(nv):
        assert pyoracle.OracleVertex(ev, v, stable=False).dump() == fixture(name, "ref", v)
        assert pyoracle.OracleVertex(ev, v, stable=True).dump() == fixture(name, "stable", v)


def test_oracle_matches_reference_central_event():
    info = central()
    with tempfile.TemporaryDirectory() as tmp:
        p = os.path.join(tmp, "c.txt")
        itsca.write_events(info["preset"], info["seed"], 1, p)
        assert hashlib.sha256(open(p, "rb").read()).hexdigest() == info["events_sha256"]
        ev = itsca.load_events(p)[0]
    for stable, key in ((False, "ref_sha256"), (True, "stable_sha256")):
        text = pyoracle.OracleVertex(ev, 0, stable=stable).dump()
        assert hashlib.sha256(text.encode()).hexdigest() == info[key]


def test_atan2f_matches_glibc():
    # the kernels' own atan2f (csrc/itsca_math.h) against glibc atan2f
    assert pyoracle.check_atan2f(4_000_000, 12345) == 0


# --- closed forms and brute force -------------------------------------------

TWO_PI = np.float32(2.0) * np.float32(3.14159265359)


def _clusters(o, layer):
    c = o.records(0, layer)
    return {"id": c[:, 0], "phi": c[:, 1].view(np.float32), "r": c[:, 2].view(np.float32), "bin": c[:, 3],
            "x": c[:, 4].view(np.float32), "y": c[:, 5].view(np.float32), "z": c[:, 6].view(np.float32)}


@pytest.mark.parametrize("seed", [11, 12])
def test_index_table_is_scan_of_bin_histogram(seed):
    ev = itsca.generate_events(1, seed)[0]
    o = pyoracle.OracleVertex(ev, 0, stable=True)
    d = __import__("dumps").parse(o.dump())
    for l in range(1, 7):
        bins = _clusters(o, l)["bin"]
        assert np.all(np.diff(bins) >= 0)
        hist = np.bincount(bins, minlength=401)[:401]
        expect = np.concatenate([[0], np.cumsum(hist)[:-1]])
        assert [int(v) for v in d[("table", l)]] == expect.tolist()


@pytest.mark.parametrize("seed", [21, 22])
def test_tracklet_set_equals_all_pairs_brute_force(seed):
    """The binned window is a superset of the cuts: the distinct accepted pairs
    equal all cluster pairs passing the tracklet cuts (Tracker.cxx:93-100)."""
    ev = itsca.generate_events(1, seed)[0]
    o = pyoracle.OracleVertex(ev, 0, stable=True)
    vz = np.float32(ev.vertices[0][2])
    thr = [0.1, 0.1, 0.3, 0.3, 0.3, 0.3]
    for l in range(6):
        a, b = _clusters(o, l), _clusters(o, l + 1)
        tl = ((a["z"] - vz) / a["r"]).astype(np.float32)[:, None]
        dz = np.abs(tl * (b["r"][None, :] - a["r"][:, None]) + a["z"][:, None] - b["z"][None, :])
        dphi = np.abs(a["phi"][:, None] - b["phi"][None, :])
        ok = (dz < np.float32(thr[l])) & ((dphi < np.float32(0.3)) | (np.abs(dphi - TWO_PI) < np.float32(0.3)))
        brute = set(zip(*np.nonzero(ok)))
        t = o.records(1, l)
        assert set(zip(t[:, 0].tolist(), t[:, 1].tolist())) == {(int(i), int(j)) for i, j in brute}


def test_cell_curvature_is_circumcircle():
    """With the vertex on the beam axis the cell's conformal-plane circle is the
    circumcircle of its three clusters in x-y (Tracker.cxx:175-218)."""
    ev = itsca.generate_events(1, 31)[0]
    o = pyoracle.OracleVertex(ev, 0, stable=True)
    checked = 0
    for l in range(5):
        cl = [_clusters(o, l + k) for k in range(3)]
        for row in o.records(2, l):
            p = [(float(cl[k]["x"][row[k]]), float(cl[k]["y"][row[k]])) for k in range(3)]
            (ax, ay), (bx, by), (cx, cy) = p
            dd = 2 * (ax * (by - cy) + bx * (cy - ay) + cx * (ay - by))
            ux = ((ax * ax + ay * ay) * (by - cy) + (bx * bx + by * by) * (cy - ay) + (cx * cx + cy * cy) * (ay - by)) / dd
            uy = ((ax * ax + ay * ay) * (cx - bx) + (bx * bx + by * by) * (ax - cx) + (cx * cx + cy * cy) * (bx - ax)) / dd
            radius = np.hypot(ax - ux, ay - uy)
            curv = row[8:9].view(np.float32)[0]
            assert abs(1.0 / radius - curv) <= 2e-3 * abs(curv) + 1e-6
            dca = abs(radius - np.hypot(ux, uy))
            assert dca <= [0.05, 0.04, 0.05, 0.2, 0.4][l] * 1.01 + 1e-4
            checked += 1
    assert checked > 300


def test_levels_and_roads_are_longest_chains():
    """level(n) = 1 + max level of its neighbours, and the roads are exactly the
    5-cell chains with levels 1..5, ordered by (c4, c3, c2, c1, c0)."""
    dumps = __import__("dumps")
    ev = itsca.generate_events(1, 41)[0]
    d = dumps.parse(pyoracle.OracleVertex(ev, 0, stable=True).dump())
    level = {l: [int(r[9]) for r in d[("cells", l)]] for l in range(5)}
    preds = {l: [[int(v) for v in r[1:]] for r in d[("neighbours", l)]] for l in range(4)}
    assert all(v == 1 for v in level[0])
    for l in range(4):
        for n, ps in enumerate(preds[l]):
            assert level[l + 1][n] == 1 + max([level[l][p] for p in ps], default=0)
            assert ps == sorted(ps)
    chains = []
    for c4, lv in enumerate(level[4]):
        if lv != 5:
            continue

        def walk(layer, cell, acc):
            if layer == 0:
                chains.append(tuple(reversed(acc)))
                return
            for p in preds[layer - 1][cell]:
                if level[layer - 1][p] == layer:
                    walk(layer - 1, p, acc + [p])

        walk(4, c4, [c4])
    roads = [tuple(int(v) for v in r[:5]) for r in d["roads"]]
    assert roads == sorted(chains, key=lambda c: c[::-1])
    assert len(roads) > 50


# --- the two sort orders of the reference -----------------------------------


@pytest.mark.parametrize("name,path,nv", CASES, ids=[c[0] for c in CASES])
def test_sort_orders_differ_only_in_inclusive_bound_duplicates(name, path, nv):
    """The reference as-is (std::sort) and the stable-order build give the same
    cluster-id-keyed tracklets, cells and roads once multiplicities are
    dropped.  Without the extra candidate of the inclusive bound
    (Tracker.cxx:84-89) both orders give that same set with every tracklet and
    cell once, so every multiplicity difference is a duplicate of that extra
    candidate (PrimaryVertexContext.cxx:52-54 decides which cluster it is)."""
    from dumps import keyed

    ev = itsca.load_events(path)[0]
    for v in range(nv):
        ref, stable = keyed(fixture(name, "ref", v)), keyed(fixture(name, "stable", v))
        for sort_stable in (False, True):
            once = keyed(pyoracle.OracleVertex(ev, v, stable=sort_stable, exclusive_bound=True).dump())
            for stage in ref:
                assert set(ref[stage]) == set(stable[stage]) == set(once[stage]), (name, v, stage)
                if stage != "roads":
                    assert max(once[stage].values(), default=1) == 1, (name, v, stage)
                # the duplicates only add records
                assert all(ref[stage][k] >= 1 and stable[stage][k] >= 1 for k in once[stage])


def test_central_event_sort_orders_have_the_same_keyed_sets():
    from dumps import distinct_digest, duplicates

    info = central()
    with tempfile.TemporaryDirectory() as tmp:
        p = os.path.join(tmp, "c.txt")
        itsca.write_events(info["preset"], info["seed"], 1, p)
        ev = itsca.load_events(p)[0]
    assert info["ref_distinct_sha256"] == info["stable_distinct_sha256"]
    once = pyoracle.OracleVertex(ev, 0, stable=False, exclusive_bound=True).dump()
    assert distinct_digest(once) == info["ref_distinct_sha256"]
    assert all(n == 0 for n in duplicates(once).values())


def test_dca_rounding_bound_holds():
    """The first-order bound on the float32 rounding of the cell DCA chain
    (tests/fp_bounds.py), which the contraction study uses to judge DCA
    disagreements, against float32 evaluations of the chain: one rounding
    per operation, and every multiply-add fused; both stay within the bound
    of the float64 value for the cells of a central event."""
    from fp_bounds import dca_float32, dca_with_bound

    ev = itsca.generate_events(2, 20000, 1)[0]
    o = pyoracle.OracleVertex(ev, 0, stable=True)
    cl = [_clusters(o, l) for l in range(7)]
    worst = [0.0, 0.0]
    n = 0
    for l in range(5):
        cells = o.records(2, l)
        for row in cells[:: max(1, len(cells) // 300)]:
            xs = [float(cl[l + k]["x"][row[k]]) for k in range(3)]
            ys = [float(cl[l + k]["y"][row[k]]) for k in range(3)]
            v, e, _ = dca_with_bound(xs, ys, 0.0, 0.0)
            if not (np.isfinite(v) and np.isfinite(e)):
                continue
            n += 1
            for k, fused in enumerate((False, True)):
                d = abs(dca_float32(xs, ys, 0.0, 0.0, fused) - v)
                worst[k] = max(worst[k], d / e)
    assert n > 1000
    assert worst[0] <= 1.0 and worst[1] <= 1.0, worst
