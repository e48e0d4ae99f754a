"""FP contraction study.  The kernels built with FP contraction on
(lib/libitsca_hip_fma.so: the same sources, a*b+c fused into one rounding)
run beside the product library (contraction off, which the parity tests hold
bit-exact with the reference's CPU tracker) on the same central events.

Tracklets, cells, neighbour edges and roads are compared as multisets of input
cluster ids, which do not depend on sort positions.  Every disagreement is
either inherited (an element it is built from already differs, or for a road
one of its cells or edges) or direct.  A direct one must lie near a cut
boundary: for some cut of the element, |q - threshold| <= k * ulp(s), with q
the cut's quantity evaluated in float64 from either run's float32 inputs and s
the largest magnitude among the terms of that evaluation (the scale at which
float32 rounds it).  k = 1 for every cut but the cell DCA.

The DCA is |R - |centre||, the difference of two terms of up to ~1e4 cm, each
the end of a chain of about 40 float32 operations from the cluster
coordinates (r = sqrtf(dx*dx + dy*dy), the plane normal, its norm, a divide,
two square roots; Tracker.cxx:175-212, Cluster.cxx:39).  There the bound is
derived per case (dca_error_bound): the rounding error of the float32 chain,
propagated to first order operation by operation in float64 (each result
within half an ulp, 2^-24 relative, of its exact value, plus its inputs'
errors times the partial derivatives), starting from the raw x, y that both
builds share.  Contraction only removes roundings from that chain, so each
build's DCA lies within the bound of the exact value; the two builds can then
disagree about the cut only if |DCA_exact - threshold| <= bound, which is
checked for every direct DCA disagreement.  How many lie within 1, 2 and 3
float32 ulps of max(R, |centre|) is printed, against the 1-ulp target.

Cut quantities (reference Tracker.cxx):
  tracklets :93-100  |tanL_vtx*(r_j - r_i) + z_i - z_j| < dZ[l];  |phi_i - phi_j| < 0.3 or its 2pi wrap
  cells     :161-216 |dtanL| < 0.025; |dphi| < 0.14 or wrap; |-avgTanL*r1 + z1 - vz| < dZ[l]; DCA <= dca[l]
  edges     :448-459 |n_c - n_n|^2 < dN[l]; |k_c - k_n| < dK[l]
"""
import os
import sys

import numpy as np
import pytest

ROOT = os.path.abspath(os.path.join(os.path.dirname(__file__), ".."))
sys.path[:0] = [os.path.join(ROOT, "tracking-itsu_amd")]
import ctypes  # noqa: E402

import itsca  # noqa: E402

pytestmark = pytest.mark.gpu

F = np.float32
TWO_PI = float(F(2.0) * F(3.14159265359))
TRK_DZ = [float(F(v)) for v in (0.1, 0.1, 0.3, 0.3, 0.3, 0.3)]
PHI_CUT = float(F(0.3))
CELL_DTANL, CELL_DPHI = float(F(0.025)), float(F(0.14))
CELL_DZ = [float(F(v)) for v in (0.2, 0.4, 0.5, 0.6, 3.0)]
CELL_DCA = [float(F(v)) for v in (0.05, 0.04, 0.05, 0.2, 0.4)]
NB_DN = [float(F(v)) for v in (0.002, 0.009, 0.002, 0.005)]
NB_DK = [float(F(v)) for v in (0.008, 0.0025, 0.003, 0.0035)]
EVENTS = 8  # central events (8000 tracks), seeds 20000..20007


def say(msg):
    print(msg, flush=True)


def ulp(x):
    return np.spacing(np.abs(np.asarray(x, dtype=np.float64)).astype(np.float32)).astype(np.float64)


def f32col(a, k):
    return np.ascontiguousarray(a[:, k]).view(np.float32).astype(np.float64)


def collect(lib, events):
    """Per item: clusters by id, and tracklets, cells, edges, roads keyed by cluster ids."""
    t = itsca.GpuTracker(device=0, lib=lib)
    t.set_events(events)
    t.run()
    items = []
    for w, ev in enumerate(events):
        n = t.counts(w)
        cl = []
        for l in range(itsca.LAYERS):
            c = t.download(w, itsca.CLUSTERS, l, n.clusters[l], 8)
            cl.append({"id": c[:, 0].copy(), "phi": f32col(c, 1), "r": f32col(c, 2), "bin": c[:, 3].copy(),
                       "x": f32col(c, 4), "y": f32col(c, 5), "z": f32col(c, 6),
                       "phibits": c[:, 1].copy(), "rbits": c[:, 2].copy()})
            order = np.argsort(cl[-1]["id"], kind="stable")
            cl[-1]["sorted_ids"], cl[-1]["order"] = cl[-1]["id"][order], order
        trk, tval = [], []
        for l in range(itsca.TRACKLET_LAYERS):
            a = t.download(w, itsca.TRACKLETS, l, n.tracklets[l], 4)
            keys = list(zip(cl[l]["id"][a[:, 0]].tolist(), cl[l + 1]["id"][a[:, 1]].tolist()))
            trk.append(keys)
            vals = list(zip(f32col(a, 2).tolist(), f32col(a, 3).tolist()))
            tval.append(dict(zip(reversed(keys), reversed(vals))))  # first occurrence of a key
        cells, cval = [], []
        for l in range(itsca.CELL_LAYERS):
            a = t.download(w, itsca.CELLS, l, n.cells[l], 10)
            keys = list(zip(cl[l]["id"][a[:, 0]].tolist(), cl[l + 1]["id"][a[:, 1]].tolist(),
                            cl[l + 2]["id"][a[:, 2]].tolist()))
            cells.append(keys)
            vals = list(np.stack([f32col(a, q) for q in (5, 6, 7, 8)], axis=1)) if len(a) else []
            cval.append(dict(zip(reversed(keys), reversed(vals))))
        edges = []
        for l in range(itsca.CELL_LAYERS - 1):
            cn = n.cells[l + 1]
            off = t.download(w, itsca.NEIGHBOUR_OFFSETS, l, cn + 1, 1)
            lst = t.download(w, itsca.NEIGHBOURS, l, n.neighbours[l], 1)
            keys = []
            lst, off = lst.tolist(), off.tolist()
            for m in range(cn):
                for c in lst[off[m]:off[m + 1]]:
                    keys.append((cells[l][c], cells[l + 1][m]))
            edges.append(keys)
        items.append({"cl": cl, "trk": trk, "tval": tval, "cells": cells, "cval": cval, "edges": edges,
                      "roads": itsca.road_tracks(t, w), "vertex": [float(v) for v in ev.vertices[0]]})
    t.close()
    return items


def mdiff(a, b):
    """Keys whose multiplicity differs: {key: (count in a, count in b)}."""
    from collections import Counter
    ca, cb = Counter(a), Counter(b)
    return {k: (ca[k], cb[k]) for k in set(ca) | set(cb) if ca[k] != cb[k]}


def rows_by_id(layer):
    return layer["sorted_ids"], layer["order"]


def cluster_values(run, l, cid):
    ids, order = rows_by_id(run["cl"][l])
    p = order[np.searchsorted(ids, cid)]
    c = run["cl"][l]
    return {k: float(c[k][p]) for k in ("x", "y", "z", "r", "phi")}


def tracklet_margin(run, l, key):
    """Smallest |q - thr| / ulp(scale) over the tracklet cuts, from this run's inputs."""
    vz = run["vertex"][2]
    a, b = cluster_values(run, l, key[0]), cluster_values(run, l + 1, key[1])
    tan_l = (a["z"] - vz) / a["r"]
    term = tan_l * (b["r"] - a["r"])
    qz = term + a["z"] - b["z"]
    mz = abs(abs(qz) - TRK_DZ[l]) / ulp(max(abs(term), abs(a["z"]), abs(b["z"])))
    dphi = abs(a["phi"] - b["phi"])
    mphi = min(abs(dphi - PHI_CUT) / ulp(max(a["phi"], b["phi"])), abs(abs(dphi - TWO_PI) - PHI_CUT) / ulp(TWO_PI))
    return min((float(mz), "tracklet dz"), (float(mphi), "tracklet dphi"))


def cell_margin(run, l, key):
    vx, vy, vz = run["vertex"]
    t, u = run["tval"][l].get(key[:2]), run["tval"][l + 1].get(key[1:])
    if t is None or u is None:
        return (np.inf, "missing input")
    c1, c2, c3 = (cluster_values(run, l + k, key[k]) for k in range(3))
    out = []
    dtl = abs(t[0] - u[0])
    out.append((abs(dtl - CELL_DTANL) / ulp(max(abs(t[0]), abs(u[0]))), "cell dtanl"))
    dph = abs(t[1] - u[1])
    out.append((min(abs(dph - CELL_DPHI) / ulp(max(abs(t[1]), abs(u[1]))),
                    abs(abs(dph - TWO_PI) - CELL_DPHI) / ulp(TWO_PI)), "cell dphi"))
    avg = 0.5 * (t[0] + u[0])
    q = -avg * c1["r"] + c1["z"] - vz
    out.append((abs(abs(q) - CELL_DZ[l]) / ulp(max(abs(avg * c1["r"]), abs(c1["z"]), abs(vz))), "cell dz"))
    # conformal fit, Tracker.cxx:175-216, in float64
    r1sq, r2sq, r3sq = c1["r"] ** 2, c2["r"] ** 2, c3["r"] ** 2
    d1 = np.array([c2["x"] - c1["x"], c2["y"] - c1["y"], r2sq - r1sq])
    d2 = np.array([c3["x"] - c1["x"], c3["y"] - c1["y"], r3sq - r1sq])
    nv = np.cross(d1, d2)
    h = nv / np.linalg.norm(nv)
    pd = -h[0] * (c2["x"] - vx) - (h[1] * c2["y"] - vy) - h[2] * r2sq
    rad = np.sqrt((1.0 - h[2] ** 2 - 4.0 * pd * h[2]) / (4.0 * h[2] ** 2))
    centre = np.hypot(-0.5 * h[0] / h[2], -0.5 * h[1] / h[2])
    if np.isfinite(rad):
        out.append((abs(abs(rad - centre) - CELL_DCA[l]) / ulp(max(rad, centre)), "cell dca"))
    return min(out)


def dca_error_bound(run, l, key):
    """(DCA in float64 from the raw coordinates, first-order bound on the
    float32 rounding error of the reference's DCA chain, max(R, |centre|)) of
    the cell `key` (fp_bounds.dca_with_bound)."""
    from fp_bounds import dca_with_bound
    vx, vy, _ = run["vertex"]
    c = [cluster_values(run, l + k, key[k]) for k in range(3)]
    return dca_with_bound([ci["x"] for ci in c], [ci["y"] for ci in c], vx, vy)


def edge_margin(run, l, key):
    c, n = run["cval"][l].get(key[0]), run["cval"][l + 1].get(key[1])
    if c is None or n is None:
        return (np.inf, "missing input")
    dn = float(np.sum((c[:3] - n[:3]) ** 2))
    dk = abs(c[3] - n[3])
    return min((abs(dn - NB_DN[l]) / ulp(max(np.max(np.abs(c[:3])), np.max(np.abs(n[:3])))), "edge normal"),
               (abs(dk - NB_DK[l]) / ulp(max(abs(c[3]), abs(n[3]))), "edge curvature"))


def test_contraction_disagreements_lie_at_cut_boundaries():
    path = os.path.join(itsca.LIB_DIR, "libitsca_hip_fma.so")
    if not os.path.exists(path):
        raise RuntimeError("lib/libitsca_hip_fma.so is missing: build with `make -C tracking-itsu_amd fma`")
    say(f"contraction study: {EVENTS} central events, seeds 20000..{20000 + EVENTS - 1}")
    events = itsca.generate_events(2, 20000, EVENTS)
    off = collect(itsca.hip_lib(), events)
    on = collect(itsca.bind_hip(ctypes.CDLL(path)), events)
    stats = {k: [0, 0] for k in ("tracklets", "cells", "edges", "roads")}  # [total off, differing]
    inherited = {k: 0 for k in stats}
    clusters_changed = clusters_total = 0
    worst = {}
    direct = []
    for w in range(EVENTS):
        A, B = off[w], on[w]
        for l in range(itsca.LAYERS):
            a, b = A["cl"][l], B["cl"][l]
            ia, oa = rows_by_id(a)
            ib, ob = rows_by_id(b)
            assert np.array_equal(ia, ib)
            clusters_total += len(ia)
            clusters_changed += int(np.sum((a["phibits"][oa] != b["phibits"][ob]) | (a["rbits"][oa] != b["rbits"][ob])))
        tdiff = [mdiff(A["trk"][l], B["trk"][l]) for l in range(itsca.TRACKLET_LAYERS)]
        cdiff = [mdiff(A["cells"][l], B["cells"][l]) for l in range(itsca.CELL_LAYERS)]
        ediff = [mdiff(A["edges"][l], B["edges"][l]) for l in range(itsca.CELL_LAYERS - 1)]
        rdiff = mdiff(A["roads"], B["roads"])
        for l, d in enumerate(tdiff):
            stats["tracklets"][0] += len(A["trk"][l])
            for key in d:
                stats["tracklets"][1] += 1
                m = min(tracklet_margin(A, l, key), tracklet_margin(B, l, key))
                direct.append(("tracklets", w, l, key, m))
        for l, d in enumerate(cdiff):
            stats["cells"][0] += len(A["cells"][l])
            for key in d:
                stats["cells"][1] += 1
                if key[:2] in tdiff[l] or key[1:] in tdiff[l + 1]:
                    inherited["cells"] += 1
                    continue
                m = min(cell_margin(A, l, key), cell_margin(B, l, key))
                direct.append(("cells", w, l, key, m))
        for l, d in enumerate(ediff):
            stats["edges"][0] += len(A["edges"][l])
            for key in d:
                stats["edges"][1] += 1
                if key[0] in cdiff[l] or key[1] in cdiff[l + 1]:
                    inherited["edges"] += 1
                    continue
                m = min(edge_margin(A, l, key), edge_margin(B, l, key))
                direct.append(("edges", w, l, key, m))
        stats["roads"][0] += len(A["roads"])
        for road in rdiff:
            stats["roads"][1] += 1
            cells = [road[k:k + 3] for k in range(5)]
            edges = [(cells[k], cells[k + 1]) for k in range(4)]
            if any(c in cdiff[k] for k, c in enumerate(cells)) or any(e in ediff[k] for k, e in enumerate(edges)):
                inherited["roads"] += 1
            else:
                direct.append(("roads", w, 0, road, (np.inf, "road without a differing cell or edge")))
    say("clusters with other phi or r bits: %d of %d" % (clusters_changed, clusters_total))
    for k, (tot, dif) in stats.items():
        d = [x for x in direct if x[0] == k]
        worst[k] = max((x[4][0] for x in d), default=0.0)
        by_cut = {}
        for x in d:
            by_cut[x[4][1]] = by_cut.get(x[4][1], 0) + 1
        say("%-9s %9d off, %6d differ: %6d inherited, %6d direct %s; largest margin %.3g ulp"
            % (k, tot, dif, inherited[k], len(d), by_cut, worst[k]))
    # the criterion separates: elements both runs agree on lie far from every cut
    rng = np.random.default_rng(1)
    A = off[0]
    for l in (0, 3):
        keys = [A["trk"][l][k] for k in rng.choice(len(A["trk"][l]), 200, replace=False)]
        med = float(np.median([tracklet_margin(A, l, k)[0] for k in keys]))
        say("agreeing tracklets of layer %d: median margin %.3g ulp" % (l, med))
        assert med > 1e3
    for l in (0, 2):
        keys = [A["cells"][l][k] for k in rng.choice(len(A["cells"][l]), 200, replace=False)]
        med = float(np.median([cell_margin(A, l, k)[0] for k in keys]))
        say("agreeing cells of layer %d: median margin %.3g ulp" % (l, med))
        assert med > 1e2
    dca = [x for x in direct if x[4][1] == "cell dca"]
    m = [x[4][0] for x in dca]
    say("cell dca: %d within 1 ulp, %d within 2 ulp, %d within 3 ulp of %d (target: 1 ulp)"
        % (sum(v <= 1 for v in m), sum(v <= 2 for v in m), sum(v <= 3 for v in m), len(m)))
    # every DCA case against its own derived rounding bound (both runs' inputs are the raw coordinates)
    ratios = []
    for x in dca:
        _, w, l, key, _ = x
        v, e, scale = dca_error_bound(off[w], l, key)
        ratios.append((abs(v - CELL_DCA[l]) / e, e / ulp(scale), x))
    if ratios:
        say("cell dca: |DCA - cut| / derived bound: max %.3g, median %.3g; bound in ulps of max(R, |centre|): "
            "median %.3g, max %.3g" % (max(r[0] for r in ratios), float(np.median([r[0] for r in ratios])),
                                       float(np.median([r[1] for r in ratios])), max(r[1] for r in ratios)))
    beyond = [r for r in ratios if not r[0] <= 1.0]
    for r in beyond[:20]:
        say("dca beyond its bound: event %d layer %d %s: %.3g of the bound" % (r[2][1], r[2][2], r[2][3], r[0]))
    bad = [x for x in direct if x[4][1] != "cell dca" and not x[4][0] <= 1.0]
    for x in bad[:20]:
        say("too far: %s event %d layer %d %s: %.3g ulp (%s)" % (x[0], x[1], x[2], x[3], x[4][0], x[4][1]))
    assert not bad, "%d disagreements lie further from every cut than 1 ulp" % len(bad)
    assert not beyond, "%d DCA disagreements lie further from the cut than their rounding bound" % len(beyond)
