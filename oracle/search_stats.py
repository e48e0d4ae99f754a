#!/usr/bin/env python3
"""Candidate tests of the tracklet window search, counted on the CPU from the
oracle's sorted clusters and index tables (TEST INFRASTRUCTURE: a measurement
of the search, not a check).

For every cluster of layer l the reference tests every cluster of its window
rows in the z bins [x, z] (TrackingUtils.cxx:30-48, Tracker.cxx:73-100); the
gfx950 search tests, per window row, only the z bins that can hold a passing
candidate (the z-range skip).  This counts, per layer: the reference's tests,
the tests left after the z-range skip, the candidates among those that pass
the phi cut alone, and the tracklets.  The share of tested candidates that
fail the phi cut bounds what phi sub-rows of the search index could save.

usage: python oracle/search_stats.py [preset seed]   (default: one central event, preset 2 seed 20000)
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path[:0] = [HERE, os.path.join(HERE, "..", "tracking-itsu_amd")]
import itsca  # noqa: E402
import pyoracle  # noqa: E402

F = np.float32
PI = F(3.14159265359)
TWO_PI = F(2.0) * PI
HALF_Z = [F(16.333)] * 3 + [F(42.140)] * 2 + [F(73.745)] * 2
LAYER_R = [F(v) for v in (2.33959, 3.14076, 3.91924, 19.6213, 24.5597, 34.388, 39.3329)]
INV_Z = [F(0.5 * 20 / np.float64(h)) for h in HALF_Z]
INV_PHI = F(20) / TWO_PI
DZ = [F(v) for v in (0.1, 0.1, 0.3, 0.3, 0.3, 0.3)]


def zbin(l, z):
    return np.trunc((z + HALF_Z[l]) * INV_Z[l]).astype(np.int64)


def norm_phi(p):
    return np.where(p < 0, p + TWO_PI, np.where(p > TWO_PI, p - TWO_PI, p)).astype(F)


def main(preset=2, seed=20000):
    ev = itsca.generate_events(preset, seed, 1)[0]
    o = pyoracle.OracleVertex(ev, 0, stable=True)
    vz = F(ev.vertices[0][2])
    cl = []
    for l in range(7):
        c = o.records(0, l)
        cl.append({"phi": c[:, 1].view(F), "r": c[:, 2].view(F), "z": c[:, 6].view(F), "bin": c[:, 3]})
    print("layer  reference_tests  after_z_skip  pass_phi_cut  tracklets  phi_fail_share")
    for l in range(6):
        a, b = cl[l], cl[l + 1]
        table = np.searchsorted(b["bin"], np.arange(401), side="left")
        table[0] = 0
        nb = len(b["r"])
        rmin, rmax = b["r"].min(), b["r"].max()
        tanl = ((a["z"] - vz) / a["r"]).astype(F)
        dirz = (tanl * (LAYER_R[l + 1] - a["r"]) + a["z"]).astype(F)
        zmin, zmax = (dirz - F(1.0)).astype(F), (dirz + F(1.0)).astype(F)
        pmin, pmax = norm_phi(a["phi"] - F(0.3)), norm_phi(a["phi"] + F(0.3))
        hz = HALF_Z[l + 1]
        empty = (zmax < -hz) | (zmin > hz) | (zmin > zmax)
        x = np.maximum(0, zbin(l + 1, zmin))
        z = np.minimum(19, zbin(l + 1, zmax))
        y = np.trunc(pmin * INV_PHI).astype(np.int64)
        w = np.trunc(pmax * INV_PHI).astype(np.int64)
        za = (tanl * (rmin - a["r"]) + a["z"]).astype(F)
        zb = (tanl * (rmax - a["r"]) + a["z"]).astype(F)
        zlo, zhi = np.minimum(za, zb) - DZ[l] - F(0.01), np.maximum(za, zb) + DZ[l] + F(0.01)
        zf, zl = np.maximum(x, zbin(l + 1, zlo)), np.minimum(z, zbin(l + 1, zhi))
        ref = skip = phi_ok = 0
        for i in np.nonzero(~empty)[0]:
            rows = (w[i] - y[i] + 1) % 20 or 20
            row = y[i]
            for _ in range(rows):
                base = row * 20
                j0, j1 = table[min(base + x[i], 400)], table[min(base + z[i] + 1, 400)]
                ref += min(j1 + 1, nb) - j0
                if zf[i] <= zl[i]:
                    k0, k1 = table[min(base + zf[i], 400)], min(table[min(base + zl[i] + 1, 400)], nb)
                    skip += max(0, k1 - k0)
                    dphi = np.abs(a["phi"][i] - b["phi"][k0:k1])
                    phi_ok += int(np.sum((dphi < F(0.3)) | (np.abs(dphi - TWO_PI) < F(0.3))))
                row = (row + 1) % 20
        ntrk = o.count(1, l)
        print("%5d  %15d  %12d  %12d  %9d  %14.3f" % (l, ref, skip, phi_ok, ntrk, 1 - phi_ok / max(skip, 1)))


if __name__ == "__main__":
    args = [int(v) for v in sys.argv[1:3]]
    main(*args)
