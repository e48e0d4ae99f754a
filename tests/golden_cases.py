"""Index of the committed golden fixtures (written by tests/golden/make_golden.py
from the reference CPU tracker)."""
import glob
import json
import os

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def small_cases():
    """[(case name, events path, number of vertices)]"""
    out = []
    for ev in sorted(glob.glob(os.path.join(GOLDEN, "*.events.txt"))):
        name = os.path.basename(ev)[: -len(".events.txt")]
        nv = len(glob.glob(os.path.join(GOLDEN, name + ".ref.v*.txt")))
        out.append((name, ev, nv))
    return out


def fixture(name, kind, vertex):
    with open(os.path.join(GOLDEN, "%s.%s.v%d.txt" % (name, kind, vertex))) as f:
        return f.read()


def central():
    with open(os.path.join(GOLDEN, "central_s2.json")) as f:
        return json.load(f)


def dense_spot_event(n_dense=80):
    """A small event (100 tracks, seed 1) plus a dense spot: one straight track
    from its vertex at phi = 1, tan(lambda) = 0.1, with n_dense clusters close
    together on layers 1 and 3 and one on the other layers.  The cluster of
    layer 0 (and of layer 2) then has n_dense tracklets, more than the 64
    slots that the tracklet count keeps, so the overflow paths run."""
    import numpy as np

    import itsca

    base = itsca.generate_events(1, 1, 1)[0]
    vx, vy, vz = (float(a) for a in base.vertices[0])
    radii = [2.34, 3.15, 3.93, 19.6, 24.55, 34.39, 39.34]
    cols = {k: [] for k in ("x", "y", "z", "mc")}
    offs = [0]
    for layer in range(itsca.LAYERS):
        sl = base.layer(layer)
        cols["x"] += list(base.x[sl])
        cols["y"] += list(base.y[sl])
        cols["z"] += list(base.z[sl])
        cols["mc"] += list(base.mc[sl])
        n = n_dense if layer in (1, 3) else 1
        for k in range(n):
            phi = 1.0 + (k - n // 2) * 2e-4
            r = radii[layer]
            cols["x"].append(vx + r * np.cos(phi))
            cols["y"].append(vy + r * np.sin(phi))
            cols["z"].append(vz + 0.1 * r + (k - n // 2) * 1e-4)
            cols["mc"].append(900000 + k)
        offs.append(len(cols["x"]))
    n_all = offs[-1]
    return itsca.Event(np.array(cols["x"], dtype=np.float32), np.array(cols["y"], dtype=np.float32),
                       np.array(cols["z"], dtype=np.float32), np.arange(n_all, dtype=np.int32),
                       np.array(cols["mc"], dtype=np.int32), np.array(offs, dtype=np.int32),
                       base.vertices.copy())
