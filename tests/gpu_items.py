"""Per-work-item output arrays of a GPU run, for comparing runs with each other."""
import hashlib

import numpy as np

import itsca


def item_arrays(t, w):
    """Counts, tracklets, cells, neighbour lists and roads of work item w."""
    n = t.counts(w)
    out = [np.array(list(n.clusters) + list(n.tracklets) + list(n.cells) + list(n.neighbours) + [n.roads])]
    out += [t.download(w, itsca.TRACKLETS, l, n.tracklets[l], 4) for l in range(itsca.TRACKLET_LAYERS)]
    out += [t.download(w, itsca.CELLS, l, n.cells[l], 10) for l in range(itsca.CELL_LAYERS)]
    for l in range(itsca.CELL_LAYERS - 1):
        out.append(t.download(w, itsca.NEIGHBOUR_OFFSETS, l, n.cells[l + 1] + 1, 1))
        out.append(t.download(w, itsca.NEIGHBOURS, l, n.neighbours[l], 1))
    out.append(t.roads(w))
    return out


def item_digest(t, w):
    h = hashlib.sha256()
    for a in item_arrays(t, w):
        h.update(np.ascontiguousarray(a, dtype=np.int32).tobytes())
        h.update(b"|")
    return h.hexdigest()


def gate_mask(phi_i, phi_j, gate):
    """The phi gate of the learned form's tracklet search on tracklets with
    cluster phis phi_i, phi_j (float32): the reference's phi cut comparisons
    (Tracker.cxx:96-98) with the gate as their bound."""
    f32 = np.float32
    d = np.abs(np.asarray(phi_i, f32) - np.asarray(phi_j, f32))
    two_pi = f32(2.0) * f32(3.14159265359)
    return (d < f32(gate)) | (np.abs(d - two_pi) < f32(gate))


def gated_tracklets(t, w, plain_trk, gates):
    """The plain tracklets of item w (per layer, (n, >=2) arrays of first and
    second cluster positions) that pass the phi gates, in the plain order;
    the cluster phis are read from the tracker's sorted clusters."""
    n = t.counts(w)
    phi = [t.download(w, itsca.CLUSTERS, l, n.clusters[l], 8)[:, 1].copy().view(np.float32)
           for l in range(itsca.LAYERS)]
    out = []
    for l in range(itsca.TRACKLET_LAYERS):
        p = plain_trk[l]
        if len(p) == 0:
            out.append(p[:, :2])
            continue
        m = gate_mask(phi[l][p[:, 0]], phi[l + 1][p[:, 1]], gates[l])
        out.append(p[m, :2])
    return out
