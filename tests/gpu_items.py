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
