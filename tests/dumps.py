"""Helpers to read canonical dumps and compare them as multisets keyed by the
stable cluster id (the reference's sorted positions depend on std::sort)."""
from collections import Counter


def parse(text):
    """-> dict section -> list of rows (each a list of tokens)"""
    out, key = {}, None
    lines = text.splitlines()
    i = 0
    while i < len(lines):
        parts = lines[i].split()
        i += 1
        if not parts:
            continue
        head = parts[0]
        if head == "vertex":
            out["vertex"] = parts[1:]
            continue
        if head in ("clusters", "tracklets", "cells", "neighbours") and len(parts) == 3:
            key = (head, int(parts[1]))
            n = int(parts[2])
            out[key] = [lines[i + k].split() for k in range(n)]
            i += n
        elif head in ("table", "tlookup", "clookup") and len(parts) == 3:
            out[(head, int(parts[1]))] = lines[i].split()
            i += 1
        elif head == "roads":
            n = int(parts[1])
            out["roads"] = [lines[i + k].split() for k in range(n)]
            i += n
    return out


def keyed(text):
    """Multisets of tracklets, cells and roads with positions replaced by
    cluster ids: order-independent, comparable across sort orders."""
    d = parse(text)
    ids = {l: [row[0] for row in d[("clusters", l)]] for l in range(7)}
    res = {}
    for l in range(6):
        res[("tracklets", l)] = Counter((ids[l][int(r[0])], ids[l + 1][int(r[1])], r[2], r[3])
                                        for r in d[("tracklets", l)])
    cells = {}
    for l in range(5):
        cells[l] = [(ids[l][int(r[0])], ids[l + 1][int(r[1])], ids[l + 2][int(r[2])]) for r in d[("cells", l)]]
        res[("cells", l)] = Counter(tuple(c) + tuple(r[5:10]) for c, r in zip(cells[l], d[("cells", l)]))
    res["roads"] = Counter(tuple(cells[k][int(r[k])] for k in range(5)) + (r[6], r[7]) for r in d["roads"])
    return res


def distinct_digest(text):
    """SHA-256 of the distinct cluster-id-keyed tracklets, cells and roads
    (multiplicities dropped), stage by stage in sorted order."""
    import hashlib

    k = keyed(text)
    h = hashlib.sha256()
    for stage in sorted(k, key=str):
        h.update(repr(stage).encode())
        for item in sorted(k[stage]):
            h.update(repr(item).encode())
    return h.hexdigest()


def duplicates(text):
    """Per stage, the number of records beyond the first of each keyed record."""
    return {str(stage): sum(c.values()) - len(c) for stage, c in keyed(text).items()}
