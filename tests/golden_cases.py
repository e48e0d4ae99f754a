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
