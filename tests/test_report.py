"""The road report (correct / duplicate / fake roads against MC labels,
IOUtils.cxx:80-170): the C++ API's classification, through its C entry
point, on the oracle's roads, against the reference's own report files
(tests/golden/report_*, written by the reference built from its sources)."""
import os

import pytest

import itsca
import pyoracle

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
FILES = ("CorrectRoads.txt", "DuplicateRoads.txt", "FakeRoads.txt")


def per_event_counts(case, kind):
    """Rows per event of each report file: Correct and Fake have a -1 line per
    event; Duplicate has none, so only its total is known."""
    out = {}
    for f in ("CorrectRoads.txt", "FakeRoads.txt"):
        blocks = []
        for line in open(os.path.join(GOLDEN, "%s.%s.%s" % (case, kind, f))).read().splitlines():
            if line.strip() == "-1":
                blocks.append(0)
            else:
                blocks[-1] += 1
        out[f] = blocks
    out["DuplicateRoads.txt"] = len(open(os.path.join(GOLDEN, "%s.%s.DuplicateRoads.txt" % (case, kind))).readlines())
    return out


@pytest.mark.parametrize("case", ["report_s21", "report_pileup3"])
@pytest.mark.parametrize("stable", [True, False], ids=["stable", "as-is"])
def test_report_counts_match_reference(case, stable):
    ev_path = os.path.join(GOLDEN, case + ".ev.txt")
    events = itsca.load_events(ev_path)
    roads = [[pyoracle.OracleVertex(ev, v, stable=stable).records(3) for v in range(len(ev.vertices))]
             for ev in events]
    got = itsca.roads_report(ev_path, os.path.join(GOLDEN, case + ".labels.txt"), roads)
    want = per_event_counts(case, "stable" if stable else "ref")
    assert [c[0] for c in got] == want["CorrectRoads.txt"]
    assert [c[2] for c in got] == want["FakeRoads.txt"]
    assert sum(c[1] for c in got) == want["DuplicateRoads.txt"]
    assert sum(c[0] for c in got) > 100
