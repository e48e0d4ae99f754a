"""The measured workloads at their full sizes, asserted.

- 256 events of log-uniform multiplicity in [50, 8000] tracks in one launch:
  every event's outputs equal the event run on its own, and a sample that
  includes the largest and the smallest event equals the oracle.
- 256 held-out central events: the learned filter's tracks are a subset of
  the plain tracks (with multiplicity) and recover at least 99.5 % of them.
- 64 pile-up events (4 vertices each) in one launch: every work item equals
  its event run on its own; one event's four vertices equal the oracle; the
  learned tracks are a subset of the plain ones.
"""
import sys

import numpy as np
import pytest

import itsca
import pyoracle
from gpu_items import item_digest
from itsca import filter_training as ft

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _batch_vs_single(t, evs):
    t.set_events(evs)
    t.run()
    batch = [item_digest(t, w) for w in range(t.num_items())]
    w = 0
    for e, ev in enumerate(evs):
        t.set_events([ev])
        t.run()
        for v in range(len(ev.vertices)):
            assert item_digest(t, v) == batch[w], (e, v)
            w += 1
    assert w == len(batch)


def test_mixed_256_event_batch():
    say("gpu full size: 256 mixed-multiplicity events")
    evs = itsca.generate_events(3, 3000, 256)
    sizes = [len(ev.x) for ev in evs]
    t = itsca.GpuTracker(device=0)
    _batch_vs_single(t, evs)
    t.set_events(evs)
    t.run()
    order = np.argsort(sizes)
    for i in sorted({int(order[-1]), int(order[0]), 17, 101, 230}):
        assert itsca.dump_item(t, i, evs[i].vertices[0]) == pyoracle.OracleVertex(evs[i], 0, stable=True).dump(), i
    t.close()


def test_central_256_event_batch_learned_within_plain():
    say("gpu full size: 256 held-out central events, learned vs plain")
    evs = itsca.generate_events(2, 20000, 256)
    t = itsca.GpuTracker(device=0)
    threshold = ft.prepare(t)  # the shipped weights and phi gates
    q = ft.evaluate(t, evs, threshold)
    t.close()
    say("  recall %.6f of %d plain tracks, %d outside" % (q["recall"], q["plain_tracks"], q["extra"]))
    assert q["extra"] == 0
    assert q["recall"] >= 0.995
    assert q["plain_tracks"] > 2_000_000


def test_pileup_64_event_batch():
    say("gpu full size: 64 pile-up events")
    evs = itsca.generate_events(5, 50000, 64)
    assert all(len(ev.vertices) == 4 for ev in evs)
    t = itsca.GpuTracker(device=0)
    _batch_vs_single(t, evs)
    big = int(np.argmax([len(ev.x) for ev in evs]))
    t.set_events([evs[big]])
    t.run()
    for v in range(4):
        assert itsca.dump_item(t, v, evs[big].vertices[v]) == pyoracle.OracleVertex(evs[big], v, stable=True).dump(), v
    threshold = ft.prepare(t)
    q = ft.evaluate(t, evs, threshold)
    t.close()
    say("  recall %.6f of %d plain tracks, %d outside" % (q["recall"], q["plain_tracks"], q["extra"]))
    assert q["extra"] == 0
    assert q["recall"] >= 0.99
