"""itsca_run_noalloc, the entry point that allocates nothing, against
itsca_run on the same events through the C ABI: two calls in a row, the
second with a smaller event; the learned filter; the error paths."""
import numpy as np
import pytest

import itsca
from itsca import filter_training as ft

pytestmark = pytest.mark.gpu


def _device(torch, ev):
    arrs = [torch.from_numpy(np.ascontiguousarray(a)).to("cuda:0")
            for a in (ev.x, ev.y, ev.z, ev.cluster_id.astype(np.int32), ev.mc.astype(np.int32))]
    torch.cuda.synchronize()
    return arrs


def _noalloc(torch, t, ev, cap):
    arrs = _device(torch, ev)
    rows = torch.full((max(cap, 1), 8), -7, dtype=torch.int32, device="cuda:0")
    torch.cuda.synchronize()
    n = t.run_noalloc([a.data_ptr() for a in arrs], ev.layer_offsets, ev.vertices[0], rows.data_ptr(), cap)
    out = rows.cpu().numpy()
    assert (out[n:] == -7).all()  # nothing past the roads
    return out[:n]


def test_noalloc_equals_itsca_run():
    torch = pytest.importorskip("torch")
    print("gpu no-alloc: central event then a small event", flush=True)
    big = itsca.generate_events(2, 20000, 1)[0]
    small = itsca.generate_events(1, 3, 1)[0]
    t = itsca.GpuTracker(device=0)
    for ev in (big, small):  # the second call has a smaller input
        t.set_events([ev])
        t.run()
        want = t.roads(0)
        got = _noalloc(torch, t, ev, len(want) + 100)
        assert len(want) > 50
        assert np.array_equal(got, want)
    # the learned filter, same weights and threshold on both entry points
    threshold = ft.prepare(t)  # the shipped weights and phi gates
    t.set_filter(True, threshold)
    for ev in (big, small):
        t.set_events([ev])
        t.run()
        want = t.roads(0)
        assert np.array_equal(_noalloc(torch, t, ev, len(want) + 100), want)
    t.set_filter(False)
    # too few road rows: a capacity error that reports the number of roads
    t.set_events([small])
    t.run()
    n = t.counts(0).roads
    with pytest.raises(itsca.ItscaError, match="roads %d" % n):
        _noalloc(torch, t, small, n - 1)
    t.close()


def test_noalloc_refuses_events_above_its_limit():
    torch = pytest.importorskip("torch")
    print("gpu no-alloc: size limit", flush=True)
    ev = itsca.generate_events(1, 3, 1)[0]
    arrs = _device(torch, ev)
    t = itsca.GpuTracker(device=0)
    offs = [0, 20000, 40000, 60000, 80000, 100000, 120000, 140000]  # more clusters than allowed; never read
    with pytest.raises(itsca.ItscaError, match="more than the 131072"):
        t.run_noalloc([a.data_ptr() for a in arrs], offs, ev.vertices[0], 0, 0)
    t.close()
