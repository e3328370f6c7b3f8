"""The render server's loop on the CPU, with stub scenes and batches and a
fake clock: HBM-budget admission and rejection, the batching window, the
bounded queue's drop count, cancellation, and the handling of a batch whose
bin storage overflowed (requeued) or that failed otherwise (raised)."""
import numpy as np
import pytest

from axi_visualizer_amd import scenes
from axi_visualizer_amd._native import AXR_ECAP, AXR_EHIP, AxrError
from axi_visualizer_amd.server import (CANCELLED, DONE, DROPPED, REJECTED, RenderServer,
                                       Request, request_bytes)

W, H = 320, 240


class Clock:
    """Seconds that advance by `tick` at every reading."""

    def __init__(self, tick=1e-5):
        self.t, self.tick = 0.0, tick

    def __call__(self):
        self.t += self.tick
        return self.t


class StubScene:
    def __init__(self, ntri):
        self.ntri = ntri
        self.view = scenes.make_camera(W, H, [0, 0, 0], 0.0, 0.0)


class StubCtx:
    def __init__(self, scene):
        self.scene = scene

    def device_bytes(self):
        return 10_000_000

    def close(self):
        pass


class StubBatch:
    """Renders in `latency` seconds of the server clock; `errors` are raised
    by the next waits, in order."""
    log = []

    def __init__(self, clock, latency, errors):
        self.clock, self.latency, self.errors = clock, latency, errors
        self.jobs, self.ready = [], 0.0

    def render(self, jobs):
        self.jobs = list(jobs)
        self.ready = self.clock.t + self.latency
        StubBatch.log.append(len(jobs))

    def done(self):
        return self.clock() >= self.ready

    def wait(self):
        if self.errors:
            raise self.errors.pop(0)
        return 0.0

    def rgba8(self, k):
        j = self.jobs[k]
        return np.full((j.height, j.width), k, np.uint32)

    def close(self):
        pass


def server(n_scenes=2, ntri=1000, budget=1e12, slots=2, max_batch=8, window_ms=2.0,
           latency=1e-3, errors=None, tick=1e-5, **kw):
    clock = Clock(tick)
    errs = list(errors or [])
    srv = RenderServer([StubScene(ntri) for _ in range(n_scenes)], hbm_budget=budget,
                       slots=slots, max_batch=max_batch, window_ms=window_ms,
                       context_factory=StubCtx,
                       batch_factory=lambda b, g: StubBatch(clock, latency, errs),
                       clock=clock, **kw)
    return srv, clock


def reqs(n, n_scenes=2, dt=1e-4, w=W, h=H):
    cam = scenes.make_camera(w, h, [0, 0, 0], 0.0, 0.0)
    return [Request(rid=i, scene=i % n_scenes, camera=cam, width=w, height=h, arrival=i * dt)
            for i in range(n)]


def test_every_request_completes_within_the_budget():
    need = request_bytes(1000, W, H)
    srv, _ = server(budget=2 * 10_000_000 + 2.5 * need)
    rep = srv.run(reqs(40))
    assert rep["completed"] == 40 and rep["rejected"] == rep["dropped"] == 0
    assert 0 < srv.max_in_flight <= 2.5 * need        # two requests fit at once
    assert all(r.state == DONE and r.batch_size <= 2 for r in srv.completed)


def test_requests_that_can_never_fit_are_rejected():
    srv, _ = server(budget=2 * 10_000_000 + request_bytes(1000, W, H))
    big = reqs(1, w=1920, h=1080)[0]
    assert not srv.submit(big)
    assert big.state == REJECTED and srv.rejected == [big]
    rep = srv.run(reqs(3))
    assert rep["completed"] == 3 and rep["rejected"] == 1


def test_young_queue_waits_for_the_window_while_a_batch_runs():
    srv, clock = server(slots=2, max_batch=8, window_ms=2.0, latency=1.0, tick=0.0)
    first = reqs(1)[0]
    srv.submit(first)
    srv.step()                               # idle device: launches at once
    assert srv.slots[0].busy and first.batch_size == 1
    for r in reqs(3)[1:]:
        r.arrival = clock.t
        srv.submit(r)
    clock.t += 1.5e-3
    srv.step()
    assert not srv.slots[1].busy and len(srv.queue) == 2   # 1.5 ms old: still waiting
    clock.t += 0.6e-3
    srv.step()
    assert srv.slots[1].busy and len(srv.queue) == 0       # 2.1 ms old: launched


def test_a_full_queue_drops_and_counts():
    srv, _ = server(max_queue=4)
    rs = reqs(6)
    ok = [srv.submit(r) for r in rs]
    assert ok == [True] * 4 + [False] * 2
    assert [r.state for r in rs[4:]] == [DROPPED, DROPPED]
    rep = srv.run([])
    assert rep["completed"] == 4 and rep["dropped"] == 2


def test_cancel_a_queued_and_a_running_request():
    srv, clock = server(slots=1, max_batch=2, latency=1.0, tick=0.0)
    rs = reqs(4)
    for r in rs:
        srv.submit(r)
    srv.step()                               # rs[0], rs[1] on the device
    assert srv.cancel(rs[3].rid)             # queued: removed
    assert srv.cancel(rs[1].rid)             # running: result discarded
    assert not srv.cancel(99)
    clock.tick = 1e-2
    srv.drain()
    assert [r.rid for r in srv.completed] == [0, 2]
    assert sorted(r.rid for r in srv.cancelled) == [1, 3]
    assert rs[1].state == CANCELLED and rs[1].rgba is None


def test_a_batch_that_lost_bin_entries_is_requeued_and_rendered_again():
    lost = AxrError("axr_batch_wait failed (-4): spill list", AXR_ECAP)
    srv, _ = server(slots=1, max_batch=8, errors=[lost], readback=True)
    StubBatch.log = []
    rep = srv.run(reqs(5, dt=0.0))
    assert rep["completed"] == 5 and rep["requeued_batches"] == 1
    assert StubBatch.log[:2] == [5, 5]       # the same five requests again
    assert all(r.attempts == 2 and r.rgba is not None for r in srv.completed)
    assert [r.rid for r in srv.completed] == [0, 1, 2, 3, 4]  # order kept


def test_other_batch_errors_propagate():
    srv, _ = server(slots=1, errors=[AxrError("axr_batch_wait failed (-2): hip", AXR_EHIP)])
    with pytest.raises(AxrError) as e:
        srv.run(reqs(3, dt=0.0))
    assert e.value.code == AXR_EHIP
