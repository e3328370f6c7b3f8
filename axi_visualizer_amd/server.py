"""Render server: resident scenes, HBM-budget admission, continuous batching.

Scenes are made resident once (``SceneContext``: geometry, textures, shader
binding, lights and their shadow maps stay in HBM).  Requests -- a scene, a
camera, a resolution, shadows on or off -- are queued as they arrive.  The
serving loop keeps several batch slots (each its own stream and buffers, so
batches overlap on the device); whenever a slot is free it forms a batch
from the queue, oldest first, admitting a request only while the device
bytes of the resident scenes, the batches in flight and the new batch stay
within the HBM budget, and launches it without waiting.  Finished slots are
polled, their requests completed and their memory released, so requests
that arrive while others render are taken as slots and memory free up.

The queue is bounded: a request that arrives to a full queue is dropped and
counted, as the reference engine counts the frames its full output queue
refuses (Compute.py:590-593).  A request can be cancelled while queued (it
is removed) or in flight (its result is discarded when the batch ends).  A
batch whose bin storage overflowed (AXR_ECAP: the storage has grown) goes
back to the head of the queue and renders again; any other error of a
batch propagates.

A request's pixels do not depend on what else is in its batch: every job is
bit-identical to the same view rendered alone.
"""
from __future__ import annotations

import collections
import dataclasses
import time
from typing import Callable, Dict, List, Optional, Sequence

import numpy as np

from ._native import AXR_ECAP, AxrError
from .batch import Batch, Job, SceneContext, settle
from .scenes import Camera, make_camera

# Default bucket capacities of a 32-px bin (entries) used for admission estimates.
_CAP_SMALL, _CAP_LARGE = 4096, 1024

# Resolutions of the serving workload's requests: 480p, 720p, 1080p.
SERVING_SIZES = [(848, 480), (1280, 720), (1920, 1080)]

# Request states
QUEUED, RUNNING, DONE, DROPPED, REJECTED, CANCELLED = (
    "queued", "running", "done", "dropped", "rejected", "cancelled")


@dataclasses.dataclass
class Request:
    rid: int
    scene: int
    camera: Camera
    width: int
    height: int
    shadows: bool = False
    pcf_taps: int = 9
    arrival: float = 0.0       # seconds on the server clock
    start: float = 0.0         # launched
    done: float = 0.0          # completion observed
    batch_size: int = 0
    rgba: Optional[np.ndarray] = None
    state: str = QUEUED
    attempts: int = 0          # launches (more than one after a requeue)

    @property
    def latency_ms(self) -> float:
        return (self.done - self.arrival) * 1e3


def request_bytes(ntri: int, w: int, h: int, bin_size: int = 32) -> int:
    """Device bytes a request holds while in flight: its setup records
    (36 B per scene triangle), outputs (12 B per pixel) and bin buckets."""
    nb = ((w + bin_size - 1) // bin_size) * ((h + bin_size - 1) // bin_size)
    return ntri * 36 + w * h * 12 + nb * (_CAP_SMALL + _CAP_LARGE) * 4 + nb * 20


class _Slot:
    def __init__(self, batch):
        self.batch = batch
        self.reqs: List[Request] = []
        self.bytes = 0

    @property
    def busy(self) -> bool:
        return bool(self.reqs)


def _default_batch(bin_size: int, graph: bool):
    return Batch(bin_size=bin_size, graph=graph)


class RenderServer:
    """Continuous-batching render loop over resident scenes.

    context_factory(scene) makes a scene resident (default SceneContext);
    batch_factory(bin_size, graph) makes a batch slot (default Batch);
    clock() gives seconds (default time.perf_counter)."""

    def __init__(self, scenes: Sequence, hbm_budget: float = 200e9, slots: int = 3,
                 max_batch: int = 64, window_ms: float = 2.0, bin_size: int = 32,
                 readback: bool = False, graph: bool = False, max_queue: int = 8192,
                 context_factory: Optional[Callable] = None,
                 batch_factory: Optional[Callable] = None,
                 clock: Callable[[], float] = time.perf_counter):
        make_ctx = context_factory or SceneContext
        make_batch = batch_factory or _default_batch
        self.contexts = [make_ctx(s) for s in scenes]
        self.ntri = [c.scene.ntri for c in self.contexts]
        self.resident = sum(c.device_bytes() for c in self.contexts)
        self.budget = float(hbm_budget)
        if self.resident > self.budget:
            raise ValueError(f"resident scenes need {self.resident} B, budget {self.budget:.0f}")
        self.slots = [_Slot(make_batch(bin_size, graph)) for _ in range(slots)]
        self.max_batch = max_batch
        self.max_queue = max_queue
        self.window = window_ms * 1e-3
        self.bin_size = bin_size
        self.readback = readback
        self.clock = clock
        self.queue: collections.deque = collections.deque()
        self.completed: List[Request] = []
        self.rejected: List[Request] = []
        self.dropped: List[Request] = []
        self.cancelled: List[Request] = []
        self.requeued = 0          # batches rendered again after a capacity loss
        self.max_in_flight = 0     # bytes, for the admission check
        self.max_busy = 0          # batches on the device at once
        self.t0 = self.clock()

    def reserve(self, width: int, height: int, shadows: bool = True):
        """Grows every slot's device pools to a full batch of the largest
        scene at width x height before serving.  A pool that grows while
        batches are in flight frees its old buffers, and a free waits for
        the whole device, which would stall every slot."""
        big = max(range(len(self.contexts)), key=lambda i: self.ntri[i])
        ctx = self.contexts[big]
        sc = ctx.scene
        cam = make_camera(width, height, sc.view.pos, 0.0, 0.0)
        cam.pos, cam.vm = sc.view.pos, sc.view.vm
        jobs = [Job(ctx, cam, width, height, shadows=shadows) for _ in range(self.max_batch)]
        for s in self.slots:
            settle(s.batch, jobs)

    # -- clock
    def now(self) -> float:
        return self.clock() - self.t0

    def in_flight_bytes(self) -> int:
        return sum(s.bytes for s in self.slots)

    def _need(self, r: Request) -> int:
        return request_bytes(self.ntri[r.scene], r.width, r.height, self.bin_size)

    def submit(self, req: Request) -> bool:
        """Queues a request; False if it is rejected (it can never fit the
        budget) or dropped (the queue is full)."""
        if self.resident + self._need(req) > self.budget:
            req.state = REJECTED
            self.rejected.append(req)
            return False
        if len(self.queue) >= self.max_queue:
            req.state = DROPPED
            self.dropped.append(req)
            return False
        req.state = QUEUED
        self.queue.append(req)
        return True

    def cancel(self, rid: int) -> bool:
        """Cancels request rid: removed from the queue, or, if its batch is
        on the device, its result is discarded when the batch ends.  False
        if no queued or running request has that id."""
        for r in self.queue:
            if r.rid == rid:
                self.queue.remove(r)
                r.state = CANCELLED
                self.cancelled.append(r)
                return True
        for s in self.slots:
            for r in s.reqs:
                if r.rid == rid and r.state == RUNNING:
                    r.state = CANCELLED
                    return True
        return False

    # -- loop pieces
    def _poll(self):
        for s in self.slots:
            if s.busy and s.batch.done():
                self._finish(s)

    def _finish(self, s: _Slot):
        reqs, s.reqs, s.bytes = s.reqs, [], 0
        try:
            s.batch.wait()
        except AxrError as e:
            if e.code != AXR_ECAP:
                raise
            # bin entries were lost and the storage has grown: the batch
            # goes back to the head of the queue, oldest first
            self.requeued += 1
            for r in reversed(reqs):
                if r.state == CANCELLED:
                    self.cancelled.append(r)
                    continue
                r.state = QUEUED
                self.queue.appendleft(r)
            return
        t = self.now()
        for k, r in enumerate(reqs):
            if r.state == CANCELLED:
                self.cancelled.append(r)
                continue
            if self.readback:
                r.rgba = s.batch.rgba8(k)
            r.done = t
            r.state = DONE
            self.completed.append(r)

    def _launch(self, s: _Slot) -> bool:
        """Forms a batch from the queue head under the budget; False if none."""
        if not self.queue:
            return False
        oldest = self.queue[0].arrival
        if len(self.queue) < self.max_batch and self.now() - oldest < self.window and \
                any(x.busy for x in self.slots):
            return False  # let the batch fill while the device is busy
        free = self.budget - self.resident - self.in_flight_bytes()
        take, used = [], 0
        while self.queue and len(take) < self.max_batch:
            need = self._need(self.queue[0])
            if used + need > free:
                break
            take.append(self.queue.popleft())
            used += need
        if not take:
            return False
        jobs = [Job(self.contexts[r.scene], r.camera, r.width, r.height, shadows=r.shadows,
                    pcf_taps=r.pcf_taps) for r in take]
        t = self.now()
        for r in take:
            r.start = t
            r.batch_size = len(take)
            r.state = RUNNING
            r.attempts += 1
        s.batch.render(jobs)
        s.reqs, s.bytes = take, used
        self.max_in_flight = max(self.max_in_flight, self.in_flight_bytes())
        self.max_busy = max(self.max_busy, sum(1 for x in self.slots if x.busy))
        return True

    def step(self):
        self._poll()
        for s in self.slots:
            if not s.busy:
                self._launch(s)

    def drain(self):
        while self.queue or any(s.busy for s in self.slots):
            self.step()

    def run(self, requests: Sequence[Request]):
        """Serves requests whose .arrival are seconds from now; returns when
        every request is complete (or dropped, rejected or cancelled)."""
        pending = sorted(requests, key=lambda r: r.arrival)
        i = 0
        self.t0 = self.clock()
        while i < len(pending) or self.queue or any(s.busy for s in self.slots):
            t = self.now()
            while i < len(pending) and pending[i].arrival <= t:
                self.submit(pending[i])
                i += 1
            self.step()
        return self.report()

    def report(self) -> Dict:
        lat = np.array([r.latency_ms for r in self.completed]) if self.completed else np.zeros(1)
        span = max((r.done for r in self.completed), default=0.0) - \
            min((r.arrival for r in self.completed), default=0.0)
        sizes = [r.batch_size for r in self.completed]
        return {"completed": len(self.completed), "rejected": len(self.rejected),
                "dropped": len(self.dropped), "cancelled": len(self.cancelled),
                "requeued_batches": self.requeued,
                "p50_ms": float(np.percentile(lat, 50)), "p99_ms": float(np.percentile(lat, 99)),
                "mean_batch": float(np.mean(sizes)) if sizes else 0.0,
                "fps": len(self.completed) / span if span > 0 else 0.0,
                "resident_bytes": int(self.resident),
                "max_in_flight_bytes": int(self.max_in_flight),
                "max_concurrent_batches": int(self.max_busy)}

    def close(self):
        for s in self.slots:
            s.batch.close()
        for c in self.contexts:
            c.close()


def poisson_requests(n_scenes: int, rate: float, duration: float, sizes, seed: int = 1,
                     shadow_frac: float = 0.5, camera_for=None) -> List[Request]:
    """Seeded Poisson arrivals at `rate` per second for `duration` seconds;
    each picks a scene, a resolution and a camera (camera_for(scene, w, h,
    rng)), with shadows on a `shadow_frac` share of requests."""
    rng = np.random.default_rng(seed)
    t, out = 0.0, []
    while True:
        t += rng.exponential(1.0 / rate)
        if t >= duration:
            break
        s = int(rng.integers(n_scenes))
        w, h = sizes[int(rng.integers(len(sizes)))]
        out.append(Request(rid=len(out), scene=s, camera=camera_for(s, w, h, rng), width=w,
                           height=h, shadows=bool(rng.random() < shadow_frac), arrival=t))
    return out
