"""Batched rendering: many views of resident scenes per launch sequence.

A ``SceneContext`` keeps a Scene resident on the device (a CLDraw holding its
groups, textures, shader binding, lights and rendered shadow maps).  A
``Batch`` renders a list of ``Job`` s -- one camera at one resolution on one
resident scene each -- with one setup launch, one bin-order launch and one
persistent raster launch for all of them (libaxr's ``axr_batch_*``).  Each
job's pixels are bit-identical to the same view drawn alone with
``CLDraw.drawAll``.  ``autotune`` picks the bin size (8/16/32 px) and the
bucket capacities by timing the actual batch.
"""
from __future__ import annotations

import ctypes
import dataclasses
import time
from typing import List, Optional, Sequence

import numpy as np

from ._native import (AXR_ECAP, OUT_HDR, OUT_IDS, OUT_RGBA8, AxrError, AxrJob, check, lib,
                      ptr)
from .ops import CLDraw
from .render import load_scene
from .scenes import Camera, Scene, make_camera


class SceneContext:
    """A Scene made resident for batches (never drawn through its own frame)."""

    def __init__(self, scene: Scene):
        self.scene = sc = scene
        self.draw = d = CLDraw(1, len(sc.groups), 16, 16, 1)
        load_scene(d, sc)
        cast = [g.cast for g in sc.groups]
        shaders = [g.flags for g in sc.groups]
        for i, c in enumerate(sc.cascades):
            d.clearShadowMap(i)
            d.shadowMap(i, cast, shaders, c.bias)
        d.setShaders(shaders, [g.skip for g in sc.groups])
        d.sync()

    @property
    def handle(self):
        return self.draw._ctx

    def device_bytes(self) -> int:
        return self.draw.deviceBytes()

    def close(self):
        self.draw.close()


@dataclasses.dataclass
class Job:
    scene: SceneContext
    camera: Camera
    width: int
    height: int
    shadows: bool = False
    pcf_taps: int = 5
    exposure: float = 1.0

    def to_c(self) -> AxrJob:
        j = AxrJob()
        j.scene = self.scene.handle
        c = self.camera
        j.pos[:] = [float(v) for v in np.asarray(c.pos, np.float32).reshape(3)]
        j.vm[:] = [float(v) for v in np.asarray(c.vm, np.float32).reshape(9)]
        j.scale, j.cax, j.cay = float(c.scale), float(c.cax), float(c.cay)
        j.width, j.height = int(self.width), int(self.height)
        j.shadows = 1 if self.shadows else 0
        j.pcf_taps = int(self.pcf_taps)
        j.exposure = float(self.exposure)
        return j


def job_camera(width: int, height: int, pos, target, fovx: float = 80.0) -> Camera:
    """Camera at pos looking at target, scale/cone for this resolution."""
    from .scenes import look_at

    a, b = look_at(pos, target)
    return make_camera(width, height, pos, a, b, fovx)


def job_view_scene(job: Job) -> Scene:
    """A job's view of its scene as a plain Scene (what CLDraw draws for
    it): the job's camera, resolution, shadow switch, PCF and exposure."""
    from .scenes import SH_SHADOW

    base = job.scene.scene
    groups = base.groups
    cascades = base.cascades
    if not job.shadows:
        groups = [dataclasses.replace(g, flags=g.flags & ~SH_SHADOW) for g in groups]
        cascades = []
    return dataclasses.replace(base, W=job.width, H=job.height, view=job.camera, groups=groups,
                               cascades=cascades, pcf_taps=job.pcf_taps,
                               exposure=job.exposure, name=f"view_{job.width}x{job.height}")


MULTIVIEW_SIZES = [(320, 240), (640, 480), (1280, 720), (1920, 1080)]


def multiview_jobs(ctx: SceneContext, n_jobs: int = 64, seed: int = 1,
                   sizes=MULTIVIEW_SIZES) -> List[Job]:
    """The batched multi-view workload's views of a spheres scene: seeded
    camera poses around and inside the 200^3 box, each looking at a point
    near the centre, resolutions drawn uniformly from `sizes`."""
    rng = np.random.default_rng(seed)
    jobs = []
    for _ in range(n_jobs):
        w, h = sizes[int(rng.integers(len(sizes)))]
        ang = rng.uniform(0, 2 * np.pi)
        dist = rng.uniform(60.0, 330.0)
        pos = np.array([dist * np.cos(ang), rng.uniform(-60.0, 120.0), dist * np.sin(ang)])
        target = rng.uniform(-60.0, 60.0, 3)
        jobs.append(Job(ctx, job_camera(w, h, pos, target), w, h))
    return jobs


class Batch:
    """One batch slot: its own stream and buffers (several may be in flight)."""

    def __init__(self, bin_size: int = 32, cap_small: int = 0, cap_large: int = 0,
                 graph: bool = False):
        self._L = lib()
        b = ctypes.c_void_p()
        check(self._L.axr_batch_create(ctypes.byref(b)), "axr_batch_create")
        self._b = b
        self.jobs: List[Job] = []
        self.ids = False
        self.configure(bin_size, cap_small, cap_large, graph)

    def configure(self, bin_size=32, cap_small=0, cap_large=0, graph=False):
        check(self._L.axr_batch_configure(self._b, int(bin_size), int(cap_small), int(cap_large),
                                          1 if graph else 0), "axr_batch_configure")
        self.bin_size = bin_size
        self.cap_small, self.cap_large, self.graph = cap_small, cap_large, graph

    def set_timing(self, enable: bool = True):
        """Per-kernel HIP-event timing of the following renders (which then
        run without graph replay)."""
        check(self._L.axr_batch_set_timing(self._b, 1 if enable else 0), "axr_batch_set_timing")

    def kernel_timing(self) -> dict:
        """setup / order / raster milliseconds of the last (timed) render and
        the bin entries its setup appended."""
        ms = (ctypes.c_float * 3)()
        n = ctypes.c_uint32()
        check(self._L.axr_batch_get_timing(self._b, ms, ctypes.byref(n)), "axr_batch_get_timing")
        return {"setup_ms": ms[0], "order_ms": ms[1], "raster_ms": ms[2],
                "bin_entries": int(n.value)}

    def visible_triangles(self, jobs) -> np.ndarray:
        """Renders the jobs once as a statistics render and returns the
        triangles each job's setup accepted."""
        check(self._L.axr_batch_set_stats(self._b, 1), "axr_batch_set_stats")
        try:
            self.render(jobs)
            out = np.zeros(len(self.jobs), np.uint32)
            check(self._L.axr_batch_get_visible(self._b, ptr(out), len(self.jobs)),
                  "axr_batch_get_visible")
        finally:
            check(self._L.axr_batch_set_stats(self._b, 0), "axr_batch_set_stats")
        return out

    def set_spill_capacity(self, spill: int):
        """Entries of the spill list from the next render on (it grows by
        itself after a render that lost entries)."""
        check(self._L.axr_batch_set_spill_capacity(self._b, int(spill)),
              "axr_batch_set_spill_capacity")

    def render(self, jobs: Sequence[Job], ids: bool = False, hdr: bool = False):
        """Enqueues the jobs and returns at once: RGBA8 frames, or winner ids
        (ids=True), or HDR u16 planes (hdr=True), each with depth."""
        self.jobs = list(jobs)
        self.ids = ids
        arr = (AxrJob * max(1, len(self.jobs)))(*[j.to_c() for j in self.jobs])
        out = OUT_IDS if ids else OUT_HDR if hdr else OUT_RGBA8
        check(self._L.axr_batch_render(self._b, arr, len(self.jobs), out), "axr_batch_render")

    def done(self) -> bool:
        return check(self._L.axr_batch_query(self._b), "axr_batch_query") == 1

    def wait(self) -> float:
        """Waits for the last render; returns its GPU time in ms."""
        ms = ctypes.c_float()
        check(self._L.axr_batch_wait(self._b, ctypes.byref(ms)), "axr_batch_wait")
        return float(ms.value)

    def _get(self, fn, j, dtype):
        job = self.jobs[j]
        out = np.empty((job.height, job.width), dtype)
        check(fn(self._b, int(j), ptr(out)), "batch readback")
        return out

    def rgba8(self, j: int) -> np.ndarray:
        """Job j's RGBA8 frame as uint32 (r | g<<8 | b<<16 | a<<24) [H, W]."""
        return self._get(self._L.axr_batch_get_rgba8, j, np.uint32)

    def depth(self, j: int) -> np.ndarray:
        return self._get(self._L.axr_batch_get_depth, j, np.float32)

    def ids_of(self, j: int) -> np.ndarray:
        return self._get(self._L.axr_batch_get_ids, j, np.uint32)

    def hdr(self, j: int):
        """Job j's HDR planes (r, g, b) u16 [H, W] of a render with hdr=True."""
        job = self.jobs[j]
        out = [np.empty((job.height, job.width), np.uint16) for _ in range(3)]
        check(self._L.axr_batch_get_hdr(self._b, int(j), *[ptr(o) for o in out]),
              "batch readback")
        return tuple(out)

    def render_cube(self, scene, name, pos, size: int, shadows: bool = True, pcf_taps: int = 5):
        """The cube map pass: `scene` (a CLDraw) seen from pos as six size^2
        HDR faces, stored as its reflection texture `name` in the strip
        layout setReflTex takes.  Enqueued; the scene's later draws wait for
        it."""
        d = scene.draw if hasattr(scene, "draw") else scene
        slot = d._refl.get(name, -1)
        p = (ctypes.c_float * 3)(*[float(v) for v in np.asarray(pos, np.float32).reshape(3)])
        slot = check(self._L.axr_render_cube(self._b, d._ctx, slot, p, int(size),
                                             1 if shadows else 0, int(pcf_taps)),
                     "axr_render_cube")
        d._refl[name] = slot
        d.reflTexSize[name] = np.int32(int(size) // 2)
        self.jobs = [Job(scene=None, camera=None, width=size, height=size) for _ in range(6)]
        self.ids = False
        return slot

    def device_bytes(self) -> int:
        n = ctypes.c_uint64()
        check(self._L.axr_batch_get_bytes(self._b, ctypes.byref(n)), "axr_batch_get_bytes")
        return int(n.value)

    def close(self):
        if getattr(self, "_b", None):
            self._L.axr_batch_destroy(self._b)
            self._b = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def settle(b: "Batch", jobs: Sequence[Job], tries: int = 6) -> float:
    """Renders until no bin entry is lost (each loss grows the storage that
    overflowed); returns the GPU ms of the complete render."""
    for _ in range(tries):
        b.render(jobs)
        try:
            return b.wait()
        except AxrError as e:
            if e.code != AXR_ECAP:
                raise
    raise AxrError("bin storage did not settle", AXR_ECAP)


def autotune(jobs: Sequence[Job], sizes=(8, 16, 32), cap_scales=(1, 4), reps: int = 3,
             graph: bool = True, batch: Optional[Batch] = None) -> dict:
    """Times the batch under each bin size and bucket-capacity scale and
    returns the fastest configuration with every timing (GPU ms, median of
    reps after one warm-up).  Capacities scale the per-size defaults
    (4096 small / 1024 large entries for a 32-px bin, by area for smaller
    bins); a configuration that spills still renders correctly, only slower."""
    own = batch is None
    b = batch or Batch()
    results = []
    try:
        for size in sizes:
            area = (size / 32.0) ** 2
            for sc in cap_scales:
                cs = max(64, int(4096 * area * sc))
                cl = max(64, int(1024 * area * sc))
                b.configure(size, cs, cl, graph)
                settle(b, jobs)
                ts = []
                for _ in range(reps):
                    b.render(jobs)
                    ts.append(b.wait())
                results.append({"bin_size": size, "cap_small": cs, "cap_large": cl,
                                "gpu_ms": float(np.median(ts))})
        best = min(results, key=lambda r: r["gpu_ms"])
        return {"best": best, "all": results}
    finally:
        if own:
            b.close()
