"""Benchmark of the MI355X rendering hot path.

Default run (one JSON line):
* the plain path at scale (the headline `value`): 1920x1080, 1,000,000
  triangles (2,000 seeded 500-triangle icospheres in a 200^3 box), 8
  mipmapped 1024^2 textures, two 2048^2 shadow cascades with 3x3 PCF,
  trilinear + Blinn-Phong.  A step is one full frame: clearShadowMap +
  shadowMap per cascade, clearZBuffer, drawAll, gamma -- the drawAll
  sequence of the reference's engine (Compute.py:308-339) with its shadow
  pass (Ops.py:1089-1141).  Scene data is resident in HBM before timing
  starts; readback is excluded and timed separately.  A second camera pose
  close to a sphere inside the box (large triangles: the tiled path) is
  timed beside it as extra fields;
* `batched_multiview`: 64 views of a 500k-triangle scene at 320x240 to
  1920x1080 per batch (setup, order, persistent raster; autotuned bins,
  graph replay) against the same views through CLDraw back to back;
* `serving`: 16 resident scenes, Poisson requests at 2,000 req/s and a
  saturated burst, against the same requests through CLDraw back to back
  and one at a time.
Each carries its frames/s, p50/p99, the roofline of its dominant kernel
(HIP events on the stream the kernel runs on) and the oracle's frames/s on
a bounded sample of the same workload on the host CPU.

`--config c1|c2|c3|c4|c5` prints one workload's line alone (c2 without the
sub-objects).
"""
from __future__ import annotations

import argparse
import dataclasses
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
# threads the CPU baseline may use: the host cores a GPU is allotted
CPU_SHARE_MAX = 16

# HBM bytes per launch of the frame kernels, measured with two rocprofv3
# counter runs (FETCH_SIZE, then WRITE_SIZE) over
# `bench.py --config c2 --steps 5 --warmup 2 --latency-frames 2
# --no-cpu-baseline --calibrate-copy 256`, each scaled by what the same
# counter reads for the plain 256 MiB device copy (fetch x 1.9997, write
# x 1.0000), headline pose; setup_bin is k_setup + k_coarse.  Valid for
# the kernels as they are; re-measure after changing one.
MEASURED_TRAFFIC = {
    "c2": {"shadow_raster": 97748146, "raster_shade": 134181922},
    # the batch raster of the bench's own batch (the 64 views; for serving
    # the median over the saturated burst's 64-request batches), same
    # calibration, from
    # `bench.py --config c4|c5 --steps 5 --warmup 2 --latency-frames 2
    # --serve-seconds 0.3 --no-cpu-baseline --calibrate-copy 256`
    "c4": {"batch_raster": 2254967526},
    "c5": {"batch_raster": 5329793301},
    # the opaque HDR raster of the 1440p full frame, from
    # `bench.py --config c3 --steps 5 --warmup 2 --latency-frames 2
    # --no-cpu-baseline --calibrate-copy 256`
    "c3": {"raster_hdr": 158039272},
}


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


# ------------------------------------------------------------ roofline
def algorithmic_bytes(sc, st):
    """Compulsory HBM bytes per frame and per kernel launch.

    Per-unit figures: a triangle carries 132 B (3 vertices x pos 12 +
    normal 12 + uv 8 + light 12); a pixel 8 B out (RGBA8 + f32 depth); a
    texture 3 B per texel over its mip chain; a shadow cascade re-reads
    36 B of positions per triangle and writes and reads its map (4 B/texel).
    Per kernel: a setup record is 36 B, a bin entry 4 B, a class or caster
    flag 1 B."""
    W, H, T = sc.W, sc.H, sc.ntri
    tex = sum(3 * g.tex.shape[0] for g in sc.groups)
    frame = T * 132 + W * H * 8 + tex
    for c in sc.cascades:
        frame += T * 36 + 2 * c.dim * c.dim * 4
    visible = st["small"] + st["large"]
    entries = st["bin_entries"]
    casters = st.get("cascade_casters", [])
    sh_entries = st.get("cascade_entries", [])
    per_kernel = {
        # bin entries and the setup records they name; RGBA8 + depth out
        # (8 B/px); the winning triangles' vertex data (132 B each, bounded
        # by the visible set)
        "raster_shade": entries * (4 + 36) + W * H * 8 + visible * 132,
        # one launch for the draw and every cascade: positions read once
        # (36 B/tri); class byte, record and entries of the draw; caster
        # flag, record and entries per cascade
        "setup_bin": T * (36 + 1) + visible * 36 + entries * 4
        + sum(T + c * 36 + e * 4 for c, e in zip(casters, sh_entries)),
        # one launch for every cascade: caster records via their bin
        # entries, each map written once
        "shadow_raster": sum(e * (36 + 4) for e in sh_entries)
        + sum(c.dim * c.dim * 4 for c in sc.cascades),
    }
    return frame, per_kernel


def roofline(kernel, algo_bytes, avg_us, traffic=None):
    ach = algo_bytes / (avg_us * 1e-6) / 1e9
    return {"bound": "hbm", "achieved": round(ach, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(ach / HBM_PEAK_GBS, 5), "traffic": traffic, "kernel": kernel,
            "algorithmic_bytes": int(algo_bytes), "avg_us": round(avg_us, 2)}


def calibrate_copy(mb):
    """One device-to-device copy of `mb` MiB through the HIP runtime: a plain
    copy kernel of known traffic (mb MiB read + mb MiB written) that gives
    the scale of the HBM byte counters (FETCH_SIZE, WRITE_SIZE) in a
    profiled run."""
    import ctypes

    hip = ctypes.CDLL("libamdhip64.so")
    n = mb << 20
    a, b = ctypes.c_void_p(), ctypes.c_void_p()
    assert hip.hipMalloc(ctypes.byref(a), ctypes.c_size_t(n)) == 0
    assert hip.hipMalloc(ctypes.byref(b), ctypes.c_size_t(n)) == 0
    assert hip.hipMemcpy(b, a, ctypes.c_size_t(n), 3) == 0  # device to device
    assert hip.hipDeviceSynchronize() == 0
    hip.hipFree(a)
    hip.hipFree(b)


# -------------------------------------------------------- CPU baseline
def cpu_info():
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        share = len(os.sched_getaffinity(0))
    except AttributeError:
        share = os.cpu_count() or 1
    return {"cpu_model": model, "host_cpus": os.cpu_count() or 1, "cpu_share": share}


def cpu_baseline(frames, what, budget_s=10.0, one_thread_s=5.0):
    """The C++ oracle (the reference's OpenCL kernels restated) rendering the
    scenes in `frames` (cycled) on the host: with one worker per allotted
    core for about budget_s, then on one thread for about one_thread_s (at
    least one frame)."""
    from oracle import pyoracle

    info = cpu_info()
    threads = max(1, min(CPU_SHARE_MAX, info["cpu_share"]))

    def rate(nthreads, budget):
        t0 = time.perf_counter()
        n = 0
        while True:
            pyoracle.render(frames[n % len(frames)], nthreads=nthreads)
            n += 1
            if time.perf_counter() - t0 >= budget:
                break
        return n, time.perf_counter() - t0

    n, dt = rate(threads, budget_s)
    n1, dt1 = rate(1, one_thread_s)
    out = {"value": round(n / dt, 3), "unit": "frames/s", "cores": threads, "kind": "port",
           "sample": f"{n} frames of {what} by the C++ oracle on {threads} threads in "
                     f"{dt:.1f} s; {n1} frames on 1 thread in {dt1:.1f} s",
           "one_thread_value": round(n1 / dt1, 4)}
    out.update(info)
    return out


# ------------------------------------------------------- plain path (C2)
def build_scene(name):
    from axi_visualizer_amd import scenes

    if name == "c2":
        return scenes.spheres_scene(), "plain_1080p_1M_tris_2x2048_shadow_pcf3x3"
    if name == "c1":
        return scenes.terrain_scene(), "plain_640x480_terrain_32258_tris"
    raise SystemExit(f"unknown config {name}")


def time_plain(r, steps, latency_frames, timing_frames):
    """Frames/s (steps back to back), per-frame latency (each submitted and
    waited for) and per-kernel HIP-event durations of a Renderer."""
    d = r.draw
    t0 = time.perf_counter()
    for _ in range(steps):
        r.frame()
    d.sync()
    dt = time.perf_counter() - t0
    lat = []
    for _ in range(latency_frames):
        t1 = time.perf_counter()
        r.frame()
        d.sync()
        lat.append((time.perf_counter() - t1) * 1e3)
    d.setTiming(True)
    for _ in range(timing_frames):
        r.frame()
    tm = d.timing()
    d.setTiming(False)
    tot = sum(v[0] for v in tm.values())
    kern = {k: {"avg_us": v[0] / v[1] * 1e3, "launches": v[1],
                "share": v[0] / tot if tot else 0.0} for k, v in tm.items()}
    return steps / dt, dt, np.array(lat if lat else [0.0]), kern


def probe_stats(r):
    """One frame's counters: visible triangles, bin entries, and per cascade
    its casters and shadow-bin entries."""
    d = r.draw
    per_cascade = []
    r.frame(after_cascade=lambda i: per_cascade.append(d.stats()))
    st = d.stats()
    st["cascade_casters"] = [p["casters"] for p in per_cascade]
    st["cascade_entries"] = [p["shadow_bin_entries"] for p in per_cascade]
    if st["bin_overflow"]:
        raise SystemExit(f"bin capacity exceeded: {st}")
    ids = d.getIds()
    st["coverage"] = round(float((ids != 0xFFFFFFFF).mean()), 5)
    return st


def kernels_json(kern):
    return {k: {"avg_us": round(v["avg_us"], 2), "share": round(v["share"], 4)}
            for k, v in sorted(kern.items(), key=lambda kv: -kv[1]["share"])}


def run_plain(args, config):
    from axi_visualizer_amd import scenes
    from axi_visualizer_amd.render import Renderer

    sc, workload = build_scene(config)
    log(f"scene {sc.name}: {sc.ntri} tris {sc.W}x{sc.H}")
    r = Renderer(sc)
    d = r.draw
    for _ in range(args.warmup):
        r.frame()
    d.sync()
    st = probe_stats(r)
    fps, dt, lat, kern = time_plain(r, args.steps, args.latency_frames,
                                    max(20, args.steps // 4))
    log(f"{args.steps} frames in {dt:.3f} s")
    dom = max(kern, key=lambda k: kern[k]["avg_us"] * kern[k]["launches"])

    # readback, timed separately (PCIe-inclusive; never part of value):
    # median of 10 reads of the finished frame into CLDraw's page-locked
    # arrays (allocated by the first, untimed read)
    def med_ms(fn, n=10):
        fn()
        ts = []
        for _ in range(n):
            t2 = time.perf_counter()
            fn()
            ts.append((time.perf_counter() - t2) * 1e3)
        return float(np.median(ts))

    readback_ms = med_ms(d.getRGBA8)
    getframe_ms = med_ms(d.getFrame)
    # the reference engine's loop (Compute.py:308-390): each frame rendered,
    # gamma'd and read back with getFrame before the next starts
    ne = max(20, args.steps // 4)
    t2 = time.perf_counter()
    for _ in range(ne):
        r.frame()
        d.getFrame()
    engine_fps = ne / (time.perf_counter() - t2)

    frame_bytes, per_kernel = algorithmic_bytes(sc, st)
    roof = None
    if dom in per_kernel:
        roof = roofline(dom, per_kernel[dom], kern[dom]["avg_us"],
                        MEASURED_TRAFFIC.get(config, {}).get(dom))
    line = {
        "metric": "frames/s", "value": round(fps, 2), "unit": "frames/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(1e3 / fps, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic (seeded scene generators)",
        "config": {"workload": workload, "W": sc.W, "H": sc.H, "triangles": sc.ntri,
                   "groups": len(sc.groups), "shadow_cascades": len(sc.cascades),
                   "shadow_dim": sc.cascades[0].dim if sc.cascades else 0,
                   "pcf_taps": sc.pcf_taps},
        "mtris_per_s": round(sc.ntri * fps / 1e6, 2),
        "mpix_per_s": round(sc.W * sc.H * fps / 1e6, 2),
        "latency_ms": {"p50": round(float(np.percentile(lat, 50)), 4),
                       "p99": round(float(np.percentile(lat, 99)), 4)},
        "readback_ms": {"rgba8": round(readback_ms, 3), "getFrame": round(getframe_ms, 3)},
        "engine_loop_fps": round(engine_fps, 2),
        "frame_algorithmic_bytes": int(frame_bytes),
        "frame_roofline_frac": round(frame_bytes * fps / 1e9 / HBM_PEAK_GBS, 5),
        "visible": st,
        "kernels": kernels_json(kern),
        "roofline": roof,
    }
    cpu_frames = [sc]
    if config == "c2":
        # the same scene from a camera inside the box, 2 units in front of a
        # large sphere: hundreds of large triangles on the tiled path and
        # most of the frame shaded
        sc2 = dataclasses.replace(sc, view=scenes.spheres_closeup_camera(sc.W, sc.H, gap=2.0),
                                  name="c2_closeup")
        r2 = Renderer(sc2)
        for _ in range(args.warmup):
            r2.frame()
        r2.draw.sync()
        st2 = probe_stats(r2)
        fps2, _, lat2, kern2 = time_plain(r2, max(20, args.steps // 2),
                                          max(10, args.latency_frames // 2),
                                          max(20, args.steps // 4))
        _, pk2 = algorithmic_bytes(sc2, st2)
        line["closeup_pose"] = {
            "fps": round(fps2, 2), "mtris_per_s": round(sc2.ntri * fps2 / 1e6, 2),
            "latency_ms": {"p50": round(float(np.percentile(lat2, 50)), 4),
                           "p99": round(float(np.percentile(lat2, 99)), 4)},
            "visible": st2, "kernels": kernels_json(kern2),
            "roofline_raster": roofline("raster_shade", pk2["raster_shade"],
                                        kern2["raster_shade"]["avg_us"]),
        }
        r2.close()
        cpu_frames.append(sc2)
    if not args.no_cpu_baseline:
        line["cpu_baseline"] = cpu_baseline([sc], f"the {workload} workload (shadow cascades, "
                                            "drawAll, gamma)")
        if config == "c2":
            from oracle import pyoracle
            ref = pyoracle.render(cpu_frames[-1], nthreads=CPU_SHARE_MAX)
            covered = max(1, int((ref["id"] != 0xFFFFFFFF).sum()))
            line["closeup_pose"]["fragments_per_covered_px"] = round(ref["stats"][2] / covered, 3)
    else:
        line["cpu_baseline"] = None
    r.close()
    return line


# ------------------------------------------------ batched multi-view (C4)
def plain_views(jobs, shaders_of, reps):
    """The jobs' views through CLDraw, one CLDraw per (scene, resolution),
    back to back without host syncs (the engine's per-frame calls:
    setScaleCull, setPos, setVM, clearZBuffer, drawAll, gamma); the scene's
    shadow maps rendered once, as the batch keeps them resident."""
    from axi_visualizer_amd.ops import CLDraw
    from axi_visualizer_amd.render import load_scene

    draws = {}
    for j in jobs:
        key = (id(j.scene), j.width, j.height)
        if key in draws:
            continue
        sc = j.scene.scene
        d = CLDraw(1, len(sc.groups), j.width, j.height, 1)
        load_scene(d, sc)
        cast = [g.cast for g in sc.groups]
        for i, c in enumerate(sc.cascades):
            d.clearShadowMap(i)
            d.shadowMap(i, cast, shaders_of(sc, True), c.bias)
        draws[key] = d

    def one_pass():
        for j in jobs:
            d = draws[(id(j.scene), j.width, j.height)]
            c = j.camera
            d.setScaleCull(c.scale, c.cax, c.cay)
            d.setPos(c.pos)
            d.setVM(c.vm)
            d.clearZBuffer()
            sc = j.scene.scene
            d.drawAll(shaders_of(sc, j.shadows), shadowIds=list(range(len(sc.cascades)))
                      if j.shadows else [], pcf=j.pcf_taps if j.shadows else 5)
            d.gamma(j.exposure)

    one_pass()
    for d in draws.values():
        d.sync()
    t0 = time.perf_counter()
    for _ in range(reps):
        one_pass()
    for d in draws.values():
        d.sync()
    fps = reps * len(jobs) / (time.perf_counter() - t0)
    for d in draws.values():
        d.close()
    return fps


def shader_flags_of(sc, shadows):
    from axi_visualizer_amd import scenes
    return [g.flags if shadows else g.flags & ~scenes.SH_SHADOW for g in sc.groups]


def batch_kernel_roofline(b, jobs, traffic=None, reps=5):
    """Per-kernel HIP-event times of the batch (eager launches) and the
    roofline of its dominant kernel, the persistent raster, by the
    single-frame raster's per-unit figures: 40 B per bin entry (4 B entry +
    36 B record), 8 B per pixel written, 132 B per visible triangle (counted
    by one statistics render before the timed ones)."""
    visible = int(b.visible_triangles(jobs).sum())
    b.set_timing(True)
    ts = []
    for _ in range(reps):
        b.render(jobs)
        ts.append(b.kernel_timing())
    b.set_timing(False)
    med = {k: float(np.median([t[k] for t in ts])) for k in ("setup_ms", "order_ms", "raster_ms")}
    entries = ts[-1]["bin_entries"]
    px = sum(j.width * j.height for j in jobs)
    algo = entries * 40 + px * 8 + visible * 132
    roof = roofline("batch_raster", algo, med["raster_ms"] * 1e3, traffic)
    roof.update(bin_entries=entries, pixels=px, visible_triangles=visible)
    return med, roof


def run_multiview(args, cpu=True):
    from axi_visualizer_amd import scenes
    from axi_visualizer_amd.batch import (Batch, SceneContext, autotune, job_view_scene,
                                          multiview_jobs, settle)

    sc = scenes.spheres_scene(W=320, H=240, n_inst=1000, shadows=False, name="c4_spheres")
    ctx = SceneContext(sc)
    jobs = multiview_jobs(ctx)
    tune = autotune(jobs, sizes=(8, 16, 32), cap_scales=(1, 4), reps=3, graph=True)
    best = tune["best"]
    log(f"c4 autotune: {best}")
    b = Batch(best["bin_size"], best["cap_small"], best["cap_large"], graph=True)
    settle(b, jobs)
    for _ in range(max(2, args.warmup)):
        b.render(jobs)
    b.wait()
    steps = max(5, args.steps // 10)
    t0 = time.perf_counter()
    for _ in range(steps):
        b.render(jobs)
    b.wait()
    dt = time.perf_counter() - t0
    fps = steps * len(jobs) / dt
    lat = []
    for _ in range(max(5, args.latency_frames // 10)):
        t1 = time.perf_counter()
        b.render(jobs)
        b.wait()
        lat.append((time.perf_counter() - t1) * 1e3)
    kms, roof = batch_kernel_roofline(b, jobs, MEASURED_TRAFFIC["c4"]["batch_raster"])
    b.close()
    plain = plain_views(jobs, shader_flags_of, reps=max(1, steps // 2))
    px = sum(j.width * j.height for j in jobs)
    out = {
        "workload": "batched_64_views_500k_tris_mixed_320x240_to_1080p",
        "jobs": len(jobs), "triangles": sc.ntri, "pixels_per_batch": px,
        "bin_size": best["bin_size"], "graph": True,
        "fps": round(fps, 2), "steps": steps, "ms_per_batch": round(dt / steps * 1e3, 4),
        "mpix_per_s": round(fps * px / len(jobs) / 1e6, 2),
        "batch_latency_ms": {"p50": round(float(np.percentile(lat, 50)), 3),
                             "p99": round(float(np.percentile(lat, 99)), 3)},
        "plain_fps_same_views": round(plain, 2),
        "speedup_vs_plain_same_views": round(fps / plain, 3),
        "kernels_ms": {k: round(v, 4) for k, v in kms.items()},
        "roofline": roof,
    }
    if cpu and not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline([job_view_scene(j) for j in jobs[:16]],
                                           "the batch's views (its first 16 jobs)",
                                           budget_s=8.0, one_thread_s=4.0)
    else:
        out["cpu_baseline"] = None
    ctx.close()
    return out


# ------------------------------------------------------------ serving (C5)
def run_serving(args, cpu=True):
    from axi_visualizer_amd import scenes
    from axi_visualizer_amd.batch import Batch, Job, job_view_scene
    from axi_visualizer_amd.server import SERVING_SIZES, RenderServer, poisson_requests

    mix = scenes.serving_scenes()
    kinds = [k for _, k in mix]
    log(f"c5: {len(mix)} scenes, {sum(s.ntri for s, _ in mix)} triangles")
    srv = RenderServer([s for s, _ in mix], hbm_budget=200e9, slots=3, max_batch=64,
                       window_ms=2.0)

    def cam(s, w, h, rng):
        return scenes.serving_camera(kinds[s], w, h, rng.random(4))

    # warm-up: every slot's buffers at their largest, then one short burst
    srv.reserve(*SERVING_SIZES[-1])
    warm = poisson_requests(len(mix), 1e6, 200e-6, SERVING_SIZES, seed=99, camera_for=cam)
    srv.run(warm)
    srv.completed.clear()
    offered = poisson_requests(len(mix), 2000.0, args.serve_seconds, SERVING_SIZES, seed=1,
                               camera_for=cam)
    at_load = srv.run(offered)
    log(f"c5 at 2000 req/s: {at_load}")
    srv.completed.clear()
    burst = poisson_requests(len(mix), 1e6, 2e-3, SERVING_SIZES, seed=2, camera_for=cam)
    for r in burst:
        r.arrival = 0.0
    sat = srv.run(burst)
    log(f"c5 saturated: {sat}")
    # the dominant kernel of a full batch of the burst's requests
    jobs64 = [Job(srv.contexts[r.scene], r.camera, r.width, r.height, shadows=r.shadows,
                  pcf_taps=r.pcf_taps) for r in burst[:64]]
    kb = Batch()
    kb.render(jobs64)
    kb.wait()
    kms, roof = batch_kernel_roofline(kb, jobs64, MEASURED_TRAFFIC["c5"]["batch_raster"])
    # the same requests one at a time, each waited for
    sample = burst[:min(len(burst), 300)]
    t0 = time.perf_counter()
    for r in sample:
        kb.render([Job(srv.contexts[r.scene], r.camera, r.width, r.height, shadows=r.shadows,
                       pcf_taps=r.pcf_taps)])
        kb.wait()
    single_fps = len(sample) / (time.perf_counter() - t0)
    kb.close()
    sample_jobs = [Job(srv.contexts[r.scene], r.camera, r.width, r.height, shadows=r.shadows,
                       pcf_taps=r.pcf_taps) for r in sample]
    plain = plain_views(sample_jobs, shader_flags_of, reps=2)
    out = {
        "workload": "serving_16_scenes_poisson_2000rps_480p_720p_1080p",
        "scenes": len(mix), "triangles": [s.ntri for s, _ in mix],
        "hbm_budget_bytes": 200e9, "slots": 3, "max_batch": 64, "window_ms": 2.0,
        "fps": round(sat["fps"], 2), "requests": sat["completed"],
        "at_offered_load": {"offered_rate": 2000.0, "fps": round(at_load["fps"], 2),
                            "p50_ms": round(at_load["p50_ms"], 3),
                            "p99_ms": round(at_load["p99_ms"], 3),
                            "mean_batch": round(at_load["mean_batch"], 2),
                            "completed": at_load["completed"]},
        "saturated": {"p50_ms": round(sat["p50_ms"], 3), "p99_ms": round(sat["p99_ms"], 3),
                      "mean_batch": round(sat["mean_batch"], 2),
                      "max_concurrent_batches": sat["max_concurrent_batches"]},
        "plain_fps_same_requests": round(plain, 2),
        "speedup_vs_plain_same_requests": round(sat["fps"] / plain, 3),
        "one_at_a_time_fps": round(single_fps, 2),
        "speedup_vs_one_at_a_time": round(sat["fps"] / single_fps, 3),
        "resident_bytes": sat["resident_bytes"],
        "kernels_ms_64_request_batch": {k: round(v, 4) for k, v in kms.items()},
        "roofline": roof,
    }
    if cpu and not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline([job_view_scene(j) for j in sample_jobs[:24]],
                                           "the burst's requests (its first 24; the oracle "
                                           "renders a request's shadow maps with it)",
                                           budget_s=8.0, one_thread_s=4.0)
    else:
        out["cpu_baseline"] = None
    srv.close()
    return out


# --------------------------------------------------------- full frame (C3)
def run_full_frame(args):
    """Full frame at 2560x1440: the config-1 terrain upscaled to 388^2
    (299,538 triangles) plus 400 config-2 spheres (200,000 triangles),
    two 2048^2 shadow cascades with 3x3 PCF, the cube map pass (6 x 512^2
    HDR faces rendered from the scene), the SSR water plane (about 30% of
    the screen, 64 march steps) over the opaque frame, 5-level fp16 bloom,
    gamma.  A step is one whole frame."""
    from axi_visualizer_amd import scenes
    from axi_visualizer_amd.render import FullFrame

    sc = scenes.full_frame_scene()
    log(f"scene {sc.name}: {sc.ntri} tris {sc.W}x{sc.H}")
    ff = FullFrame(sc, scenes.full_frame_probe(sc), cube_size=512)
    d = ff.draw
    for _ in range(args.warmup):
        ff.frame()
    d.sync()
    ff.cube.wait()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        ff.frame()
    d.sync()
    dt = time.perf_counter() - t0
    fps = args.steps / dt
    log(f"{args.steps} frames in {dt:.3f} s")
    lat = []
    for _ in range(args.latency_frames):
        t1 = time.perf_counter()
        ff.frame()
        d.sync()
        lat.append((time.perf_counter() - t1) * 1e3)
    lat = np.array(lat if lat else [0.0])
    # the opaque drawAll's counters (bin entries, visible triangles) for the
    # raster's algorithmic bytes: one drawAll of the frame without its
    # water groups (the SSR pass sets up and draws them on its own)
    water_mask = [m or isinstance(s, dict) for m, s in zip(ff.mask, ff.shaders)]
    d.setPos(sc.view.pos)
    d.setVM(sc.view.vm)
    d.clearZBuffer()
    d.drawAll([0 if isinstance(s, dict) else s for s in ff.shaders], mask=water_mask,
              shadowIds=list(range(len(sc.cascades))), pcf=sc.pcf_taps)
    st = d.stats()
    d.sync()
    # per-pass times with the cube pass finished before the drawAll is
    # issued (the timed frames above overlap the two)
    d.setTiming(True)
    nt = max(10, args.steps // 4)
    cube_ms = []
    for _ in range(nt):
        ff.frame(serial=True)
        cube_ms.append(ff.cube.wait())
    tm = d.timing()
    d.setTiming(False)
    kern = {k: {"avg_us": v[0] / v[1] * 1e3, "launches": v[1]} for k, v in tm.items()}
    kern["cube_pass"] = {"avg_us": float(np.mean(cube_ms)) * 1e3, "launches": nt}
    tot = sum(v["avg_us"] * v["launches"] for v in kern.values())
    for v in kern.values():
        v["share"] = v["avg_us"] * v["launches"] / tot if tot else 0.0
    water = sum(g.ntri for g in sc.groups if g.ssr >= 0)
    W, H = sc.W, sc.H
    raster = next((k for k in kern if k.startswith("raster") and "ssr" not in k), None)
    roof = None
    if raster:
        # HDR raster: bin entries (4 B) and their records (36 B), the visible
        # triangles' vertex data (132 B), depth and three u16 planes out
        # (10 B/px); counters of the frame's last setup
        algo = st["bin_entries"] * 40 + (st["small"] + st["large"]) * 132 + W * H * 10
        roof = roofline(raster, algo, kern[raster]["avg_us"],
                        MEASURED_TRAFFIC["c3"].get(raster))
    line = {
        "metric": "frames/s (full frame)", "value": round(fps, 2), "unit": "frames/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(1e3 / fps, 4), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic (seeded scene generators)",
        "config": {"workload": "full_frame_1440p_500k_tris_shadow_cube6x512_ssr64_bloom5",
                   "W": W, "H": H, "triangles": sc.ntri, "water_triangles": water,
                   "shadow_cascades": len(sc.cascades), "shadow_dim": sc.cascades[0].dim,
                   "cube_face": 512, "ssr_steps": sc.ssr_steps, "bloom_levels": 5},
        "mtris_per_s": round(sc.ntri * fps / 1e6, 2),
        "mpix_per_s": round(W * H * fps / 1e6, 2),
        "latency_ms": {"p50": round(float(np.percentile(lat, 50)), 4),
                       "p99": round(float(np.percentile(lat, 99)), 4)},
        "kernels": kernels_json(kern),
        "visible": st,
        "roofline": roof, "cpu_baseline": None,
    }
    ff.close()
    return line


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--config", default="all",
                    help="all (default: c2 with the batched and serving sub-objects), "
                         "c1, c2, c3, c4 or c5 alone")
    ap.add_argument("--latency-frames", type=int, default=100)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--serve-seconds", type=float, default=3.0,
                    help="serving: seconds of Poisson arrivals")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="also run one plain device copy of MIB MiB (counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU only (asked for --gpus {args.gpus}); "
              "multi-GPU sharding is a design note, not a measured path", file=sys.stderr)
        sys.exit(2)

    if args.config == "c3":
        line = run_full_frame(args)
    elif args.config == "c4":
        line = run_multiview(args)
    elif args.config == "c5":
        line = run_serving(args)
    elif args.config in ("c1", "c2"):
        line = run_plain(args, args.config)
    elif args.config == "all":
        line = run_plain(args, "c2")
        line["metric"] = ("frames/s at 1080p / 1M tris (plain path); batched-serving vs "
                          "plain fps in batched_multiview and serving")
        mv = run_multiview(args)
        sv = run_serving(args)
        for sub in (mv, sv):
            sub["speedup_vs_plain_headline"] = round(sub["fps"] / line["value"], 3)
        line["batched_multiview"] = mv
        line["serving"] = sv
    else:
        raise SystemExit(f"unknown config {args.config}")
    if args.calibrate_copy:
        calibrate_copy(args.calibrate_copy)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
