"""Benchmark of the MI355X rendering hot path.

Default workload (the plain path at scale): 1920x1080, 1,000,000 triangles
(2,000 seeded 500-triangle icospheres in a 200^3 box), 8 mipmapped 1024^2
textures, two 2048^2 shadow cascades with 3x3 PCF, trilinear + Blinn-Phong.
A step is one full frame: clearShadowMap + shadowMap per cascade,
clearZBuffer, drawAll, gamma -- the drawAll sequence of the reference's
engine (Compute.py:308-339) with its shadow pass (Ops.py:1089-1141).
Scene data is resident in HBM before timing starts; readback is excluded
and timed separately.

Prints one JSON line with frames/s, Mtris/s, Mpix/s, per-frame p50/p99
latency, the roofline of the dominant kernel (HIP events on the render
stream) and the oracle's CPU frames/s on a bounded sample.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0

# HBM bytes per launch of the frame kernels, measured with two rocprofv3
# counter runs (FETCH_SIZE, then WRITE_SIZE) over
# `bench.py --steps 5 --warmup 2 --latency-frames 2 --no-cpu-baseline
# --calibrate-copy 256`, each scaled by what the same counter reads for the
# plain 256 MiB device copy (fetch x 1.9997, write x 1.0000); setup_bin is
# k_setup + k_coarse.  Valid for the
# kernels as they are; re-measure after changing one.
MEASURED_TRAFFIC = {
    "c2": {"shadow_raster": 92224574, "raster_shade": 134691073, "setup_bin": 134339832},
}


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
    casters = st["cascade_casters"]
    sh_entries = st["cascade_entries"]
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


def cpu_baseline(sc, budget_s=12.0):
    from oracle import pyoracle

    cores = min(16, os.cpu_count() or 1)
    t0 = time.perf_counter()
    n = 0
    while True:
        pyoracle.render(sc, nthreads=cores)
        n += 1
        if time.perf_counter() - t0 >= budget_s:
            break
    dt = time.perf_counter() - t0
    return {"value": n / dt, "unit": "frames/s", "cores": cores, "kind": "port",
            "sample": f"{n} full frames of the same workload (shadow cascades, drawAll, "
                      f"gamma) by the C++ oracle, {dt:.1f} s"}


def build_scene(name):
    from axi_visualizer_amd import scenes

    if name == "c2":
        return scenes.spheres_scene(), "plain_1080p_1M_tris_2x2048_shadow_pcf3x3"
    if name == "c1":
        return scenes.terrain_scene(), "plain_640x480_terrain_32258_tris"
    raise SystemExit(f"unknown config {name}")


def bench_c4(args):
    """Batched multi-view: 64 jobs of a 500k-triangle scene per batch,
    autotuned bin size and capacity, graph-replayed; against the same 64
    views drawn one at a time through CLDraw (the plain path)."""
    from axi_visualizer_amd import scenes
    from axi_visualizer_amd.batch import (MULTIVIEW_SIZES, Batch, SceneContext, autotune,
                                          multiview_jobs, settle)
    from axi_visualizer_amd.ops import CLDraw
    from axi_visualizer_amd.render import load_scene

    sc = scenes.spheres_scene(W=320, H=240, n_inst=1000, shadows=False, name="c4_spheres")
    ctx = SceneContext(sc)
    jobs = multiview_jobs(ctx)
    tune = autotune(jobs, sizes=(8, 16, 32), cap_scales=(1, 4), reps=3, graph=True)
    best = tune["best"]
    print(f"[bench] c4 autotune: {tune}", file=sys.stderr, flush=True)
    b = Batch(best["bin_size"], best["cap_small"], best["cap_large"], graph=True)
    settle(b, jobs)
    for _ in range(max(2, args.warmup)):
        b.render(jobs)
    b.wait()
    steps = max(1, args.steps // 20)
    t0 = time.perf_counter()
    for _ in range(steps):
        b.render(jobs)
    b.wait()
    dt = time.perf_counter() - t0
    fps = steps * len(jobs) / dt
    lat, gpu = [], []
    for _ in range(max(5, args.latency_frames // 10)):
        t1 = time.perf_counter()
        b.render(jobs)
        gpu.append(b.wait())
        lat.append((time.perf_counter() - t1) * 1e3)
    # plain path: the same views one at a time, one CLDraw per resolution
    draws = {}
    for w, h in MULTIVIEW_SIZES:
        d = CLDraw(1, len(sc.groups), w, h, 1)
        load_scene(d, sc)
        draws[w, h] = d
    shaders = [g.flags for g in sc.groups]

    def plain_pass():
        for j in jobs:
            d = draws[j.width, j.height]
            d.setScaleCull(j.camera.scale, j.camera.cax, j.camera.cay)
            d.setPos(j.camera.pos)
            d.setVM(j.camera.vm)
            d.clearZBuffer()
            d.drawAll(shaders, shadowIds=[])
            d.gamma(1.0)

    plain_pass()
    for d in draws.values():
        d.sync()
    t0 = time.perf_counter()
    reps = max(1, steps // 2)
    for _ in range(reps):
        plain_pass()
    for d in draws.values():
        d.sync()
    plain_fps = reps * len(jobs) / (time.perf_counter() - t0)
    px = sum(j.width * j.height for j in jobs)
    algo = sum(sc.ntri * 36 + j.width * j.height * 8 for j in jobs)
    gpu_ms = float(np.median(gpu))
    ach = algo / (gpu_ms * 1e-3) / 1e9
    line = {
        "metric": "frames/s (batched, 64 views per launch sequence)", "value": round(fps, 2),
        "unit": "frames/s", "n_gpus": 1, "steps": steps, "warmup": args.warmup,
        "ms_per_step": round(dt / steps * 1e3, 4), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic (seeded scene generators)",
        "config": {"workload": "batched_64_views_500k_tris_mixed_320x240_to_1080p",
                   "jobs": len(jobs), "triangles": sc.ntri, "pixels_per_batch": px,
                   "bin_size": best["bin_size"], "cap_small": best["cap_small"],
                   "cap_large": best["cap_large"], "graph": True},
        "plain_fps_same_views": round(plain_fps, 2),
        "speedup_vs_plain": round(fps / plain_fps, 3),
        "mpix_per_s": round(fps * px / len(jobs) / 1e6, 2),
        "batch_latency_ms": {"p50": round(float(np.percentile(lat, 50)), 3),
                             "p99": round(float(np.percentile(lat, 99)), 3)},
        "batch_gpu_ms": round(gpu_ms, 3),
        "autotune": tune["all"],
        "roofline": {"bound": "hbm", "achieved": round(ach, 2), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(ach / HBM_PEAK_GBS, 5), "traffic": None,
                     "kernel": "batch (setup + order + raster)", "algorithmic_bytes": int(algo),
                     "avg_us": round(gpu_ms * 1e3, 2)},
        "cpu_baseline": None,
    }
    b.close()
    for d in draws.values():
        d.close()
    ctx.close()
    print(json.dumps(line))


def bench_c5(args):
    """Serving: 16 resident scenes (50k-1M triangles, terrain and spheres,
    1-4 textures, two shadow cascades each), Poisson arrivals at 2,000 req/s,
    random scene / camera / resolution (480p, 720p, 1080p), shadows on half
    the requests, 200 GB HBM admission budget, batches formed every <= 2 ms.
    Reports latency at the offered load, the sustained (saturated) frames/s,
    and the same requests rendered one at a time for comparison."""
    from axi_visualizer_amd import scenes
    from axi_visualizer_amd.batch import Batch, Job
    from axi_visualizer_amd.server import SERVING_SIZES as C5_SIZES
    from axi_visualizer_amd.server import RenderServer, poisson_requests

    mix = scenes.serving_scenes()
    kinds = [k for _, k in mix]
    print(f"[bench] c5: {len(mix)} scenes, {sum(s.ntri for s, _ in mix)} triangles",
          file=sys.stderr, flush=True)
    srv = RenderServer([s for s, _ in mix], hbm_budget=200e9, slots=3, max_batch=64,
                       window_ms=2.0)

    def cam(s, w, h, rng):
        return scenes.serving_camera(kinds[s], w, h, rng.random(4))

    # warm-up: every slot's buffers at their largest, then one short burst
    srv.reserve(*C5_SIZES[-1])
    warm = poisson_requests(len(mix), 1e6, 200e-6, C5_SIZES, seed=99, camera_for=cam)
    srv.run(warm)
    srv.completed.clear()
    offered = poisson_requests(len(mix), 2000.0, args.serve_seconds, C5_SIZES, seed=1,
                               camera_for=cam)
    at_load = srv.run(offered)
    print(f"[bench] c5 at 2000 req/s: {at_load}", file=sys.stderr, flush=True)
    srv.completed.clear()
    burst = poisson_requests(len(mix), 1e6, 2e-3, C5_SIZES, seed=2, camera_for=cam)
    for r in burst:
        r.arrival = 0.0
    sat = srv.run(burst)
    print(f"[bench] c5 saturated: {sat}", file=sys.stderr, flush=True)
    # the same kind of requests one at a time, each waited for
    one = Batch()
    sample = burst[:min(len(burst), 300)]
    t0 = time.perf_counter()
    for r in sample:
        one.render([Job(srv.contexts[r.scene], r.camera, r.width, r.height, shadows=r.shadows,
                        pcf_taps=r.pcf_taps)])
        one.wait()
    single_fps = len(sample) / (time.perf_counter() - t0)
    one.close()
    line = {
        "metric": "frames/s (serving, sustained)", "value": round(sat["fps"], 2),
        "unit": "frames/s", "n_gpus": 1, "steps": sat["completed"], "warmup": len(warm),
        "ms_per_step": round(1e3 / sat["fps"], 4) if sat["fps"] else None,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic (seeded scene generators)",
        "config": {"workload": "serving_16_scenes_poisson_2000rps_480p_720p_1080p",
                   "scenes": len(mix), "triangles": [s.ntri for s, _ in mix],
                   "hbm_budget_bytes": 200e9, "slots": 3, "max_batch": 64, "window_ms": 2.0},
        "offered_rate": 2000.0,
        "at_offered_load": {"fps": round(at_load["fps"], 2), "p50_ms": round(at_load["p50_ms"], 3),
                            "p99_ms": round(at_load["p99_ms"], 3),
                            "mean_batch": round(at_load["mean_batch"], 2),
                            "completed": at_load["completed"]},
        "saturated": {"p50_ms": round(sat["p50_ms"], 3), "p99_ms": round(sat["p99_ms"], 3),
                      "mean_batch": round(sat["mean_batch"], 2),
                      "max_concurrent_batches": sat["max_concurrent_batches"]},
        "one_at_a_time_fps": round(single_fps, 2),
        "speedup_vs_one_at_a_time": round(sat["fps"] / single_fps, 3) if single_fps else None,
        "resident_bytes": sat["resident_bytes"],
        "roofline": None, "cpu_baseline": None,
    }
    srv.close()
    print(json.dumps(line))


def bench_c3(args):
    """Full frame at 2560x1440: the config-1 terrain upscaled to 388^2
    (299,538 triangles) plus 400 config-2 spheres (200,000 triangles),
    two 2048^2 shadow cascades with 3x3 PCF, the cube map pass (6 x 512^2
    HDR faces rendered from the scene), the SSR water plane (about 30% of
    the screen, 64 march steps) over the opaque frame, 5-level fp16 bloom,
    gamma.  A step is one whole frame."""
    from axi_visualizer_amd import scenes
    from axi_visualizer_amd.render import FullFrame

    sc = scenes.full_frame_scene()
    print(f"[bench] scene {sc.name}: {sc.ntri} tris {sc.W}x{sc.H}", file=sys.stderr, flush=True)
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
    print(f"[bench] {args.steps} frames in {dt:.3f} s", file=sys.stderr, flush=True)
    lat = []
    for _ in range(args.latency_frames):
        t1 = time.perf_counter()
        ff.frame()
        d.sync()
        lat.append((time.perf_counter() - t1) * 1e3)
    lat = np.array(lat)
    # per-kernel durations: HIP events on the scene's stream, the cube pass
    # by the batch's own events (waited for after each frame)
    d.setTiming(True)
    nt = max(10, args.steps // 4)
    cube_ms = []
    for _ in range(nt):
        ff.frame()
        cube_ms.append(ff.cube.wait())
    tm = d.timing()
    d.setTiming(False)
    kern = {k: {"avg_us": v[0] / v[1] * 1e3, "launches": v[1]} for k, v in tm.items()}
    kern["cube_pass"] = {"avg_us": float(np.mean(cube_ms)) * 1e3, "launches": nt}
    tot = sum(v["avg_us"] * v["launches"] for v in kern.values())
    for v in kern.values():
        v["share"] = v["avg_us"] * v["launches"] / tot if tot else 0.0
    st = d.stats()
    water = sum(g.ntri for g in sc.groups if g.ssr >= 0)
    W, H = sc.W, sc.H
    # algorithmic bytes of the kernels a roofline is stated for: bloom reads
    # and writes the frame's planes (12 B/px) and writes each fp16 level and
    # its upsampled sum once and reads them back (5 x 6 B per level texel,
    # 1/3 texel per frame pixel over the levels)
    per_kernel = {"bloom": W * H * 12 + int(5 * 6 * W * H / 3)}
    dom = max((k for k in kern if k in per_kernel), key=lambda k: kern[k]["avg_us"])
    ach = per_kernel[dom] / (kern[dom]["avg_us"] * 1e-6) / 1e9
    roof = {"bound": "hbm", "achieved": round(ach, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(ach / HBM_PEAK_GBS, 5), "traffic": None, "kernel": dom,
            "algorithmic_bytes": int(per_kernel[dom]), "avg_us": round(kern[dom]["avg_us"], 2)}
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
        "kernels": {k: {"avg_us": round(v["avg_us"], 2), "share": round(v["share"], 4)}
                    for k, v in sorted(kern.items(), key=lambda kv: -kv[1]["share"])},
        "visible": st,
        "roofline": roof, "cpu_baseline": None,
    }
    ff.close()
    print(json.dumps(line))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--config", default="c2")
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
        return bench_c3(args)
    if args.config == "c4":
        return bench_c4(args)
    if args.config == "c5":
        return bench_c5(args)

    from axi_visualizer_amd.render import Renderer

    sc, workload = build_scene(args.config)
    print(f"[bench] scene {sc.name}: {sc.ntri} tris {sc.W}x{sc.H}", file=sys.stderr, flush=True)
    r = Renderer(sc)
    d = r.draw
    for _ in range(args.warmup):
        r.frame()
    d.sync()
    # one probe frame for the per-pass counters (casters and shadow-bin
    # entries of each cascade, screen-bin entries of the draw)
    per_cascade = []
    r.frame(after_cascade=lambda i: per_cascade.append(d.stats()))
    st = d.stats()
    st["cascade_casters"] = [p["casters"] for p in per_cascade]
    st["cascade_entries"] = [p["shadow_bin_entries"] for p in per_cascade]
    if st["bin_overflow"]:
        raise SystemExit(f"bin capacity exceeded: {st}")

    # throughput: K frames back to back
    t0 = time.perf_counter()
    for _ in range(args.steps):
        r.frame()
    d.sync()
    dt = time.perf_counter() - t0
    fps = args.steps / dt
    print(f"[bench] {args.steps} frames in {dt:.3f} s", file=sys.stderr, flush=True)

    # latency: each frame submitted alone and waited for
    lat = []
    for _ in range(args.latency_frames):
        t1 = time.perf_counter()
        r.frame()
        d.sync()
        lat.append((time.perf_counter() - t1) * 1e3)
    lat = np.array(lat)

    # per-kernel durations with HIP events on the render stream
    d.setTiming(True)
    nt = max(20, args.steps // 4)
    for _ in range(nt):
        r.frame()
    tm = d.timing()
    d.setTiming(False)
    kern = {k: {"avg_us": v[0] / v[1] * 1e3, "launches": v[1], "share": 0.0}
            for k, v in tm.items()}
    tot = sum(v[0] for v in tm.values())
    for k, v in tm.items():
        kern[k]["share"] = v[0] / tot if tot else 0.0
    dom = max(tm, key=lambda k: tm[k][0])

    # readback, timed separately
    t2 = time.perf_counter()
    d.getRGBA8()
    readback_ms = (time.perf_counter() - t2) * 1e3

    frame_bytes, per_kernel = algorithmic_bytes(sc, st)
    dom_us = kern[dom]["avg_us"]
    dom_bytes = per_kernel.get(dom)
    roof = None
    if dom_bytes:
        ach = dom_bytes / (dom_us * 1e-6) / 1e9
        traffic = MEASURED_TRAFFIC.get(args.config, {}).get(dom)
        roof = {"bound": "hbm", "achieved": round(ach, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(ach / HBM_PEAK_GBS, 5), "traffic": traffic, "kernel": dom,
                "algorithmic_bytes": int(dom_bytes), "avg_us": round(dom_us, 2)}
    frame_gbs = frame_bytes * fps / 1e9

    if args.calibrate_copy:
        calibrate_copy(args.calibrate_copy)
    cpu = None if args.no_cpu_baseline else cpu_baseline(sc)
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
        "readback_ms": round(readback_ms, 3),
        "frame_algorithmic_bytes": int(frame_bytes),
        "frame_roofline_frac": round(frame_gbs / HBM_PEAK_GBS, 5),
        "visible": st,
        "kernels": {k: {"avg_us": round(v["avg_us"], 2), "share": round(v["share"], 4)}
                    for k, v in sorted(kern.items(), key=lambda kv: -kv[1]["share"])},
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    r.close()
    print(json.dumps(line))


if __name__ == "__main__":
    main()
