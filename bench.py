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
# plain 256 MiB device copy (fetch x 1.9995, write x 1.0000).  Valid for the
# kernels as they are; re-measure after changing one.
MEASURED_TRAFFIC = {
    "c2": {"shadow_raster": 105085628, "raster_shade": 133268406, "setup_bin": 138315164},
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


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--config", default="c2")
    ap.add_argument("--latency-frames", type=int, default=100)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="also run one plain device copy of MIB MiB (counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU only (asked for --gpus {args.gpus}); "
              "multi-GPU sharding is a design note, not a measured path", file=sys.stderr)
        sys.exit(2)

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
