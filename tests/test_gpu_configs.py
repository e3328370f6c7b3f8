"""The full-frame, batched multi-view and serving workloads at their full
sizes against the oracle: the 2560x1440 full frame (500k triangles, two
2048^2 cascades, the 6 x 512^2 cube pass, SSR water, 5-level bloom); the
64-view batch of a 500k-triangle scene at 320x240 to 1920x1080 with
autotuned bins and graph replay; a saturated burst of the 16-scene serving
mix at 480p, 720p and 1080p, shadows on and off.  Depth, shadow maps and
winner ids bit-exact; RGB within 1/255 on >= 99.9% of pixels with the worst
error printed; every batched or served frame identical to the same view
rendered alone."""
import numpy as np
import pytest

from axi_visualizer_amd import scenes
from oracle import pyoracle
from tests.parity import RGB_MIN_FRAC, RGB_TOL, compare_rgba, job_scene

pytestmark = pytest.mark.gpu


def test_full_frame_at_1440p_matches_oracle():
    from axi_visualizer_amd.render import FullFrame
    from tests.test_gpu_ssr import oracle_cube

    sc = scenes.full_frame_scene()
    assert (sc.W, sc.H) == (2560, 1440) and sc.ntri >= 500_000
    probe = scenes.full_frame_probe(sc)
    ff = FullFrame(sc, probe, cube_size=512)
    try:
        ff.frame(gamma=True)
        got = [p.copy() for p in ff.draw.getFrame()]
        maps = [ff.draw.getSHM(i) for i in range(len(sc.cascades))]
        cube_gpu = ff.draw.getReflTex(0)
        ff.frame(gamma=True)
        again = ff.draw.getFrame()
    finally:
        ff.close()
    cube = oracle_cube(sc, probe, 512)
    for c in range(3):
        bad = np.argwhere(cube_gpu[c] != cube[..., c])
        assert bad.size == 0, f"cube channel {c}: {len(bad)} texels differ"
    sc.refl = [cube]
    ex = sc.exposure
    sc.exposure = 0.0
    ref = pyoracle.render(sc, with_shadows_out=True)
    for i, (a, b) in enumerate(zip(maps, ref["shadow"])):
        assert np.array_equal(a.view(np.uint32), b.view(np.uint32)), f"shadow map {i}"
    planes = pyoracle.bloom(ref["r"], ref["g"], ref["b"], levels=5, threshold=1.0,
                            white=9835.0, strength=1.0)
    want = pyoracle.gamma(*planes, ex)
    bad = np.argwhere(got[3].view(np.uint32) != ref["depth"].view(np.uint32))
    assert bad.size == 0, f"depth differs at {len(bad)} px"
    diff = np.abs(np.stack(got[:3], -1).astype(np.int64) -
                  np.stack(want, -1).astype(np.int64)).max(-1)
    frac = float((diff <= RGB_TOL).mean())
    water = float((ref["depth"] < 65535.0).mean())
    print(f"[parity] full frame 2560x1440: rgb<=1/255 on {frac:.6f}, exact "
          f"{float((diff == 0).mean()):.6f}, worst {int(diff.max())}, covered {water:.4f}",
          flush=True)
    assert frac >= RGB_MIN_FRAC
    for c in range(3):
        assert np.array_equal(again[c], got[c]), "a second frame differs"


def test_multiview_batch_at_full_size_matches_single_frames_and_oracle():
    from axi_visualizer_amd.batch import (MULTIVIEW_SIZES, Batch, SceneContext, autotune,
                                          multiview_jobs, settle)
    from axi_visualizer_amd.ops import CLDraw
    from axi_visualizer_amd.render import load_scene

    sc = scenes.spheres_scene(W=320, H=240, n_inst=1000, shadows=False, name="c4_spheres")
    assert sc.ntri == 500_000
    ctx = SceneContext(sc)
    draws = {}
    b = None
    try:
        jobs = multiview_jobs(ctx)
        assert len(jobs) == 64 and {(j.width, j.height) for j in jobs} == set(MULTIVIEW_SIZES)
        tune = autotune(jobs, sizes=(8, 16, 32), cap_scales=(1, 4), reps=2, graph=True)
        best = tune["best"]
        print(f"[batch] autotune best {best}", flush=True)
        b = Batch(best["bin_size"], best["cap_small"], best["cap_large"], graph=True)
        settle(b, jobs)
        b.render(jobs)  # a graph replay
        b.wait()
        rgba = [b.rgba8(j) for j in range(len(jobs))]
        depth = [b.depth(j) for j in range(len(jobs))]
        b.render(jobs, ids=True)
        b.wait()
        ids = [b.ids_of(j) for j in range(len(jobs))]
        for w, h in MULTIVIEW_SIZES:
            d = CLDraw(1, len(sc.groups), w, h, 1)
            load_scene(d, sc)
            draws[w, h] = d
        shaders = [g.flags for g in sc.groups]
        for k, j in enumerate(jobs):
            d = draws[j.width, j.height]
            d.setScaleCull(j.camera.scale, j.camera.cax, j.camera.cay)
            d.setPos(j.camera.pos)
            d.setVM(j.camera.vm)
            d.clearZBuffer()
            d.drawAll(shaders, shadowIds=[])
            d.gamma(1.0)
            single = np.ascontiguousarray(d.getRGBA8()).view(np.uint32).reshape(j.height, j.width)
            assert np.array_equal(single, rgba[k]), f"job {k} ({j.width}x{j.height}) rgba"
            assert np.array_equal(d.getIds(), ids[k]), f"job {k} ids"
            _, _, _, dd = d.getFrame()
            assert np.array_equal(dd.view(np.uint32), depth[k].view(np.uint32)), f"job {k} depth"
        for w, h in [(1920, 1080), (1280, 720)]:
            k = next(i for i, j in enumerate(jobs) if (j.width, j.height) == (w, h))
            ref = pyoracle.render(job_scene(sc, jobs[k]))
            compare_rgba(f"batch job {k} {w}x{h}", rgba[k], depth[k], ref, ids=ids[k])
    finally:
        if b is not None:
            b.close()
        for d in draws.values():
            d.close()
        ctx.close()


def test_serving_burst_at_full_size_matches_alone_and_oracle():
    from axi_visualizer_amd.batch import Batch, Job
    from axi_visualizer_amd.server import SERVING_SIZES, RenderServer, poisson_requests

    mix = scenes.serving_scenes()
    kinds = [k for _, k in mix]
    assert len(mix) == 16

    def cam(s, w, h, rng):
        return scenes.serving_camera(kinds[s], w, h, rng.random(4))

    srv = RenderServer([s for s, _ in mix], hbm_budget=200e9, slots=3, max_batch=64,
                       window_ms=2.0, readback=True)
    alone = Batch()
    try:
        srv.reserve(*SERVING_SIZES[-1])
        burst = poisson_requests(len(mix), 1e6, 160e-6, SERVING_SIZES, seed=5, camera_for=cam)
        for r in burst:
            r.arrival = 0.0
        rep = srv.run(burst)
        print(f"[serve] {rep}", flush=True)
        assert rep["completed"] == len(burst) >= 100 and rep["rejected"] == 0
        assert rep["mean_batch"] > 16
        picked = {}
        for r in srv.completed:
            alone.render([Job(srv.contexts[r.scene], r.camera, r.width, r.height,
                              shadows=r.shadows, pcf_taps=r.pcf_taps)])
            alone.wait()
            assert np.array_equal(alone.rgba8(0), r.rgba), f"request {r.rid} differs alone"
            picked.setdefault((r.width, r.height, r.shadows), r)
        assert len(picked) == 2 * len(SERVING_SIZES), sorted(picked)
        for (w, h, sh), r in sorted(picked.items()):
            job = Job(srv.contexts[r.scene], r.camera, w, h, shadows=sh, pcf_taps=r.pcf_taps)
            alone.render([job])
            alone.wait()
            depth = alone.depth(0)
            alone.render([job], ids=True)
            alone.wait()
            ids = alone.ids_of(0)
            ref = pyoracle.render(job_scene(srv.contexts[r.scene].scene, job))
            compare_rgba(f"served {kinds[r.scene]} scene {r.scene} {w}x{h} shadows={sh}",
                         r.rgba, depth, ref, ids=ids)
    finally:
        alone.close()
        srv.close()
