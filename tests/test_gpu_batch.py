"""Batched rendering on the GPU: every job of a batch is bit-identical to the
same view drawn alone through CLDraw (itself parity-checked against the
oracle), whatever the bin size, the rest of the batch, or graph replay."""
import dataclasses

import numpy as np
import pytest

from axi_visualizer_amd import scenes
from axi_visualizer_amd._native import AxrError
from axi_visualizer_amd.batch import Batch, Job, SceneContext, autotune, job_camera
from axi_visualizer_amd.render import Renderer
from oracle import pyoracle
from tests.parity import RGB_MIN_FRAC, RGB_TOL, job_scene

pytestmark = pytest.mark.gpu


def base_spheres():
    return scenes.spheres_scene(W=320, H=240, n_inst=200, tex_side=64, shadow_dim=256,
                                cam_dist=150.0, name="batch_spheres")


def base_terrain(cam_dist=40.0, grid=32):
    return scenes.terrain_scene(W=320, H=240, grid=grid, tex_side=64, cam_dist=cam_dist,
                                name="batch_terrain")


def scene_view(base, w, h):
    """The base scene's own camera (yaw pi/4, pitch pi/6) at resolution w x h."""
    return scenes.make_camera(w, h, base.view.pos, np.pi / 4, np.pi / 6)


view_scene = job_scene


def single(view):
    r = Renderer(view)
    try:
        r.frame()
        rgba = r.draw.getRGBA8().copy()
        ids = r.draw.getIds().copy()
        _, _, _, depth = r.draw.getFrame()
        return {"rgba": rgba, "ids": ids, "depth": depth.copy()}
    finally:
        r.close()


def make_jobs(ctx, sizes, seed=5, shadows=(True, False), pcf=(9, 5)):
    rng = np.random.default_rng(seed)
    jobs = []
    for k, (w, h) in enumerate(sizes):
        ang = rng.uniform(0, 2 * np.pi)
        dist = rng.uniform(110.0, 170.0)
        pos = np.array([dist * np.cos(ang), rng.uniform(20, 90), dist * np.sin(ang)])
        target = rng.uniform(-20, 20, 3)
        jobs.append(Job(ctx, job_camera(w, h, pos, target), w, h,
                        shadows=shadows[k % len(shadows)], pcf_taps=pcf[k % len(pcf)]))
    return jobs


def rgba8_of(batch, j):
    v = batch.rgba8(j)
    return np.stack([(v >> s) & 0xFF for s in (0, 8, 16)], -1).astype(np.uint8)


@pytest.fixture(scope="module")
def spheres_ctx():
    ctx = SceneContext(base_spheres())
    yield ctx
    ctx.close()


SIZES = [(320, 240), (128, 96), (256, 160), (64, 48), (320, 176)]


def test_batch_jobs_match_single_frames(spheres_ctx):
    jobs = make_jobs(spheres_ctx, SIZES)
    b = Batch()
    try:
        b.render(jobs, ids=True)
        b.wait()
        got_ids = [b.ids_of(j) for j in range(len(jobs))]
        got_depth = [b.depth(j) for j in range(len(jobs))]
        b.render(jobs)
        b.wait()
        for j, job in enumerate(jobs):
            ref = single(view_scene(spheres_ctx.scene, job))
            assert np.array_equal(got_ids[j], ref["ids"]), f"job {j}: ids differ"
            assert np.array_equal(got_depth[j].view(np.uint32), ref["depth"].view(np.uint32)), \
                f"job {j}: depth differs"
            assert np.array_equal(rgba8_of(b, j), ref["rgba"][..., :3]), f"job {j}: colour differs"
            assert np.array_equal(b.depth(j).view(np.uint32), ref["depth"].view(np.uint32))
    finally:
        b.close()


def test_batch_job_matches_oracle(spheres_ctx):
    job = make_jobs(spheres_ctx, [(320, 240)], seed=9, shadows=(True,), pcf=(9,))[0]
    b = Batch()
    try:
        b.render([job], ids=True)
        b.wait()
        ids, depth = b.ids_of(0), b.depth(0)
        b.render([job])
        b.wait()
        rgb = rgba8_of(b, 0).astype(np.int64)
    finally:
        b.close()
    ref = pyoracle.render(view_scene(spheres_ctx.scene, job))
    assert np.array_equal(ids, ref["id"])
    assert np.array_equal(depth.view(np.uint32), ref["depth"].view(np.uint32))
    diff = np.abs(rgb - np.stack([ref["r"], ref["g"], ref["b"]], -1).astype(np.int64)).max(-1)
    assert (diff <= RGB_TOL).mean() >= RGB_MIN_FRAC, f"worst {diff.max()}"


@pytest.mark.parametrize("bin_size", [8, 16])
def test_bin_size_does_not_change_pixels(spheres_ctx, bin_size):
    jobs = make_jobs(spheres_ctx, SIZES, seed=7)
    ref_b, b = Batch(bin_size=32), Batch(bin_size=bin_size)
    try:
        ref_b.render(jobs)
        b.render(jobs)
        ref_b.wait()
        b.wait()
        for j in range(len(jobs)):
            assert np.array_equal(b.rgba8(j), ref_b.rgba8(j)), f"job {j}"
            assert np.array_equal(b.depth(j).view(np.uint32), ref_b.depth(j).view(np.uint32))
    finally:
        ref_b.close()
        b.close()


def test_large_triangles_in_small_bins_match_single_frames():
    """Close-up terrain: large triangles take the 16-px tile path, which an
    8-px bin clips and a 32-px bin holds four of; at widths that are not
    multiples of 16 the tiles fold as the reference's tiled draw folds them."""
    ctx = SceneContext(base_terrain(cam_dist=18.0, grid=24))
    try:
        jobs = [Job(ctx, scene_view(ctx.scene, w, h), w, h)
                for w, h in [(320, 240), (160, 112), (200, 117), (37, 29)]]
        for size in (8, 16, 32):
            b = Batch(bin_size=size)
            try:
                b.render(jobs, ids=True)
                b.wait()
                for j, job in enumerate(jobs):
                    ref = single(view_scene(ctx.scene, job))
                    assert np.array_equal(b.ids_of(j), ref["ids"]), f"bin {size} job {j}"
                    assert (ref["ids"] != 0xFFFFFFFF).mean() > 0.3
            finally:
                b.close()
    finally:
        ctx.close()


def test_job_same_alone_and_in_the_mix(spheres_ctx):
    terrain = SceneContext(base_terrain())
    try:
        target = make_jobs(spheres_ctx, [(256, 160)], seed=21)[0]
        others = make_jobs(spheres_ctx, SIZES, seed=3)
        mix = others[:2] + [Job(terrain, scene_view(terrain.scene, 320, 240), 320, 240)] + \
            [target] + others[2:]
        alone, mixed = Batch(), Batch(bin_size=16)
        try:
            alone.render([target])
            mixed.render(mix)
            alone.wait()
            mixed.wait()
            k = 3  # position of target in mix
            assert np.array_equal(alone.rgba8(0), mixed.rgba8(k))
            assert np.array_equal(alone.depth(0).view(np.uint32), mixed.depth(k).view(np.uint32))
        finally:
            alone.close()
            mixed.close()
    finally:
        terrain.close()


def test_graph_replays_match_eager(spheres_ctx):
    """Graph replay with new cameras each time, eager work in between; every
    replay is checked against an eager render of the same jobs."""
    g, e = Batch(graph=True), Batch(graph=False)
    try:
        for rep in range(3):
            jobs = make_jobs(spheres_ctx, SIZES, seed=100 + rep)
            g.render(jobs)
            g.wait()
            e.render(jobs)
            e.wait()
            for j in range(len(jobs)):
                assert np.array_equal(g.rgba8(j), e.rgba8(j)), f"replay {rep} job {j}"
    finally:
        g.close()
        e.close()


def test_autotune_returns_a_measured_best(spheres_ctx):
    jobs = make_jobs(spheres_ctx, SIZES, seed=4)
    res = autotune(jobs, sizes=(16, 32), cap_scales=(1,), reps=2)
    assert res["best"]["bin_size"] in (16, 32)
    assert all(r["gpu_ms"] > 0 for r in res["all"])


def test_spilling_batch_matches_roomy_batch(spheres_ctx):
    """Bucket capacities far below the bins' entry counts: the overflow goes
    to the spill list and is still drawn; the wait reads the spill counter
    from the render's own copy, grows the buckets and clears the counters,
    and the next render (no spill) draws the same pixels."""
    jobs = make_jobs(spheres_ctx, SIZES, seed=7)
    tiny, roomy = Batch(cap_small=64, cap_large=64), Batch()
    try:
        roomy.render(jobs)
        roomy.wait()
        held = []
        for rep in range(3):
            tiny.render(jobs)
            tiny.wait()
            held.append(tiny.device_bytes())
            for j in range(len(jobs)):
                assert np.array_equal(tiny.rgba8(j), roomy.rgba8(j)), f"render {rep} job {j}"
                assert np.array_equal(tiny.depth(j).view(np.uint32),
                                      roomy.depth(j).view(np.uint32)), f"render {rep} job {j}"
        assert held[1] > held[0], held  # the first render spilled: its buckets grew
    finally:
        tiny.close()
        roomy.close()


def test_lost_entries_fail_every_readback_until_the_next_render(spheres_ctx):
    """Tiny buckets and a tiny spill list: the render loses bin entries.
    The wait reports it (AXR_ECAP) and grows the list; every later wait and
    readback of that render fails too; the next render is whole."""
    jobs = make_jobs(spheres_ctx, SIZES, seed=8)
    tiny, roomy = Batch(cap_small=4, cap_large=4), Batch()
    try:
        roomy.render(jobs)
        roomy.wait()
        tiny.set_spill_capacity(16)
        tiny.render(jobs)
        with pytest.raises(AxrError, match="spill list"):
            tiny.wait()
        with pytest.raises(AxrError, match="render the batch again"):
            tiny.rgba8(0)
        with pytest.raises(AxrError, match="render the batch again"):
            tiny.wait()
        for _ in range(8):  # the list doubles per loss
            tiny.render(jobs)
            try:
                tiny.wait()
                break
            except AxrError as e:
                assert "spill list" in str(e)
        for j in range(len(jobs)):
            assert np.array_equal(tiny.rgba8(j), roomy.rgba8(j)), f"job {j}"
    finally:
        tiny.close()
        roomy.close()
