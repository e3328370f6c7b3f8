"""Render server on the GPU: continuous batching over resident scenes, each
result identical alone and in the mix, admission within the HBM budget."""
import numpy as np
import pytest

from axi_visualizer_amd import scenes
from axi_visualizer_amd.batch import Batch, Job
from axi_visualizer_amd.server import RenderServer, poisson_requests, request_bytes

pytestmark = pytest.mark.gpu

SIZES = [(320, 240), (256, 160), (128, 96)]


def small_scenes():
    sp = scenes.spheres_scene(W=320, H=240, n_inst=120, n_tex=2, tex_side=64, shadow_dim=256,
                              name="serve_spheres")
    te = scenes.terrain_scene(W=320, H=240, grid=40, tex_side=64, n_tex=2, shadow_dim=256,
                              name="serve_terrain")
    return [(sp, "spheres"), (te, "terrain")]


def requests(kinds, n_rate=4000.0, duration=0.02, seed=3):
    return poisson_requests(len(kinds), n_rate, duration, SIZES, seed=seed,
                            camera_for=lambda s, w, h, rng: scenes.serving_camera(
                                kinds[s], w, h, rng.random(4)))


def test_each_result_is_the_same_alone_and_in_the_mix():
    sc = small_scenes()
    kinds = [k for _, k in sc]
    srv = RenderServer([s for s, _ in sc], slots=3, max_batch=16, window_ms=1.0,
                       readback=True)
    try:
        srv.reserve(*SIZES[0])   # pools grown before serving, as the benchmark does
        reqs = requests(kinds)
        assert len(reqs) >= 40
        rep = srv.run(reqs)
        assert rep["completed"] == len(reqs) and rep["rejected"] == 0
        assert rep["mean_batch"] > 1.0, rep           # requests were batched together
        alone = Batch()
        try:
            for r in srv.completed:
                alone.render([Job(srv.contexts[r.scene], r.camera, r.width, r.height,
                                  shadows=r.shadows, pcf_taps=r.pcf_taps)])
                alone.wait()
                assert np.array_equal(alone.rgba8(0), r.rgba), f"request {r.rid}"
        finally:
            alone.close()
        print(f"[server] {rep}", flush=True)
    finally:
        srv.close()


def test_admission_keeps_in_flight_bytes_within_the_budget():
    sc = small_scenes()
    kinds = [k for _, k in sc]
    probe = RenderServer([s for s, _ in sc], slots=1)
    resident = probe.resident
    probe.close()
    big = max(request_bytes(s.ntri, w, h) for s, _ in sc for w, h in SIZES)
    budget = resident + 2.5 * big
    srv = RenderServer([s for s, _ in sc], hbm_budget=budget, slots=3, max_batch=16,
                       window_ms=1.0)
    try:
        reqs = requests(kinds, seed=8)
        rep = srv.run(reqs)
        assert rep["completed"] == len(reqs)
        assert 0 < srv.max_in_flight <= budget - resident
        # every batch's estimated bytes fit what the budget left
        assert all(r.batch_size * min(request_bytes(s.ntri, w, h) for s, _ in sc
                                      for w, h in SIZES) <= budget - resident
                   for r in srv.completed)
    finally:
        srv.close()
