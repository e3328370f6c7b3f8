"""Several drawAll calls per clearZBuffer (Ops.py:716-946: drawAll draws
into the current buffers; only clearZBuffer resets them): every later draw
tests its fragments against the depth the earlier ones left.  Checked
bit-exactly (winner ids, depth, HDR planes) against the oracle rendering
each draw onto the previous one's output."""
import dataclasses

import numpy as np
import pytest

from axi_visualizer_amd import scenes
from axi_visualizer_amd.render import Renderer
from oracle import pyoracle

pytestmark = pytest.mark.gpu


def masked(sc, draw, view=None):
    """The scene as one drawAll sees it: groups outside `draw` skipped."""
    groups = [dataclasses.replace(g, skip=k not in draw) for k, g in enumerate(sc.groups)]
    return dataclasses.replace(sc, groups=groups, view=view or sc.view, exposure=0.0)


def compose_case(sc, draws, views=None, between=None, label=""):
    """draws: group sets of consecutive drawAll calls after one clear;
    between(d, k): called after draw k on the CLDraw; returns 'gamma' when
    it applied gamma (the oracle then gammas its planes in place)."""
    views = views or [sc.view] * len(draws)
    r = Renderer(sc)
    d = r.draw
    try:
        for i, c in enumerate(sc.cascades):
            d.clearShadowMap(i)
            d.shadowMap(i, r.cast, r.shaders, c.bias)
        d.clearZBuffer()
        gammas = []
        for k, draw in enumerate(draws):
            d.setPos(views[k].pos)
            d.setVM(views[k].vm)
            d.drawAll(r.shaders, mask=[i not in draw for i in range(len(sc.groups))],
                      shadowIds=list(range(len(sc.cascades))), pcf=sc.pcf_taps)
            gammas.append(between(d, k) if between else None)
        ids = d.getIds().copy()
        rr, gg, bb, depth = (x.copy() for x in d.getFrame())
    finally:
        r.close()
    ref = None
    for k, draw in enumerate(draws):
        ref = pyoracle.render(masked(sc, draw, views[k]), onto=ref)
        if gammas[k] == "gamma":
            ref["r"], ref["g"], ref["b"] = pyoracle.gamma(ref["r"], ref["g"], ref["b"],
                                                          sc.exposure)
    for name, got, want in (("ids", ids, ref["id"]), ("depth", depth.view(np.uint32),
                                                        ref["depth"].view(np.uint32)),
                            ("r", rr, ref["r"]), ("g", gg, ref["g"]), ("b", bb, ref["b"])):
        bad = np.argwhere(got != want)
        assert bad.size == 0, f"{label}: {name} differs at {len(bad)} px, first {bad[:3].tolist()}"
    covered = float((ref["id"] != 0xFFFFFFFF).mean())
    print(f"[parity] compose {label}: bit-exact, covered {covered:.3f}", flush=True)
    return ref


@pytest.fixture(scope="module")
def spheres():
    return scenes.spheres_scene(W=320, H=240, n_inst=300, n_tex=4, tex_side=64, shadow_dim=256,
                                cam_dist=40.0, name="compose_spheres")


def test_second_draw_onto_the_first(spheres):
    ref = compose_case(spheres, [{0, 1}, {2, 3}], label="two draws")
    ids = ref["id"][ref["id"] != 0xFFFFFFFF]
    assert len(np.unique(ids)) > 1000


def test_three_draws_with_the_same_groups_again(spheres):
    compose_case(spheres, [{0}, {1, 2}, {0, 3}], label="three draws")


def test_camera_moves_between_draws(spheres):
    v2 = scenes.make_camera(spheres.W, spheres.H, np.array([30.0, 20.0, -80.0]), 0.1, 0.2)
    compose_case(spheres, [{0, 1, 2, 3}, {1, 3}], views=[spheres.view, v2], label="moved camera")


def test_readback_and_gamma_between_draws(spheres):
    def between(d, k):
        if k == 0:
            d.getFrame()  # the first draw shaded to HDR before the second
            return None
        if k == 1:
            d.gamma(spheres.exposure)  # planes gamma'd in place, then drawn onto
            return "gamma"
        return None
    compose_case(spheres, [{0, 2}, {1}, {3}], between=between, label="readback and gamma")


def test_large_triangles_composed():
    """Close-up terrain (tiled path) drawn in two halves of its groups."""
    sc = scenes.terrain_scene(W=256, H=192, grid=12, n_tex=2, cam_dist=14.0, tex_side=64,
                              name="compose_terrain")
    compose_case(sc, [{0}, {1}], label="terrain halves")
