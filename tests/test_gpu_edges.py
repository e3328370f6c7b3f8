"""HIP path vs oracle on edge cases of the inputs: frame sizes that are not
multiples of the bin size, hostile geometry (NaN and infinite vertices,
zero-area and single-point triangles, triangles crossing or behind the eye
plane, enormous, distant and sub-pixel ones), and a view with nothing in
it.  The same contract as the other parity tests: winner ids and depth
bit-exact, RGB within 1/255 on >= 99.9% of pixels."""
import dataclasses

import numpy as np
import pytest

from axi_visualizer_amd import scenes
from tests.parity import check_scene

pytestmark = pytest.mark.gpu

NONE = 0xFFFFFFFF


def hostile_scene(W=256, H=192):
    """Random triangles whose first ten are replaced by hostile ones."""
    sc = scenes.random_tris_scene(W=W, H=H, ntri=600, seed=9, name="hostile")
    g = sc.groups[0]
    xyz = g.xyz.reshape(-1, 3, 3).copy()
    bad = []
    t = xyz[0].copy(); t[1, 0] = np.nan; bad.append(t)            # NaN vertex
    t = xyz[1].copy(); t[2, 2] = np.inf; bad.append(t)            # +inf depth
    t = xyz[2].copy(); t[0, 1] = -np.inf; bad.append(t)           # -inf coordinate
    t = xyz[3].copy(); t[2] = 0.5 * (t[0] + t[1]); bad.append(t)  # collinear
    t = xyz[4].copy(); t[1] = t[0]; t[2] = t[0]; bad.append(t)    # one point
    t = xyz[5].copy(); t[0, 2] = -3.0; bad.append(t)              # crosses the eye plane
    t = xyz[6].copy(); t[:, 2] = -10.0; bad.append(t)             # behind the eye
    t = xyz[7].copy() * 1e6; t[:, 2] = np.abs(t[:, 2]) + 1.0; bad.append(t)  # enormous
    t = xyz[8].copy(); t[:, 2] = 1e7; bad.append(t)               # far away
    t = xyz[9].copy(); c = t.mean(0); t = c + (t - c) * 1e-4; bad.append(t)  # sub-pixel
    xyz[:len(bad)] = np.array(bad, np.float32)
    groups = [dataclasses.replace(g, xyz=xyz.reshape(-1, 3).astype(np.float32))] + sc.groups[1:]
    return dataclasses.replace(sc, groups=groups), len(bad)


@pytest.mark.parametrize("W,H", [(208, 143), (48, 17), (16, 1)])
def test_ragged_frame_sizes(W, H):
    sc = scenes.random_tris_scene(W=W, H=H, ntri=900, seed=13, name=f"ragged_{W}x{H}")
    check_scene(sc)


def test_hostile_geometry():
    sc, nbad = hostile_scene()
    rep = check_scene(sc)
    assert rep["covered"] > 0.5
    # NaN, infinite and behind-the-eye triangles never win a pixel (both paths)
    from tests.parity import render_gpu
    ids = render_gpu(sc)["id"]
    for k in (0, 1, 2, 6):
        assert not np.any(ids == k), f"hostile triangle {k} drew pixels"


def test_hostile_geometry_with_shadows():
    sc, _ = hostile_scene(W=192, H=144)
    sp = scenes.spheres_scene(W=192, H=144, n_inst=20, tex_side=32, shadow_dim=256,
                              cam_dist=60.0, name="hostile_shadow_src")
    sc = dataclasses.replace(sc, cascades=sp.cascades, pcf_taps=9,
                             groups=[dataclasses.replace(g, flags=g.flags | scenes.SH_SHADOW)
                                     for g in sc.groups], name="hostile_shadows")
    check_scene(sc)


def test_nothing_in_view():
    """The terrain camera turned to the sky: every triangle is culled, every
    pixel gets the clear colour and depth."""
    sc = scenes.terrain_scene(W=320, H=240, grid=32, tex_side=64, name="sky_only")
    sc = dataclasses.replace(sc, view=scenes.make_camera(320, 240, sc.view.pos, 0.0, -1.4))
    rep = check_scene(sc)
    assert rep["covered"] == 0.0
