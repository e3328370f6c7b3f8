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
from tests.edge_scenes import SLIVER_TRIS, hostile_scene, sliver_scene
from tests.parity import check_scene

pytestmark = pytest.mark.gpu

NONE = 0xFFFFFFFF


@pytest.mark.parametrize("W,H", [(208, 143), (48, 17), (16, 1), (200, 143), (1000, 61),
                                 (24, 160), (37, 29)])
def test_ragged_frame_sizes(W, H):
    """Heights and widths that are not multiples of the bin or tile size; at
    widths that are not multiples of 16 the tiled pass walks the reference's
    folded tile boxes (Templates.py:23-24)."""
    sc = scenes.random_tris_scene(W=W, H=H, ntri=900, seed=13, name=f"ragged_{W}x{H}")
    check_scene(sc)


@pytest.mark.parametrize("W,H", [(200, 80), (24, 160), (192, 80)])
def test_folded_tiles_of_a_ragged_width(W, H):
    """A rectangle of large triangles over the whole view: at W = 200 the
    tiles the reference's tiled draw never walks stay undrawn, bit-exact with
    the oracle (whose tile pattern test_oracle pins by hand)."""
    from tests.test_oracle import rect_scene
    check_scene(rect_scene(W, H), hdr=True)  # the scene has no gamma (exposure 0)


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


def test_coarse_slivers():
    """Large triangles two screen-y steps tall and 14,000 px wide in an
    8192 x 48 frame: their coarse rows include ones whose first tile column
    saturates to INT_MAX (checked on the CPU by test_coarse_host); the
    coarse launch must bin exactly the oracle's tiles and never outside the
    grid."""
    sc = sliver_scene()
    rep = check_scene(sc)
    from tests.parity import render_gpu
    ids = render_gpu(sc)["id"]
    for k in range(len(SLIVER_TRIS)):
        print(f"[parity] sliver {k}: {int((ids == k).sum())} px", flush=True)
    assert rep["covered"] > 0.02  # the fillers; the slivers cover no pixel centre


def test_nothing_in_view():
    """The terrain camera turned to the sky: every triangle is culled, every
    pixel gets the clear colour and depth."""
    sc = scenes.terrain_scene(W=320, H=240, grid=32, tex_side=64, name="sky_only")
    sc = dataclasses.replace(sc, view=scenes.make_camera(320, 240, sc.view.pos, 0.0, -1.4))
    rep = check_scene(sc)
    assert rep["covered"] == 0.0
