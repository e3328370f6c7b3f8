"""Post passes on the GPU against the oracle, bit-exact: CLDraw.blur
(Ops.py:1009-1020; Shaders/Post/blur.c, blur2.c) and the fp16 bloom, on
rendered frames (HDR and after gamma) and on cleared frames of edge sizes."""
import numpy as np
import pytest

from axi_visualizer_amd import scenes
from axi_visualizer_amd._native import AxrError
from axi_visualizer_amd.ops import CLDraw
from axi_visualizer_amd.render import Renderer
from oracle import pyoracle

pytestmark = pytest.mark.gpu


def spheres(W=320, H=256):
    return scenes.spheres_scene(W=W, H=H, n_inst=60, tex_side=64, shadow_dim=256,
                                cam_dist=150.0, name="post_spheres")


def same(got, want, what):
    for c, (p, q) in enumerate(zip(got, want)):
        bad = np.argwhere(p != q)
        assert bad.size == 0, f"{what} channel {c}: {len(bad)} px differ, first {bad[:4].tolist()}"


def test_blur_on_pending_hdr_frame():
    r = Renderer(spheres())
    try:
        r.frame(gamma=False)
        before = [p.copy() for p in r.draw.getFrame()[:3]]
        r.frame(gamma=False)  # same frame again, now blurred straight from the draw
        r.draw.blur()
        after = r.draw.getFrame()[:3]
        same(after, pyoracle.blur(*before), "blur")
        assert any((a != b).any() for a, b in zip(after, before))
    finally:
        r.close()


def test_blur_after_gamma_and_twice():
    r = Renderer(spheres())
    try:
        r.frame(gamma=True)
        p0 = [p.copy() for p in r.draw.getFrame()[:3]]
        r.draw.blur()
        p1 = [p.copy() for p in r.draw.getFrame()[:3]]
        same(p1, pyoracle.blur(*p0), "blur after gamma")
        r.draw.blur()
        same(r.draw.getFrame()[:3], pyoracle.blur(*p1), "second blur")
    finally:
        r.close()


@pytest.mark.parametrize("levels,threshold,strength", [(5, 1.0, 1.0), (3, 0.25, 0.5), (1, 0.0, 2.0)])
def test_bloom_on_rendered_frame(levels, threshold, strength):
    r = Renderer(spheres())
    try:
        r.frame(gamma=False)
        before = [p.copy() for p in r.draw.getFrame()[:3]]
        r.draw.bloom(levels=levels, threshold=threshold, white=9835.0, strength=strength)
        after = r.draw.getFrame()[:3]
        want = pyoracle.bloom(*before, levels=levels, threshold=threshold, white=9835.0,
                              strength=strength)
        same(after, want, f"bloom L{levels}")
    finally:
        r.close()


@pytest.mark.parametrize("W,H", [(16, 8), (16, 10), (48, 66)])
def test_blur_cleared_frame_edge_sizes(W, H):
    d = CLDraw(1, 1, W, H, 0)
    try:
        d.clearZBuffer()
        before = [p.copy() for p in d.getFrame()[:3]]
        d.clearZBuffer()
        d.blur()
        same(d.getFrame()[:3], pyoracle.blur(*before), f"blur {W}x{H}")
    finally:
        d.close()


def test_bloom_size_and_state_errors():
    d = CLDraw(1, 1, 48, 40, 0)
    try:
        with pytest.raises(AxrError):
            d.blur()  # nothing cleared yet
        d.clearZBuffer()
        with pytest.raises(AxrError):
            d.bloom(levels=4)  # 40 is not a multiple of 16
        with pytest.raises(AxrError):
            d.bloom(levels=3, white=0.0)
        d.bloom(levels=3, threshold=0.5)
        d.getFrame()
    finally:
        d.close()
    for W, H in ((16, 9), (16, 6)):  # odd height, fewer than 8 rows
        d = CLDraw(1, 1, W, H, 0)
        try:
            d.clearZBuffer()
            with pytest.raises(AxrError):
                d.blur()
        finally:
            d.close()
