"""The no-allocation bloom entry point (axr_bloom_planes_noalloc) against
the allocating one (axr_bloom) on the same HDR frames: bit-identical planes
on two calls in a row, the second on a smaller frame, and a clear error for
a frame beyond the scratch array."""
import ctypes

import numpy as np
import pytest

from axi_visualizer_amd import scenes
from axi_visualizer_amd._native import AxrError, check, lib
from axi_visualizer_amd.render import Renderer

pytestmark = pytest.mark.gpu

BLOOM = dict(levels=5, threshold=0.25, white=9835.0, strength=1.0)


def _hip():
    h = ctypes.CDLL("libamdhip64.so")
    h.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    h.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    h.hipFree.argtypes = [ctypes.c_void_p]
    return h


def frames(W, H):
    """An HDR frame of the terrain scene and its bloom by axr_bloom."""
    sc = scenes.terrain_scene(W=W, H=H, grid=48, tex_side=64, cam_dist=60.0)
    r = Renderer(sc)
    try:
        r.frame(gamma=False)
        before = [p.copy() for p in r.draw.getFrame()[:3]]
        r.draw.bloom(**BLOOM)
        after = [p.copy() for p in r.draw.getFrame()[:3]]
    finally:
        r.close()
    return before, after


def bloom_noalloc(planes, W, H):
    hip = _hip()
    n = W * H * 2
    bufs = []
    try:
        for p in planes:
            d = ctypes.c_void_p()
            assert hip.hipMalloc(ctypes.byref(d), n) == 0
            bufs.append(d)
            host = np.ascontiguousarray(p, np.uint16)
            assert hip.hipMemcpy(d, host.ctypes.data, n, 1) == 0  # host to device
        check(lib().axr_bloom_planes_noalloc(bufs[0], bufs[1], bufs[2], W, H, BLOOM["levels"],
                                             BLOOM["threshold"], BLOOM["white"],
                                             BLOOM["strength"], None),
              "axr_bloom_planes_noalloc")
        out = []
        for d in bufs:
            host = np.empty((H, W), np.uint16)
            assert hip.hipMemcpy(host.ctypes.data, d, n, 2) == 0  # device to host (syncs)
            out.append(host)
        return out
    finally:
        for d in bufs:
            hip.hipFree(d)


def test_noalloc_bloom_equals_bloom_twice_then_smaller():
    for W, H in [(640, 480), (320, 192)]:
        before, want = frames(W, H)
        got = bloom_noalloc(before, W, H)
        for c in range(3):
            bad = np.argwhere(got[c] != want[c])
            assert bad.size == 0, f"{W}x{H} channel {c}: {len(bad)} px differ"
        assert any(not np.array_equal(a, b) for a, b in zip(before, want))  # bloom did something
        print(f"[noalloc] {W}x{H}: bit-identical to axr_bloom", flush=True)


def test_noalloc_bloom_rejects_frames_beyond_its_scratch():
    with pytest.raises(AxrError, match="scratch"):
        check(lib().axr_bloom_planes_noalloc(ctypes.c_void_p(16), ctypes.c_void_p(16),
                                             ctypes.c_void_p(16), 8192, 4096, 1, 0.5, 1.0,
                                             1.0, None), "axr_bloom_planes_noalloc")
