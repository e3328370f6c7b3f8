"""SSR water (Shaders_src/av_SSR.c via drawAll's "SSR" shader entries,
Ops.py:788, 859-869, 910-924) and reflection cube maps (setReflTex,
Ops.py:458-467) on the GPU against the oracle: depth bit-exact, colour
bit-exact in HDR (the tolerance contract is reported beside it)."""
import numpy as np
import pytest

from axi_visualizer_amd import scenes
from axi_visualizer_amd._native import AxrError
from axi_visualizer_amd.render import Renderer
from oracle import pyoracle

pytestmark = pytest.mark.gpu


def gpu_frame(sc, gamma=False):
    r = Renderer(sc)
    try:
        r.frame(gamma=gamma)
        return r.draw.getFrame()
    finally:
        r.close()


def check(sc, label):
    sc.exposure = 0.0
    rr, gg, bb, d = gpu_frame(sc)
    ref = pyoracle.render(sc)
    dg, dr = d.view(np.uint32), ref["depth"].view(np.uint32)
    bad = np.argwhere(dg != dr)
    assert bad.size == 0, f"{label}: depth differs at {len(bad)} px, first {bad[:3].tolist()}"
    got = np.stack([rr, gg, bb], -1).astype(np.int64)
    want = np.stack([ref["r"], ref["g"], ref["b"]], -1).astype(np.int64)
    diff = np.abs(got - want).max(-1)
    water = (ref["depth"] != pyoracle.render(_no_water(sc))["depth"])
    print(f"[ssr] {label}: water px {int(water.sum())}, exact {float((diff == 0).mean()):.6f}, "
          f"worst {int(diff.max())}", flush=True)
    bad = np.argwhere(diff != 0)
    assert bad.size == 0, f"{label}: colour differs at {len(bad)} px, first {bad[:3].tolist()}"
    assert water.mean() > 0.1
    return water


def _no_water(sc):
    import copy
    s2 = copy.copy(sc)
    s2.groups = [copy.copy(g) for g in sc.groups]
    for g in s2.groups:
        if g.ssr >= 0:
            g.skip = True
    return s2


def test_water_matches_oracle():
    check(scenes.water_scene(), "water")


def test_water_with_shadows_and_64_steps():
    check(scenes.water_scene(W=256, H=192, shadow_dim=256, ssr_steps=64), "water_shadow_64")


@pytest.mark.parametrize("W,H", [(208, 117), (48, 37), (201, 117), (37, 29)])
def test_water_ragged_frame(W, H):
    """Frames whose 32-px bins are cut at the right and bottom edges (the
    march and the snapshot of the opaque frame run to the last pixel; a
    pixel count that is not a multiple of 8 takes the snapshot's tail)."""
    check(scenes.water_scene(W=W, H=H, ssr_steps=64), f"water_{W}x{H}")


def test_metal_variant_step_two():
    sc = scenes.water_scene(W=192, H=160, n_inst=10)
    sc.ssr_step, sc.ssr_steps, sc.ssr_absorb, sc.ssr_fresnel = 2, 800, 0.4, False
    check(sc, "metal")


def test_masked_water_is_not_drawn_and_refl_round_trip():
    sc = scenes.water_scene(W=160, H=128)
    sc.exposure = 0.0
    r = Renderer(sc)
    try:
        got = r.draw.getReflTex(0)
        for c in range(3):
            assert np.array_equal(got[c], sc.refl[0][..., c])
        r.mask = [False] * (len(sc.groups) - 1) + [True]
        r.frame(gamma=False)
        frame = r.draw.getFrame()
        ids = r.draw.getIds()  # no SSR draw ran: the opaque ids stay available
    finally:
        r.close()
    ref = pyoracle.render(_no_water(sc))
    assert np.array_equal(frame[3].view(np.uint32), ref["depth"].view(np.uint32))
    assert np.array_equal(ids, ref["id"])


def test_ssr_errors():
    sc = scenes.water_scene(W=160, H=128, n_inst=0)
    r = Renderer(sc)
    try:
        d = r.draw
        d.setPos(sc.view.pos)
        d.setVM(sc.view.vm)
        d.clearZBuffer()
        d.drawAll([g.flags for g in sc.groups[:-1]] + [{"SSR": "nope"}], shadowIds=[])
    except KeyError:
        pass
    else:
        raise AssertionError("unknown reflection texture accepted")
    finally:
        r.close()
    r = Renderer(sc)
    try:
        r.frame(gamma=True)
        with pytest.raises(AxrError):
            r.draw.getIds()  # the SSR draw replaced the bins
        ng = len(sc.groups)
        grp = np.array([ng - 1], np.int32)
        slot = np.array([0], np.int32)
        from axi_visualizer_amd._native import ptr
        rc = r.draw._L.axr_draw_ssr(r.draw._ctx, ptr(grp), ptr(slot), 1, 400, 1, 0.3, 1)
        assert rc < 0  # after gamma
        slot[0] = 5
        r.draw.clearZBuffer()
        rc = r.draw._L.axr_draw_ssr(r.draw._ctx, ptr(grp), ptr(slot), 1, 400, 1, 0.3, 1)
        assert rc < 0  # no such cube map
        grp[0] = 0  # the terrain's texture is a mip chain
        slot[0] = 0
        rc = r.draw._L.axr_draw_ssr(r.draw._ctx, ptr(grp), ptr(slot), 1, 400, 1, 0.3, 1)
        assert rc < 0
    finally:
        r.close()


def oracle_cube(sc, pos, size):
    """The cube pass restated: six oracle renders of the scene (water
    groups not drawn) from pos, laid side by side."""
    import copy
    strip = np.empty((size, 6 * size, 3), np.uint16)
    for f in range(6):
        s2 = _no_water(sc)
        s2.W = s2.H = size
        s2.view = scenes.cube_face_camera(f, pos, size)
        s2.exposure = 0.0
        o = pyoracle.render(s2)
        for c, k in enumerate("rgb"):
            strip[:, f * size:(f + 1) * size, c] = o[k]
    return strip


def _cube_probe(sc):
    ex = float(sc.groups[0].xyz[:, 0].max())
    return np.array([ex * 0.5, 12.0, ex * 0.5], np.float32)


def test_cube_pass_matches_oracle():
    from axi_visualizer_amd.batch import Batch
    sc = scenes.water_scene(W=160, H=128)
    pos, size = _cube_probe(sc), 64
    r = Renderer(sc)
    b = Batch()
    try:
        d = r.draw
        d.setShaders([g.flags for g in sc.groups], mask=[g.ssr >= 0 for g in sc.groups])
        b.render_cube(d, "probe", pos, size, shadows=False)
        got = d.getReflTex("probe")
    finally:
        b.close()
        r.close()
    want = oracle_cube(sc, pos, size)
    for c in range(3):
        bad = np.argwhere(got[c] != want[..., c])
        assert bad.size == 0, f"cube channel {c}: {len(bad)} texels differ, first {bad[:3].tolist()}"
    assert len(np.unique(want[..., 0])) > 50  # the faces see the scene, not only sky


def test_ssr_reflects_the_rendered_cube():
    from axi_visualizer_amd.batch import Batch
    sc = scenes.water_scene(W=192, H=144, ssr_steps=64)
    pos, size = _cube_probe(sc), 32
    r = Renderer(sc)
    b = Batch()
    try:
        d = r.draw
        d.setShaders([g.flags for g in sc.groups], mask=[g.ssr >= 0 for g in sc.groups])
        b.render_cube(d, 0, pos, size, shadows=False)
        r.frame(gamma=False)
        rr, gg, bb, dd = d.getFrame()
    finally:
        b.close()
        r.close()
    sc.refl = [oracle_cube(sc, pos, size)]
    sc.exposure = 0.0
    ref = pyoracle.render(sc)
    assert np.array_equal(dd.view(np.uint32), ref["depth"].view(np.uint32))
    for c, p in zip("rgb", (rr, gg, bb)):
        bad = np.argwhere(p != ref[c])
        assert bad.size == 0, f"{c}: {len(bad)} px differ, first {bad[:3].tolist()}"


def test_full_frame_pipeline_matches_oracle_composition():
    """Shadows, cube pass, opaque + SSR drawAll, bloom, gamma at a small
    size: every stage restated by the oracle, compared bit-exactly."""
    from axi_visualizer_amd.render import FullFrame
    sc = scenes.full_frame_scene(W=320, H=192, grid=64, n_inst=40, tex_side=64, terrain_tex=64,
                                 shadow_dim=256, refl_side=32)
    probe = scenes.full_frame_probe(sc)
    ff = FullFrame(sc, probe, cube_size=32)
    try:
        ff.frame(gamma=True)
        got = [p.copy() for p in ff.draw.getFrame()]
        ff.frame(gamma=True)  # a second frame: same pixels
        again = ff.draw.getFrame()
    finally:
        ff.close()
    cube = oracle_cube(sc, probe, 32)
    sc.refl = [cube]
    ex = sc.exposure
    sc.exposure = 0.0
    ref = pyoracle.render(sc)
    planes = pyoracle.bloom(ref["r"], ref["g"], ref["b"], levels=5, threshold=1.0,
                            white=9835.0, strength=1.0)
    want = pyoracle.gamma(*planes, ex)
    assert np.array_equal(got[3].view(np.uint32), ref["depth"].view(np.uint32))
    for c in range(3):
        bad = np.argwhere(got[c] != want[c])
        assert bad.size == 0, f"channel {c}: {len(bad)} px differ, first {bad[:3].tolist()}"
        assert np.array_equal(again[c], got[c])
