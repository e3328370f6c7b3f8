"""HIP path vs oracle: winner ids and depth bit-exact, RGB within 1/255 on
>= 99.9% of pixels (worst error printed), shadow maps bit-exact."""
import pytest

from axi_visualizer_amd import scenes
from tests.parity import check_scene

pytestmark = pytest.mark.gpu


def test_smoke_terrain_small():
    sc = scenes.terrain_scene(W=320, H=240, grid=48, tex_side=64, cam_dist=60.0, name="smoke")
    check_scene(sc)


def test_config1_terrain_full():
    """Config 1: 128x128 terrain (32,258 tris), 640x480, trilinear + Blinn-Phong."""
    sc = scenes.terrain_scene()
    assert sc.ntri == 32258
    check_scene(sc)


def test_terrain_closeup_large_triangles():
    """Camera close to a coarse grid: large triangles take the tiled path."""
    sc = scenes.terrain_scene(W=640, H=480, grid=24, cam_dist=18.0, name="closeup")
    rep = check_scene(sc)
    assert rep["covered"] > 0.5


def test_random_tris_mixed_paths():
    sc = scenes.random_tris_scene(W=512, H=384, ntri=2000, name="random_mixed")
    check_scene(sc)


@pytest.mark.parametrize("flags", [0, scenes.SH_PHONG, scenes.SH_CULL | scenes.SH_MIP])
def test_random_tris_shader_variants(flags):
    sc = scenes.random_tris_scene(W=320, H=240, ntri=800, seed=11, flags=flags,
                                  mip=bool(flags & scenes.SH_MIP), name=f"random_f{flags}")
    check_scene(sc)


def test_hdr_frame_without_gamma():
    sc = scenes.random_tris_scene(W=256, H=192, ntri=500, seed=5, name="random_hdr")
    check_scene(sc, hdr=True)


def test_spheres_shadows_small():
    sc = scenes.spheres_scene(W=640, H=480, n_inst=200, tex_side=128, shadow_dim=512,
                              cam_dist=200.0, name="spheres_small")
    check_scene(sc)


def test_config2_spheres_full():
    """Config 2: 1,000,000 tris, 1920x1080, 2 x 2048^2 cascades, 3x3 PCF."""
    sc = scenes.spheres_scene()
    assert sc.ntri == 1_000_000
    check_scene(sc)
