"""The setup's per-chunk culling tests (axr_tri.h: chunk_outside_view,
chunk_outside_ortho) and their camera constants (axr_common.h: make_cull),
built for the host and checked on the CPU.

A chunk is skipped only when no point of its bounding sphere can pass the
setup's accept test: trisetup's (Pipe/trisetup.c:45-47: positive depth and
|dx/dd| < caX, |dy/dd| < caY) for the view, trisetupOrtho's
(Pipe/trisetupOrtho.c:49-57: positive biased depth and pixel coordinates
within the 120-pixel guard band) for a shadow camera.  Every sphere the
tests skip must be separated from the accepted region, in float64 with the
exact thresholds, by at least half the margin the kernels keep (1e-4 of the
coordinate magnitudes), so the float32 evaluation cannot skip a chunk the
setup would have drawn.  Spheres well inside the view must be kept, and a
good share of the random ones must be skipped (the tests are not vacuous)."""
import ctypes
import os
import shutil
import subprocess

import numpy as np
import pytest

from axi_visualizer_amd import scenes

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
MARGIN = 0.5e-4  # half the kernels' relative margin


@pytest.fixture(scope="module")
def host(tmp_path_factory):
    if not os.path.exists(HIPCC) and not shutil.which("hipcc"):
        pytest.skip("hipcc not available")
    out = str(tmp_path_factory.mktemp("cull") / "libcull_host.so")
    cmd = [HIPCC if os.path.exists(HIPCC) else "hipcc", "-x", "hip", "--cuda-host-only", "-O2",
           "-std=c++17", "-fPIC", "-shared", "-ffp-contract=off",
           "-I", os.path.join(ROOT, "axi_visualizer_amd", "csrc"), "-I", os.path.join(ROOT, "include"),
           os.path.join(HERE, "cull_host.cpp"), "-o", out]
    res = subprocess.run(cmd, capture_output=True, text=True)
    assert res.returncode == 0, res.stderr
    L = ctypes.CDLL(out)
    p, i, f = ctypes.c_void_p, ctypes.c_int, ctypes.c_float
    L.host_outside_view.argtypes = [p, i, p, p, f, f, f, p]
    L.host_outside_ortho.argtypes = [p, i, p, p, f, i, f, p]
    return L


def _f32(a):
    return np.ascontiguousarray(a, np.float32)


def outside_view(L, sph, cam):
    sph = _f32(sph)
    out = np.empty(len(sph), np.int32)
    L.host_outside_view(sph.ctypes.data, len(sph), _f32(cam.pos).ctypes.data,
                        _f32(cam.vm).ctypes.data, cam.scale, cam.cax, cam.cay, out.ctypes.data)
    return out.astype(bool)


def outside_ortho(L, sph, cas):
    sph = _f32(sph)
    out = np.empty(len(sph), np.int32)
    c = cas.cam
    L.host_outside_ortho(sph.ctypes.data, len(sph), _f32(c.pos).ctypes.data,
                         _f32(c.vm).ctypes.data, c.scale, cas.dim, cas.bias, out.ctypes.data)
    return out.astype(bool)


def _mag(c, eye, r):
    x = c - eye
    return (np.abs(c).sum(1) + np.abs(eye).sum() + np.abs(x).sum(1) + 2 * r)


def view_separation(sph, cam):
    """For each sphere, how far (in the magnitude's units) it lies beyond the
    best rejecting plane of trisetup's accept region, float64 and exact
    thresholds; positive means every point is rejected."""
    s = sph.astype(np.float64)
    c, r = s[:, :3], s[:, 3]
    eye = np.asarray(cam.pos, np.float64)
    vm = np.asarray(cam.vm, np.float64).reshape(3, 3)
    F, R, U = vm[0], vm[1], vm[2]
    x = c - eye
    seps = [-(x @ F + r * np.linalg.norm(F))]  # behind: max depth <= 0
    for n in (R - cam.cax * F, -R - cam.cax * F, U - cam.cay * F, -U - cam.cay * F):
        seps.append(x @ n - r * np.linalg.norm(n))  # beyond a side plane
    return np.max(seps, axis=0) / _mag(c, eye, r)


def ortho_separation(sph, cas):
    s = sph.astype(np.float64)
    c, r = s[:, :3], s[:, 3]
    cam = cas.cam
    eye = np.asarray(cam.pos, np.float64)
    vm = np.asarray(cam.vm, np.float64).reshape(3, 3)
    F, R, U = vm[0], vm[1], vm[2]
    x = c - eye
    sc, half, dim = abs(cam.scale), float(cas.dim // 2), cas.dim
    mag = _mag(c, eye, r)
    seps = [-(x @ F + r * np.linalg.norm(F) + cas.bias) / mag]
    # projected pixel ranges; accepted only inside (-121, dim + 120)
    px = x @ R * cam.scale + half
    py = x @ U * -cam.scale + half
    rx, ry = r * np.linalg.norm(R) * sc, r * np.linalg.norm(U) * sc
    pmag = mag * sc + dim  # pixels
    seps += [(-121 - (px + rx)) / pmag, ((px - rx) - (dim + 120)) / pmag,
             (-121 - (py + ry)) / pmag, ((py - ry) - (dim + 120)) / pmag]
    return np.max(seps, axis=0)


def _spheres(rng, n, centre, spread, rmax):
    c = centre + rng.uniform(-spread, spread, (n, 3))
    r = rng.uniform(0, rmax, n) ** 2 / rmax
    return np.concatenate([c, r[:, None]], 1).astype(np.float32)


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_view_culling_is_conservative(host, seed):
    rng = np.random.default_rng(seed)
    for k in range(6):
        a, b = rng.uniform(-np.pi, np.pi), rng.uniform(-1.2, 1.2)
        W, H = [(1920, 1080), (640, 480), (320, 240)][k % 3]
        eye = rng.uniform(-300, 300, 3)
        cam = scenes.make_camera(W, H, eye, a, b, fovx=float(rng.uniform(40, 110)))
        sph = _spheres(rng, 4000, eye, 400.0, 60.0)
        out = outside_view(host, sph, cam)
        sep = view_separation(sph, cam)
        bad = out & (sep < MARGIN)
        assert not bad.any(), f"skipped spheres the setup could draw: {sph[bad][:3]} {sep[bad][:3]}"
        assert 0.2 < out.mean() < 0.98, out.mean()  # not vacuous
        # spheres well inside the view cone are kept
        fwd = np.asarray(cam.vm, np.float64).reshape(3, 3)[0]
        inside = np.concatenate([eye + np.outer(rng.uniform(20, 300, 200), fwd),
                                 rng.uniform(0.1, 5, (200, 1))], 1).astype(np.float32)
        assert not outside_view(host, inside, cam).any()


@pytest.mark.parametrize("seed", [1, 2])
def test_ortho_culling_is_conservative(host, seed):
    rng = np.random.default_rng(seed)
    for k in range(6):
        dim = [2048, 512, 256][k % 3]
        centre = rng.uniform(-100, 100, 3)
        extent = float(rng.uniform(50, 400))
        cas = scenes.make_cascade(dim, centre, extent, float(rng.uniform(-np.pi, np.pi)),
                                  float(rng.uniform(0.2, 1.4)))
        sph = _spheres(rng, 4000, centre, 1.5 * extent, 0.2 * extent)
        out = outside_ortho(host, sph, cas)
        sep = ortho_separation(sph, cas)
        bad = out & (sep < MARGIN)
        assert not bad.any(), f"skipped spheres the ortho setup could draw: {sph[bad][:3]}"
        assert 0.2 < out.mean() < 0.98, out.mean()
        # the cascade's centre is drawn
        assert not outside_ortho(host, np.array([[*centre, 1.0]], np.float32), cas).any()


def test_zero_constants_cull_nothing(host):
    """A camera whose rows are zero gives zero planes: nothing is skipped."""
    from axi_visualizer_amd.scenes import Camera
    cam = Camera(pos=np.zeros(3, np.float32), vm=np.zeros((3, 3), np.float32), scale=1.0,
                 cax=1.0, cay=1.0)
    sph = np.array([[5, 0, 0, 1], [-5, 0, 0, 1], [0, 0, 0, 0]], np.float32)
    assert not outside_view(host, sph, cam).any()
