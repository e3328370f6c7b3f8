"""Coarse binning's tile enumeration (axr_tri.h: coarse_rows,
coarse_block_tiles, coarse_walk) built for the host and checked on the CPU
against the oracle's Pipe/coarse.c restatement (orc_coarse_tri) and the
tile grid's bounds, without a GPU.  The kernels k_coarse (a block per large
triangle) and the batched setup (one thread per triangle) run exactly these
templates."""
import ctypes
import os
import shutil
import subprocess

import numpy as np
import pytest

from oracle import pyoracle as po
from tests.edge_scenes import SLIVER_H, SLIVER_TRIS, SLIVER_W, screen_to_world, sliver_camera

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
INT_MAX = 2**31 - 1


@pytest.fixture(scope="module")
def host(tmp_path_factory):
    if not os.path.exists(HIPCC) and not shutil.which("hipcc"):
        pytest.skip("hipcc not available")
    out = str(tmp_path_factory.mktemp("coarse") / "libcoarse_host.so")
    cmd = [HIPCC if os.path.exists(HIPCC) else "hipcc", "-x", "hip", "--cuda-host-only", "-O2",
           "-std=c++17", "-fPIC", "-shared", "-ffp-contract=off",
           "-I", os.path.join(ROOT, "axi_visualizer_amd", "csrc"), "-I", os.path.join(ROOT, "include"),
           os.path.join(HERE, "coarse_host.cpp"), "-o", out]
    res = subprocess.run(cmd, capture_output=True, text=True)
    assert res.returncode == 0, res.stderr
    L = ctypes.CDLL(out)
    p, i = ctypes.c_void_p, ctypes.c_int
    L.host_coarse_rows.argtypes = [p, p, i, i, p, p, p, i]
    L.host_coarse_block.argtypes = [p, p, i, i, i, p, i]
    L.host_coarse_walk.argtypes = [p, p, i, i, p, i]
    return L


def _xy(P):
    return (np.ascontiguousarray(P[:, 0], np.float32), np.ascontiguousarray(P[:, 1], np.float32))


def block_tiles(L, P, WC, HC, nthreads=256, cap=1 << 20):
    px, py = _xy(P)
    out = np.empty(cap, np.int32)
    n = L.host_coarse_block(px.ctypes.data, py.ctypes.data, WC, HC, nthreads, out.ctypes.data, cap)
    assert 0 <= n <= cap
    return out[:n]


def walk_tiles(L, P, WC, HC, cap=1 << 20):
    px, py = _xy(P)
    out = np.empty(cap, np.int32)
    n = L.host_coarse_walk(px.ctypes.data, py.ctypes.data, WC, HC, out.ctypes.data, cap)
    assert 0 <= n <= cap
    return out[:n]


def rows(L, P, WC, HC, cap=256):
    px, py = _xy(P)
    cy = np.empty(cap, np.int32)
    first = np.empty(cap, np.int32)
    xe = np.empty(cap, np.float32)
    n = L.host_coarse_rows(px.ctypes.data, py.ctypes.data, WC, HC, cy.ctypes.data,
                           first.ctypes.data, xe.ctypes.data, cap)
    assert 0 <= n <= cap
    return list(zip(cy[:n].tolist(), first[:n].tolist(), xe[:n].tolist()))


def check_tri(L, P, WC, HC):
    ref = po.coarse(P.reshape(6), WC, HC)
    blk = block_tiles(L, P, WC, HC)
    assert np.all((blk >= 0) & (blk < WC * HC)), f"tile outside the grid: {blk.min()}..{blk.max()}"
    assert sorted(blk.tolist()) == sorted(ref.tolist())
    assert walk_tiles(L, P, WC, HC).tolist() == ref.tolist()
    return ref


@pytest.mark.parametrize("name", list(SLIVER_TRIS))
def test_sliver_rows_saturate_and_tiles_stay_in_grid(host, name):
    W, H = SLIVER_W, SLIVER_H
    WC, HC = (W + 15) // 16, (H + 15) // 16
    xyz = screen_to_world(SLIVER_TRIS[name])
    vn = np.tile(np.array([0.0, 0.0, -1.0], np.float32), (3, 1))
    P, _, cls = po.trisetup(xyz, vn, sliver_camera(), W, H)
    assert cls[0] == 2, "the sliver must be a large triangle"
    # the camera maps depth-1 vertices exactly
    assert P[0].tolist() == [[float(np.float32(x)), float(np.float32(y))]
                             for x, y in SLIVER_TRIS[name]]
    r = rows(host, P[0], WC, HC)
    assert all(0 <= cy < HC and first >= 0 and xe <= WC - 1 for cy, first, xe in r)
    if not name.endswith("mirror"):
        # the case the strided row walk must survive: first = INT_MAX
        assert any(first == INT_MAX for _, first, _ in r), r
    check_tri(host, P[0], WC, HC)


@pytest.mark.parametrize("W,H,seed", [(640, 480, 1), (1920, 1080, 2), (208, 143, 3),
                                      (4096, 32, 4)])
def test_block_tiles_equal_the_oracle_on_random_large_triangles(host, W, H, seed):
    """Random triangles of 96 px to 2W across, partly off screen, in every
    orientation: the block enumeration emits the serial walk's multiset and
    nothing outside the grid."""
    WC, HC = (W + 15) // 16, (H + 15) // 16
    rng = np.random.default_rng(seed)
    n = 0
    for _ in range(300):
        c = rng.uniform([-0.2 * W, -0.2 * H], [1.2 * W, 1.2 * H])
        ext = rng.uniform(50, 1.0 * W)
        P = (c + rng.uniform(-ext, ext, (3, 2))).astype(np.float32)
        if rng.uniform() < 0.2:  # a flat or steep sliver
            P[rng.integers(3), 1] = P[(rng.integers(1, 3) + 0) % 3, 1]
        ref = check_tri(host, P, WC, HC)
        n += len(ref)
    assert n > 0
