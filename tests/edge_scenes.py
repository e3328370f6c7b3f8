"""Scenes for the edge-case parity tests: hostile geometry and coarse-binning
slivers.  Shared by the CPU checks of the coarse enumeration and the GPU
parity tests."""
import dataclasses

import numpy as np

from axi_visualizer_amd import scenes


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


# ------------------------------------------------------- coarse slivers
# An 8192 x 48 frame seen by a camera at the origin looking down +z with
# scale 1, so a vertex at depth 1 projects to sx = x + W/2, sy = H/2 - y
# exactly.  Near sy = 24 (tile row 1 of Pipe/coarse.c's P/16 - 0.5 grid)
# screen y moves in steps of 2^-19, so a triangle 14,000 px wide and two
# such steps tall has edge slopes near 3.7e9 tiles per tile row: two rows
# away from its vertices an edge position passes 2^31 and its floorf
# saturates to INT_MAX.  The serial walk emits nothing in such a row; a
# strided row walk that adds a thread offset to that INT_MAX wraps to
# negative tile columns.
SLIVER_W, SLIVER_H = 8192, 48
_E = 2.0 ** -19
# (sx, sy) per vertex; the first and third reach the saturating row in the
# first (y1 -> y2) and second (y3 -> y2) halves of the walk, the others are
# their mirror images.
SLIVER_TRIS = {
    "first_half": [(-2000.0, 24 - _E), (12000.0, 24 + _E), (12040.0, 24 + _E)],
    "first_half_mirror": [(12000.0, 24 - _E), (-2000.0, 24 + _E), (-2040.0, 24 + _E)],
    "second_half": [(12000.0, 24 - _E), (12040.0, 24 - _E), (-2000.0, 24 + _E)],
    "second_half_mirror": [(-2000.0, 24 - _E), (-2040.0, 24 - _E), (12000.0, 24 + _E)],
}


def sliver_camera(W=SLIVER_W, H=SLIVER_H):
    return scenes.Camera(pos=np.zeros(3, np.float32),
                         vm=np.array([[0, 0, 1], [-1, 0, 0], [0, -1, 0]], np.float32),
                         scale=1.0, cax=1e4, cay=1e4)


def screen_to_world(pts, z=1.0, W=SLIVER_W, H=SLIVER_H):
    """World vertices that sliver_camera projects to the screen points (at
    depth 1 exactly; at other depths up to rounding)."""
    return np.array([[(sx - W / 2) * z, (H / 2 - sy) * z, z] for sx, sy in pts], np.float32)


def sliver_scene(nfill=300, seed=21):
    """The four slivers drawn over random filler triangles (1-200 px, depth
    1.5-20) in one group, with a mipmapped noise texture."""
    W, H = SLIVER_W, SLIVER_H
    rng = np.random.default_rng(seed)
    tris = [screen_to_world(p) for p in SLIVER_TRIS.values()]
    for _ in range(nfill):
        c = (rng.uniform(0, W), rng.uniform(0, H))
        size = rng.choice([3.0, 20.0, 200.0])
        pts = [(c[0] + size * rng.uniform(-1, 1), c[1] + size * rng.uniform(-1, 1))
               for _ in range(3)]
        tris.append(screen_to_world(pts, z=float(rng.uniform(1.5, 20.0))))
    xyz = np.concatenate(tris).astype(np.float32)
    n = xyz.shape[0]
    vn = np.tile(np.array([0.0, 0.0, -1.0], np.float32), (n, 1))
    uv = rng.uniform(0, 1, (n, 2)).astype(np.float32)
    tex, side = scenes.texture_for(scenes.noise_texture(seed, 600, 64), True)
    g = scenes.Group(xyz=xyz, vn=vn, uv=uv, tex=tex, tex_side=side, mip=True)
    return scenes.Scene(name="coarse_slivers", W=W, H=H, groups=[g], view=sliver_camera(),
                        light_int=np.array([0.9, 0.85, 0.8], np.float32),
                        light_dir=scenes.view_vec(0.4, 0.7).astype(np.float32), amb=0.2)
