"""ctypes wrapper of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, where it is the checker (or the timed CPU
baseline), never the product path.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

ORC_MIP, ORC_SHADOW, ORC_PHONG, ORC_CULL = 1, 2, 4, 8


class OrcGroup(ctypes.Structure):
    _fields_ = [("ntri", ctypes.c_int), ("xyz", ctypes.c_void_p), ("vn", ctypes.c_void_p),
                ("uv", ctypes.c_void_p), ("li", ctypes.c_void_p), ("tr", ctypes.c_void_p),
                ("tg", ctypes.c_void_p), ("tb", ctypes.c_void_p), ("tex_size", ctypes.c_int),
                ("flags", ctypes.c_uint32), ("mip_bias", ctypes.c_float),
                ("spec_pow", ctypes.c_int), ("skip", ctypes.c_int), ("cast", ctypes.c_int),
                ("ssr", ctypes.c_int)]


class OrcCam(ctypes.Structure):
    _fields_ = [("pos", ctypes.c_float * 3), ("vm", ctypes.c_float * 9),
                ("scale", ctypes.c_float), ("cax", ctypes.c_float), ("cay", ctypes.c_float)]


class OrcShadowCfg(ctypes.Structure):
    _fields_ = [("dim", ctypes.c_int), ("cam", OrcCam), ("bias", ctypes.c_float)]


class OrcFrame(ctypes.Structure):
    _fields_ = [("W", ctypes.c_int), ("H", ctypes.c_int), ("view", OrcCam),
                ("lint", ctypes.c_float * 3), ("ldir", ctypes.c_float * 3),
                ("amb", ctypes.c_float), ("nshadow", ctypes.c_int),
                ("shadow", OrcShadowCfg * 2), ("pcf_taps", ctypes.c_int),
                ("exposure", ctypes.c_float), ("clear_rgb", ctypes.c_uint16 * 3),
                ("nthreads", ctypes.c_int),
                ("refl_r", ctypes.c_void_p * 4), ("refl_g", ctypes.c_void_p * 4),
                ("refl_b", ctypes.c_void_p * 4), ("refl_size", ctypes.c_int * 4),
                ("ssr_steps", ctypes.c_int), ("ssr_step", ctypes.c_int),
                ("ssr_absorb", ctypes.c_float), ("ssr_fresnel", ctypes.c_int),
                ("init_depth", ctypes.c_void_p), ("init_id", ctypes.c_void_p),
                ("init_r", ctypes.c_void_p), ("init_g", ctypes.c_void_p),
                ("init_b", ctypes.c_void_p)]


class OrcOut(ctypes.Structure):
    _fields_ = [("depth", ctypes.c_void_p), ("id", ctypes.c_void_p), ("r", ctypes.c_void_p),
                ("g", ctypes.c_void_p), ("b", ctypes.c_void_p),
                ("shadow", ctypes.c_void_p * 2), ("stats", ctypes.c_int64 * 8)]


_lib = None


def build() -> str:
    """Builds liboracle.so with its Makefile (g++, -ffp-contract=off)."""
    res = subprocess.run(["make", "-C", HERE, "liboracle.so"], capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"oracle build failed:\n{res.stdout}\n{res.stderr}")
    return LIB_PATH


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        p, i, f = ctypes.c_void_p, ctypes.c_int, ctypes.c_float
        L.orc_render.argtypes = [p, i, p, p]
        L.orc_trisetup.argtypes = [p, p, i, p, i, i, i, p, p, p]
        L.orc_shadow_setup.argtypes = [p, i, p, i, f, p, p, p]
        L.orc_coarse_tri.argtypes = [p, i, i, p, i]
        L.orc_fragments.argtypes = [i, p, p, i, i, i, p, p, p, p, i]
        L.orc_gamma.argtypes = [p, p, p, ctypes.c_int64, f]
        L.orc_create_mips.argtypes = [p, i, p]
        L.orc_blur.argtypes = [p, p, p, i, i]
        L.orc_math.argtypes = [i, p, p, ctypes.c_int64]
        L.orc_vertlight.argtypes = [p, p, i, p, p, i, p, p, i, p, p, p, i, f, p]
        for fn in ("orc_render", "orc_trisetup", "orc_shadow_setup", "orc_coarse_tri",
                   "orc_fragments", "orc_gamma", "orc_create_mips", "orc_vertlight", "orc_blur", "orc_math"):
            getattr(L, fn).restype = ctypes.c_int
        _lib = L
    return _lib


def _c(a, dtype):
    return np.ascontiguousarray(np.asarray(a, dtype=dtype))


def make_cam(cam) -> OrcCam:
    c = OrcCam()
    c.pos[:] = [float(x) for x in np.asarray(cam.pos, np.float32)]
    c.vm[:] = [float(x) for x in np.asarray(cam.vm, np.float32).reshape(9)]
    c.scale = float(np.float32(cam.scale))
    c.cax = float(np.float32(cam.cax))
    c.cay = float(np.float32(cam.cay))
    return c


def vertlight(xyz, vn, dirI, dirD, amb, pointI=None, pointP=None, spotI=None, spotD=None,
              spotP=None):
    """Pipe/vert.c vertL over (nv, 3) vertices."""
    xyz = _c(xyz, np.float32).reshape(-1, 3)
    vn = _c(vn, np.float32).reshape(-1, 3)
    out = np.empty_like(xyz)
    dI = _c(dirI, np.float32).reshape(-1, 3)
    dD = _c(dirD, np.float32).reshape(-1, 3)
    z = np.zeros((0, 3), np.float32)
    pI = z if pointI is None else _c(pointI, np.float32).reshape(-1, 3)
    pP = z if pointP is None else _c(pointP, np.float32).reshape(-1, 3)
    sI = z if spotI is None else _c(spotI, np.float32).reshape(-1, 3)
    sD = z if spotD is None else _c(spotD, np.float32).reshape(-1, 3)
    sP = z if spotP is None else _c(spotP, np.float32).reshape(-1, 3)
    lib().orc_vertlight(xyz.ctypes.data, vn.ctypes.data, xyz.shape[0], dI.ctypes.data,
                        dD.ctypes.data, dI.shape[0], pI.ctypes.data, pP.ctypes.data,
                        pI.shape[0], sI.ctypes.data, sD.ctypes.data, sP.ctypes.data,
                        sI.shape[0], float(amb), out.ctypes.data)
    return out


def scene_vertlight(scene, xyz, vn):
    """vertL over a group's vertices with the scene's lights: the primary
    light first (skipped by vert.c), then its extra directional, point and
    spot lights when the scene has them."""
    dI = [np.asarray(scene.light_int, np.float32).reshape(1, 3)]
    dD = [np.asarray(scene.light_dir, np.float32).reshape(1, 3)]
    extra = getattr(scene, "extra_dirs", None)
    if extra is not None:
        dI.append(np.asarray(extra[0], np.float32).reshape(-1, 3))
        dD.append(np.asarray(extra[1], np.float32).reshape(-1, 3))
    pts = getattr(scene, "points", None) or (None, None)
    sps = getattr(scene, "spots", None) or (None, None, None)
    return vertlight(xyz, vn, np.concatenate(dI), np.concatenate(dD), scene.amb,
                     pointI=pts[0], pointP=pts[1], spotI=sps[0], spotD=sps[1], spotP=sps[2])


def render(scene, nthreads=0, with_shadows_out=False, onto=None):
    """Renders a scenes.Scene: shadow cascades, clear, drawAll, gamma.
    Returns dict(depth, id, r, g, b, stats[, shadow]).  onto: a previous
    render's output (rendered with exposure 0) that this drawAll draws onto
    instead of a cleared frame."""
    keep = []
    groups = (OrcGroup * len(scene.groups))()
    for k, g in enumerate(scene.groups):
        xyz = _c(g.xyz, np.float32)
        vn = _c(g.vn, np.float32)
        uv = _c(g.uv, np.float32)
        li = scene_vertlight(scene, xyz, vn)
        tr = _c(g.tex[:, 0], np.uint16)
        tg = _c(g.tex[:, 1], np.uint16)
        tb = _c(g.tex[:, 2], np.uint16)
        keep += [xyz, vn, uv, li, tr, tg, tb]
        G = groups[k]
        G.ntri = g.ntri
        G.xyz, G.vn, G.uv, G.li = xyz.ctypes.data, vn.ctypes.data, uv.ctypes.data, li.ctypes.data
        G.tr, G.tg, G.tb = tr.ctypes.data, tg.ctypes.data, tb.ctypes.data
        G.tex_size = g.tex_size
        G.flags = int(g.flags)
        G.mip_bias = float(g.mip_bias)
        G.spec_pow = int(g.spec_pow)
        G.skip = 1 if g.skip else 0
        G.cast = 1 if g.cast else 0
        G.ssr = int(getattr(g, "ssr", -1))
    fr = OrcFrame()
    for k, strip in enumerate(getattr(scene, "refl", [])[:4]):
        planes = [_c(strip[..., c], np.uint16) for c in range(3)]
        keep += planes
        fr.refl_r[k], fr.refl_g[k], fr.refl_b[k] = (x.ctypes.data for x in planes)
        fr.refl_size[k] = strip.shape[0]
    fr.ssr_steps = int(getattr(scene, "ssr_steps", 400))
    fr.ssr_step = int(getattr(scene, "ssr_step", 1))
    fr.ssr_absorb = float(np.float32(getattr(scene, "ssr_absorb", 0.3)))
    fr.ssr_fresnel = 1 if getattr(scene, "ssr_fresnel", True) else 0
    fr.W, fr.H = scene.W, scene.H
    fr.view = make_cam(scene.view)
    fr.lint[:] = [float(x) for x in np.asarray(scene.light_int, np.float32)]
    fr.ldir[:] = [float(x) for x in np.asarray(scene.light_dir, np.float32)]
    fr.amb = float(np.float32(scene.amb))
    fr.nshadow = len(scene.cascades)
    for i, c in enumerate(scene.cascades):
        fr.shadow[i].dim = c.dim
        fr.shadow[i].cam = make_cam(c.cam)
        fr.shadow[i].bias = float(np.float32(c.bias))
    fr.pcf_taps = scene.pcf_taps
    fr.exposure = float(scene.exposure)
    fr.clear_rgb[:] = list(scene.clear_rgb)
    fr.nthreads = nthreads
    if onto is not None:
        init = [_c(onto[k], t) for k, t in (("depth", np.float32), ("id", np.uint32),
                                            ("r", np.uint16), ("g", np.uint16), ("b", np.uint16))]
        keep += init
        fr.init_depth, fr.init_id, fr.init_r, fr.init_g, fr.init_b = (x.ctypes.data for x in init)
    n = scene.W * scene.H
    out = {"depth": np.empty((scene.H, scene.W), np.float32),
           "id": np.empty((scene.H, scene.W), np.uint32),
           "r": np.empty((scene.H, scene.W), np.uint16),
           "g": np.empty((scene.H, scene.W), np.uint16),
           "b": np.empty((scene.H, scene.W), np.uint16)}
    o = OrcOut()
    o.depth, o.id = out["depth"].ctypes.data, out["id"].ctypes.data
    o.r, o.g, o.b = out["r"].ctypes.data, out["g"].ctypes.data, out["b"].ctypes.data
    if with_shadows_out:
        out["shadow"] = [np.empty((c.dim, c.dim), np.float32) for c in scene.cascades]
        for i, s in enumerate(out["shadow"]):
            o.shadow[i] = s.ctypes.data
    rc = lib().orc_render(ctypes.addressof(groups), len(scene.groups), ctypes.byref(fr),
                          ctypes.byref(o))
    if rc != 0:
        raise RuntimeError(f"orc_render failed: {rc}")
    out["stats"] = list(o.stats)
    del keep, n
    return out


def trisetup(xyz, vn, cam, W, H, cull=False):
    xyz = _c(xyz, np.float32).reshape(-1, 3)
    vn = _c(vn, np.float32).reshape(-1, 3)
    T = xyz.shape[0] // 3
    P = np.empty((T, 3, 2), np.float32)
    Z = np.empty((T, 3), np.float32)
    cls = np.empty(T, np.int8)
    c = make_cam(cam)
    lib().orc_trisetup(xyz.ctypes.data, vn.ctypes.data, T, ctypes.byref(c), W, H,
                       1 if cull else 0, P.ctypes.data, Z.ctypes.data, cls.ctypes.data)
    return P, Z, cls


def coarse(P6, WC, HC, cap=1 << 16):
    P6 = _c(P6, np.float32).reshape(6)
    out = np.empty(cap, np.int32)
    n = lib().orc_coarse_tri(P6.ctypes.data, WC, HC, out.ctypes.data, cap)
    if n < 0:
        raise RuntimeError("coarse capacity")
    return out[:n]


def fragments(mode, P6, Z3, W, H, tile=0, cap=1 << 20):
    """Fragments of one triangle: mode 0 small path, 1 tiled path (one
    tile), 2 shadow path.  Returns (x, y, tz, half) arrays in order."""
    P6 = _c(P6, np.float32).reshape(6)
    Z3 = _c(Z3, np.float32).reshape(3)
    xs = np.empty(cap, np.int32)
    ys = np.empty(cap, np.int32)
    tz = np.empty(cap, np.float32)
    hf = np.empty(cap, np.int8)
    n = lib().orc_fragments(mode, P6.ctypes.data, Z3.ctypes.data, W, H, tile, xs.ctypes.data,
                            ys.ctypes.data, tz.ctypes.data, hf.ctypes.data, cap)
    if n < 0:
        raise RuntimeError("fragment capacity")
    return xs[:n], ys[:n], tz[:n], hf[:n]


def gamma(r, g, b, ex):
    r, g, b = (_c(x, np.uint16).copy() for x in (r, g, b))
    lib().orc_gamma(r.ctypes.data, g.ctypes.data, b.ctypes.data, r.size, float(ex))
    return r, g, b


def blur(r, g, b):
    """CLDraw.blur on (H, W) u16 planes; returns new planes."""
    r, g, b = (_c(x, np.uint16).copy() for x in (r, g, b))
    H, W = r.shape
    if lib().orc_blur(r.ctypes.data, g.ctypes.data, b.ctypes.data, W, H) < 0:
        raise ValueError("blur: W >= 4, H >= 8, both even")
    return r, g, b


def log2_cw(x):
    """The decreed log2 (mip level selection) over float32 values."""
    x = _c(x, np.float32).reshape(-1)
    y = np.empty_like(x)
    lib().orc_math(0, x.ctypes.data, y.ctypes.data, x.size)
    return y


def exp_cw(x):
    """The decreed exp (SSR transmittance) over float32 values."""
    x = _c(x, np.float32).reshape(-1)
    y = np.empty_like(x)
    lib().orc_math(1, x.ctypes.data, y.ctypes.data, x.size)
    return y


def _bilerp_up(src, h, w):
    """Bilinear upsample of a (h/2, w/2) float32 level to (h, w): sample at
    ((x + 0.5) / 2 - 0.5, (y + 0.5) / 2 - 0.5), edges clamped, weights in
    float32, terms added left to right."""
    sh, sw = src.shape
    fx = (np.arange(w, dtype=np.float32) + np.float32(0.5)) / np.float32(2) - np.float32(0.5)
    fy = (np.arange(h, dtype=np.float32) + np.float32(0.5)) / np.float32(2) - np.float32(0.5)
    fx = np.maximum(fx, np.float32(0))
    fy = np.maximum(fy, np.float32(0))
    x0 = np.minimum(fx.astype(np.int64), sw - 1)
    y0 = np.minimum(fy.astype(np.int64), sh - 1)
    ax = fx - x0.astype(np.float32)
    ay = fy - y0.astype(np.float32)
    x1 = np.minimum(x0 + 1, sw - 1)
    y1 = np.minimum(y0 + 1, sh - 1)
    bx, by = np.float32(1) - ax, np.float32(1) - ay
    s00, s10 = src[y0][:, x0], src[y0][:, x1]
    s01, s11 = src[y1][:, x0], src[y1][:, x1]
    return ((by[:, None] * bx[None, :] * s00 + by[:, None] * ax[None, :] * s10)
            + ay[:, None] * bx[None, :] * s01) + ay[:, None] * ax[None, :] * s11


def bloom(r, g, b, levels=5, threshold=1.0, white=9835.0, strength=1.0):
    """Multi-level bloom on an fp16 pyramid (the full-frame pipeline's
    post pass; the reference's blur is single-level).  Per channel:
    level 1 = max(avg2x2(frame / white) - threshold, 0); level l+1 =
    avg2x2(level l); going up, up_l = level_l + bilinear(up_{l+1}); frame
    += white * strength * bilinear(up_1), stored saturated toward zero.
    Every level is stored as float16 (round to nearest even); arithmetic is
    float32 in the order written."""
    f32, f16 = np.float32, np.float16
    out = []
    for plane in (r, g, b):
        P = _c(plane, np.uint16)
        H, W = P.shape
        x = P.astype(f32) / f32(white)
        def avg(a):
            return ((a[0::2, 0::2] + a[0::2, 1::2]) + a[1::2, 0::2] + a[1::2, 1::2]) * f32(0.25)
        lv = [np.maximum(avg(x) - f32(threshold), f32(0)).astype(f16)]
        for _ in range(levels - 1):
            lv.append(avg(lv[-1].astype(f32)).astype(f16))
        up = lv[-1]
        for l in range(levels - 2, -1, -1):
            h, w = lv[l].shape
            up = (lv[l].astype(f32) + _bilerp_up(up.astype(f32), h, w)).astype(f16)
        add = _bilerp_up(up.astype(f32), H, W) * (f32(white) * f32(strength))
        v = P.astype(f32) + add
        out.append(np.clip(np.trunc(np.maximum(v, f32(0))), 0, 65535).astype(np.uint16))
    return tuple(out)


def create_mips(plane):
    plane = _c(plane, np.uint16)
    N = plane.shape[0]
    out = np.empty(((4 * N * N - 1) // 3,), np.uint16)
    n = lib().orc_create_mips(plane.ctypes.data, N, out.ctypes.data)
    if n < 0:
        raise RuntimeError("create_mips: N must be a power of two")
    return out[:n]


def shadow_setup(xyz, cam, dim, bias):
    xyz = _c(xyz, np.float32).reshape(-1, 3)
    T = xyz.shape[0] // 3
    XY = np.empty((T, 3, 2), np.int32)
    Z = np.empty((T, 3), np.float32)
    ok = np.empty(T, np.int8)
    c = make_cam(cam)
    lib().orc_shadow_setup(xyz.ctypes.data, T, ctypes.byref(c), dim, float(bias),
                           XY.ctypes.data, Z.ctypes.data, ok.ctypes.data)
    return XY, Z, ok
