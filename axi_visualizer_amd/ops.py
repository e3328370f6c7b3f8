"""CLDraw: the reference's draw-call and shader-binding API over libaxr.

Method names, argument order and argument meaning follow
``PyOpenCL27C/Ops.py`` class ``CLDraw`` (file:line cited per method); the
device work behind each call is a hand-written gfx950 HIP kernel reached
through the C ABI in include/axr.h.  Errors raise ``AxrError`` (a
RuntimeError), as PyOpenCL raised on the reference side.

Scope: the plain raster path (triangle setup, small-triangle scanline,
16x16 tile binning and tiled scanline, shading with trilinear mips,
Blinn-Phong and PCF shadows, the ortho shadow-map depth pass, gamma).
Shader keywords outside it -- alpha, multi, sky, SSR, refl, sub, emissive,
normal, fog, 2d -- raise NotImplementedError instead of rendering
something else.
"""
from __future__ import annotations

import ctypes
from typing import Optional, Sequence

import numpy as np

from ._native import AxrError, check, lib, pinned_array, ptr
from .scenes import SH_CULL, SH_MIP, SH_PHONG, SH_SHADOW, view_mat

_UNSUPPORTED = ("alpha", "multi", "sky", "SSR", "refl", "sub", "emissive", "normal",
                "fog", "2d")


def is_ssr(spec) -> bool:
    """A water shader entry: a dict with an "SSR" key naming the reflection
    cube map (shaders[tn]["SSR"], Ops.py:859-860)."""
    return isinstance(spec, dict) and "SSR" in spec


def shader_flags(spec) -> int:
    """Maps a reference shader string (matched by substring, as Ops.py:768-805
    does) or an int of AXR_SH_* bits to flag bits."""
    if isinstance(spec, (int, np.integer)):
        return int(spec)
    if isinstance(spec, dict):
        spec = " ".join(str(k) for k in spec)
    s = str(spec)
    for kw in _UNSUPPORTED:
        if kw in s:
            raise NotImplementedError(f"shader keyword {kw!r} is outside the raster hot path")
    f = 0
    if "mip" in s:
        f |= SH_MIP
    if "shadow" in s:
        f |= SH_SHADOW
    if "phong" in s:
        f |= SH_PHONG
    if "cull" in s:
        f |= SH_CULL
    return f


def _f32(a, shape=None) -> np.ndarray:
    a = np.ascontiguousarray(np.asarray(a, dtype=np.float32))
    if shape is not None:
        a = a.reshape(shape)
    return a


class CLDraw:
    """Ops.py:127-229.  One render target of w x h with its groups."""

    def __init__(self, size_sky, max_uv, w, h, max_particles):
        self.W = np.int32(w)
        self.H = np.int32(h)
        self.A = int(w) * int(h)
        self._L = lib()
        c = ctypes.c_void_p()
        check(self._L.axr_create(int(w), int(h), ctypes.byref(c)), "axr_create")
        self._ctx = c
        self.gSize = []
        self.texSize = []
        self._isMip = []
        self._texels = []
        self.SHADOWMAP = {}
        self.ambLight = np.float32(0.2)
        # point and spot lights of the last vertLight that set them (a call
        # that passes None keeps them, as Ops.py:564-600 keeps its buffers)
        self._points = (0, [None, None])
        self._spots = (0, [None, None, None])
        self._flags_buf = None
        self._hframe = None   # getFrame's arrays (page-locked, reused)
        self._hrgba = None    # getRGBA8's array
        self._mask_buf = None
        self._sid_buf = np.zeros(2, np.int32)
        self.reflTexSize = {}
        self._refl = {}
        # REFL_LENGTH, REFL_STEP, ABSORB and the Fresnel term of the water
        # shader (Shaders_src/av_SSR.c:6-13, 52-53)
        self.ssrParams = dict(steps=400, step=1, absorb=0.3, fresnel=True)

    # -- lifetime
    def close(self):
        if getattr(self, "_ctx", None):
            self._L.axr_destroy(self._ctx)
            self._ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _chk(self, rc, what):
        return check(rc, what)

    # -- scene setup
    def setScaleCull(self, s, cx, cy):
        """Ops.py:231-233."""
        self.sScale = np.float32(s)
        self.caX, self.caY = np.float32(cx), np.float32(cy)
        self._chk(self._L.axr_set_scale_cull(self._ctx, float(self.sScale), float(self.caX),
                                             float(self.caY)), "setScaleCull")

    def setSkyTex(self, r, g, b, size):
        """Ops.py:469-473.  Accepted for API compatibility; sky drawing is
        outside the raster hot path, so the texture is not used."""
        self.skyTexSize = np.int32(size)

    def addTextureGroup(self, xyz, uv, vn, r, g, b, mip=False, build_mips=False):
        """Ops.py:331-378.  xyz, vn: (3T, 3); uv: (T, 3, 2) or (3T, 2);
        r, g, b: N x N planes, or mip chains when mip=N.  build_mips (with
        mip=N): r, g, b are the N x N base level and the chain is built on
        the device (Utils.createMips' 2x2 integer box filter)."""
        p = _f32(xyz, (-1, 3))
        n = _f32(vn, (-1, 3))
        t = _f32(uv, (-1, 2))
        if p.shape[0] % 3 or n.shape != p.shape or t.shape[0] != p.shape[0]:
            raise ValueError("xyz, vn and uv must describe the same 3T vertices")
        rr = np.ascontiguousarray(np.asarray(r).astype(np.uint16).reshape(-1))
        gg = np.ascontiguousarray(np.asarray(g).astype(np.uint16).reshape(-1))
        bb = np.ascontiguousarray(np.asarray(b).astype(np.uint16).reshape(-1))
        if mip:
            tex_size = int(np.log2(mip))
        else:
            tex_size = int(np.asarray(r).shape[0])
        tn = ctypes.c_int()
        self._chk(self._L.axr_add_texture_group(
            self._ctx, ptr(p), ptr(t), ptr(n), p.shape[0] // 3, ptr(rr), ptr(gg), ptr(bb),
            rr.shape[0], tex_size, (2 if build_mips else 1) if mip else 0, ctypes.byref(tn)),
            "addTextureGroup")
        self.gSize.append(np.int32(p.shape[0]))
        self.texSize.append(np.int32(tex_size))
        self._texels.append(((4 ** (tex_size + 1)) - 1) // 3 if mip else rr.shape[0])
        self._isMip.append(bool(mip))
        return tn.value

    def setGroupMaterial(self, tn, mip_bias=0.0, spec_pow=16):
        """Material constants the reference passes as shader arguments
        (mipBias, av_draw_normmap_mip.c:21; specPow, av_draw_normmap.c:282)."""
        self._chk(self._L.axr_set_group_material(self._ctx, int(tn), float(mip_bias),
                                                 int(spec_pow)), "setGroupMaterial")

    def setPrimaryLight(self, dirI, dirD):
        """Ops.py:1143-1151.  dirI, dirD: (1, 3)."""
        i = _f32(dirI, (-1, 3))[0].copy()
        d = _f32(dirD, (-1, 3))[0].copy()
        self._chk(self._L.axr_set_primary_light(self._ctx, ptr(i), ptr(d)), "setPrimaryLight")

    def vertLight(self, mask, dirI, dirD, pointI=None, pointP=None,
                  spotI=None, spotD=None, spotP=None):
        """Ops.py:552-612 -> Pipe/vert.c vertL.  mask[tn] True lights group
        tn.  pointI (spotI) = 1 removes the point (spot) lights; None keeps
        those of the previous call (Ops.py:564-600: the light buffers
        persist, and Compute.py:429 relies on it); an array sets them."""
        m = np.ascontiguousarray(np.asarray(mask, dtype=bool).astype(np.uint8))
        di = _f32(dirI, (-1, 3))
        dd = _f32(dirD, (-1, 3))

        def lights(kept, i, *rest):
            if i is None:
                return kept
            if isinstance(i, int) and i == 1:
                return 0, [None] * (1 + len(rest))
            a = [_f32(i, (-1, 3))] + [_f32(x, (-1, 3)) for x in rest]
            if any(x.shape != a[0].shape for x in a):
                raise ValueError("light arrays must have the same number of lights")
            return a[0].shape[0], a

        self._points = lights(self._points, pointI, pointP)
        self._spots = lights(self._spots, spotI, spotD, spotP)
        npt, (pI, pP) = self._points
        nsp, (sI, sD, sP) = self._spots
        self._chk(self._L.axr_vert_light(
            self._ctx, ptr(m), m.shape[0], ptr(di), ptr(dd), di.shape[0],
            ptr(pI), ptr(pP), npt, ptr(sI), ptr(sD), ptr(sP), nsp,
            float(self.ambLight)), "vertLight")

    # -- per frame
    def setPos(self, vc):
        """Ops.py:475-476."""
        p = _f32(vc, (3,))
        self._chk(self._L.axr_set_pos(self._ctx, ptr(p)), "setPos")

    def setVM(self, vM):
        """Ops.py:477-479.  Rows (forward, horizontal, vertical)."""
        v = _f32(vM, (3, 3))
        self._chk(self._L.axr_set_vm(self._ctx, ptr(v)), "setVM")

    def clearZBuffer(self):
        """Ops0.py:650-653 (Ops.py:939-946 also clears colour)."""
        self._chk(self._L.axr_clear_zbuffer(self._ctx), "clearZBuffer")

    def setClearColor(self, rgb):
        c = np.ascontiguousarray(np.asarray(rgb, dtype=np.uint16).reshape(3))
        self._chk(self._L.axr_set_clear_color(self._ctx, ptr(c)), "setClearColor")

    def drawAll(self, shaders, mask=None, shadowIds=[0, 1], useOpacitySM=False, pcf=5):
        """Ops.py:716-937.  shaders[tn]: shader string (keywords matched by
        substring) or AXR_SH_* int; mask[tn] True skips group tn."""
        ng = len(self.gSize)
        if self._flags_buf is None or self._flags_buf.shape[0] != ng:
            self._flags_buf = np.zeros(ng, np.uint32)
            self._mask_buf = np.zeros(ng, np.uint8)
        water = []
        for tn in range(ng):
            masked = mask is not None and mask[tn]
            if is_ssr(shaders[tn]):
                # drawn after the opaque groups, over their frame
                self._flags_buf[tn] = 0
                self._mask_buf[tn] = 1
                if not masked:
                    water.append(tn)
                continue
            self._flags_buf[tn] = shader_flags(shaders[tn])
            self._mask_buf[tn] = 1 if masked else 0
        sids = [s for s in (shadowIds or []) if s in self.SHADOWMAP][:2]
        ns = len(sids)
        for k, s in enumerate(sids):
            self._sid_buf[k] = s
        self._chk(self._L.axr_draw_all(self._ctx, self._flags_buf.ctypes.data,
                                       self._mask_buf.ctypes.data, ng,
                                       self._sid_buf.ctypes.data, ns, int(pcf)), "drawAll")
        if water:
            names = [shaders[tn]["SSR"] for tn in water]
            missing = [n for n in names if n not in self._refl]
            if missing:
                raise KeyError(f"drawAll: no reflection texture {missing[0]!r} (setReflTex)")
            grp = np.array(water, np.int32)
            slot = np.array([self._refl[n] for n in names], np.int32)
            p = self.ssrParams
            self._chk(self._L.axr_draw_ssr(self._ctx, ptr(grp), ptr(slot), len(water),
                                           int(p["steps"]), int(p["step"]), float(p["absorb"]),
                                           1 if p["fresnel"] else 0), "drawAll (SSR)")

    def setReflTex(self, name, r, g, b, size):
        """Ops.py:458-467: a reflection cube strip (faces of side `size`,
        size rows of 6*size texels, Cubemap.py setupTex2's layout) under
        `name`; SSR groups name it in their shader entry."""
        n = 6 * int(size) * int(size)
        planes = [np.ascontiguousarray(np.asarray(x).reshape(-1)[:n].astype(np.uint16))
                  for x in (r, g, b)]
        if any(p.size != n for p in planes):
            raise ValueError(f"setReflTex: planes need 6*size*size = {n} texels")
        slot = self._refl.get(name, -1)
        slot = self._chk(self._L.axr_set_refl_tex(self._ctx, slot, *[ptr(p) for p in planes],
                                                  int(size)), "setReflTex")
        self._refl[name] = slot
        self.reflTexSize[name] = np.int32(int(size) // 2)
        return len(self.reflTexSize) - 1

    def getReflTex(self, name):
        """The cube strip `name` as stored on the device: (r, g, b) planes of
        shape (size, 6 * size)."""
        S = 2 * int(self.reflTexSize[name])
        out = [np.empty((S, 6 * S), np.uint16) for _ in range(3)]
        self._chk(self._L.axr_get_refl_tex(self._ctx, self._refl[name], *[ptr(o) for o in out]),
                  "getReflTex")
        return tuple(out)

    def gamma(self, ex):
        """Ops.py:947-957 -> Shaders/Post/gamma.c."""
        self._chk(self._L.axr_gamma(self._ctx, float(ex)), "gamma")

    def blur(self):
        """Ops.py:1009-1020 -> Shaders/Post/blur.c, blur2.c: brightness-
        weighted half-resolution copy blurred 19 taps each way and added
        back to the frame."""
        self._chk(self._L.axr_blur(self._ctx), "blur")

    def bloom(self, levels=5, threshold=1.0, white=9835.0, strength=1.0):
        """Multi-level fp16 bloom on the frame: max(frame / white -
        threshold, 0) downsampled `levels` times, upsampled back and added
        with weight white * strength.  W and H must divide by 2**levels."""
        self._chk(self._L.axr_bloom(self._ctx, int(levels), float(threshold), float(white),
                                    float(strength)), "bloom")

    def getFrame(self):
        """Ops.py:1224-1230: (R, G, B) u16 and depth f32, each (H, W).  As in
        the reference (Ops.py:216-219), the same four arrays come back from
        every call, overwritten by the next one; they live in page-locked
        memory, so the copy runs at the link's DMA rate."""
        if self._hframe is None:
            H, W = int(self.H), int(self.W)
            self._hframe = tuple(pinned_array((H, W), t)
                                 for t in (np.uint16, np.uint16, np.uint16, np.float32))
        r, g, b, d = self._hframe
        self._chk(self._L.axr_get_frame(self._ctx, ptr(r), ptr(g), ptr(b), ptr(d)), "getFrame")
        return self._hframe

    def getRGBA8(self):
        """The gamma'd frame as (H, W, 4) uint8: one page-locked array,
        overwritten by the next call (as getFrame's are)."""
        if self._hrgba is None:
            self._hrgba = pinned_array((int(self.H), int(self.W)), np.uint32)
        self._chk(self._L.axr_get_rgba8(self._ctx, ptr(self._hrgba)), "getRGBA8")
        return self._hrgba.view(np.uint8).reshape(int(self.H), int(self.W), 4)

    def getIds(self):
        """Depth-winning global triangle index per pixel (0xFFFFFFFF: none)."""
        out = np.empty((int(self.H), int(self.W)), np.uint32)
        self._chk(self._L.axr_get_ids(self._ctx, ptr(out)), "getIds")
        return out

    # -- shadow maps (Ops.py:1022-1141)
    def addShadowMap(self, i, size, scale, ambLight=None, useGI=False):
        if useGI:
            raise NotImplementedError("GI shadow maps are outside the raster hot path")
        amb = -1.0 if ambLight is None else float(ambLight)
        if ambLight is not None:
            self.ambLight = np.float32(ambLight)
        self._chk(self._L.axr_add_shadow_map(self._ctx, int(i), int(size), float(scale), amb),
                  "addShadowMap")
        self.SHADOWMAP[i] = {"dim": np.int32(size), "dim2": np.int32(size // 2),
                             "scale": np.float32(scale)}

    def placeShadowMap(self, i, pos, facing, ambLight=None):
        """facing: (a, b) angles -> viewMat (Ops.py:1078-1087), or a 3x3 basis."""
        p = _f32(pos, (3,))
        f = np.asarray(facing, dtype=np.float64)
        vm = _f32(view_mat(*f) if f.shape == (2,) else f, (3, 3))
        amb = -1.0 if ambLight is None else float(ambLight)
        if ambLight is not None:
            self.ambLight = np.float32(ambLight)
        self._chk(self._L.axr_place_shadow_map(self._ctx, int(i), ptr(p), ptr(vm), amb),
                  "placeShadowMap")
        self.SHADOWMAP[i]["vecnp"] = vm

    def clearShadowMap(self, i):
        self._chk(self._L.axr_clear_shadow_map(self._ctx, int(i)), "clearShadowMap")

    def shadowMap(self, i, whichCast, shaders, bias, proj="ortho"):
        if proj != "ortho":
            raise NotImplementedError("only ortho shadow projection is in the raster hot path")
        for tn, cast in enumerate(whichCast):
            if cast and "alpha" in str(shaders[tn]):
                raise NotImplementedError("alpha-tested casters are outside the hot path")
        m = np.ascontiguousarray(np.asarray(whichCast, dtype=bool).astype(np.uint8))
        self._chk(self._L.axr_shadow_map(self._ctx, int(i), ptr(m), m.shape[0], float(bias)),
                  "shadowMap")

    def getSHM(self, i):
        """Ops.py:1216-1222."""
        n = int(self.SHADOWMAP[i]["dim"])
        s = np.empty((n, n), np.float32)
        self._chk(self._L.axr_get_shadow_map(self._ctx, int(i), ptr(s)), "getSHM")
        return s

    # -- extras
    def sync(self):
        self._chk(self._L.axr_sync(self._ctx), "sync")

    def stats(self):
        s = np.zeros(6, np.uint32)
        self._chk(self._L.axr_get_stats(self._ctx, ptr(s)), "stats")
        return {"small": int(s[0]), "large": int(s[1]), "casters": int(s[2]),
                "bin_entries": int(s[3]), "bin_overflow": int(s[4]),
                "shadow_bin_entries": int(s[5])}

    def setShaders(self, shaders, mask=None):
        """drawAll's shader binding without drawing: prepares this context
        as a resident scene for batches."""
        ng = len(self.gSize)
        f = np.array([shader_flags(shaders[tn]) for tn in range(ng)], np.uint32)
        m = np.array([1 if (mask is not None and mask[tn]) else 0 for tn in range(ng)], np.uint8)
        self._chk(self._L.axr_set_shaders(self._ctx, ptr(f), ptr(m), ng), "setShaders")

    def getTexels(self, tn):
        """Group tn's texels as stored on the device: (r, g, b) u16 planes."""
        n = int(self._texels[tn])
        out = [np.empty(n, np.uint16) for _ in range(3)]
        self._chk(self._L.axr_get_texels(self._ctx, int(tn), *[ptr(o) for o in out]),
                  "getTexels")
        return tuple(out)

    def getVertLight(self, tn):
        """Group tn's per-vertex light I (the LI buffer vertL writes,
        Ops.py:605-611) as stored on the device: (3T, 3) float32."""
        out = np.empty((int(self.gSize[tn]), 3), np.float32)
        self._chk(self._L.axr_get_vertex_light(self._ctx, int(tn), ptr(out)), "getVertLight")
        return out

    def deviceBytes(self) -> int:
        n = ctypes.c_uint64()
        self._chk(self._L.axr_get_bytes(self._ctx, ctypes.byref(n)), "deviceBytes")
        return int(n.value)

    def setBinCapacity(self, small=0, large=0, shadow=0, spill=0):
        """Bucket capacities (entries per bin) and spill-list size; 0 keeps."""
        self._chk(self._L.axr_set_bin_capacity(self._ctx, int(small), int(large), int(shadow),
                                               int(spill)), "setBinCapacity")

    def setTiming(self, on: bool):
        self._chk(self._L.axr_set_timing(self._ctx, 1 if on else 0), "setTiming")

    def timing(self):
        n, nl = 64, 32
        names = ctypes.create_string_buffer(n * nl)
        ms = np.zeros(n, np.float64)
        cnt = np.zeros(n, np.int64)
        k = self._chk(self._L.axr_get_timing(self._ctx, names, nl, ptr(ms), ptr(cnt), n),
                      "timing")
        out = {}
        for j in range(k):
            name = names.raw[j * nl:(j + 1) * nl].split(b"\0", 1)[0].decode()
            out[name] = (float(ms[j]), int(cnt[j]))
        return out
