"""Drives a Scene through CLDraw the way the reference engine does.

``Renderer.frame()`` issues the per-frame call sequence of
``ThreeDBackend.render`` (Compute.py:308-339) with the shadow pass of
``CLDraw.shadowMap`` (Ops.py:1089-1141) in front and ``gamma`` as the
post-process: clearShadowMap + shadowMap per cascade, setPos/setVM,
clearZBuffer, drawAll, gamma.
"""
from __future__ import annotations

import numpy as np

from .ops import CLDraw
from .scenes import Scene


def load_scene(d: CLDraw, sc: Scene):
    """Uploads a Scene into a CLDraw the way the reference engine's start-up
    does: groups with their textures, materials, the primary light, vertex
    light, shadow maps placed."""
    d.setScaleCull(sc.view.scale, sc.view.cax, sc.view.cay)
    for g in sc.groups:
        planes = [g.tex[:, c] if g.mip else g.tex[:, c].reshape(g.tex_side, g.tex_side)
                  for c in range(3)]
        tn = d.addTextureGroup(g.xyz, g.uv, g.vn, *planes,
                               mip=g.tex_side if g.mip else False)
        d.setGroupMaterial(tn, g.mip_bias, g.spec_pow)
    d.ambLight = np.float32(sc.amb)
    d.setPrimaryLight(sc.light_int[None], sc.light_dir[None])
    # light 0 is shaded per pixel; the others light the vertices (with the
    # primary light alone vertL leaves I at the ambient floor)
    L = sc.vert_lights()
    d.vertLight([True] * len(sc.groups), L["dirI"], L["dirD"],
                1 if L["pointI"] is None else L["pointI"], L["pointP"],
                1 if L["spotI"] is None else L["spotI"], L["spotD"], L["spotP"])
    for i, c in enumerate(sc.cascades):
        d.addShadowMap(i, c.dim, c.cam.scale, sc.amb)
        d.placeShadowMap(i, c.cam.pos, c.cam.vm, sc.amb)
    for k, strip in enumerate(sc.refl):
        d.setReflTex(k, strip[..., 0], strip[..., 1], strip[..., 2], strip.shape[0])
    d.ssrParams = dict(steps=sc.ssr_steps, step=sc.ssr_step, absorb=sc.ssr_absorb,
                       fresnel=sc.ssr_fresnel)


def scene_shaders(sc: Scene):
    """drawAll shader entries of a Scene's groups: AXR_SH_* flags, or
    {"SSR": k} for a water group reflecting cube map k."""
    return [{"SSR": g.ssr} if g.ssr >= 0 else g.flags for g in sc.groups]


class Renderer:
    def __init__(self, scene: Scene):
        self.scene = sc = scene
        self.draw = d = CLDraw(1, len(sc.groups), sc.W, sc.H, 1)
        load_scene(d, sc)
        self.shaders = scene_shaders(sc)
        self.mask = [g.skip for g in sc.groups]
        self.cast = [g.cast for g in sc.groups]

    def frame(self, gamma=True, after_cascade=None):
        """One frame; ``after_cascade(i)`` runs after shadow pass i (probes)."""
        sc, d = self.scene, self.draw
        for i, c in enumerate(sc.cascades):
            d.clearShadowMap(i)
            d.shadowMap(i, self.cast, self.shaders, c.bias)
            if after_cascade is not None:
                after_cascade(i)
        d.setPos(sc.view.pos)
        d.setVM(sc.view.vm)
        d.clearZBuffer()
        d.drawAll(self.shaders, mask=self.mask, shadowIds=list(range(len(sc.cascades))),
                  pcf=sc.pcf_taps)
        if gamma:
            d.gamma(sc.exposure)

    def close(self):
        self.draw.close()


class FullFrame(Renderer):
    """The full-frame pipeline on one scene: shadow cascades; the cube map
    pass (the scene from `probe` as six cube_size^2 HDR faces, one batch,
    into reflection cube 0); drawAll, whose SSR water groups draw over the
    opaque frame reflecting that cube; bloom on the HDR frame; gamma."""

    def __init__(self, scene: Scene, probe, cube_size: int = 512, bloom: dict = None):
        from .batch import Batch

        super().__init__(scene)
        self.probe = np.asarray(probe, np.float32)
        self.cube_size = int(cube_size)
        self.bloom = dict(levels=5, threshold=1.0, white=9835.0, strength=1.0)
        if bloom:
            self.bloom.update(bloom)
        self.cube = Batch()
        # the cube pass sees what drawAll's opaque pass draws (no water)
        water = [isinstance(s, dict) for s in self.shaders]
        self.draw.setShaders([0 if w else s for s, w in zip(self.shaders, water)],
                             mask=[m or w for m, w in zip(self.mask, water)])

    def frame(self, gamma=True, after_cascade=None, serial=False):
        """One frame.  The cube pass runs on its own stream beside the opaque
        drawAll (the SSR pass waits for it); serial=True finishes it before
        the drawAll is issued (per-pass timing without the overlap)."""
        sc, d = self.scene, self.draw
        for i, c in enumerate(sc.cascades):
            d.clearShadowMap(i)
            d.shadowMap(i, self.cast, self.shaders, c.bias)
            if after_cascade is not None:
                after_cascade(i)
        self.cube.render_cube(d, 0, self.probe, self.cube_size, shadows=bool(sc.cascades),
                              pcf_taps=sc.pcf_taps)
        if serial:
            self.cube.wait()
        d.setPos(sc.view.pos)
        d.setVM(sc.view.vm)
        d.clearZBuffer()
        d.drawAll(self.shaders, mask=self.mask, shadowIds=list(range(len(sc.cascades))),
                  pcf=sc.pcf_taps)
        d.bloom(**self.bloom)
        if gamma:
            d.gamma(sc.exposure)

    def close(self):
        self.cube.close()
        super().close()
