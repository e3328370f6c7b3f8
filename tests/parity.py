"""Shared parity checks: HIP path vs oracle on one scene."""
import dataclasses

import numpy as np

from axi_visualizer_amd.render import Renderer
from oracle import pyoracle

RGB_TOL = 1          # per channel, in 8-bit units after gamma
RGB_MIN_FRAC = 0.999  # fraction of pixels within RGB_TOL


def render_gpu(scene, hdr=False, caps=None):
    r = Renderer(scene)
    try:
        if caps:
            r.draw.setBinCapacity(**caps)
        r.frame(gamma=not hdr)
        out = {"id": r.draw.getIds()}
        rr, gg, bb, d = r.draw.getFrame()
        out.update(r=rr, g=gg, b=bb, depth=d, stats=r.draw.stats())
        if scene.cascades:
            out["shadow"] = [r.draw.getSHM(i) for i in range(len(scene.cascades))]
    finally:
        r.close()
    return out


def compare(scene, gpu, ref, label=""):
    """Asserts the bit-exact and tolerance contracts; returns a report."""
    assert gpu["stats"]["bin_overflow"] == 0, f"{label}: bin pool overflow {gpu['stats']}"
    ids_eq = np.array_equal(gpu["id"], ref["id"])
    if not ids_eq:
        bad = np.argwhere(gpu["id"] != ref["id"])
        y, x = bad[0]
        raise AssertionError(
            f"{label}: winner id differs at {len(bad)} px, first ({x},{y}): "
            f"gpu {gpu['id'][y, x]} oracle {ref['id'][y, x]}")
    dg = gpu["depth"].view(np.uint32)
    dr = ref["depth"].view(np.uint32)
    if not np.array_equal(dg, dr):
        bad = np.argwhere(dg != dr)
        y, x = bad[0]
        raise AssertionError(f"{label}: depth differs at {len(bad)} px, first ({x},{y}): "
                             f"gpu {gpu['depth'][y, x]!r} oracle {ref['depth'][y, x]!r}")
    diff = np.abs(np.stack([gpu["r"], gpu["g"], gpu["b"]], -1).astype(np.int64) -
                  np.stack([ref["r"], ref["g"], ref["b"]], -1).astype(np.int64)).max(-1)
    frac = float((diff <= RGB_TOL).mean())
    worst = int(diff.max())
    assert frac >= RGB_MIN_FRAC, f"{label}: only {frac:.6f} of px within {RGB_TOL}/255 (worst {worst})"
    if "shadow" in gpu and "shadow" in ref:
        for i, (a, b) in enumerate(zip(gpu["shadow"], ref["shadow"])):
            assert np.array_equal(a.view(np.uint32), b.view(np.uint32)), \
                f"{label}: shadow map {i} differs at {(a.view(np.uint32) != b.view(np.uint32)).sum()} texels"
    covered = float((ref["id"] != 0xFFFFFFFF).mean())
    return {"frac_within": frac, "worst": worst, "covered": covered,
            "exact_rgb": float((diff == 0).mean())}


def check_scene(scene, label=None, hdr=False, caps=None):
    label = label or scene.name
    if hdr:
        scene.exposure = 0.0
    gpu = render_gpu(scene, hdr=hdr, caps=caps)
    ref = pyoracle.render(scene, with_shadows_out=bool(scene.cascades))
    rep = compare(scene, gpu, ref, label)
    print(f"[parity] {label}: {rep}", flush=True)
    return rep


def job_scene(base, job):
    """A batch job's view of its scene as a plain Scene (for CLDraw and the
    oracle): the job's camera, resolution, shadow switch, PCF and exposure."""
    from axi_visualizer_amd import scenes

    groups = base.groups
    cascades = base.cascades
    if not job.shadows:
        groups = [dataclasses.replace(g, flags=g.flags & ~scenes.SH_SHADOW) for g in groups]
        cascades = []
    return dataclasses.replace(base, W=job.width, H=job.height, view=job.camera, groups=groups,
                               cascades=cascades, pcf_taps=job.pcf_taps,
                               exposure=job.exposure, name=f"view_{job.width}x{job.height}")


def rgba_planes(rgba):
    """uint32 RGBA8 (r | g<<8 | b<<16 | a<<24) [H, W] -> r, g, b uint16 planes."""
    return [((rgba >> s) & 0xFF).astype(np.uint16) for s in (0, 8, 16)]


def compare_rgba(label, rgba, depth, ref, ids=None):
    """A job's RGBA8 frame and depth (and ids) against an oracle render:
    ids and depth bit-exact, RGB within RGB_TOL on >= RGB_MIN_FRAC of
    pixels; returns the worst error."""
    if ids is not None:
        bad = np.argwhere(ids != ref["id"])
        assert bad.size == 0, f"{label}: winner id differs at {len(bad)} px, first {bad[:3].tolist()}"
    bad = np.argwhere(depth.view(np.uint32) != ref["depth"].view(np.uint32))
    assert bad.size == 0, f"{label}: depth differs at {len(bad)} px, first {bad[:3].tolist()}"
    got = np.stack(rgba_planes(rgba), -1).astype(np.int64)
    want = np.stack([ref["r"], ref["g"], ref["b"]], -1).astype(np.int64)
    diff = np.abs(got - want).max(-1)
    frac = float((diff <= RGB_TOL).mean())
    worst = int(diff.max())
    print(f"[parity] {label}: rgb<=1/255 on {frac:.6f}, exact {float((diff == 0).mean()):.6f}, "
          f"worst {worst}, covered {float((ref['id'] != 0xFFFFFFFF).mean()):.4f}", flush=True)
    assert frac >= RGB_MIN_FRAC, f"{label}: only {frac:.6f} within {RGB_TOL}/255 (worst {worst})"
    return worst
