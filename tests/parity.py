"""Shared parity checks: HIP path vs oracle on one scene."""
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
