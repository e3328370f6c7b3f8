"""Oracle pins for the post passes: the reference's blur (Shaders/Post/blur.c,
blur2.c via CLDraw.blur, Ops.py:1009-1020) and the 5-level fp16 bloom.

Each check is fixed by something other than the oracle's own loops: a known
answer of the reference's sample indexing, the sum of the tap table, the
separable impulse response, a ramp through the bilinear upsample, and the
constant-frame fixed point of the pyramid."""
import numpy as np
import pytest

from oracle import pyoracle

TAPS = [0.10125, 0.09815, 0.08941, 0.07654, 0.06157, 0.04654, 0.03306, 0.02207, 0.01384, 0.00816]


def planes(H, W, v=0):
    return [np.full((H, W), v, np.uint16) for _ in range(3)]


def test_blur_zero_frame_unchanged():
    out = pyoracle.blur(*planes(32, 48))
    for p in out:
        assert not p.any()


def test_blur_samples_only_even_rows():
    # blur.c:23-36 reads row 2cy at columns 2cx, 2cx+1, 2cx+1, 2cx+2 (the
    # last wraps to column 0 of row 2cy+1): odd rows, column 0 aside, never
    # reach the half-resolution copy, so such a frame comes back unchanged.
    H, W = 32, 48
    r, g, b = planes(H, W)
    for p in (r, g, b):
        p[1::2, 1:] = 60000
    out = pyoracle.blur(r, g, b)
    for p, q in zip(out, (r, g, b)):
        assert np.array_equal(p, q)
    # column 0 of an odd row does reach it, through the last column's sample
    r[1, 0] = g[1, 0] = b[1, 0] = 60000
    out = pyoracle.blur(r, g, b)
    assert (out[0].astype(np.int64) - r).sum() > 0


def _trunc_safe(x):
    f = x - np.floor(x)
    assert 1e-3 < f < 1 - 1e-3, x  # away from an integer: rounding cannot move trunc
    return int(np.floor(x))


def test_blur_uniform_frame_interior_closed_form():
    # Constant c (a multiple of 4): the 2x2 sample is c, the luminance is
    # c * (0.3626 + 0.5152 + 0.1222) = c, each blur multiplies by the tap sum / 0.9.
    H, W, c = 64, 64, 4000
    out = pyoracle.blur(*planes(H, W, c))
    s = (TAPS[0] + 2 * sum(TAPS[1:])) / 0.9
    d = _trunc_safe(c * c / 65536.0)
    t = _trunc_safe(d * s)
    u = _trunc_safe(t * s)
    for p in out:
        assert int(p[32, 32]) == c + u  # even pixel: the upsample is one texel
        assert int(p[30, 30]) == c + u


def test_blur_coloured_frame_weights_channels_by_luminance():
    # lum = 0.3626 R + 0.5152 G + 0.1222 B (blur.c:31) scales every channel
    H, W = 64, 64
    c = (4000, 2000, 8000)
    planes_ = [np.full((H, W), v, np.uint16) for v in c]
    out = pyoracle.blur(*planes_)
    lum = 0.3626 * c[0] + 0.5152 * c[1] + 0.1222 * c[2]
    s = (TAPS[0] + 2 * sum(TAPS[1:])) / 0.9
    for k in range(3):
        d = _trunc_safe(c[k] * lum / 65536.0)
        u = _trunc_safe(_trunc_safe(d * s) * s)
        assert int(out[k][32, 32]) == c[k] + u, k


def test_blur_impulse_is_separable_with_the_tap_table():
    # One white pixel at (2ry, 2rx) feeds half-res texels (ry, rx) and
    # (ry, rx-1) with d = trunc(q^2 / 65536), q = v/4; the horizontal pass
    # spreads row ry, the vertical pass then spreads each column.
    H, W, ry, rx, v = 64, 96, 12, 24, 65532
    r, g, b = planes(H, W)
    r[2 * ry, 2 * rx] = g[2 * ry, 2 * rx] = b[2 * ry, 2 * rx] = v
    out = pyoracle.blur(r, g, b)
    q = v // 4
    d = int(q * q / 65536.0)
    w = lambda k: TAPS[abs(k)] if abs(k) <= 9 else 0.0
    checked = 0
    for dy in range(-6, 7):
        for dx in range(-7, 8):
            x, y = rx + dx, ry + dy
            t_val = d * (w(x - rx) + w(x - rx + 1)) / 0.9
            if t_val - np.floor(t_val) < 1e-3 or t_val - np.floor(t_val) > 1 - 1e-3:
                continue
            t = int(np.floor(t_val))
            u_val = t * w(y - ry) / 0.9
            if u_val - np.floor(u_val) < 1e-3 or u_val - np.floor(u_val) > 1 - 1e-3:
                continue
            u = int(np.floor(u_val))
            base = int(r[2 * y, 2 * x])
            assert int(out[0][2 * y, 2 * x]) == min(base + u, 65535), (dx, dy)
            checked += 1
    assert checked > 150


@pytest.mark.parametrize("h,w", [(8, 12), (16, 16)])
def test_bilinear_upsample_of_a_ramp(h, w):
    # Sampling at ((x + 0.5) / 2 - 0.5) clamped to [0, side - 1]: a ramp in x
    # comes back as that position, and the transposed ramp as the y position.
    sh, sw = h // 2, w // 2
    ramp_x = np.tile(np.arange(sw, dtype=np.float32), (sh, 1))
    ramp_y = np.tile(np.arange(sh, dtype=np.float32)[:, None], (1, sw))
    px = np.clip(np.arange(w) / 2.0 - 0.25, 0, sw - 1)
    py = np.clip(np.arange(h) / 2.0 - 0.25, 0, sh - 1)
    assert np.array_equal(pyoracle._bilerp_up(ramp_x, h, w), np.tile(px, (h, 1)).astype(np.float32))
    assert np.array_equal(pyoracle._bilerp_up(ramp_y, h, w),
                          np.tile(py[:, None], (1, w)).astype(np.float32))


@pytest.mark.parametrize("levels", [1, 3, 5])
def test_bloom_constant_frame_closed_form(levels):
    # A constant frame keeps every level constant (averages and bilinear
    # weights of a constant are exact), so the added light is the fp16 sum
    # of `levels` copies of max(v / white - threshold, 0).
    H, W, v, white, thr, strength = 64, 96, 20000, 9835.0, 1.0, 0.5
    out = pyoracle.bloom(*planes(H, W, v), levels=levels, threshold=thr, white=white,
                         strength=strength)
    q = np.float16(np.float32(v) / np.float32(white) - np.float32(thr))
    acc = q
    for _ in range(levels - 1):
        acc = np.float16(np.float32(q) + np.float32(acc))
    add = float(acc) * white * strength
    want = min(int(np.floor(v + add)), 65535)
    assert want < 65535
    for p in out:
        assert (p == want).all(), (np.unique(p), want)


def test_bloom_below_threshold_or_zero_strength_is_identity():
    rng = np.random.default_rng(3)
    r, g, b = (rng.integers(0, 9000, (32, 64)).astype(np.uint16) for _ in range(3))
    for kw in (dict(threshold=1.0), dict(threshold=0.0, strength=0.0)):
        out = pyoracle.bloom(r, g, b, levels=4, white=9835.0, **kw)
        for p, q in zip(out, (r, g, b)):
            assert np.array_equal(p, q)


def test_bloom_light_is_local():
    # One bright texel: the added light peaks there and falls off with
    # distance; far corners get nothing.
    H, W = 128, 128
    r, g, b = planes(H, W)
    r[64:66, 64:66] = 60000
    out = pyoracle.bloom(r, g, b, levels=3, threshold=0.5)
    add = out[0].astype(np.int64) - r
    assert add[0, 0] == 0 and add[127, 127] == 0
    y, x = np.unravel_index(add.argmax(), add.shape)
    assert abs(y - 64.5) <= 3 and abs(x - 64.5) <= 3
    assert add[65, 67] > add[65, 70] > add[65, 73] > 0 == add[65, 100]
    assert not out[1].any() and not out[2].any()


def _ulps(a, b):
    a = np.asarray(a, np.float32)
    b = np.asarray(b, np.float64).astype(np.float32)
    return np.abs(a.view(np.int32).astype(np.int64) - b.view(np.int32).astype(np.int64))


def test_log2_decree_exact_at_powers_of_two_and_close_elsewhere():
    k = np.arange(-140, 128)
    x = np.ldexp(np.float32(1), k).astype(np.float32)
    assert np.array_equal(pyoracle.log2_cw(x), k.astype(np.float32))
    rng = np.random.default_rng(5)
    y = (rng.random(200000) * 60 - 30).astype(np.float32)
    x = np.exp2(y.astype(np.float64)).astype(np.float32)
    got = pyoracle.log2_cw(x)
    want = np.log2(x.astype(np.float64))
    far = np.abs(want) > 0.25  # relative ulps are meaningful away from x = 1
    assert _ulps(got[far], want[far]).max() <= 4
    assert np.abs(got[~far] - want[~far]).max() < 1e-7
    s = pyoracle.log2_cw(np.array([0.0, -1.0, np.inf, np.nan], np.float32))
    assert s[0] == -np.inf and np.isnan(s[1]) and s[2] == np.inf and np.isnan(s[3])


def test_exp_decree_close_to_exp():
    rng = np.random.default_rng(6)
    x = (rng.random(200000) * 160 - 79).astype(np.float32)
    got = pyoracle.exp_cw(x)
    assert _ulps(got, np.exp(x.astype(np.float64))).max() <= 4
    s = pyoracle.exp_cw(np.array([0.0, -80.0, -1e4, 89.0, np.nan], np.float32))
    assert s[0] == 1.0 and s[1] == 0.0 and s[2] == 0.0 and s[3] == np.inf and s[4] == 0.0
