"""Seeded synthetic scenes and camera math for the rendering hot path.

Camera conventions follow the reference: ``view_vec``/``view_mat`` restate
``Utils.viewVec``/``Utils.viewMat`` (PyOpenCL27C/Utils.py:8-20), ``look_at``
restates ``ThreeDBackend.lookAt`` (Compute.py:262-265) and ``scale_cull``
restates the fovx-derived projection scale and cull cone of
``ThreeDBackend.__init__`` (Compute.py:127-135).  Geometry conventions follow
the reference's object loaders: per-triangle vertex lists, normals pointing
"inward" (negated, VertObjects.py:655-687 for terrain), texels encoded to u16
as v*v/8 (VertObjects.py:124-126), mip chains built by ``create_mips``
(Utils.py:79-93).

All randomness comes from a counter-based SplitMix64 stream keyed by
(seed, stream, index), and the generators use only +, -, *, / and sqrt on
float64 before the final cast to float32, so a scene is the same on every
machine.
"""
from __future__ import annotations

import dataclasses
from math import atan2, cos, pi, sin, sqrt, tan
from typing import List, Optional

import numpy as np

SH_MIP = 1
SH_SHADOW = 2
SH_PHONG = 4
SH_CULL = 8

_GOLD = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


# --------------------------------------------------------------------- RNG
def _mix(x: np.ndarray) -> np.ndarray:
    with np.errstate(over="ignore"):
        x = (x ^ (x >> np.uint64(30))) * _M1
        x = (x ^ (x >> np.uint64(27))) * _M2
        return x ^ (x >> np.uint64(31))


def rand_u64(seed: int, stream: int, n: int, offset: int = 0) -> np.ndarray:
    """n SplitMix64 outputs of counter stream (seed, stream) from offset."""
    with np.errstate(over="ignore"):
        key = _mix(np.array([seed & 0xFFFFFFFFFFFFFFFF], np.uint64) * _GOLD
                   + np.array([stream & 0xFFFFFFFFFFFFFFFF], np.uint64))
        ctr = np.arange(offset, offset + n, dtype=np.uint64)
        return _mix(key + (ctr + np.uint64(1)) * _GOLD)


def rand_uniform(seed: int, stream: int, n: int, offset: int = 0) -> np.ndarray:
    """Uniform float64 in [0, 1) with 53 random bits."""
    return (rand_u64(seed, stream, n, offset) >> np.uint64(11)).astype(np.float64) * (2.0 ** -53)


# ------------------------------------------------------------ camera math
def view_vec(a: float, b: float) -> np.ndarray:
    """Utils.viewVec: forward direction for yaw a, pitch b."""
    return np.array([sin(a) * cos(b), -sin(b), cos(a) * cos(b)])


def view_mat(a: float, b: float) -> np.ndarray:
    """Utils.viewMat: rows (forward, horizontal, vertical)."""
    a2 = a + pi / 2
    b2 = b - pi / 2
    return np.array([[sin(a) * cos(b), -sin(b), cos(a) * cos(b)],
                     [-sin(a2), 0, -cos(a2)],
                     [-sin(a) * cos(b2), sin(b2), -cos(a) * cos(b2)]])


def look_at(pos, target):
    """Compute.lookAt: yaw/pitch facing target from pos."""
    d = np.asarray(target, np.float64) - np.asarray(pos, np.float64)
    a = pi / 2 - atan2(d[2], d[0])
    b = -atan2(d[1], sqrt(d[0] ** 2 + d[2] ** 2))
    return a, b


def scale_cull(W: int, H: int, fovx: float):
    """(scale, caX, caY) as ThreeDBackend.__init__ derives them from fovx."""
    W2, H2 = int(W / 2), int(H / 2)
    scale = W2 / tan(fovx * pi / 360)
    return scale, W2 / scale + 2, H2 / scale + 2


def create_mips(tex: np.ndarray) -> np.ndarray:
    """Utils.createMips: N x N x C -> (sum 4^k, C), smallest level first."""
    a = np.asarray(tex).astype(np.int64)
    n = a.shape[0]
    if a.shape[1] != n or n & (n - 1):
        raise ValueError("mip textures must be square with a power-of-two side")
    levels = [a]
    while a.shape[0] > 1:
        b = a[:-1:2] + a[1::2]
        a = (b[:, :-1:2] + b[:, 1::2]) >> 2
        levels.append(a)
    levels.reverse()
    ch = a.shape[-1]
    return np.concatenate([x.reshape(-1, ch) for x in levels], axis=0)


def encode_texels(rgb8: np.ndarray) -> np.ndarray:
    """VertObjects.py:124-126 gamma encoding: u16 = v*v/8 (truncated)."""
    t = rgb8.astype(np.float64)
    return (t * t / 8.0).astype(np.uint16)


# ------------------------------------------------------------ containers
@dataclasses.dataclass
class Group:
    """One material group, in the layout CLDraw.addTextureGroup receives."""
    xyz: np.ndarray          # (3T, 3) float32
    vn: np.ndarray           # (3T, 3) float32
    uv: np.ndarray           # (3T, 2) float32
    tex: np.ndarray          # (ntexels, 3) uint16: N*N texels or a mip chain
    tex_side: int            # N
    mip: bool
    flags: int = SH_MIP | SH_PHONG
    mip_bias: float = 0.0
    spec_pow: int = 16
    cast: bool = True
    skip: bool = False
    ssr: int = -1            # >= 0: SSR water reflecting scene.refl[ssr]

    @property
    def ntri(self) -> int:
        return self.xyz.shape[0] // 3

    @property
    def tex_size(self) -> int:
        """texSize as the reference passes it: N, or log2(N) for mips."""
        return int(self.tex_side).bit_length() - 1 if self.mip else int(self.tex_side)


@dataclasses.dataclass
class Camera:
    pos: np.ndarray          # (3,) float32
    vm: np.ndarray           # (3, 3) float32 rows forward/horizontal/vertical
    scale: float
    cax: float = 0.0
    cay: float = 0.0


@dataclasses.dataclass
class Cascade:
    dim: int
    cam: Camera
    bias: float = 0.2


@dataclasses.dataclass
class Scene:
    name: str
    W: int
    H: int
    groups: List[Group]
    view: Camera
    light_int: np.ndarray    # (3,) primary directional light intensity
    light_dir: np.ndarray    # (3,) direction the light travels
    amb: float = 0.2
    cascades: List[Cascade] = dataclasses.field(default_factory=list)
    pcf_taps: int = 5
    exposure: float = 1.0
    clear_rgb: tuple = (1200, 4033, 21675)
    refl: List[np.ndarray] = dataclasses.field(default_factory=list)  # (S, 6S, 3) u16 strips
    ssr_steps: int = 400     # REFL_LENGTH (av_SSR.c:7)
    ssr_step: int = 1        # REFL_STEP
    ssr_absorb: float = 0.3  # ABSORB
    ssr_fresnel: bool = True
    # vertex-lit lights of vertL (Pipe/vert.c:38-88) beside the primary
    # light, which is light 0 and shaded per pixel: directional lights 1..
    # as (I, D), point lights as (I, P), spot lights as (I, D, P), each
    # array (k, 3) float32
    extra_dirs: Optional[tuple] = None
    points: Optional[tuple] = None
    spots: Optional[tuple] = None

    @property
    def ntri(self) -> int:
        return sum(g.ntri for g in self.groups)

    def vert_lights(self) -> dict:
        """vertLight's arguments (Ops.py:552): dirI/dirD with the primary
        light first, then the point and spot lights (None when absent)."""
        dI = [np.asarray(self.light_int, np.float32).reshape(1, 3)]
        dD = [np.asarray(self.light_dir, np.float32).reshape(1, 3)]
        if self.extra_dirs is not None:
            dI.append(np.asarray(self.extra_dirs[0], np.float32).reshape(-1, 3))
            dD.append(np.asarray(self.extra_dirs[1], np.float32).reshape(-1, 3))
        out = {"dirI": np.concatenate(dI), "dirD": np.concatenate(dD)}
        f = lambda a: np.asarray(a, np.float32).reshape(-1, 3)  # noqa: E731
        out["pointI"], out["pointP"] = (None, None) if self.points is None else map(f, self.points)
        out["spotI"], out["spotD"], out["spotP"] = ((None, None, None) if self.spots is None
                                                    else map(f, self.spots))
        return out


def make_camera(W: int, H: int, pos, a: float, b: float, fovx: float = 80.0) -> Camera:
    s, cx, cy = scale_cull(W, H, fovx)
    return Camera(pos=np.asarray(pos, np.float64).astype(np.float32),
                  vm=view_mat(a, b).astype(np.float32),
                  scale=float(np.float32(s)), cax=float(np.float32(cx)),
                  cay=float(np.float32(cy)))


def make_cascade(dim: int, center, extent: float, light_a: float, light_b: float,
                 dist: float = 400.0, bias: float = 0.2) -> Cascade:
    """Ortho shadow camera centred on `center`, covering `extent` units."""
    d = view_vec(light_a, light_b)
    pos = np.asarray(center, np.float64) - dist * d
    cam = Camera(pos=pos.astype(np.float32), vm=view_mat(light_a, light_b).astype(np.float32),
                 scale=float(np.float32(dim / extent)))
    return Cascade(dim=dim, cam=cam, bias=bias)


def noise_texture(seed: int, stream: int, n: int) -> np.ndarray:
    """n x n x 3 RGB8 uniform noise, encoded to u16."""
    v = (rand_u64(seed, stream, n * n * 3) >> np.uint64(56)).astype(np.uint8)
    return encode_texels(v.reshape(n, n, 3))


def texture_for(tex_u16: np.ndarray, mip: bool):
    n = tex_u16.shape[0]
    if mip:
        return create_mips(tex_u16).astype(np.uint16), n
    return tex_u16.reshape(-1, 3).astype(np.uint16), n


# -------------------------------------------------------------- terrain
def value_noise(seed: int, xs: np.ndarray, zs: np.ndarray, octaves: int = 4,
                base_freq: int = 4, amplitude: float = 10.0) -> np.ndarray:
    """Seeded lattice value noise with smoothstep blending, in [-amp, amp].

    xs, zs: coordinates in [0, 1].  Octave o has (base_freq * 2^o) cells and
    weight 0.5^o; the weighted sum is normalised by the total weight."""
    total = np.zeros(np.broadcast(xs, zs).shape)
    wsum = 0.0
    for o in range(octaves):
        f = base_freq << o
        lat = rand_uniform(seed, 1000 + o, (f + 1) * (f + 1)) * 2.0 - 1.0
        lat = lat.reshape(f + 1, f + 1)
        u = np.clip(xs * f, 0, f - 1e-9)
        v = np.clip(zs * f, 0, f - 1e-9)
        iu = np.floor(u).astype(np.int64)
        iv = np.floor(v).astype(np.int64)
        fu = u - iu
        fv = v - iv
        su = fu * fu * (3.0 - 2.0 * fu)
        sv = fv * fv * (3.0 - 2.0 * fv)
        a = lat[iu, iv] * (1 - su) + lat[iu + 1, iv] * su
        b = lat[iu, iv + 1] * (1 - su) + lat[iu + 1, iv + 1] * su
        w = 0.5 ** o
        total += w * (a * (1 - sv) + b * sv)
        wsum += w
    return amplitude * total / wsum


def terrain_mesh(heights: np.ndarray, spacing: float, uvspread: float = 2.0):
    """VertTerrain.create (VertObjects.py:689-700) triangulation and UVs with
    central-difference normals in the reference's inward (-y) convention."""
    n = heights.shape[0]
    i, j = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
    pts = np.stack([i * spacing, heights, j * spacing], axis=-1)
    gx = np.gradient(heights, spacing, axis=0)
    gz = np.gradient(heights, spacing, axis=1)
    nrm = np.stack([gx, -np.ones_like(gx), gz], axis=-1)
    nrm /= np.sqrt((nrm * nrm).sum(-1, keepdims=True))
    ci, cj = np.meshgrid(np.arange(n - 1), np.arange(n - 1), indexing="ij")
    ci = ci.reshape(-1)
    cj = cj.reshape(-1)
    # wedge A: (i,j), (i+1,j), (i,j+1); wedge B: (i+1,j+1), (i,j+1), (i+1,j)
    va = [(ci, cj), (ci + 1, cj), (ci, cj + 1)]
    vb = [(ci + 1, cj + 1), (ci, cj + 1), (ci + 1, cj)]
    ta = [(ci, cj), (ci + 0.99, cj), (ci, cj + 0.99)]
    tb = [(ci + 0.99, cj + 0.99), (ci, cj + 0.99), (ci + 0.99, cj)]
    ncell = ci.shape[0]
    xyz = np.empty((ncell, 2, 3, 3))
    vn = np.empty((ncell, 2, 3, 3))
    uv = np.empty((ncell, 2, 3, 2))
    for k in range(3):
        xyz[:, 0, k] = pts[va[k][0], va[k][1]]
        xyz[:, 1, k] = pts[vb[k][0], vb[k][1]]
        vn[:, 0, k] = nrm[va[k][0], va[k][1]]
        vn[:, 1, k] = nrm[vb[k][0], vb[k][1]]
        uv[:, 0, k] = np.stack(ta[k], -1) / uvspread
        uv[:, 1, k] = np.stack(tb[k], -1) / uvspread
    uv %= 1.0  # texMode "repeat" (VertObjects.py:194-196)
    return (xyz.reshape(-1, 3).astype(np.float32), vn.reshape(-1, 3).astype(np.float32),
            uv.reshape(-1, 2).astype(np.float32))


def terrain_scene(W=640, H=480, grid=128, extent=100.0, seed=1, tex_side=256,
                  mip=True, cam_dist=80.0, cam_yaw=pi / 4, cam_pitch=pi / 6,
                  flags=SH_MIP | SH_PHONG, name="c1_terrain", n_tex=1,
                  shadow_dim=0) -> Scene:
    """Heightmap terrain lit by one directional light (config 1 at defaults:
    128x128 grid = 32,258 triangles, 640x480, 30 degree pitch).  n_tex > 1
    splits the triangles into that many groups of contiguous rows, each with
    its own texture; shadow_dim > 0 adds two shadow cascades and the shadow
    shader flag."""
    t = np.arange(grid) / (grid - 1)
    xs, zs = np.meshgrid(t, t, indexing="ij")
    h = value_noise(seed, xs, zs)
    spacing = extent / (grid - 1)
    xyz, vn, uv = terrain_mesh(h, spacing)
    if shadow_dim:
        flags |= SH_SHADOW
    T = xyz.shape[0] // 3
    cuts = [T * k // n_tex for k in range(n_tex + 1)]
    groups = []
    for k in range(n_tex):
        a, b = 3 * cuts[k], 3 * cuts[k + 1]
        tex, side = texture_for(noise_texture(seed, 77 + k, tex_side), mip)
        groups.append(Group(xyz=np.ascontiguousarray(xyz[a:b]), vn=np.ascontiguousarray(vn[a:b]),
                            uv=np.ascontiguousarray(uv[a:b]), tex=tex, tex_side=side, mip=mip,
                            flags=flags if mip else flags & ~SH_MIP))
    centre = np.array([extent / 2, float(h.mean()), extent / 2])
    pos = centre - cam_dist * view_vec(cam_yaw, cam_pitch)
    la, lb = 0.5, 0.9
    cascades = []
    if shadow_dim:
        cascades = [make_cascade(shadow_dim, centre, extent * 0.6, la, lb),
                    make_cascade(shadow_dim, centre, extent * 1.5, la, lb)]
    return Scene(name=name, W=W, H=H, groups=groups,
                 view=make_camera(W, H, pos, cam_yaw, cam_pitch),
                 light_int=np.array([0.9, 0.85, 0.8], np.float32),
                 light_dir=view_vec(la, lb).astype(np.float32), amb=0.2, cascades=cascades,
                 pcf_taps=9 if shadow_dim else 5)


# -------------------------------------------------------------- spheres
_ICO_V = None


def geodesic_sphere(freq: int = 5):
    """Unit icosphere with each icosahedron face split freq x freq:
    20*freq^2 triangles (500 at freq 5) over 10*freq^2+2 shared vertices.
    Returns (vertices (V,3) float64, faces (F,3) int)."""
    phi = (1.0 + sqrt(5.0)) / 2.0
    v = np.array([[-1, phi, 0], [1, phi, 0], [-1, -phi, 0], [1, -phi, 0],
                  [0, -1, phi], [0, 1, phi], [0, -1, -phi], [0, 1, -phi],
                  [phi, 0, -1], [phi, 0, 1], [-phi, 0, -1], [-phi, 0, 1]], np.float64)
    f = np.array([[0, 11, 5], [0, 5, 1], [0, 1, 7], [0, 7, 10], [0, 10, 11],
                  [1, 5, 9], [5, 11, 4], [11, 10, 2], [10, 7, 6], [7, 1, 8],
                  [3, 9, 4], [3, 4, 2], [3, 2, 6], [3, 6, 8], [3, 8, 9],
                  [4, 9, 5], [2, 4, 11], [6, 2, 10], [8, 6, 7], [9, 8, 1]])
    keys = {}
    verts = []
    faces = []

    def vid(a, b, c, ia, ib, ic):
        # barycentric lattice point; shared edges map to the same key
        key = tuple(sorted(((ia, a), (ib, b), (ic, c))))
        key = tuple(kv for kv in key if kv[0] > 0)
        if key not in keys:
            p = (ia * v[a] + ib * v[b] + ic * v[c]) / freq
            p = p / sqrt(float(p @ p))
            keys[key] = len(verts)
            verts.append(p)
        return keys[key]

    for a, b, c in f:
        idx = {}
        for i in range(freq + 1):
            for j in range(freq + 1 - i):
                k = freq - i - j
                idx[i, j] = vid(a, b, c, k, i, j)
        for i in range(freq):
            for j in range(freq - i):
                faces.append((idx[i, j], idx[i + 1, j], idx[i, j + 1]))
                if i + j < freq - 1:
                    faces.append((idx[i + 1, j], idx[i + 1, j + 1], idx[i, j + 1]))
    return np.array(verts), np.array(faces)


def _rotations(seed: int, n: int) -> np.ndarray:
    """Uniform random rotations from unit quaternions (rejection in the 4-ball)."""
    m = 64
    q = rand_uniform(seed, 2001, n * m * 4).reshape(n, m, 4) * 2.0 - 1.0
    r2 = (q * q).sum(-1)
    ok = (r2 <= 1.0) & (r2 > 1e-6)
    first = np.argmax(ok, axis=1)
    q = q[np.arange(n), first]
    q /= np.sqrt((q * q).sum(-1, keepdims=True))
    w, x, y, z = q.T
    R = np.empty((n, 3, 3))
    R[:, 0, 0] = 1 - 2 * (y * y + z * z)
    R[:, 0, 1] = 2 * (x * y - w * z)
    R[:, 0, 2] = 2 * (x * z + w * y)
    R[:, 1, 0] = 2 * (x * y + w * z)
    R[:, 1, 1] = 1 - 2 * (x * x + z * z)
    R[:, 1, 2] = 2 * (y * z - w * x)
    R[:, 2, 0] = 2 * (x * z - w * y)
    R[:, 2, 1] = 2 * (y * z + w * x)
    R[:, 2, 2] = 1 - 2 * (x * x + y * y)
    return R


def sphere_instances(n_inst: int, seed: int = 1, freq: int = 5, box=200.0,
                     noise: float = 0.05, smin: float = 0.5, smax: float = 3.0, centre=0.0):
    """Seeded icosphere instances: per-vertex radial noise, random scale,
    rotation and position.  Returns per-instance triangle arrays
    (n_inst, F, 3, 3) xyz / normals and (n_inst, F, 3, 2) uv."""
    verts, faces = geodesic_sphere(freq)
    V, F = verts.shape[0], faces.shape[0]
    rad = 1.0 + noise * (rand_uniform(seed, 2002, n_inst * V).reshape(n_inst, V) * 2.0 - 1.0)
    scale = smin + (smax - smin) * rand_uniform(seed, 2003, n_inst)
    pos = (rand_uniform(seed, 2004, n_inst * 3).reshape(n_inst, 3) - 0.5) * np.asarray(box) \
        + np.asarray(centre)
    R = _rotations(seed, n_inst)
    local = verts[None] * rad[..., None] * scale[:, None, None]          # (n, V, 3)
    world = np.einsum("nvk,nlk->nvl", local, R) + pos[:, None, :]
    nrm = -np.einsum("vk,nlk->nvl", verts, R)                              # inward
    uvv = np.stack([0.5 + 0.5 * verts[:, 0], 0.5 + 0.5 * verts[:, 1]], -1)
    xyz = world[:, faces]                                                  # (n, F, 3, 3)
    vn = nrm[:, faces]
    uv = np.broadcast_to(uvv[faces], (n_inst, F, 3, 2))
    return xyz, vn, uv


def spheres_scene(W=1920, H=1080, n_inst=2000, n_tex=8, tex_side=1024, seed=1,
                  shadows=True, shadow_dim=2048, pcf_taps=9, cam_dist=330.0,
                  cam_yaw=0.3, cam_pitch=0.25, flags=None, name="c2_spheres") -> Scene:
    """Config 2 at defaults: 2,000 x 500-triangle icospheres (1,000,000
    triangles) in a 200^3 box, 8 mipmapped 1024^2 textures (instance k uses
    texture k % 8), 1920x1080, two 2048^2 shadow cascades with 3x3 PCF."""
    xyz, vn, uv = sphere_instances(n_inst, seed)
    if flags is None:
        flags = SH_MIP | SH_PHONG | (SH_SHADOW if shadows else 0)
    groups = []
    for t in range(n_tex):
        sel = np.arange(t, n_inst, n_tex)
        tex, side = texture_for(noise_texture(seed, 300 + t, tex_side), True)
        groups.append(Group(xyz=xyz[sel].reshape(-1, 3).astype(np.float32),
                            vn=vn[sel].reshape(-1, 3).astype(np.float32),
                            uv=np.ascontiguousarray(uv[sel]).reshape(-1, 2).astype(np.float32),
                            tex=tex, tex_side=side, mip=True, flags=flags))
    fwd = view_vec(cam_yaw, cam_pitch)
    pos = -cam_dist * fwd
    la, lb = 0.5, 0.9
    cascades = []
    if shadows:
        near_c = pos + 180.0 * fwd
        cascades = [make_cascade(shadow_dim, near_c, 130.0, la, lb),
                    make_cascade(shadow_dim, np.zeros(3), 360.0, la, lb)]
    return Scene(name=name, W=W, H=H, groups=groups,
                 view=make_camera(W, H, pos, cam_yaw, cam_pitch),
                 light_int=np.array([0.9, 0.85, 0.8], np.float32),
                 light_dir=view_vec(la, lb).astype(np.float32), amb=0.2,
                 cascades=cascades, pcf_taps=pcf_taps)


def spheres_closeup_camera(W=1920, H=1080, n_inst=2000, seed=1, gap=4.0, yaw=0.3, pitch=0.25,
                           box=200.0, smin=0.5, smax=3.0) -> Camera:
    """A camera inside the spheres scene's box: `gap` units in front of
    the largest instance within 60 units of the centre, looking at it along
    (yaw, pitch), with the rest of the box behind it.  Its near triangles
    span hundreds of pixels (the tiled path), the far ones a few."""
    scale = smin + (smax - smin) * rand_uniform(seed, 2003, n_inst)
    pos = (rand_uniform(seed, 2004, n_inst * 3).reshape(n_inst, 3) - 0.5) * np.asarray(box)
    near = np.linalg.norm(pos, axis=1) < 60.0
    k = int(np.argmax(np.where(near, scale, -1.0)))
    fwd = view_vec(yaw, pitch)
    eye = pos[k] - (scale[k] * 1.05 + gap) * fwd
    return make_camera(W, H, eye, yaw, pitch)


# -------------------------------------------------------------- water
def water_plane(x0: float, x1: float, z0: float, z1: float, y: float, n: int = 8,
                uv_repeat: float = 4.0):
    """A flat n x n quad grid at height y with inward (-y) normals, split as
    terrain_mesh splits its cells."""
    h = np.full((n + 1, n + 1), float(y))
    xyz, vn, uv = terrain_mesh(h, 1.0, uvspread=(n / uv_repeat))
    xyz = xyz.astype(np.float64)
    xyz[:, 0] = x0 + xyz[:, 0] * ((x1 - x0) / n)
    xyz[:, 2] = z0 + xyz[:, 2] * ((z1 - z0) / n)
    return xyz.astype(np.float32), vn, uv


def cube_strip(seed: int, side: int) -> np.ndarray:
    """A seeded reflection cube strip: side rows of 6 * side u16 texels (a
    smooth sky gradient per face plus noise), as setReflTex receives it."""
    y = np.arange(side)[:, None] / max(side - 1, 1)
    base = np.empty((side, 6 * side, 3))
    for f in range(6):
        for c in range(3):
            base[:, f * side:(f + 1) * side, c] = 2000.0 + 9000.0 * (0.3 + 0.7 * y) * (
                0.5 + 0.5 * ((f + c) % 3) / 2.0)
    n = rand_uniform(seed, 4100, side * 6 * side * 3).reshape(side, 6 * side, 3)
    return np.clip(base * (0.85 + 0.3 * n), 0, 65535).astype(np.uint16)


def cube_face_camera(f: int, pos, size: int) -> Camera:
    """Camera of cube face f (0..5: +x, +y, +z, -x, -y, -z) at pos, 90 degree
    field of view on size x size: forward s e_a, horizontal row -s e_u,
    vertical row s e_v with (u, v) = (z, y), (x, z), (y, x) for a = x, y, z,
    so that texel (x, y) of the face holds the direction the SSR shader's
    cube lookup (Shaders_src/av_SSR.c:107-128) maps to it."""
    sgn = 1.0 if f < 3 else -1.0
    a = f % 3
    u, v = {0: (2, 1), 1: (0, 2), 2: (1, 0)}[a]
    vm = np.zeros((3, 3), np.float32)
    vm[0, a] = sgn
    vm[1, u] = -sgn
    vm[2, v] = sgn
    return Camera(pos=np.asarray(pos, np.float32).reshape(3), vm=vm,
                  scale=float(np.float32(size // 2)), cax=3.0, cay=3.0)


def water_scene(W=320, H=240, grid=48, extent=60.0, seed=1, n_inst=24, refl_side=32,
                shadow_dim=0, water_level=-1.0, ssr_steps=400, name="water") -> Scene:
    """Terrain with a few spheres and an SSR water plane at water_level over
    the whole terrain (the valleys below it are under water), reflecting a
    seeded cube strip."""
    t = np.arange(grid) / (grid - 1)
    xs, zs = np.meshgrid(t, t, indexing="ij")
    h = value_noise(seed, xs, zs)
    spacing = extent / (grid - 1)
    xyz, vn, uv = terrain_mesh(h, spacing)
    flags = SH_MIP | SH_PHONG | (SH_SHADOW if shadow_dim else 0)
    tex, side = texture_for(noise_texture(seed, 77, 64), True)
    groups = [Group(xyz=xyz, vn=vn, uv=uv, tex=tex, tex_side=side, mip=True, flags=flags)]
    if n_inst:
        sx, sv, su = sphere_instances(n_inst, seed, box=extent * 0.8, smin=1.0, smax=3.0)
        sx = sx.reshape(-1, 3) + np.array([extent / 2, 8.0, extent / 2])
        stex, sside = texture_for(noise_texture(seed, 300, 64), True)
        groups.append(Group(xyz=sx.astype(np.float32), vn=sv.reshape(-1, 3).astype(np.float32),
                            uv=np.ascontiguousarray(su).reshape(-1, 2).astype(np.float32),
                            tex=stex, tex_side=sside, mip=True, flags=flags))
    wx, wn, wu = water_plane(0.0, extent, 0.0, extent, water_level, n=8)
    wtex, wside = texture_for(noise_texture(seed, 90, 64), False)
    groups.append(Group(xyz=wx, vn=wn, uv=wu, tex=(wtex // 4 + 500).astype(np.uint16),
                        tex_side=wside, mip=False, flags=0, cast=False, ssr=0))
    centre = np.array([extent / 2, float(h.mean()), extent / 2])
    yaw, pitch = pi / 4, pi / 7
    pos = centre - extent * 0.9 * view_vec(yaw, pitch)
    la, lb = 0.5, 0.9
    cascades = []
    if shadow_dim:
        cascades = [make_cascade(shadow_dim, centre, extent * 0.8, la, lb),
                    make_cascade(shadow_dim, centre, extent * 1.6, la, lb)]
    return Scene(name=name, W=W, H=H, groups=groups,
                 view=make_camera(W, H, pos, yaw, pitch),
                 light_int=np.array([0.9, 0.85, 0.8], np.float32),
                 light_dir=view_vec(la, lb).astype(np.float32), amb=0.2, cascades=cascades,
                 pcf_taps=9 if shadow_dim else 5, refl=[cube_strip(seed, refl_side)],
                 ssr_steps=ssr_steps)


def full_frame_scene(W=2560, H=1440, grid=388, extent=100.0, n_inst=400, n_tex=8,
                     tex_side=1024, terrain_tex=256, shadow_dim=2048, refl_side=512,
                     water_level=-3.0, ssr_steps=64, seed=1, name="c3_full_frame") -> Scene:
    """Config 3 at defaults: the config-1 terrain over the same 100 x 100
    extent upscaled to a 388 x 388 grid (299,538 triangles) plus 400
    config-2 spheres (200,000 triangles, 8 mipmapped 1024^2 textures) above
    it, an SSR water plane at water_level over the terrain (64 march steps),
    a 512^2-per-face reflection cube, two 2048^2 shadow cascades with 3x3
    PCF, 2560x1440.  The cube strip here is a seeded placeholder: the
    pipeline renders the cube from the scene every frame."""
    t = np.arange(grid) / (grid - 1)
    xs, zs = np.meshgrid(t, t, indexing="ij")
    h = value_noise(seed, xs, zs)
    spacing = extent / (grid - 1)
    xyz, vn, uv = terrain_mesh(h, spacing)
    flags = SH_MIP | SH_PHONG | (SH_SHADOW if shadow_dim else 0)
    tex, side = texture_for(noise_texture(seed, 77, terrain_tex), True)
    groups = [Group(xyz=xyz, vn=vn, uv=uv, tex=tex, tex_side=side, mip=True, flags=flags)]
    sx, sv, su = sphere_instances(n_inst, seed, box=(2 * extent, 30.0, 2 * extent),
                                  centre=(extent / 2, 18.0, extent / 2))
    for k in range(n_tex):
        sel = np.arange(k, n_inst, n_tex)
        stex, sside = texture_for(noise_texture(seed, 300 + k, tex_side), True)
        groups.append(Group(xyz=sx[sel].reshape(-1, 3).astype(np.float32),
                            vn=sv[sel].reshape(-1, 3).astype(np.float32),
                            uv=np.ascontiguousarray(su[sel]).reshape(-1, 2).astype(np.float32),
                            tex=stex, tex_side=sside, mip=True, flags=flags))
    wx, wn, wu = water_plane(-extent, 2 * extent, -extent, 2 * extent, water_level, n=16)
    wtex, wside = texture_for(noise_texture(seed, 90, 256), False)
    groups.append(Group(xyz=wx, vn=wn, uv=wu, tex=(wtex // 4 + 500).astype(np.uint16),
                        tex_side=wside, mip=False, flags=0, cast=False, ssr=0))
    centre = np.array([extent / 2, float(h.mean()), extent / 2])
    yaw, pitch = pi / 4, pi / 6
    pos = centre - 80.0 * view_vec(yaw, pitch)
    la, lb = 0.5, 0.9
    cascades = []
    if shadow_dim:
        cascades = [make_cascade(shadow_dim, centre, extent * 0.6, la, lb),
                    make_cascade(shadow_dim, centre, extent * 1.5, la, lb)]
    return Scene(name=name, W=W, H=H, groups=groups,
                 view=make_camera(W, H, pos, yaw, pitch),
                 light_int=np.array([0.9, 0.85, 0.8], np.float32),
                 light_dir=view_vec(la, lb).astype(np.float32), amb=0.2, cascades=cascades,
                 pcf_taps=9 if shadow_dim else 5, refl=[cube_strip(seed, refl_side)],
                 ssr_steps=ssr_steps)


def full_frame_probe(sc: Scene) -> np.ndarray:
    """Cube map probe of a full-frame scene: above the terrain centre, 6
    units over the water."""
    w = [g for g in sc.groups if g.ssr >= 0][0]
    ext = float(sc.groups[0].xyz[:, 0].max())
    return np.array([ext / 2, float(w.xyz[0, 1]) + 6.0, ext / 2], np.float32)


def random_tris_scene(W=256, H=192, ntri=400, seed=3, size=30.0, depth=(5.0, 40.0),
                      tex_side=64, mip=True, flags=SH_MIP | SH_PHONG, ngroups=2,
                      name="random_tris") -> Scene:
    """Random overlapping triangles in front of a camera at the origin
    looking down +z: mixes small and large triangles, depth ties, and
    vertices off screen; used by parity tests."""
    u = rand_uniform(seed, 4001, ntri * 9).reshape(ntri, 3, 3)
    centre = np.empty((ntri, 1, 3))
    cu = rand_uniform(seed, 4002, ntri * 3).reshape(ntri, 3)
    centre[:, 0, 0] = (cu[:, 0] - 0.5) * 2.2 * size
    centre[:, 0, 1] = (cu[:, 1] - 0.5) * 1.6 * size
    centre[:, 0, 2] = depth[0] + (depth[1] - depth[0]) * cu[:, 2]
    spread = np.where(rand_uniform(seed, 4003, ntri) < 0.3, size * 0.6, size * 0.08)
    xyz = centre + (u - 0.5) * spread[:, None, None]
    xyz[:, :, 2] = np.maximum(xyz[:, :, 2], 1.0)
    # a few exact duplicates to force depth ties
    dup = ntri // 10
    xyz[ntri - dup:] = xyz[:dup]
    nv = rand_uniform(seed, 4004, ntri * 9).reshape(ntri, 3, 3) - 0.5
    nv /= np.sqrt((nv * nv).sum(-1, keepdims=True))
    uv = rand_uniform(seed, 4005, ntri * 6).reshape(ntri, 3, 2)
    groups = []
    per = ntri // ngroups
    for k in range(ngroups):
        lo, hi = k * per, (ntri if k == ngroups - 1 else (k + 1) * per)
        tex, side = texture_for(noise_texture(seed, 500 + k, tex_side), mip)
        groups.append(Group(xyz=xyz[lo:hi].reshape(-1, 3).astype(np.float32),
                            vn=nv[lo:hi].reshape(-1, 3).astype(np.float32),
                            uv=uv[lo:hi].reshape(-1, 2).astype(np.float32),
                            tex=tex, tex_side=side, mip=mip,
                            flags=flags if mip else flags & ~SH_MIP))
    return Scene(name=name, W=W, H=H, groups=groups,
                 view=make_camera(W, H, [0.0, 0.0, 0.0], 0.0, 0.0, fovx=90.0),
                 light_int=np.array([0.9, 0.85, 0.8], np.float32),
                 light_dir=view_vec(0.4, 0.7).astype(np.float32), amb=0.2)


# -------------------------------------------------------------- serving mix
def serving_scenes(n: int = 16, seed: int = 1, tex_side: int = 1024, shadow_dim: int = 2048,
                   tri_range=(50_000, 1_000_000)):
    """The serving workload's resident scenes: n scenes of 50k-1M triangles,
    alternately terrain grids and icosphere instances, each with 1-4 seeded
    mipmapped textures and two shadow cascades.  Returns [(scene, kind)]."""
    u = rand_uniform(seed, 5001, 2 * n).reshape(n, 2)
    out = []
    for i in range(n):
        target = int(tri_range[0] + u[i, 0] * (tri_range[1] - tri_range[0]))
        ntex = 1 + min(3, int(u[i, 1] * 4))
        if i % 2 == 0:
            grid = max(8, int(round(sqrt(target / 2.0))) + 1)
            sc = terrain_scene(grid=grid, tex_side=tex_side, seed=seed + i, n_tex=ntex,
                               shadow_dim=shadow_dim, name=f"serve{i}_terrain")
            out.append((sc, "terrain"))
        else:
            sc = spheres_scene(n_inst=max(1, target // 500), n_tex=ntex, tex_side=tex_side,
                               seed=seed + i, shadow_dim=shadow_dim, name=f"serve{i}_spheres")
            out.append((sc, "spheres"))
    return out


def serving_camera(kind: str, w: int, h: int, u) -> Camera:
    """A request's camera for a scene of the given kind; u: 4 uniforms."""
    if kind == "terrain":
        centre = np.array([50.0, 0.0, 50.0])
        yaw = 2 * pi * u[0]
        pitch = 0.35 + 0.4 * u[1]
        pos = centre - (40.0 + 80.0 * u[2]) * view_vec(yaw, pitch)
        return make_camera(w, h, pos, yaw, pitch)
    yaw = 2 * pi * u[0]
    pitch = -0.2 + 0.7 * u[1]
    pos = -(150.0 + 200.0 * u[2]) * view_vec(yaw, pitch)
    return make_camera(w, h, pos, yaw, pitch)
