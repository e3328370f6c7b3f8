// axr_tri.h -- per-triangle device math: setup, scanline walks, replay, shade.
//
// One definition of each walk serves two purposes: the depth pass
// enumerates fragments (ax, cy, tz, half) and issues 64-bit key minima; the
// shade pass replays the same walk for a single pixel to rebuild the
// winning fragment's interpolants.  Both execute the identical instruction
// sequence on identical inputs, so they agree bit for bit.
#pragma once
#include "axr_common.h"

namespace axr {

struct V3 {
  float x, y, z;
};
AXR_HD V3 v3(float x, float y, float z) { return {x, y, z}; }
AXR_HD V3 sub(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ V3 add(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ __forceinline__ V3 scl(V3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
AXR_HD float dot(V3 a, V3 b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
__device__ __forceinline__ V3 normalize(V3 v) { return scl(v, 1.0f / sqrtf(dot(v, v))); }
__device__ __forceinline__ float pown(float x, int n) {
  float r = 1.0f, b = x;
  unsigned e = (unsigned)(n < 0 ? 0 : n);
  while (e) {
    if (e & 1u) r = r * b;
    b = b * b;
    e >>= 1;
  }
  return r;
}
AXR_HD V3 cam_row(const AxrCam& c, int r) {
  return {c.vm[3 * r], c.vm[3 * r + 1], c.vm[3 * r + 2]};
}

// ---------------------------------------------------- chunk visibility
// Triangles are grouped in chunks of 64 consecutive global indices (one
// wave of a setup launch) with a bounding sphere (centre, radius) per chunk.
// A chunk is skipped when no triangle in it can pass the setup's accept
// test: trisetup accepts only triangles whose three vertices all lie in
// front of the eye and inside the cone |dx/dd| < caX, |dy/dd| < caY
// (Pipe/trisetup.c:45-47); the ortho setup only those inside the guard
// band with positive biased depth (Pipe/trisetupOrtho.c:49-57).  The
// sphere must lie beyond one of those planes by a margin of 1e-4 of the
// coordinate magnitudes -- far above the rounding of the setup's float
// arithmetic -- so a skipped triangle is one the setup would have rejected.

__device__ __forceinline__ float v3len(V3 a) { return sqrtf(dot(a, a)); }
// The L1 norm, an upper bound of the length: margins use it, so the tests
// take no square root per chunk.
AXR_HD float l1norm(V3 a) { return fabsf(a.x) + fabsf(a.y) + fabsf(a.z); }

// True when no point of the sphere can pass the perspective accept test.
// k: the camera's cull constants (make_cull); a block with ok == 0 skips
// nothing.
AXR_HD bool chunk_outside_view(const float4& sph, const AxrCam& cam, const AxrCull& k) {
  const float r = sph.w;
  if (!k.ok || !(r >= 0.0f)) return false;  // no constants, or no bounds (NaN input): keep
  const V3 c = v3(sph.x, sph.y, sph.z), eye = v3(cam.pos[0], cam.pos[1], cam.pos[2]);
  const V3 x = sub(c, eye);
  const float m = 1e-4f * (l1norm(c) + k.l1eye + l1norm(x) + 2.0f * r) + 1e-30f;
  if (dot(x, cam_row(cam, 0)) + r * k.nf <= -m) return true;  // behind the eye
#pragma unroll
  for (int i = 0; i < 4; i++)
    if (dot(x, v3(k.n[i][0], k.n[i][1], k.n[i][2])) - r * k.nn[i] >= m)
      return true;  // beyond a side plane
  return false;
}

// True when no point of the sphere can pass the ortho accept test.
AXR_HD bool chunk_outside_ortho(const float4& sph, const AxrCam& cam, const AxrCull& k,
                                int dim, float bias) {
  const float r = sph.w;
  if (!k.ok || !(r >= 0.0f)) return false;
  const V3 c = v3(sph.x, sph.y, sph.z), eye = v3(cam.pos[0], cam.pos[1], cam.pos[2]);
  const V3 x = sub(c, eye);
  const float mag = l1norm(c) + k.l1eye + l1norm(x) + 2.0f * r;
  const V3 F = cam_row(cam, 0), R = cam_row(cam, 1), U = cam_row(cam, 2);
  if (dot(x, F) + r * k.nf + bias <= -(1e-4f * mag + fabsf(bias) * 1e-4f + 1e-30f))
    return true;  // every point has biased depth <= 0
  const float s = fabsf(cam.scale), half = (float)(dim / 2);
  const float mp = 1.0f + 1e-4f * (mag * s + (float)dim);  // pixels
  // projected coordinate ranges of the sphere: x * scale + half, y * -scale + half
  const float cxp = dot(x, R) * cam.scale + half, rx = r * k.nrs;
  const float cyp = dot(x, U) * -cam.scale + half, ry = r * k.nus;
  const float lo = -121.0f - mp, hi = (float)(dim + 120) + mp;
  if (cxp + rx <= lo || cxp - rx >= hi) return true;
  if (cyp + ry <= lo || cyp - ry >= hi) return true;
  return false;
}

// ------------------------------------------------------------- attributes
// Perspective-premultiplied vertex attributes (AVSL.py:414-422): UV, I, N, PXYZ.
struct Attr {
  float u, v, ix, iy, iz, nx, ny, nz, px, py, pz;
};
#define AXR_ATTR_FIELDS(M) M(u) M(v) M(ix) M(iy) M(iz) M(nx) M(ny) M(nz) M(px) M(py) M(pz)

__device__ __forceinline__ Attr a_scale(const Attr& a, float s) {
  Attr r;
#define M(f) r.f = a.f * s;
  AXR_ATTR_FIELDS(M)
#undef M
  return r;
}
// (a - b) * s
__device__ __forceinline__ Attr a_slope(const Attr& a, const Attr& b, float s) {
  Attr r;
#define M(f) r.f = (a.f - b.f) * s;
  AXR_ATTR_FIELDS(M)
#undef M
  return r;
}
// -(a - b) * s
__device__ __forceinline__ Attr a_nslope(const Attr& a, const Attr& b, float s) {
  Attr r;
#define M(f) r.f = (-(a.f - b.f)) * s;
  AXR_ATTR_FIELDS(M)
#undef M
  return r;
}
// a1 + s * (a3 - a1)
__device__ __forceinline__ Attr a_lerp4(const Attr& a1, const Attr& a3, float s) {
  Attr r;
#define M(f) r.f = a1.f + s * (a3.f - a1.f);
  AXR_ATTR_FIELDS(M)
#undef M
  return r;
}
// a + d * s
__device__ __forceinline__ Attr a_mad(const Attr& a, float d, const Attr& s) {
  Attr r;
#define M(f) r.f = a.f + d * s.f;
  AXR_ATTR_FIELDS(M)
#undef M
  return r;
}
__device__ __forceinline__ void a_inc(Attr& a, const Attr& s) {
#define M(f) a.f += s.f;
  AXR_ATTR_FIELDS(M)
#undef M
}
__device__ __forceinline__ void a_dec(Attr& a, const Attr& s) {
#define M(f) a.f -= s.f;
  AXR_ATTR_FIELDS(M)
#undef M
}
template <class T>
AXR_HD void swp(T& a, T& b) {
  T t = a;
  a = b;
  b = t;
}

// Vertex k of global triangle gi, premultiplied by its reciprocal depth z.
// vpos: (x,y,z,0); vattr[2v] = (nx,ny,nz,u), vattr[2v+1] = (v,Ix,Iy,Iz).
__device__ __forceinline__ Attr load_attr(const float4* __restrict__ vpos,
                                          const float4* __restrict__ vattr, uint32_t vtx,
                                          float z) {
  float4 p = vpos[vtx];
  float4 a0 = vattr[2 * (size_t)vtx];
  float4 a1 = vattr[2 * (size_t)vtx + 1];
  Attr r;
  r.u = a0.w * z;
  r.v = a1.x * z;
  r.ix = a1.y * z;
  r.iy = a1.z * z;
  r.iz = a1.w * z;
  r.nx = a0.x * z;
  r.ny = a0.y * z;
  r.nz = a0.z * z;
  r.px = p.x * z;
  r.py = p.y * z;
  r.pz = p.z * z;
  return r;
}

// ------------------------------------------------------------ a5: setup
// d[k] = 1 / d[k] for the three vertex depths, correctly rounded: rcp_rn
// when every |d| lies in [2^-60, 2^60] (where it equals IEEE division),
// IEEE division, kept inside the branch by the empty asm, otherwise.
__device__ __forceinline__ void recip3(float (&d)[3]) {
  bool ok = true;
#pragma unroll
  for (int k = 0; k < 3; k++) {
    const float a = fabsf(d[k]);
    ok &= (a >= 0x1p-60f) & (a <= 0x1p60f);
  }
  if (ok) {
#pragma unroll
    for (int k = 0; k < 3; k++) d[k] = rcp_rn(d[k]);
  } else {
    float b0 = d[0], b1 = d[1], b2 = d[2];
    asm volatile("" : "+v"(b0), "+v"(b1), "+v"(b2));
    d[0] = 1.0f / b0;
    d[1] = 1.0f / b1;
    d[2] = 1.0f / b2;
  }
}

// Pipe/trisetup.c:28-80 (+ trisetup_cull.c:49-54).  0 reject, 1 small, 2 large.
__device__ __forceinline__ int tri_setup(const float4 (&vp4)[3],
                                         const float4* __restrict__ vattr, uint32_t gi,
                                         const AxrCam& cam, int W, int H, bool cull, float px[3],
                                         float py[3], float zz[3]) {
  const V3 vp = v3(cam.pos[0], cam.pos[1], cam.pos[2]);
  const V3 SVd = cam_row(cam, 0), SVx = cam_row(cam, 1), SVy = cam_row(cam, 2);
  V3 x[3];
  float dd[3], dx[3], dy[3];
#pragma unroll
  for (int k = 0; k < 3; k++) {
    const float4 p = vp4[k];
    x[k] = sub(v3(p.x, p.y, p.z), vp);
    dd[k] = dot(x[k], SVd);
    dx[k] = dot(x[k], SVx);
    dy[k] = dot(x[k], SVy);
  }
  recip3(dd);
#pragma unroll
  for (int k = 0; k < 3; k++) {
    dx[k] = dx[k] * dd[k];
    dy[k] = dy[k] * dd[k];
  }
  bool yes = dd[0] > 0 && dd[1] > 0 && dd[2] > 0 && fabsf(dx[0]) < cam.cax &&
             fabsf(dx[1]) < cam.cax && fabsf(dx[2]) < cam.cax && fabsf(dy[0]) < cam.cay &&
             fabsf(dy[1]) < cam.cay && fabsf(dy[2]) < cam.cay;
  if (cull) {
    bool back = true;
#pragma unroll
    for (int k = 0; k < 3; k++) {
      float4 n = vattr[2 * (3 * (size_t)gi + k)];
      if (!(dot(v3(n.x, n.y, n.z), x[k]) < -0.0f)) back = false;
    }
    if (back) yes = false;
  }
  const float hw = (float)(W / 2), hh = (float)(H / 2);
  int dix[3], diy[3];
#pragma unroll
  for (int k = 0; k < 3; k++) {
    px[k] = dx[k] * -cam.scale + hw;
    py[k] = dy[k] * cam.scale + hh;
    zz[k] = dd[k];
    dix[k] = f2i(px[k]);
    diy[k] = f2i(py[k]);
  }
  long long maxX = imax(dix[0], imax(dix[1], dix[2]));
  long long minX = imin(dix[0], imin(dix[1], dix[2]));
  long long maxY = imax(diy[0], imax(diy[1], diy[2]));
  long long minY = imin(diy[0], imin(diy[1], diy[2]));
  long long dimx = maxX - minX, dimy = maxY - minY;
  if (dimx == 0 || dimx > 2LL * W) yes = false;
  if (dimy == 0 || dimy > 2LL * H) yes = false;
  if (maxX < 0 || maxY < 0 || minX > W || minY > H) yes = false;
  if (!yes) return 0;
  return (dimx < 96 && dimy < 96) ? 1 : 2;  // SMALL_THRESH * TILE_SIZE
}

// ------------------------------------------------- sorted triangle shapes
// Sort by y with strict '>' swaps in the order (1,2), (2,3), (1,2).
// vk[] records which input vertex landed in each sorted slot.
template <class T>
struct Geom {
  T x1, y1, x2, y2, x3, y3;
  float z1, z2, z3;
  int k1, k2, k3;
  float yd, x4, z4;  // 4th vertex on the long edge at y2
};

template <class T>
__device__ __forceinline__ void geom_sort(Geom<T>& g) {
  if (g.y1 > g.y2) { swp(g.y1, g.y2); swp(g.x1, g.x2); swp(g.z1, g.z2); swp(g.k1, g.k2); }
  if (g.y2 > g.y3) { swp(g.y2, g.y3); swp(g.x2, g.x3); swp(g.z2, g.z3); swp(g.k2, g.k3); }
  if (g.y1 > g.y2) { swp(g.y1, g.y2); swp(g.x1, g.x2); swp(g.z1, g.z2); swp(g.k1, g.k2); }
}

// Small path (Templates.py:230-246): vertices snapped to int.
__device__ __forceinline__ Geom<int> small_geom(const float px[3], const float py[3],
                                                const float z[3]) {
  Geom<int> g;
  g.x1 = f2i(px[0]); g.y1 = f2i(py[0]); g.z1 = z[0]; g.k1 = 0;
  g.x2 = f2i(px[1]); g.y2 = f2i(py[1]); g.z2 = z[1]; g.k2 = 1;
  g.x3 = f2i(px[2]); g.y3 = f2i(py[2]); g.z3 = z[2]; g.k3 = 2;
  geom_sort(g);
  // integer row counts below 2^24: the Markstein quotient is IEEE's
  // (y3 > y1 for every triangle setup accepts: its int box has height)
  const float D = (float)(g.y3 - g.y1);
  g.yd = div_rn((float)(g.y2 - g.y1), D, rcp_rn(D));
  g.x4 = (float)g.x1 + g.yd * (float)(g.x3 - g.x1);
  g.z4 = g.z1 + g.yd * (g.z3 - g.z1);
  return g;
}

// Tiled path (Templates.py:54-68): float vertices.
__device__ __forceinline__ Geom<float> tile_geom(const float px[3], const float py[3],
                                                 const float z[3]) {
  Geom<float> g;
  g.x1 = px[0]; g.y1 = py[0]; g.z1 = z[0]; g.k1 = 0;
  g.x2 = px[1]; g.y2 = py[1]; g.z2 = z[1]; g.k2 = 1;
  g.x3 = px[2]; g.y3 = py[2]; g.z3 = z[2]; g.k3 = 2;
  geom_sort(g);
  g.yd = (g.y2 - g.y1) / (g.y3 - g.y1);
  g.x4 = g.x1 + g.yd * (g.x3 - g.x1);
  g.z4 = g.z1 + g.yd * (g.z3 - g.z1);
  return g;
}

// Edge slopes of one half, swapped so slope1 >= slope2 (Templates.py:95-111, 141-156).
struct Edges {
  float s1, s2, sz1, sz2, inv;
  bool swapped;
};
// 1/D for an integer row count D in [0, 2^24]: rcp_rn is IEEE's reciprocal
// there, and D = 0 (a half of one row) gives IEEE's +inf.
__device__ __forceinline__ float rcp_rows(int d) {
  const float D = (float)d;
  return d == 0 ? __builtin_inff() : rcp_rn(D);
}
__device__ __forceinline__ Edges small_edges(const Geom<int>& g, int half) {
  Edges e;
  if (half == 0) {
    e.inv = rcp_rows(g.y2 - g.y1);
    e.s1 = (float)(g.x2 - g.x1) * e.inv;
    e.s2 = (g.x4 - (float)g.x1) * e.inv;
    e.sz1 = (g.z2 - g.z1) * e.inv;
    e.sz2 = (g.z4 - g.z1) * e.inv;
  } else {
    e.inv = rcp_rows(g.y3 - g.y2);
    e.s1 = (float)(g.x3 - g.x2) * e.inv;
    e.s2 = ((float)g.x3 - g.x4) * e.inv;
    e.sz1 = (g.z3 - g.z2) * e.inv;
    e.sz2 = (g.z3 - g.z4) * e.inv;
  }
  e.swapped = e.s1 < e.s2;
  if (e.swapped) {
    swp(e.s1, e.s2);
    swp(e.sz1, e.sz2);
  }
  return e;
}
__device__ __forceinline__ Edges tile_edges(const Geom<float>& g, int half) {
  Edges e;
  if (half == 0) {
    e.inv = 1 / (g.y2 - g.y1);
    e.s1 = (g.x2 - g.x1) * e.inv;
    e.s2 = (g.x4 - g.x1) * e.inv;
    e.sz1 = (g.z2 - g.z1) * e.inv;
    e.sz2 = (g.z4 - g.z1) * e.inv;
  } else {
    e.inv = 1 / (g.y3 - g.y2);
    e.s1 = (g.x3 - g.x2) * e.inv;
    e.s2 = (g.x3 - g.x4) * e.inv;
    e.sz1 = (g.z3 - g.z2) * e.inv;
    e.sz2 = (g.z3 - g.z4) * e.inv;
  }
  e.swapped = e.s1 < e.s2;
  if (e.swapped) {
    swp(e.s1, e.s2);
    swp(e.sz1, e.sz2);
  }
  return e;
}

__device__ __forceinline__ float span_t(int ax, float cx1, float cx2) {
  float t = ((float)ax - cx2) / (cx1 - cx2);
  return omax(0.0f, omin(1.0f, t));
}
__device__ __forceinline__ float span_tz(float t, float cz1, float cz2) {
  return 1 / ((1 - t) * cz2 + t * cz1);
}

// A row whose span quotients and depth reciprocals stay where div_rn and
// rcp_rn are exact: |cx1 - cx2| in [2^-30, 2^30]; cx2 zero or |cx2| in
// [2^-7, 2^29], so every nonzero ax - cx2 (ax an integer below 2^24 in
// magnitude) lies in [2^-30, 2^30]; cz1, cz2 of one sign with magnitudes in
// [2^-59, 2^59], so their convex combinations lie in [2^-60, 2^60].
__device__ __forceinline__ bool rn_row(int xa, int xb, float cx1, float cx2, float cz1, float cz2) {
  const float d = fabsf(cx1 - cx2), c = fabsf(cx2), z1 = fabsf(cz1), z2 = fabsf(cz2);
  const bool dx = d >= 0x1p-30f && d <= 0x1p30f;
  const bool x2 = cx2 == 0.0f || (c >= 0x1p-7f && c <= 0x1p29f);
  const bool ax = xa >= -(1 << 24) && xb <= (1 << 24);
  const bool zs = (cz1 > 0.0f) == (cz2 > 0.0f) && z1 >= 0x1p-59f && z1 <= 0x1p59f &&
                  z2 >= 0x1p-59f && z2 <= 0x1p59f;
  return dx && x2 && ax && zs;
}

// The texels xa..xb of row cy with depth tz = span_tz(span_t(ax)).  Rows
// inside rn_row's ranges take the span width's reciprocal once and two
// fused steps a texel, and hand frag_key(ax, cy, ordf(tz)) the ordered key
// directly: tz there is finite, nonzero and of the sign cz1 and cz2 share,
// so ordf is one xor.  Other rows divide in IEEE arithmetic and call
// frag(ax, cy, tz).  Both give the same bits.
// `proven`: the caller has shown rn_row's ranges for every row of the
// triangle (shadow_fast), so only a zero span width needs IEEE division.
template <class F, class K>
__device__ __forceinline__ void span_run(int xa, int xb, int cy, float cx1, float cx2, float cz1,
                                         float cz2, F&& frag, K&& frag_key, bool proven = false) {
  if (xa > xb) return;
  if (proven ? cx1 != cx2 : rn_row(xa, xb, cx1, cx2, cz1, cz2)) {
    const float d = cx1 - cx2, y = rcp_rn(d);
    const uint32_t flip = cz1 > 0.0f ? 0x80000000u : 0xffffffffu;
    for (int ax = xa; ax <= xb; ax++) {
      const float t = omax(0.0f, omin(1.0f, div_rn((float)ax - cx2, d, y)));
      frag_key(ax, cy, __float_as_uint(rcp_rn((1 - t) * cz2 + t * cz1)) ^ flip);
    }
  } else {
    for (int ax = xa; ax <= xb; ax++) frag(ax, cy, span_tz(span_t(ax, cx1, cx2), cz1, cz2));
  }
}

// span_run for a row of a caster shadow_fast proves: rn_row's ranges hold,
// and a zero span width needs no division either.  IEEE gives (ax - cx2)/0
// = +inf or NaN for ax >= cx2 and -inf below, which the clamp to [0, 1]
// turns into t = 1 or 0.  Such a row (cx1 == cx2) holds at most one texel,
// so the row clamps to [s, s] with s that texel's t, which the clamp
// returns whatever q the division of a zero width left: the texel loop is
// the same straight-line code for every row.  The clamp is the hardware's
// median of three: q itself is never NaN on a row of nonzero width (the
// proven ranges), and with lo == hi the median is lo whatever q is.  It
// matches omax(lo, omin(hi, q)) up to the sign of a zero t, which the
// interpolation (1 - t)*cz2 + t*cz1 of nonzero depths does not see.
template <class K>
__device__ __forceinline__ void span_run_proven(int xa, int xb, int cy, float cx1, float cx2,
                                                float cz1, float cz2, K&& frag_key) {
  const float d = cx1 - cx2, y = rcp_rn(d);
  const float s = (float)xa - cx2 < 0.0f ? 0.0f : 1.0f;
  const bool flat = d == 0.0f;
  const float lo = flat ? s : 0.0f, hi = flat ? s : 1.0f;
  const uint32_t flip = cz1 > 0.0f ? 0x80000000u : 0xffffffffu;
  for (int ax = xa; ax <= xb; ax++) {
    const float t = __builtin_amdgcn_fmed3f(div_rn((float)ax - cx2, d, y), lo, hi);
    frag_key(ax, cy, __float_as_uint(rcp_rn((1 - t) * cz2 + t * cz1)) ^ flip);
  }
}

// ------------------------------------------------- a7: small-path walk
// Small triangles whose rows can all take div_rn and rcp_rn: vertex x in
// [1, 2^20], so every span end (rounding accumulated over fewer than 128
// rows) lies in [0.99, 2^21] and a nonzero cx1 - cx2 or ax - cx2 has
// magnitude in [2^-24, 2^22]; z of one sign,
// magnitudes in [2^-57, 2^57] within a factor 256 of each other, so the
// fewer than 128 rows' accumulated rounding keeps cz, and so the span
// interpolant w, inside [2^-58, 2^58].  Rows with cx1 == cx2 (a half's
// apex: at most one pixel) take the exact 0 or 1 a division by zero
// clamps to.
__device__ __forceinline__ bool small_fast(const Geom<int>& g) {
  const float a = fabsf(g.z1), b = fabsf(g.z2), c = fabsf(g.z3);
  bool ok = imin(g.x1, imin(g.x2, g.x3)) >= 1;
  ok &= imax(g.x1, imax(g.x2, g.x3)) <= (1 << 20);
  ok &= (a >= 0x1p-57f) & (a <= 0x1p57f) & (b >= 0x1p-57f) & (b <= 0x1p57f);
  ok &= (c >= 0x1p-57f) & (c <= 0x1p57f);
  ok &= ((g.z1 > 0.0f) == (g.z2 > 0.0f)) & ((g.z1 > 0.0f) == (g.z3 > 0.0f));
  ok &= fmaxf(a, fmaxf(b, c)) <= 256.0f * fminf(a, fminf(b, c));
  return ok;
}

// Templates.py:250-321.  The pixels inside the clip rectangle [xlo, xhi) x
// [ylo, yhi), which the caller keeps inside the screen (the reference's own
// bounds test), as frag(ax, cy, tz, half) or, on small_fast rows, as
// frag_key(ax, cy, ordf(tz), half): tz there is finite, nonzero and of the
// vertices' sign, so ordf is one xor and tz needs no NaN test.  Every row
// still accumulates, so clipping never changes a fragment's value.
template <class F, class K>
__device__ __forceinline__ void small_walk(const Geom<int>& g, int xlo, int ylo, int xhi, int yhi,
                                           F&& frag, K&& frag_key) {
  const bool fast = small_fast(g);
  const uint32_t flip = g.z1 > 0.0f ? 0x80000000u : 0xffffffffu;
  // the row's pixels xa.. with (float)ax <= xr and ax < xhi.  On a fast
  // triangle's row of zero width (each half's first row among them) the
  // interpolant clamps to [s, s], s the exact 0 or 1 IEEE's (ax - cx2)/0
  // clamps to for its one pixel, as span_run_proven does.
  auto row = [&](int xa, float xr, int cy, float cx1, float cx2, float cz1, float cz2, int half) {
    const float d = cx1 - cx2;
    if (fast) {
      const int xb = imin(xhi - 1, f2i(floorf(xr)));
      const float y = rcp_rn(d);
      const float s = (float)xa - cx2 < 0.0f ? 0.0f : 1.0f;
      const bool flat = d == 0.0f;
      const float lo = flat ? s : 0.0f, hi = flat ? s : 1.0f;
      for (int ax = xa; ax <= xb; ax++) {
        const float t = __builtin_amdgcn_fmed3f(div_rn((float)ax - cx2, d, y), lo, hi);
        frag_key(ax, cy, __float_as_uint(rcp_rn((1 - t) * cz2 + t * cz1)) ^ flip, half);
      }
    } else {
      for (int ax = xa; ax < xhi && (float)ax <= xr; ax++) {
        float t = span_t(ax, cx1, cx2);
        frag(ax, cy, span_tz(t, cz1, cz2), half);
      }
    }
  };
  {
    Edges e = small_edges(g, 0);
    float cx1 = (float)g.x1, cx2 = cx1, cz1 = g.z1, cz2 = g.z1;
    for (int cy = g.y1; cy <= g.y2; cy++) {
      if (cy >= yhi) break;
      if (cy >= ylo) row(imax(f2i(cx2), xlo), cx1, cy, cx1, cx2, cz1, cz2, 0);
      cx1 += e.s1;
      cx2 += e.s2;
      cz1 += e.sz1;
      cz2 += e.sz2;
    }
  }
  {
    Edges e = small_edges(g, 1);
    float cx1 = (float)g.x3, cx2 = cx1, cz1 = g.z3, cz2 = g.z3;
    for (int cy = g.y3; cy >= g.y2; cy--) {
      if (cy < ylo) break;
      if (cy < yhi) row(imax(f2i(cx1), xlo), cx2, cy, cx1, cx2, cz1, cz2, 1);
      cx1 -= e.s1;
      cx2 -= e.s2;
      cz1 -= e.sz1;
      cz2 -= e.sz2;
    }
  }
}

// Pixel rectangle that contains every fragment of a small triangle: rows
// y1..y3 exactly; columns within one pixel of the snapped vertices (span
// ends accumulate rounding of well under a pixel over < 96 rows, and the
// left end truncates), widened to two for margin.
__device__ __forceinline__ void small_bounds(const float px[3], const float py[3], int& x0,
                                             int& y0, int& x1, int& y1) {
  const int a = f2i(px[0]), b = f2i(px[1]), c = f2i(px[2]);
  const int d = f2i(py[0]), e = f2i(py[1]), f = f2i(py[2]);
  x0 = imin(a, imin(b, c)) - 2;
  x1 = imax(a, imax(b, c)) + 2;
  y0 = imin(d, imin(e, f));
  y1 = imax(d, imax(e, f));
}

// ------------------------------------------------- a9: tiled-path walk
struct TileBox {
  float xMin, xMax, yMin, yMax;
};
// Templates.py:23-29: the box the tiled draw's work-group bx walks.  Its
// column is bx % (int)(W/16) and its row bx / (W/16) in float arithmetic,
// truncated; for the index bx = WC*cy + ax of coarse tile (cy, ax) that is
// the tile itself when W is a multiple of 16.  For other widths W/16 has a
// fraction and a tile whose column is small against its row (ax < cy/2 at
// W/16 = k + 1/2) maps to a box of the row above: the reference walks the
// triangle there, and the oracle and the kernels do the same.
AXR_HD void walked_tile(int cy, int ax, int WC, int W, int& tx, int& ty) {
  if ((W & 15) == 0) {
    tx = ax;
    ty = cy;
    return;
  }
  const int bx = WC * cy + ax;
  const float wt = (float)W / 16.0f;
  tx = bx % (int)wt;
  ty = f2i((float)bx / wt);
}
// The box of 16-px tile (tx, ty).
AXR_HD TileBox tile_box_xy(int tx, int ty) {
  TileBox b;
  b.xMin = (float)tx * 16.0f;
  b.xMax = ((float)tx + 1.0f) * 16.0f - 1.0f;
  b.yMin = (float)ty * 16.0f;
  b.yMax = ((float)ty + 1.0f) * 16.0f - 1.0f;
  return b;
}

// Templates.py:113-182 inside one tile.  frag(ax, cy, tz, half).
template <class F>
__device__ __forceinline__ void tile_walk(const Geom<float>& g, const TileBox& b, F&& frag) {
  {
    Edges e = tile_edges(g, 0);
    int cy = f2i(oclamp(g.y1, b.yMin, b.yMax));
    float dy = (float)cy - g.y1;
    float cx1 = g.x1 + dy * e.s1, cx2 = g.x1 + dy * e.s2;
    float cz1 = g.z1 + dy * e.sz1, cz2 = g.z1 + dy * e.sz2;
    if (g.y2 > b.yMin) {
      const float yEnd = oclamp(g.y2, b.yMin, b.yMax);
      for (; (float)cy <= yEnd; cy++) {
        if (cx2 < b.xMax && cx1 > b.xMin) {
          const float xe = oclamp(cx1, b.xMin, b.xMax);
          for (int ax = f2i(oclamp(cx2, b.xMin, b.xMax)); (float)ax <= xe; ax++) {
            float t = span_t(ax, cx1, cx2);
            frag(ax, cy, span_tz(t, cz1, cz2), 0);
          }
        }
        cx1 += e.s1;
        cx2 += e.s2;
        cz1 += e.sz1;
        cz2 += e.sz2;
      }
    }
  }
  {
    Edges e = tile_edges(g, 1);
    int cy = f2i(oclamp(g.y3, b.yMin, b.yMax));
    float dy = (float)cy - g.y3;
    float cx1 = g.x3 + dy * e.s1, cx2 = g.x3 + dy * e.s2;
    float cz1 = g.z3 + dy * e.sz1, cz2 = g.z3 + dy * e.sz2;
    if (g.y2 < b.yMax) {
      const float yEnd = oclamp(g.y2, b.yMin, b.yMax);
      for (; (float)cy >= yEnd; cy--) {
        if (cx1 < b.xMax && cx2 > b.xMin) {
          const float xe = oclamp(cx2, b.xMin, b.xMax);
          for (int ax = f2i(oclamp(cx1, b.xMin, b.xMax)); (float)ax <= xe; ax++) {
            float t = span_t(ax, cx1, cx2);
            frag(ax, cy, span_tz(t, cz1, cz2), 1);
          }
        }
        cx1 -= e.s1;
        cx2 -= e.s2;
        cz1 -= e.sz1;
        cz2 -= e.sz2;
      }
    }
  }
}

// ---------------------------------------------------- a8: coarse binning
// Pipe/coarse.c:26-236 over one large triangle.  emit(tile) per entry,
// duplicates included.
template <class F>
AXR_HD void coarse_walk(const float px[3], const float py[3], int WC, int HC,
                                            F&& emit) {
  float x1 = px[0] / 16.0f - 0.5f, y1 = py[0] / 16.0f - 0.5f;
  float x2 = px[1] / 16.0f - 0.5f, y2 = py[1] / 16.0f - 0.5f;
  float x3 = px[2] / 16.0f - 0.5f, y3 = py[2] / 16.0f - 0.5f;
  if (y1 > y2) { swp(y1, y2); swp(x1, x2); }
  if (y2 > y3) { swp(y2, y3); swp(x2, x3); }
  if (y1 > y2) { swp(y1, y2); swp(x1, x2); }
  const float x4 = (x1 + ((y2 - y1) / (y3 - y1)) * (x3 - x1));
  const float y4 = y2;
  const float wmax = (float)WC - 1;
  {
    float slope1 = (x2 - x1) / (y2 - y1);
    float slope2 = (x4 - x1) / (y4 - y1);
    float cx1 = x1, cx2 = x1;
    if (slope1 < slope2) swp(slope1, slope2);
    const float sbias = 1.0f + floorf(y1) - y1;
    if (y1 != y2) {
      const float yEnd = ceilf(y2);
      for (int cy = f2i(floorf(y1)); (float)cy <= yEnd; cy++) {
        float mx1 = cx1 + sbias * slope1;
        if (slope1 < 0) mx1 -= slope1;
        float mx2 = cx2 + sbias * slope2;
        if (slope2 > 0) mx2 -= slope2;
        mx1 = omin(mx1, wmax);
        mx2 = omax(mx2, 0.0f);
        if (cy >= 0 && cy < HC) {
          const float xe = ceilf(mx1);
          for (int ax = f2i(floorf(mx2)); (float)ax <= xe; ax++) emit(WC * cy + ax);
        }
        cx1 += slope1;
        cx2 += slope2;
      }
    }
  }
  {
    float slope1 = (x3 - x2) / (y3 - y2);
    float slope2 = (x3 - x4) / (y3 - y4);
    float cx1 = x3, cx2 = x3;
    if (slope1 < slope2) swp(slope1, slope2);
    const float sbias = y3 - floorf(y3);
    if (y2 != y3) {
      const float yEnd = floorf(y2);
      for (int cy = f2i(ceilf(y3)); (float)cy >= yEnd; cy--) {
        float mx1 = cx1 - sbias * slope1;
        if (slope1 < 0) mx1 += slope1;
        float mx2 = cx2 - sbias * slope2;
        if (slope2 > 0) mx2 += slope2;
        mx2 = omin(mx2, wmax);
        mx1 = omax(mx1, 0.0f);
        if (cy >= 0 && cy < HC) {
          const float xe = ceilf(mx2);
          for (int ax = f2i(floorf(mx1)); (float)ax <= xe; ax++) emit(WC * cy + ax);
        }
        cx1 -= slope1;
        cx2 -= slope2;
      }
    }
  }
}

// The rows of coarse_walk split over nl threads: thread `lane` takes rows
// lane, lane + nl, ... of each half, reaching a row's edge positions by the
// same repeated additions as the serial walk, and calls row(cy, first, xe)
// for each row inside the grid: tiles first, first + 1, ... while
// (float)tile <= xe, as coarse_walk emits them.  Together the threads emit
// coarse_walk's entries.
template <class R>
AXR_HD void coarse_rows(const float px[3], const float py[3], int WC, int HC,
                                            int lane, int nl, R&& row) {
  float x1 = px[0] / 16.0f - 0.5f, y1 = py[0] / 16.0f - 0.5f;
  float x2 = px[1] / 16.0f - 0.5f, y2 = py[1] / 16.0f - 0.5f;
  float x3 = px[2] / 16.0f - 0.5f, y3 = py[2] / 16.0f - 0.5f;
  if (y1 > y2) { swp(y1, y2); swp(x1, x2); }
  if (y2 > y3) { swp(y2, y3); swp(x2, x3); }
  if (y1 > y2) { swp(y1, y2); swp(x1, x2); }
  const float x4 = (x1 + ((y2 - y1) / (y3 - y1)) * (x3 - x1));
  const float y4 = y2;
  const float wmax = (float)WC - 1;
  if (y1 != y2) {
    float slope1 = (x2 - x1) / (y2 - y1);
    float slope2 = (x4 - x1) / (y4 - y1);
    float cx1 = x1, cx2 = x1;
    if (slope1 < slope2) swp(slope1, slope2);
    const float sbias = 1.0f + floorf(y1) - y1;
    const float yEnd = ceilf(y2);
    int cy = f2i(floorf(y1));
    for (int k = 0; k < lane && (float)cy <= yEnd; k++, cy++) {
      cx1 += slope1;
      cx2 += slope2;
    }
    // rows only grow from here: past the last tile row none is emitted
    while ((float)cy <= yEnd && cy < HC) {
      float mx1 = cx1 + sbias * slope1;
      if (slope1 < 0) mx1 -= slope1;
      float mx2 = cx2 + sbias * slope2;
      if (slope2 > 0) mx2 -= slope2;
      mx1 = omin(mx1, wmax);
      mx2 = omax(mx2, 0.0f);
      if (cy >= 0) row(cy, f2i(floorf(mx2)), ceilf(mx1));
      for (int k = 0; k < nl && (float)cy <= yEnd; k++, cy++) {
        cx1 += slope1;
        cx2 += slope2;
      }
    }
  }
  if (y2 != y3) {
    float slope1 = (x3 - x2) / (y3 - y2);
    float slope2 = (x3 - x4) / (y3 - y4);
    float cx1 = x3, cx2 = x3;
    if (slope1 < slope2) swp(slope1, slope2);
    const float sbias = y3 - floorf(y3);
    const float yEnd = floorf(y2);
    int cy = f2i(ceilf(y3));
    for (int k = 0; k < lane && (float)cy >= yEnd; k++, cy--) {
      cx1 -= slope1;
      cx2 -= slope2;
    }
    // rows only shrink from here: above the first tile row none is emitted
    while ((float)cy >= yEnd && cy >= 0) {
      float mx1 = cx1 - sbias * slope1;
      if (slope1 < 0) mx1 += slope1;
      float mx2 = cx2 - sbias * slope2;
      if (slope2 > 0) mx2 += slope2;
      mx2 = omin(mx2, wmax);
      mx1 = omax(mx1, 0.0f);
      if (cy < HC) row(cy, f2i(floorf(mx1)), ceilf(mx2));
      for (int k = 0; k < nl && (float)cy >= yEnd; k++, cy--) {
        cx1 -= slope1;
        cx2 -= slope2;
      }
    }
  }
}

// Coarse binning of one large triangle by a block of nthreads threads
// (nthreads a power of two): each tile row of coarse_rows goes to a group of
// G threads, G the largest power of two with 2 x rows x G <= nthreads, and
// thread `tid` takes the row's columns first + sub, first + sub + G, ...
// (sub = tid % G).  batch(cy, ax, G, xe) receives up to eight of them at a
// time: columns ax + j*G, j < 8, those with (float)(ax + j*G) <= xe.
// A row's `first` can be far right of the grid: floorf of an edge position
// past 2^31 saturates to INT_MAX, with xe <= WC - 1, where the serial walk
// emits nothing.  first is clamped to WC before the stride is added, so
// first + sub cannot wrap, and every emitted column lies in [0, WC - 1].
template <class B>
AXR_HD void coarse_block_tiles(const float px[3], const float py[3], int WC, int HC, int tid,
                               int nthreads, B&& batch) {
  // tile rows the walk may visit (both halves; clamped to the grid), an
  // estimate only used to size the groups
  const float ylo = omin(py[0], omin(py[1], py[2])) / 16.0f - 0.5f;
  const float yhi = omax(py[0], omax(py[1], py[2])) / 16.0f - 0.5f;
  const int r0 = imax(0, f2i(floorf(ylo))), r1 = imin(HC - 1, f2i(ceilf(yhi)));
  const int rows = imax(1, r1 - r0 + 1);
  int G = 1;
  while (G < nthreads && rows * 2 * G <= nthreads) G *= 2;
  const int sub = tid % G;
  coarse_rows(px, py, WC, HC, tid / G, nthreads / G, [&](int cy, int first, float xe) {
    for (int ax = imin(first, WC) + sub; (float)ax <= xe; ax += 8 * G) batch(cy, ax, G, xe);
  });
}

// ------------------------------------------------ a11: shadow depth walk
// Shaders/drawmin.c:86-176 (int coordinates, strict top loop) on a dim x dim
// map, emitting only the texels inside [xlo, xhi) x [ylo, yhi): a texel as
// frag(ax, cy, tz) or as frag_key(ax, cy, ordf(tz)) (span_run).
// Casters whose rows all lie in rn_row's ranges, by the argument small_fast
// makes for screen triangles: vertex x in [1, 2^20] and an x extent below
// 96, so over fewer than 128 rows every span end stays in [0.99, 2^21] and
// a nonzero cx1 - cx2 or ax - cx2 has magnitude in [2^-24, 2^22] (floats
// from 0.5 up are multiples of 2^-24); z of one sign with magnitudes in
// [2^-57, 2^57] within a factor 256 of each other, so every row's cz and
// span interpolant stay inside [2^-59, 2^59].
__device__ __forceinline__ bool shadow_fast(int x1, int x2, int x3, int y1, int y3, float z1,
                                            float z2, float z3) {
  const int xmin = imin(x1, imin(x2, x3)), xmax = imax(x1, imax(x2, x3));
  const float a = fabsf(z1), b = fabsf(z2), c = fabsf(z3);
  bool ok = (xmin >= 1) & (xmax <= (1 << 20)) & (xmax - xmin < 96) & (y3 - y1 < 128);
  ok &= (a >= 0x1p-57f) & (a <= 0x1p57f) & (b >= 0x1p-57f) & (b <= 0x1p57f);
  ok &= (c >= 0x1p-57f) & (c <= 0x1p57f);
  ok &= ((z1 > 0.0f) == (z2 > 0.0f)) & ((z1 > 0.0f) == (z3 > 0.0f));
  ok &= fmaxf(a, fmaxf(b, c)) <= 256.0f * fminf(a, fminf(b, c));
  return ok;
}

template <class F, class K>
__device__ __forceinline__ void shadow_walk(const AxrShRec& r, int dim, int xlo, int ylo, int xhi,
                                            int yhi, F&& frag, K&& frag_key) {
  int x1 = r.x[0], y1 = r.y[0], x2 = r.x[1], y2 = r.y[1], x3 = r.x[2], y3 = r.y[2];
  float z1 = r.z[0], z2 = r.z[1], z3 = r.z[2];
  if (y1 > y2) { swp(y1, y2); swp(x1, x2); swp(z1, z2); }
  if (y2 > y3) { swp(y2, y3); swp(x2, x3); swp(z2, z3); }
  if (y1 > y2) { swp(y1, y2); swp(x1, x2); swp(z1, z2); }
  if (!(y1 < dim && y3 >= 0)) return;
  const bool proven = shadow_fast(x1, x2, x3, y1, y3, z1, z2, z3);
  // The quotients of the row counts D are integers' quotients, so every
  // integer numerator, and every z numerator rn_num accepts, divides
  // exactly through div_rn.  A zero D leaves its half one row (or none)
  // that starts at a single point, so the slopes D = 0 would make are never
  // used.
  const float D31 = (float)(y3 - y1), D21 = (float)(y2 - y1), D32 = (float)(y3 - y2);
  const float r21 = rcp_rn(D21), r32 = rcp_rn(D32);
  const float q = div_rn(D21, D31, rcp_rn(D31));
  const int x4 = f2i((float)x1 + q * (float)(x3 - x1));
  const float z4 = z1 + q * (z3 - z1);
  const float a1 = z2 - z1, a2 = z4 - z1, a3 = z3 - z2, a4 = z3 - z4;
  float zs[4] = {div_rn0(a1, D21, r21), div_rn0(a2, D21, r21), div_rn0(a3, D32, r32),
                 div_rn0(a4, D32, r32)};
  bool zok = rn_num(a1);  // no short-circuit: one branch below, not four
  zok &= rn_num(a2);
  zok &= rn_num(a3);
  zok &= rn_num(a4);
  if (!zok) {
    // IEEE division, kept inside the branch by the empty asm
    float b1 = a1, b2 = a2, b3 = a3, b4 = a4;
    asm volatile("" : "+v"(b1), "+v"(b2), "+v"(b3), "+v"(b4));
    zs[0] = b1 / D21;
    zs[1] = b2 / D21;
    zs[2] = b3 / D32;
    zs[3] = b4 / D32;
  }
  // Both halves in one row loop: rows y1.. of the top half (edges from
  // vertex 1 down, drawmin.c's first loop), then rows y3.. of the bottom
  // half (edges from vertex 3 up, its second loop), each accumulated exactly
  // as its own loop does (x - s is x + (-s)).  One loop lets a wave's lanes
  // run their rows back to back: a caster with a long top and a short
  // bottom half no longer waits for another's long bottom half.
  float t1 = div_rn((float)(x2 - x1), D21, r21), t2 = div_rn((float)(x4 - x1), D21, r21);
  float tz1 = zs[0], tz2 = zs[1];
  if (t1 < t2) {
    swp(t1, t2);
    swp(tz1, tz2);
  }
  float b1 = div_rn((float)(x3 - x2), D32, r32), b2 = div_rn((float)(x3 - x4), D32, r32);
  float bz1 = zs[2], bz2 = zs[3];
  if (b1 < b2) {
    swp(b1, b2);
    swp(bz1, bz2);
  }
  // rows of the top half before its first row at or past yhi; of the bottom
  // half before its first row above ylo
  const int n0 = imax(0, imin(y2, yhi - 1) - y1 + 1);
  const int n1 = imax(0, y3 - imax(y2 + 1, ylo) + 1);
  auto rows = [&](auto all_proven) {
    float cx1 = (float)x1, cx2 = cx1, cz1 = z1, cz2 = z1;
    float s1 = t1, s2 = t2, sz1 = tz1, sz2 = tz2;
    int cy = y1, dy = 1;
    for (int i = 0; i < n0 + n1; i++) {
      if (i == n0) {  // the bottom half starts at vertex 3
        cx1 = cx2 = (float)x3;
        cz1 = cz2 = z3;
        s1 = -b1;
        s2 = -b2;
        sz1 = -bz1;
        sz2 = -bz2;
        cy = y3;
        dy = -1;
      }
      // top-half rows lie above yhi and bottom-half rows at or below ylo
      // by n0 and n1, so one test of both bounds serves either half
      const bool top = i < n0;
      if ((unsigned)(cy - ylo) < (unsigned)(yhi - ylo)) {
        const float lo = top ? cx2 : cx1, hi = top ? cx1 : cx2;
        const int xa = imax(f2i(lo), xlo), xb = imin(f2i(hi), xhi - 1);
        if (decltype(all_proven)::value)
          span_run_proven(xa, xb, cy, cx1, cx2, cz1, cz2, frag_key);
        else
          span_run(xa, xb, cy, cx1, cx2, cz1, cz2, frag, frag_key, proven);
      }
      cx1 += s1;
      cx2 += s2;
      cz1 += sz1;
      cz2 += sz2;
      cy += dy;
    }
  };
  // a wave whose casters are all proven takes the one-path row loop
  if (__ballot(!proven) == 0)
    rows(std::true_type{});
  else
    rows(std::false_type{});
}

// Texel rectangle containing every fragment of a caster: rows y1..y3;
// columns within one texel of the vertices (the int x4 and the truncated
// span ends), widened to two.
__device__ __forceinline__ void shadow_bounds(const AxrShRec& r, int& x0, int& y0, int& x1,
                                              int& y1) {
  x0 = imin(r.x[0], imin(r.x[1], r.x[2])) - 2;
  x1 = imax(r.x[0], imax(r.x[1], r.x[2])) + 2;
  y0 = imin(r.y[0], imin(r.y[1], r.y[2]));
  y1 = imax(r.y[0], imax(r.y[1], r.y[2]));
}

// Work class of a caster, by its rows y_max - y_min + 1: 1-2, 3, 4-6, 7 and
// more give 0..3.  Setup stores it in the top two bits of the caster's
// shadow-bin entries (triangle indices stay below 2^30) so that the walk
// can order a bin's casters by size without loading their records.
#ifndef AXR_SHC_LO
#define AXR_SHC_LO 3
#endif
#ifndef AXR_SHC_MID
#define AXR_SHC_MID 4
#endif
#ifndef AXR_SHC_HI
#define AXR_SHC_HI 7
#endif
#define AXR_SH_ENTRY_MASK 0x3fffffffu
__device__ __forceinline__ uint32_t shadow_class(const AxrShRec& r) {
  const int rows = imax(r.y[0], imax(r.y[1], r.y[2])) - imin(r.y[0], imin(r.y[1], r.y[2])) + 1;
  return (uint32_t)(rows >= AXR_SHC_LO) + (uint32_t)(rows >= AXR_SHC_MID) +
         (uint32_t)(rows >= AXR_SHC_HI);
}

// Pipe/trisetupOrtho.c:20-59.
__device__ __forceinline__ bool shadow_setup(const float4 (&vp4)[3], const AxrCam& cam,
                                             int dim, float bias, AxrShRec& r) {
  const V3 vp = v3(cam.pos[0], cam.pos[1], cam.pos[2]);
  const V3 SVd = cam_row(cam, 0), SVx = cam_row(cam, 1), SVy = cam_row(cam, 2);
  const float half = (float)(dim / 2);
  bool yes = true;
  float dd[3], dx[3], dy[3];
#pragma unroll
  for (int k = 0; k < 3; k++) {
    const float4 p = vp4[k];
    V3 x = sub(v3(p.x, p.y, p.z), vp);
    dd[k] = dot(x, SVd);
    dx[k] = dot(x, SVx);
    dy[k] = dot(x, SVy);
    dd[k] += bias;
  }
  recip3(dd);
#pragma unroll
  for (int k = 0; k < 3; k++) {
    dx[k] = dx[k] * cam.scale + half;
    dy[k] = dy[k] * -cam.scale + half;
    r.x[k] = f2i(dx[k]);
    r.y[k] = f2i(dy[k]);
    r.z[k] = dd[k];
    if (!(dd[k] > 0)) yes = false;
    if (!(r.x[k] >= -120 && r.x[k] < dim + 120 && r.y[k] >= -120 && r.y[k] < dim + 120))
      yes = false;
  }
  return yes;
}

// ------------------------------------------------------------- shading
struct MipInfo {
  int mip;
  float fmip;
  int lenMip, lenMip2, startMip, startMip2;
};

// av_draw_normmap_mip.c:26-59 on sorted coordinates; mutates a1..a3 through
// the reference's /= z, *= z round trip (after the 4th vertex is formed).
template <class T>
__device__ __forceinline__ MipInfo mip_setup(Attr& a1, Attr& a2, Attr& a3, const Geom<T>& g,
                                             int numMip, float bias) {
  a1.u /= g.z1; a1.v /= g.z1; a2.u /= g.z2; a2.v /= g.z2; a3.u /= g.z3; a3.v /= g.z3;
  a1.px /= g.z1; a1.py /= g.z1; a1.pz /= g.z1;
  a2.px /= g.z2; a2.py /= g.z2; a2.pz /= g.z2;
  a3.px /= g.z3; a3.py /= g.z3; a3.pz /= g.z3;
  const float dU1x = a2.u - a1.u, dU1y = a2.v - a1.v;
  const float dU2x = a3.u - a1.u, dU2y = a3.v - a1.v;
  const float d1 = (float)((g.y2 - g.y1) * (g.y2 - g.y1) + (g.x2 - g.x1) * (g.x2 - g.x1));
  const float d2 = (float)((g.y3 - g.y1) * (g.y3 - g.y1) + (g.x3 - g.x1) * (g.x3 - g.x1));
  const float d3 = (float)((g.y3 - g.y2) * (g.y3 - g.y2) + (g.x3 - g.x2) * (g.x3 - g.x2));
  const float s1x = fabsf(dU1x / d1), s1y = fabsf(dU1y / d1);
  const float s2x = fabsf(dU2x / d2), s2y = fabsf(dU2y / d2);
  const float s3x = fabsf((a3.u - a2.u) / d3), s3y = fabsf((a3.v - a2.v) / d3);
  float dt = omax(s1x, omax(s1y, omax(s2x, s2y)));
  dt = omax(dt, omax(s3x, s3y));
  float fmip = omax(0.0f, omin((float)(numMip - 1), fabsf(log2_cw(dt)) / 2 + bias));
  MipInfo m;
  m.mip = f2i(fmip);
  m.fmip = fmip - (float)m.mip;
  m.lenMip = 1 << m.mip;
  m.lenMip2 = m.lenMip << 1;
  m.startMip = (int)(((1u << (2 * m.mip)) - 1u) / 3u);  // sum of 4^i, i < mip
  m.startMip2 = m.startMip + m.lenMip * m.lenMip;
  a1.u *= g.z1; a1.v *= g.z1; a2.u *= g.z2; a2.v *= g.z2; a3.u *= g.z3; a3.v *= g.z3;
  a1.px *= g.z1; a1.py *= g.z1; a1.pz *= g.z1;
  a2.px *= g.z2; a2.py *= g.z2; a2.pz *= g.z2;
  a3.px *= g.z3; a3.py *= g.z3; a3.pz *= g.z3;
  return m;
}

// Bilinear at one level (av_draw_base.c:29-42, av_draw_normmap_mip.c:73-93).
__device__ __forceinline__ void bilerp(const ushort4* __restrict__ tex, float tu, float tv, int L,
                                       int start, float o[3]) {
  int tex1 = f2i(tu);
  tu -= (float)tex1;
  tex1 = abs_mask(tex1, L - 1);
  int tex2 = f2i(tv);
  tv -= (float)tex2;
  tex2 = abs_mask(tex2, L - 1);
  const int xn = imin(tex1 + 1, L - 1), yn = imin(tex2 + 1, L - 1);
  const ushort4 t00 = tex[start + tex1 + L * tex2];
  const ushort4 t10 = tex[start + xn + L * tex2];
  const ushort4 t01 = tex[start + tex1 + L * yn];
  const ushort4 t11 = tex[start + xn + L * yn];
  const float texr1 = tu, texr2 = tv;
  const float texi1 = 1 - texr1, texi2 = 1 - texr2;
  const float w00 = texi1 * texi2, w10 = texr1 * texi2, w01 = texi1 * texr2, w11 = texr1 * texr2;
  o[0] = w00 * (float)t00.x + w10 * (float)t10.x + w01 * (float)t01.x + w11 * (float)t11.x;
  o[1] = w00 * (float)t00.y + w10 * (float)t10.y + w01 * (float)t01.y + w11 * (float)t11.y;
  o[2] = w00 * (float)t00.z + w10 * (float)t10.z + w01 * (float)t01.z + w11 * (float)t11.z;
}

// PCF (drawreflcubemapfresneltexsh.c:162-182 cross, or a 3x3 grid), summed
// over the bound cascades and capped (av_alpha_max.c:53-84).
__device__ __forceinline__ float pcf_light(const AxrDrawConst& dc, V3 wpos) {
  const int K = dc.pcf_taps == 9 ? 9 : 5;
  const int d = 2;  // shDist
  int shadow = 0;
#pragma unroll
  for (int c = 0; c < AXR_SHADOW_BIND; c++) {
    if (c >= dc.nshadow) break;
    const AxrShadowDev& sm = dc.sh[c];
    const V3 pos = sub(wpos, v3(sm.cam.pos[0], sm.cam.pos[1], sm.cam.pos[2]));
    const float depth = dot(pos, cam_row(sm.cam, 0));
    const int wS = sm.wS;
    const int sx = f2i(dot(pos, cam_row(sm.cam, 1)) * sm.cam.scale) + wS;
    const int sy = f2i(dot(pos, cam_row(sm.cam, 2)) * -sm.cam.scale) + wS;
    if (sx >= d && sx < 2 * wS - d && sy >= d && sy < 2 * wS - d) {
      const uint32_t* SD = sm.map;
      const long long st = 2LL * wS;
      if (K == 5) {
        if (unordf(SD[st * sy + sx]) < depth) shadow += 1;
        if (unordf(SD[st * (sy + d) + sx]) < depth) shadow += 1;
        if (unordf(SD[st * (sy - d) + sx]) < depth) shadow += 1;
        if (unordf(SD[st * sy + (sx + d)]) < depth) shadow += 1;
        if (unordf(SD[st * sy + (sx - d)]) < depth) shadow += 1;
      } else {
#pragma unroll
        for (int oy = -d; oy <= d; oy += d)
#pragma unroll
          for (int ox = -d; ox <= d; ox += d)
            if (unordf(SD[st * (sy + oy) + (sx + ox)]) < depth) shadow += 1;
      }
    }
  }
  shadow = imin(K, shadow);
  float light = (float)shadow * dc.amb + (float)(K - shadow);
  if (K == 5)
    light *= 0.2f;
  else
    light = light / 9.0f;
  return light;
}

// One fragment's HDR colour (u16 per channel).
__device__ __forceinline__ void shade(const AxrDrawConst& dc, const AxrGroupDev& grp,
                                      const MipInfo& mi, float t, float tz, const Attr& c1,
                                      const Attr& c2, uint16_t rgb[3]) {
  const float u = 1 - t;
  float L[3];
  const float fu = (u * c2.u + t * c1.u) * tz;
  const float fv = (u * c2.v + t * c1.v) * tz;
  if (grp.flags & AXR_SH_MIP) {
    float o1[3], o2[3];
    bilerp(grp.tex, fu * (float)mi.lenMip, fv * (float)mi.lenMip, mi.lenMip, mi.startMip, o1);
    bilerp(grp.tex, fu * (float)mi.lenMip2, fv * (float)mi.lenMip2, mi.lenMip2, mi.startMip2, o2);
#pragma unroll
    for (int c = 0; c < 3; c++) L[c] = (1 - mi.fmip) * o1[c] + mi.fmip * o2[c];
  } else {
    bilerp(grp.tex, fu * (float)grp.tex_size, fv * (float)grp.tex_size, grp.tex_size, 0, L);
  }
  const V3 wpos = v3((u * c2.px + t * c1.px) * tz, (u * c2.py + t * c1.py) * tz,
                     (u * c2.pz + t * c1.pz) * tz);
  float light = 1;
  if ((grp.flags & AXR_SH_SHADOW) && dc.nshadow > 0) light = pcf_light(dc, wpos);
  const float col[3] = {(u * c2.ix + t * c1.ix) * tz, (u * c2.iy + t * c1.iy) * tz,
                        (u * c2.iz + t * c1.iz) * tz};
  const V3 nrm = normalize(v3((u * c2.nx + t * c1.nx) * tz, (u * c2.ny + t * c1.ny) * tz,
                              (u * c2.nz + t * c1.nz) * tz));
  const V3 LDir = v3(dc.ldir[0], dc.ldir[1], dc.ldir[2]);
  const float nl = omax(0.0f, dot(nrm, LDir));
  const float dcol[3] = {dc.lint[0] * nl, dc.lint[1] * nl, dc.lint[2] * nl};
  if (grp.flags & AXR_SH_PHONG) {
    // av_draw_normmap.c:276-284, 303-310 (without the environment terms)
    const V3 a = normalize(sub(wpos, v3(dc.view.pos[0], dc.view.pos[1], dc.view.pos[2])));
    const V3 h = normalize(add(a, LDir));
    const float theta = omax(0.0f, dot(h, nrm));
    const float sp = pown(theta, grp.spec_pow) * 256 * 4.0f * light;
    const float k = (float)(grp.spec_pow + 2) / (8 * 3.1416f);
#pragma unroll
    for (int c = 0; c < 3; c++) {
      const float spec = (dc.lint[c] * sp) * k;
      rgb[c] = f2u16(spec + L[c] * (light * dcol[c] + col[c]));
    }
  } else {
#pragma unroll
    for (int c = 0; c < 3; c++)
      rgb[c] = f2u16(L[c] * (light * (dcol[c] + col[c]) + (1 - light) * col[c]));
  }
}

// Shaders/Post/gamma.c:13-29 on one pixel.
__device__ __forceinline__ void gamma_px(float ex, uint16_t rgb[3]) {
  const float k = 8 * ex;
  float j0 = sqrtf(k * (float)rgb[0]) / 1.1f;
  float j1 = sqrtf(k * (float)rgb[1]) / 1.1f;
  float j2 = sqrtf(k * (float)rgb[2]) / 1.1f;
  const float m = omax(j0, omax(j1, j2)) / 255.0f;
  if (m > 1.0f) {
    const float s = sqrtf(m);
    j0 = j0 / s;
    j1 = j1 / s;
    j2 = j2 / s;
  }
  rgb[0] = (uint16_t)omin(255.0f, j0);
  rgb[1] = (uint16_t)omin(255.0f, j1);
  rgb[2] = (uint16_t)omin(255.0f, j2);
}

}  // namespace axr
