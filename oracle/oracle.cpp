// oracle.cpp -- CPU restatement of the AXI Visualizer rendering hot path.
//
// TEST INFRASTRUCTURE ONLY (the parity checker; see oracle.h for the list of
// reference kernels restated and the arithmetic decrees).  Built with
// g++ -O2 -ffp-contract=off -fno-fast-math so every float operation is one
// IEEE binary32 operation in the order written.
//
// Structure: the reference's serial semantics are executed literally -- a
// depth buffer F, per-fragment `F >= tz` tests in pass order, shading of
// every passing fragment.  Parallelism never reorders fragments of one
// pixel: each worker owns a band of 16 screen rows (one tile row) and walks
// every triangle touching the band in pass order.
#include "oracle.h"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <thread>
#include <vector>

namespace {

// ---------------------------------------------------------------- arithmetic
struct f2 {
  float x, y;
};
struct f3 {
  float x, y, z;
};

inline f3 sub3(f3 a, f3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline f3 add3(f3 a, f3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline f3 scl3(f3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
inline float dot3(f3 a, f3 b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
inline float omin(float x, float y) { return y < x ? y : x; }
inline float omax(float x, float y) { return x < y ? y : x; }
inline int imin(int x, int y) { return y < x ? y : x; }
inline int imax(int x, int y) { return x < y ? y : x; }
// OpenCL clamp: fmin(fmax(x, lo), hi); identical to omin(omax()) off NaN.
inline float oclamp(float x, float lo, float hi) { return omin(omax(x, lo), hi); }
inline int f2i(float x) {
  if (x != x) return 0;
  if (x >= 2147483648.0f) return 2147483647;
  if (x < -2147483648.0f) return (-2147483647 - 1);
  return (int)x;
}
inline uint16_t f2u16(float x) {
  if (!(x > 0.0f)) return 0;
  if (x >= 65535.0f) return 65535;
  return (uint16_t)x;
}
inline f3 normalize3(f3 v) {
  float s = 1.0f / std::sqrt(dot3(v, v));
  return scl3(v, s);
}
inline float pown_(float x, int n) {
  float r = 1.0f, b = x;
  unsigned e = (unsigned)(n < 0 ? 0 : n);
  while (e) {
    if (e & 1u) r = r * b;
    b = b * b;
    e >>= 1;
  }
  return r;
}
// log2(x): x = m 2^e with m in [sqrt(1/2), sqrt(2)); log2 x = e + ln(m)/ln 2,
// ln m = 2 atanh(s), s = (m - 1)/(m + 1), by its odd series through s^9.
// Decree for OpenCL's log2 (the vendor's): exact at powers of two, within a
// few ulp elsewhere; -inf at 0, NaN below 0 or at NaN, +inf at +inf.
inline float log2_cw(float x) {
  if (x != x || x < 0.0f) return NAN;
  if (x == 0.0f) return -INFINITY;
  if (x == INFINITY) return x;
  int e = 0;
  float m = std::frexp(x, &e);
  if (m < 0.70710678f) {
    m = m * 2.0f;
    e -= 1;
  }
  const float s = (m - 1.0f) / (m + 1.0f);
  const float s2 = s * s;
  float q = 0.11111111f;
  q = q * s2 + 0.14285715f;
  q = q * s2 + 0.2f;
  q = q * s2 + 0.33333334f;
  q = q * s2 + 1.0f;
  return (float)e + ((2.0f * s) * q) * 1.44269504f;
}
inline int abs_mask(int v, int m) {
  unsigned u = v < 0 ? 0u - (unsigned)v : (unsigned)v;
  return (int)(u & (unsigned)m);
}

// Perspective-premultiplied vertex attributes: UV(2) I(3) N(3) PXYZ(3).
constexpr int NA = 11;
struct Attr {
  float v[NA];
};
inline Attr a_scale(const Attr& a, float s) {
  Attr r;
  for (int i = 0; i < NA; i++) r.v[i] = a.v[i] * s;
  return r;
}
inline Attr a_sub(const Attr& a, const Attr& b) {
  Attr r;
  for (int i = 0; i < NA; i++) r.v[i] = a.v[i] - b.v[i];
  return r;
}
inline Attr a_add(const Attr& a, const Attr& b) {
  Attr r;
  for (int i = 0; i < NA; i++) r.v[i] = a.v[i] + b.v[i];
  return r;
}
inline Attr a_neg(const Attr& a) {
  Attr r;
  for (int i = 0; i < NA; i++) r.v[i] = -a.v[i];
  return r;
}
// a1 + s * (a3 - a1)
inline Attr a_lerp4(const Attr& a1, const Attr& a3, float s) {
  Attr r;
  for (int i = 0; i < NA; i++) r.v[i] = a1.v[i] + s * (a3.v[i] - a1.v[i]);
  return r;
}
// a + d * s
inline Attr a_mad(const Attr& a, float d, const Attr& s) {
  Attr r;
  for (int i = 0; i < NA; i++) r.v[i] = a.v[i] + d * s.v[i];
  return r;
}
inline void a_inc(Attr& a, const Attr& s) {
  for (int i = 0; i < NA; i++) a.v[i] += s.v[i];
}
inline void a_dec(Attr& a, const Attr& s) {
  for (int i = 0; i < NA; i++) a.v[i] -= s.v[i];
}

// ------------------------------------------------------------------ context
struct MipInfo {
  int mip = 0;
  float fmip = 0.0f;
  int lenMip = 0, lenMip2 = 0, startMip = 0, startMip2 = 0;
};

struct ShadowMap {
  int dim = 0;
  int wS = 0;  // dim/2
  OrcCam cam{};
  std::vector<float> F;
};

struct Ctx {
  const OrcGroup* groups;
  int ngroups;
  const OrcFrame* fr;
  std::vector<int64_t> base;  // first global triangle index per group
  ShadowMap sm[2];
  int nsm = 0;
  // frame
  std::vector<float> F;
  std::vector<uint32_t> id;
  std::vector<uint16_t> R, G, B;
  // setup
  std::vector<f2> P;
  std::vector<float> Z;
  std::vector<int8_t> cls;
};

inline f3 ld3(const float* p, int64_t i) { return {p[3 * i], p[3 * i + 1], p[3 * i + 2]}; }

// Raw (not premultiplied) attributes of global vertex v of group g.
inline Attr raw_attr(const OrcGroup& g, int64_t v) {
  Attr a;
  a.v[0] = g.uv[2 * v];
  a.v[1] = g.uv[2 * v + 1];
  for (int k = 0; k < 3; k++) {
    a.v[2 + k] = g.li[3 * v + k];
    a.v[5 + k] = g.vn[3 * v + k];
    a.v[8 + k] = g.xyz[3 * v + k];
  }
  return a;
}

// ------------------------------------------------------------ a5: trisetup
// Pipe/trisetup.c:28-80 (+ trisetup_cull.c:49-54 when cull).
int setup_one(const float* xyz, const float* vn, int64_t i, const OrcCam& cam, int W,
              int H, int cull, f2 Pout[3], float Zout[3]) {
  f3 vp = {cam.pos[0], cam.pos[1], cam.pos[2]};
  f3 SVd = {cam.vm[0], cam.vm[1], cam.vm[2]};
  f3 SVx = {cam.vm[3], cam.vm[4], cam.vm[5]};
  f3 SVy = {cam.vm[6], cam.vm[7], cam.vm[8]};
  f3 x[3];
  float dd[3], dx[3], dy[3];
  for (int k = 0; k < 3; k++) {
    x[k] = sub3(ld3(xyz, 3 * i + k), vp);
    dd[k] = dot3(x[k], SVd);
    dx[k] = dot3(x[k], SVx);
    dy[k] = dot3(x[k], SVy);
  }
  for (int k = 0; k < 3; k++) dd[k] = 1.0f / dd[k];
  for (int k = 0; k < 3; k++) {
    dx[k] = dx[k] * dd[k];
    dy[k] = dy[k] * dd[k];
  }
  bool yes = false;
  if (dd[0] > 0 && dd[1] > 0 && dd[2] > 0 && std::fabs(dx[0]) < cam.cax &&
      std::fabs(dx[1]) < cam.cax && std::fabs(dx[2]) < cam.cax &&
      std::fabs(dy[0]) < cam.cay && std::fabs(dy[1]) < cam.cay && std::fabs(dy[2]) < cam.cay)
    yes = true;
  if (cull) {
    bool back = true;
    for (int k = 0; k < 3; k++)
      if (!(dot3(ld3(vn, 3 * i + k), x[k]) < -0.0f)) back = false;
    if (back) yes = false;
  }
  const float hw = (float)(W / 2), hh = (float)(H / 2);
  int dix[3], diy[3];
  for (int k = 0; k < 3; k++) {
    dx[k] = dx[k] * -cam.scale + hw;
    dy[k] = dy[k] * cam.scale + hh;
    dix[k] = f2i(dx[k]);
    diy[k] = f2i(dy[k]);
  }
  int64_t maxX = imax(dix[0], imax(dix[1], dix[2]));
  int64_t minX = imin(dix[0], imin(dix[1], dix[2]));
  int64_t maxY = imax(diy[0], imax(diy[1], diy[2]));
  int64_t minY = imin(diy[0], imin(diy[1], diy[2]));
  int64_t dimx = maxX - minX, dimy = maxY - minY;
  if (dimx == 0 || dimx > 2 * (int64_t)W) yes = false;
  if (dimy == 0 || dimy > 2 * (int64_t)H) yes = false;
  if (maxX < 0 || maxY < 0 || minX > W || minY > H) yes = false;
  bool small = std::max(dimx, dimy) < 96;  // SMALL_THRESH * TILE_SIZE
  for (int k = 0; k < 3; k++) {
    Pout[k] = {dx[k], dy[k]};
    Zout[k] = dd[k];
  }
  if (!yes) return 0;
  return small ? 1 : 2;
}

// ---------------------------------------------------- a8: coarse binning
// Pipe/coarse.c:4-127.  Calls emit(tile) per (possibly duplicate) entry.
template <class Emit>
void coarse_walk(const f2 P[3], int WC, int HC, Emit&& emit) {
  float x1 = P[0].x / 16.0f - 0.5f, y1 = P[0].y / 16.0f - 0.5f;
  float x2 = P[1].x / 16.0f - 0.5f, y2 = P[1].y / 16.0f - 0.5f;
  float x3 = P[2].x / 16.0f - 0.5f, y3 = P[2].y / 16.0f - 0.5f;
  float t;
  if (y1 > y2) { t = y1; y1 = y2; y2 = t; t = x1; x1 = x2; x2 = t; }
  if (y2 > y3) { t = y2; y2 = y3; y3 = t; t = x2; x2 = x3; x3 = t; }
  if (y1 > y2) { t = y1; y1 = y2; y2 = t; t = x1; x1 = x2; x2 = t; }
  float x4 = (x1 + ((y2 - y1) / (y3 - y1)) * (x3 - x1));
  float y4 = y2;
  float slope1 = (x2 - x1) / (y2 - y1);
  float slope2 = (x4 - x1) / (y4 - y1);
  float cx1 = x1, cx2 = x1;
  if (slope1 < slope2) { t = slope1; slope1 = slope2; slope2 = t; }
  const float wmax = (float)WC - 1;
  float sbias = 1.0f + std::floor(y1) - y1;
  if (y1 != y2) {
    for (int cy = f2i(std::floor(y1)); (float)cy <= std::ceil(y2); cy++) {
      float mx1 = cx1 + sbias * slope1;
      if (slope1 < 0) mx1 -= slope1;
      float mx2 = cx2 + sbias * slope2;
      if (slope2 > 0) mx2 -= slope2;
      mx1 = omin(mx1, wmax);
      mx2 = omax(mx2, 0.0f);
      if (cy >= 0 && cy < HC)
        for (int ax = f2i(std::floor(mx2)); (float)ax <= std::ceil(mx1); ax++) emit(WC * cy + ax);
      cx1 += slope1;
      cx2 += slope2;
    }
  }
  slope1 = (x3 - x2) / (y3 - y2);
  slope2 = (x3 - x4) / (y3 - y4);
  cx1 = x3;
  cx2 = x3;
  if (slope1 < slope2) { t = slope1; slope1 = slope2; slope2 = t; }
  sbias = y3 - std::floor(y3);
  if (y2 != y3) {
    for (int cy = f2i(std::ceil(y3)); (float)cy >= std::floor(y2); cy--) {
      float mx1 = cx1 - sbias * slope1;
      if (slope1 < 0) mx1 += slope1;
      float mx2 = cx2 - sbias * slope2;
      if (slope2 > 0) mx2 += slope2;
      mx2 = omin(mx2, wmax);
      mx1 = omax(mx1, 0.0f);
      if (cy >= 0 && cy < HC)
        for (int ax = f2i(std::floor(mx1)); (float)ax <= std::ceil(mx2); ax++) emit(WC * cy + ax);
      cx1 -= slope1;
      cx2 -= slope2;
    }
  }
}

// ------------------------------------------------------ mip level (a10)
// Shaders_src/av_draw_normmap_mip.c:26-58 on sorted vertex coordinates
// (int in the small path, float in the tiled path).  Mutates the
// premultiplied UV and PXYZ through the reference's /= z, *= z round trip.
template <class T>
MipInfo mip_setup(Attr a[3], const float z[3], T x1, T y1, T x2, T y2, T x3, T y3, int numMip,
                  float mipBias) {
  for (int k = 0; k < 3; k++) {
    a[k].v[0] /= z[k];
    a[k].v[1] /= z[k];
    for (int c = 8; c < 11; c++) a[k].v[c] /= z[k];
  }
  float dU1x = a[1].v[0] - a[0].v[0], dU1y = a[1].v[1] - a[0].v[1];
  float dU2x = a[2].v[0] - a[0].v[0], dU2y = a[2].v[1] - a[0].v[1];
  float d1 = (float)((y2 - y1) * (y2 - y1) + (x2 - x1) * (x2 - x1));
  float d2 = (float)((y3 - y1) * (y3 - y1) + (x3 - x1) * (x3 - x1));
  float d3 = (float)((y3 - y2) * (y3 - y2) + (x3 - x2) * (x3 - x2));
  float s1x = std::fabs(dU1x / d1), s1y = std::fabs(dU1y / d1);
  float s2x = std::fabs(dU2x / d2), s2y = std::fabs(dU2y / d2);
  float s3x = std::fabs((a[2].v[0] - a[1].v[0]) / d3);
  float s3y = std::fabs((a[2].v[1] - a[1].v[1]) / d3);
  float dt = omax(s1x, omax(s1y, omax(s2x, s2y)));
  dt = omax(dt, omax(s3x, s3y));
  float fmip = omax(0.0f, omin((float)(numMip - 1), std::fabs(log2_cw(dt)) / 2 + mipBias));
  MipInfo m;
  m.mip = f2i(fmip);
  m.fmip = fmip - (float)m.mip;
  m.lenMip = 1 << m.mip;
  m.lenMip2 = m.lenMip << 1;
  m.startMip = 0;
  for (int i = 0; i < m.mip; i++) m.startMip += (1 << i) * (1 << i);
  m.startMip2 = m.startMip + m.lenMip * m.lenMip;
  for (int k = 0; k < 3; k++) {
    a[k].v[0] *= z[k];
    a[k].v[1] *= z[k];
    for (int c = 8; c < 11; c++) a[k].v[c] *= z[k];
  }
  return m;
}

// ---------------------------------------------------------- a10: shading
// Bilinear taps at one level (av_draw_base.c:29-42 / av_draw_normmap_mip.c:73-93).
inline void bilerp(const OrcGroup& g, float tu, float tv, int L, int start, float out[3]) {
  int tex1 = f2i(tu);
  tu -= (float)tex1;
  tex1 = abs_mask(tex1, L - 1);
  int tex2 = f2i(tv);
  tv -= (float)tex2;
  tex2 = abs_mask(tex2, L - 1);
  int i00 = start + tex1 + L * tex2;
  int i10 = start + imin(tex1 + 1, L - 1) + L * tex2;
  int i01 = start + tex1 + L * imin(tex2 + 1, L - 1);
  int i11 = start + imin(tex1 + 1, L - 1) + L * imin(tex2 + 1, L - 1);
  float texr1 = tu, texr2 = tv;
  float texi1 = 1 - texr1, texi2 = 1 - texr2;
  const uint16_t* pl[3] = {g.tr, g.tg, g.tb};
  for (int c = 0; c < 3; c++) {
    const uint16_t* T = pl[c];
    out[c] = texi1 * texi2 * (float)T[i00] + texr1 * texi2 * (float)T[i10] +
             texi1 * texr2 * (float)T[i01] + texr1 * texr2 * (float)T[i11];
  }
}

// PCF: 5-tap cross (drawreflcubemapfresneltexsh.c:162-182) or build-defined
// 3x3 grid, summed over cascades and capped (av_alpha_max.c:53-84).
float pcf_light(const Ctx& cx, f3 wpos) {
  const OrcFrame& fr = *cx.fr;
  const int K = fr.pcf_taps == 9 ? 9 : 5;
  const int d = 2;  // shDist
  int shadow = 0;
  for (int c = 0; c < cx.nsm; c++) {
    const ShadowMap& sm = cx.sm[c];
    f3 SP = {sm.cam.pos[0], sm.cam.pos[1], sm.cam.pos[2]};
    f3 SVd = {sm.cam.vm[0], sm.cam.vm[1], sm.cam.vm[2]};
    f3 SVx = {sm.cam.vm[3], sm.cam.vm[4], sm.cam.vm[5]};
    f3 SVy = {sm.cam.vm[6], sm.cam.vm[7], sm.cam.vm[8]};
    f3 pos = sub3(wpos, SP);
    float depth = dot3(pos, SVd);
    int wS = sm.wS;
    int sx = f2i(dot3(pos, SVx) * sm.cam.scale) + wS;
    int sy = f2i(dot3(pos, SVy) * -sm.cam.scale) + wS;
    if (sx >= d && sx < 2 * wS - d && sy >= d && sy < 2 * wS - d) {
      const float* SD = sm.F.data();
      const int64_t st = 2 * wS;
      if (K == 5) {
        if (SD[st * sy + sx] < depth) shadow += 1;
        if (SD[st * (sy + d) + sx] < depth) shadow += 1;
        if (SD[st * (sy - d) + sx] < depth) shadow += 1;
        if (SD[st * sy + (sx + d)] < depth) shadow += 1;
        if (SD[st * sy + (sx - d)] < depth) shadow += 1;
      } else {
        for (int oy = -d; oy <= d; oy += d)
          for (int ox = -d; ox <= d; ox += d)
            if (SD[st * (sy + oy) + (sx + ox)] < depth) shadow += 1;
      }
    }
  }
  shadow = imin(K, shadow);
  float light = (float)shadow * fr.amb + (float)(K - shadow);
  if (K == 5)
    light *= 0.2f;
  else
    light = light / 9.0f;
  return light;
}

// One fragment's colour.  c1/c2 are the current premultiplied edge
// attributes; t the span parameter; tz the fragment depth.
void shade(const Ctx& cx, const OrcGroup& g, const MipInfo& mi, float t, float tz,
           const Attr& c1, const Attr& c2, uint16_t rgb[3]) {
  const OrcFrame& fr = *cx.fr;
  const float u = 1 - t;
  float L[3];
  float fu = (u * c2.v[0] + t * c1.v[0]) * tz;
  float fv = (u * c2.v[1] + t * c1.v[1]) * tz;
  if (g.flags & ORC_MIP) {
    float o1[3], o2[3];
    bilerp(g, fu * (float)mi.lenMip, fv * (float)mi.lenMip, mi.lenMip, mi.startMip, o1);
    bilerp(g, fu * (float)mi.lenMip2, fv * (float)mi.lenMip2, mi.lenMip2, mi.startMip2, o2);
    for (int c = 0; c < 3; c++) L[c] = (1 - mi.fmip) * o1[c] + mi.fmip * o2[c];
  } else {
    bilerp(g, fu * (float)g.tex_size, fv * (float)g.tex_size, g.tex_size, 0, L);
  }
  f3 wpos = {(u * c2.v[8] + t * c1.v[8]) * tz, (u * c2.v[9] + t * c1.v[9]) * tz,
             (u * c2.v[10] + t * c1.v[10]) * tz};
  float light = 1;
  if ((g.flags & ORC_SHADOW) && cx.nsm > 0) light = pcf_light(cx, wpos);
  f3 col = {(u * c2.v[2] + t * c1.v[2]) * tz, (u * c2.v[3] + t * c1.v[3]) * tz,
            (u * c2.v[4] + t * c1.v[4]) * tz};
  f3 nrm = normalize3({(u * c2.v[5] + t * c1.v[5]) * tz, (u * c2.v[6] + t * c1.v[6]) * tz,
                       (u * c2.v[7] + t * c1.v[7]) * tz});
  f3 LDir = {fr.ldir[0], fr.ldir[1], fr.ldir[2]};
  f3 LInt = {fr.lint[0], fr.lint[1], fr.lint[2]};
  f3 dirCol = scl3(LInt, omax(0.0f, dot3(nrm, LDir)));
  const float* dc = &dirCol.x;
  const float* cc = &col.x;
  if (g.flags & ORC_PHONG) {
    // av_draw_normmap.c:276-284, 303-310 (without the environment terms)
    f3 VP = {fr.view.pos[0], fr.view.pos[1], fr.view.pos[2]};
    f3 a = normalize3(sub3(wpos, VP));
    f3 h = normalize3(add3(a, LDir));
    float theta = omax(0.0f, dot3(h, nrm));
    float sp = pown_(theta, g.spec_pow) * 256 * 4.0f * light;
    f3 spec = scl3(LInt, sp);
    float k = (float)(g.spec_pow + 2) / (8 * 3.1416f);
    spec = scl3(spec, k);
    const float* sc = &spec.x;
    for (int c = 0; c < 3; c++) rgb[c] = f2u16(sc[c] + L[c] * (light * dc[c] + cc[c]));
  } else {
    // av_draw_base.c:44-55
    for (int c = 0; c < 3; c++)
      rgb[c] = f2u16(L[c] * (light * (dc[c] + cc[c]) + (1 - light) * cc[c]));
  }
}

// exp(x): Cody-Waite reduction x = k ln2 + r (ln 2 split into a 16-bit head
// and a tail) and the degree-7 Taylor polynomial of e^r in Horner form,
// times 2^k; 0 for x <= -80 (and NaN), +inf above 88.
float exp_cw(float x) {
  if (!(x > -80.0f)) return 0.0f;
  if (x > 88.0f) return INFINITY;
  const float k = std::rint(x * 1.44269504f);
  const float r = (x - k * 0.693145751953125f) - k * 1.428606765330187e-6f;
  float p = 0.0001984127f;
  p = p * r + 0.0013888889f;
  p = p * r + 0.008333334f;
  p = p * r + 0.041666668f;
  p = p * r + 0.16666667f;
  p = p * r + 0.5f;
  p = p * r + 1.0f;
  p = p * r + 1.0f;
  return std::ldexp(p, (int)k);
}

// The opaque frame as the SSR pass reads it.
struct Snap {
  std::vector<float> F;
  std::vector<uint16_t> R, G, B;
};

// Shaders_src/av_SSR.c:37-148 (water: REFL_STEP, REFL_LENGTH and ABSORB
// from the frame, REFL_DEPTH 2, REFL_VSIZE 2, REFL_DBIAS 0.5, REFL_FADE 80,
// maxattn 0.1).  Frame reads (own pixel, march, hit) are of the snapshot.
// Decrees: fast_normalize = normalize; exp = exp_cw; float->ushort stores
// saturate toward zero; if no cube face matches (NaN direction) face 0 is
// sampled at the clamped (ax, cy) the shader's rxy still holds.
void shade_ssr(const Ctx& cx, const OrcGroup& g, const Snap& S, int ax, int cy, float t,
               float tz, const Attr& c1, const Attr& c2, uint16_t rgb[3]) {
  const OrcFrame& fr = *cx.fr;
  const int wF = fr.W, hF = fr.H;
  const int64_t p = (int64_t)wF * cy + ax;
  const float u = 1 - t;
  const f3 pos = {(u * c2.v[8] + t * c1.v[8]) * tz, (u * c2.v[9] + t * c1.v[9]) * tz,
                  (u * c2.v[10] + t * c1.v[10]) * tz};
  const f3 norm = normalize3({(u * c2.v[5] + t * c1.v[5]) * tz, (u * c2.v[6] + t * c1.v[6]) * tz,
                              (u * c2.v[7] + t * c1.v[7]) * tz});
  const int lenT = g.tex_size;
  float texr1 = ((u * c2.v[0] + t * c1.v[0]) * tz) * (float)lenT;
  int tex1 = abs_mask(f2i(texr1), lenT - 1);
  float texr2 = ((float)lenT * (u * c2.v[1] + t * c1.v[1])) * tz;
  int tex2 = abs_mask(f2i(texr2), lenT - 1);
  const int tex = tex1 + lenT * tex2;
  const f3 vp = {fr.view.pos[0], fr.view.pos[1], fr.view.pos[2]};
  float frs = 1.0f;
  if (fr.ssr_fresnel) {
    frs = 1 - omax(0.0f, dot3(normalize3(sub3(pos, vp)), norm)) * 0.1f;
    for (int k = 0; k < 6; k++) frs *= frs;
  }
  const float scatter = exp_cw(fr.ssr_absorb * (tz - S.F[p]));
  const float tsr = scatter * (float)S.R[p] + (1 - scatter) * (float)g.tr[tex];
  const float tsg = scatter * (float)S.G[p] + (1 - scatter) * (float)g.tg[tex];
  const float tsb = scatter * (float)S.B[p] + (1 - scatter) * (float)g.tb[tex];

  const f3 a = sub3(pos, vp);
  const float nd = dot3(a, norm);
  const f3 refl = scl3(normalize3(sub3(a, scl3(norm, 2 * nd))), 2.0f);
  const bool test = refl.y < 0;
  const f3 SVd = {fr.view.vm[0], fr.view.vm[1], fr.view.vm[2]};
  const f3 SVx = {fr.view.vm[3], fr.view.vm[4], fr.view.vm[5]};
  const f3 SVy = {fr.view.vm[6], fr.view.vm[7], fr.view.vm[8]};
  float rxyx = (float)ax, rxyy = (float)cy;
  const f3 ar = add3(a, refl);
  const float rpz = dot3(ar, SVd);
  const float rpx = dot3(ar, SVx) * -fr.view.scale / rpz + (float)(wF / 2);
  const float rpy = dot3(ar, SVy) * fr.view.scale / rpz + (float)(hF / 2);
  const float dt = omax(std::fabs(rpx - rxyx), std::fabs(rpy - rxyy));
  const int vx = fr.ssr_step;
  const float slopex = (rpx - rxyx) / dt * (float)vx;
  const float slopey = (rpy - rxyy) / dt * (float)vx;
  const float slopez = (1.f / rpz - 1.f / tz) / dt * (float)vx;
  float sy = (float)cy, sx = (float)ax, sz = 1.f / tz, sd = 2.f;
  int rn = 0, hit = 0, lx = 0, ly = 0;
  for (rn = 0; (hit == 0) && (rn < fr.ssr_steps) && (sx >= 0) && (sx < (float)wF) && (sy >= 0) &&
               (sy < (float)hF) && (sz > 0);
       rn += vx) {
    const float f = S.F[(int64_t)wF * f2i(sy) + f2i(sx)];
    if ((f < 1.f / sz) && (f > 1.f / sz - sd)) {
      lx = f2i(sx);
      ly = f2i(sy);
      hit = 1;
    }
    sx += slopex;
    sy += slopey;
    sz += slopez;
    sd = std::fabs(1.f / sz - 1.f / (sz - slopez)) + 0.5f;
  }
  int face = 0;
  const float m = omax(std::fabs(refl.x), omax(std::fabs(refl.y), std::fabs(refl.z)));
  if (m == std::fabs(refl.x)) {
    face = (refl.x > 0) ? 0 : 3;
    rxyx = refl.z / refl.x;
    rxyy = refl.y / refl.x;
  }
  if (m == std::fabs(refl.y)) {
    face = (refl.y > 0) ? 1 : 4;
    rxyx = refl.x / refl.y;
    rxyy = refl.z / refl.y;
  }
  if (m == std::fabs(refl.z)) {
    face = (refl.z > 0) ? 2 : 5;
    rxyx = refl.y / refl.z;
    rxyy = refl.x / refl.z;
  }
  const int dimR = fr.refl_size[g.ssr] / 2;
  rxyx = omin((rxyx + 1) * (float)dimR, (float)(dimR * 2 - 1));
  rxyy = omin((rxyy + 1) * (float)dimR, (float)(dimR * 2 - 1));
  const int lenR = 2 * dimR;
  const int rex1 = f2i(rxyx) + face * lenR;
  const int rex2 = f2i(rxyy);
  const int64_t rex = rex1 + 6 * (int64_t)lenR * rex2;
  const float RRc = fr.refl_r[g.ssr][rex], RGc = fr.refl_g[g.ssr][rex],
              RBc = fr.refl_b[g.ssr][rex];
  const float fade = omin(1.f, (float)(fr.ssr_steps - rn) / (float)80);
  if (test) frs = 0.f;
  if (hit) {
    const int64_t h = (int64_t)wF * ly + lx;
    const float ox = S.R[h], oy = S.G[h], oz = S.B[h];
    rgb[0] = f2u16((1 - frs) * tsr + frs * (fade * ox + (1 - fade) * RRc));
    rgb[1] = f2u16((1 - frs) * tsg + frs * (fade * oy + (1 - fade) * RGc));
    rgb[2] = f2u16((1 - frs) * tsb + frs * (fade * oz + (1 - fade) * RBc));
  } else {
    rgb[0] = f2u16((1 - frs) * tsr + frs * RRc);
    rgb[1] = f2u16((1 - frs) * tsg + frs * RGc);
    rgb[2] = f2u16((1 - frs) * tsb + frs * RBc);
  }
}

// --------------------------------------------------- a7: small-path raster
// Templates.py:210-323.  frag(ax, cy, t, tz, c1, c2, half) for on-screen
// pixels in evaluation order.
struct SmallTri {
  int x1, y1, x2, y2, x3, y3;
  float z1, z2, z3;
  Attr a1, a2, a3, a4;
};

// Templates.py:241-246 + AVSL.py:457-462: the 4th vertex's attributes are
// formed after the sort and before any shader setup code runs.
inline void small_fourth(SmallTri& s) {
  float ydiff1 = (float)(s.y2 - s.y1) / (float)(s.y3 - s.y1);
  s.a4 = a_lerp4(s.a1, s.a3, ydiff1);
}

inline void small_sort(SmallTri& s) {
  int it;
  float ft;
  Attr at;
  if (s.y1 > s.y2) {
    it = s.y1; s.y1 = s.y2; s.y2 = it; it = s.x1; s.x1 = s.x2; s.x2 = it;
    ft = s.z1; s.z1 = s.z2; s.z2 = ft; at = s.a1; s.a1 = s.a2; s.a2 = at;
  }
  if (s.y2 > s.y3) {
    it = s.y2; s.y2 = s.y3; s.y3 = it; it = s.x2; s.x2 = s.x3; s.x3 = it;
    ft = s.z2; s.z2 = s.z3; s.z3 = ft; at = s.a2; s.a2 = s.a3; s.a3 = at;
  }
  if (s.y1 > s.y2) {
    it = s.y1; s.y1 = s.y2; s.y2 = it; it = s.x1; s.x1 = s.x2; s.x2 = it;
    ft = s.z1; s.z1 = s.z2; s.z2 = ft; at = s.a1; s.a1 = s.a2; s.a2 = at;
  }
}

// Visits rows [rlo, rhi) only; all rows still accumulate.
template <class Frag>
void raster_small(const SmallTri& s, int W, int H, int rlo, int rhi, Frag&& frag) {
  const int x1 = s.x1, y1 = s.y1, x2 = s.x2, y2 = s.y2, x3 = s.x3, y3 = s.y3;
  const float z1 = s.z1, z2 = s.z2, z3 = s.z3;
  float ydiff1 = (float)(y2 - y1) / (float)(y3 - y1);
  float x4 = (float)x1 + ydiff1 * (float)(x3 - x1);
  float z4 = z1 + ydiff1 * (z3 - z1);
  const Attr& a4 = s.a4;
  // bottom half
  ydiff1 = 1 / (float)(y2 - y1);
  float slope1 = (float)(x2 - x1) * ydiff1;
  float slope2 = (x4 - (float)x1) * ydiff1;
  float slopez1 = (z2 - z1) * ydiff1;
  float slopez2 = (z4 - z1) * ydiff1;
  Attr sa1 = a_scale(a_sub(s.a2, s.a1), ydiff1);
  Attr sa2 = a_scale(a_sub(a4, s.a1), ydiff1);
  float cx1 = (float)x1, cx2 = (float)x1, cz1 = z1, cz2 = z1;
  Attr c1 = s.a1, c2 = s.a1;
  if (slope1 < slope2) {
    std::swap(slope1, slope2);
    std::swap(slopez1, slopez2);
    std::swap(sa1, sa2);
  }
  for (int cy = y1; cy <= y2; cy++) {
    if (cy >= rlo && cy < rhi) {
      for (int ax = f2i(cx2); (float)ax <= cx1; ax++) {
        if (cy >= 0 && cy < H && ax >= 0 && ax < W) {
          float t = ((float)ax - cx2) / (cx1 - cx2);
          t = omax(0.0f, omin(1.0f, t));
          float tz = 1 / ((1 - t) * cz2 + t * cz1);
          frag(ax, cy, t, tz, c1, c2, 0);
        }
      }
    }
    cx1 += slope1;
    cx2 += slope2;
    cz1 += slopez1;
    cz2 += slopez2;
    a_inc(c1, sa1);
    a_inc(c2, sa2);
  }
  // top half
  ydiff1 = 1 / (float)(y3 - y2);
  slope1 = (float)(x3 - x2) * ydiff1;
  slope2 = ((float)x3 - x4) * ydiff1;
  slopez1 = (z3 - z2) * ydiff1;
  slopez2 = (z3 - z4) * ydiff1;
  sa1 = a_scale(a_neg(a_sub(s.a2, s.a3)), ydiff1);
  sa2 = a_scale(a_neg(a_sub(a4, s.a3)), ydiff1);
  cx1 = (float)x3;
  cx2 = (float)x3;
  cz1 = z3;
  cz2 = z3;
  c1 = s.a3;
  c2 = s.a3;
  if (slope1 < slope2) {
    std::swap(slope1, slope2);
    std::swap(slopez1, slopez2);
    std::swap(sa1, sa2);
  }
  for (int cy = y3; cy >= y2; cy--) {
    if (cy >= rlo && cy < rhi) {
      for (int ax = f2i(cx1); (float)ax <= cx2; ax++) {
        if (cy >= 0 && cy < H && ax >= 0 && ax < W) {
          float t = ((float)ax - cx2) / (cx1 - cx2);
          t = omax(0.0f, omin(1.0f, t));
          float tz = 1 / ((1 - t) * cz2 + t * cz1);
          frag(ax, cy, t, tz, c1, c2, 1);
        }
      }
    }
    cx1 -= slope1;
    cx2 -= slope2;
    cz1 -= slopez1;
    cz2 -= slopez2;
    a_dec(c1, sa1);
    a_dec(c2, sa2);
  }
}

// ---------------------------------------------------- a9: tiled raster
// Templates.py:16-185 for one tile bx of a W-wide screen.
struct TileTri {
  float x1, y1, x2, y2, x3, y3;
  float z1, z2, z3;
  Attr a1, a2, a3, a4;
};

inline void tile_fourth(TileTri& s) {
  float ydiff1 = (s.y2 - s.y1) / (s.y3 - s.y1);
  s.a4 = a_lerp4(s.a1, s.a3, ydiff1);
}

inline void tile_sort(TileTri& s) {
  float ft;
  Attr at;
  if (s.y1 > s.y2) {
    ft = s.y1; s.y1 = s.y2; s.y2 = ft; ft = s.x1; s.x1 = s.x2; s.x2 = ft;
    ft = s.z1; s.z1 = s.z2; s.z2 = ft; at = s.a1; s.a1 = s.a2; s.a2 = at;
  }
  if (s.y2 > s.y3) {
    ft = s.y2; s.y2 = s.y3; s.y3 = ft; ft = s.x2; s.x2 = s.x3; s.x3 = ft;
    ft = s.z2; s.z2 = s.z3; s.z3 = ft; at = s.a2; s.a2 = s.a3; s.a3 = at;
  }
  if (s.y1 > s.y2) {
    ft = s.y1; s.y1 = s.y2; s.y2 = ft; ft = s.x1; s.x1 = s.x2; s.x2 = ft;
    ft = s.z1; s.z1 = s.z2; s.z2 = ft; at = s.a1; s.a1 = s.a2; s.a2 = at;
  }
}

struct TileBox {
  float xMin, xMax, yMin, yMax;
};
inline TileBox tile_box(int bx, int W) {
  int tileX = bx % (int)((float)W / 16.0f);
  int tileY = f2i((float)bx / ((float)W / 16.0f));
  TileBox b;
  b.xMin = (float)tileX * 16.0f;
  b.xMax = ((float)tileX + 1.0f) * 16.0f - 1.0f;
  b.yMin = (float)tileY * 16.0f;
  b.yMax = ((float)tileY + 1.0f) * 16.0f - 1.0f;
  return b;
}

template <class Frag>
void raster_tile(const TileTri& s, const TileBox& tb, Frag&& frag) {
  const float x1 = s.x1, y1 = s.y1, x2 = s.x2, y2 = s.y2, x3 = s.x3, y3 = s.y3;
  const float z1 = s.z1, z2 = s.z2, z3 = s.z3;
  const float xMin = tb.xMin, xMax = tb.xMax, yMin = tb.yMin, yMax = tb.yMax;
  float ydiff1 = (y2 - y1) / (y3 - y1);
  float x4 = x1 + ydiff1 * (x3 - x1);
  float z4 = z1 + ydiff1 * (z3 - z1);
  const Attr& a4 = s.a4;
  // bottom half
  ydiff1 = 1 / (y2 - y1);
  float slope1 = (x2 - x1) * ydiff1;
  float slope2 = (x4 - x1) * ydiff1;
  float slopez1 = (z2 - z1) * ydiff1;
  float slopez2 = (z4 - z1) * ydiff1;
  Attr sa1 = a_scale(a_sub(s.a2, s.a1), ydiff1);
  Attr sa2 = a_scale(a_sub(a4, s.a1), ydiff1);
  if (slope1 < slope2) {
    std::swap(slope1, slope2);
    std::swap(slopez1, slopez2);
    std::swap(sa1, sa2);
  }
  int cy = f2i(oclamp(y1, yMin, yMax));
  float dy = (float)cy - y1;
  float cx1 = x1 + dy * slope1;
  float cx2 = x1 + dy * slope2;
  float cz1 = z1 + dy * slopez1;
  float cz2 = z1 + dy * slopez2;
  Attr c1 = a_mad(s.a1, dy, sa1);
  Attr c2 = a_mad(s.a1, dy, sa2);
  if (y2 > yMin) {
    for (; (float)cy <= oclamp(y2, yMin, yMax); cy++) {
      if (cx2 < xMax && cx1 > xMin) {
        for (int ax = f2i(oclamp(cx2, xMin, xMax)); (float)ax <= oclamp(cx1, xMin, xMax); ax++) {
          float t = ((float)ax - cx2) / (cx1 - cx2);
          t = omax(0.0f, omin(1.0f, t));
          float tz = 1 / ((1 - t) * cz2 + t * cz1);
          frag(ax, cy, t, tz, c1, c2, 0);
        }
      }
      cx1 += slope1;
      cx2 += slope2;
      cz1 += slopez1;
      cz2 += slopez2;
      a_inc(c1, sa1);
      a_inc(c2, sa2);
    }
  }
  // top half
  ydiff1 = 1 / (y3 - y2);
  slope1 = (x3 - x2) * ydiff1;
  slope2 = (x3 - x4) * ydiff1;
  slopez1 = (z3 - z2) * ydiff1;
  slopez2 = (z3 - z4) * ydiff1;
  sa1 = a_scale(a_neg(a_sub(s.a2, s.a3)), ydiff1);
  sa2 = a_scale(a_neg(a_sub(a4, s.a3)), ydiff1);
  if (slope1 < slope2) {
    std::swap(slope1, slope2);
    std::swap(slopez1, slopez2);
    std::swap(sa1, sa2);
  }
  cy = f2i(oclamp(y3, yMin, yMax));
  dy = (float)cy - y3;
  cx1 = x3 + dy * slope1;
  cx2 = x3 + dy * slope2;
  cz1 = z3 + dy * slopez1;
  cz2 = z3 + dy * slopez2;
  c1 = a_mad(s.a3, dy, sa1);
  c2 = a_mad(s.a3, dy, sa2);
  if (y2 < yMax) {
    for (; (float)cy >= oclamp(y2, yMin, yMax); cy--) {
      if (cx1 < xMax && cx2 > xMin) {
        for (int ax = f2i(oclamp(cx1, xMin, xMax)); (float)ax <= oclamp(cx2, xMin, xMax); ax++) {
          float t = ((float)ax - cx2) / (cx1 - cx2);
          t = omax(0.0f, omin(1.0f, t));
          float tz = 1 / ((1 - t) * cz2 + t * cz1);
          frag(ax, cy, t, tz, c1, c2, 1);
        }
      }
      cx1 -= slope1;
      cx2 -= slope2;
      cz1 -= slopez1;
      cz2 -= slopez2;
      a_dec(c1, sa1);
      a_dec(c2, sa2);
    }
  }
}

// ------------------------------------------------------ a11: shadow pass
// Pipe/trisetupOrtho.c:20-59.
bool shadow_setup_one(const float* xyz, int64_t i, const OrcCam& cam, int dim, float bias,
                      int ix[3], int iy[3], float zz[3]) {
  f3 vp = {cam.pos[0], cam.pos[1], cam.pos[2]};
  f3 SVd = {cam.vm[0], cam.vm[1], cam.vm[2]};
  f3 SVx = {cam.vm[3], cam.vm[4], cam.vm[5]};
  f3 SVy = {cam.vm[6], cam.vm[7], cam.vm[8]};
  float dd[3], dx[3], dy[3];
  for (int k = 0; k < 3; k++) {
    f3 x = sub3(ld3(xyz, 3 * i + k), vp);
    dd[k] = dot3(x, SVd);
    dx[k] = dot3(x, SVx);
    dy[k] = dot3(x, SVy);
  }
  const float half = (float)(dim / 2);
  for (int k = 0; k < 3; k++) {
    dd[k] += bias;
    dd[k] = 1 / dd[k];
    dx[k] = dx[k] * cam.scale + half;
    dy[k] = dy[k] * -cam.scale + half;
    ix[k] = f2i(dx[k]);
    iy[k] = f2i(dy[k]);
    zz[k] = dd[k];
  }
  bool yes = true;
  for (int k = 0; k < 3; k++) {
    if (!(dd[k] > 0)) yes = false;
    if (!(ix[k] >= -120 && ix[k] < dim + 120 && iy[k] >= -120 && iy[k] < dim + 120)) yes = false;
  }
  return yes;
}

struct ShTri {
  int x1, y1, x2, y2, x3, y3;
  float z1, z2, z3;
};

// Shaders/drawmin.c:86-176.  frag(ax, cy, tz, half).
template <class Frag>
void raster_shadow(ShTri s, int wF, int hF, int rlo, int rhi, Frag&& frag) {
  int it;
  float zt;
  if (s.y1 > s.y2) {
    it = s.y1; s.y1 = s.y2; s.y2 = it; it = s.x1; s.x1 = s.x2; s.x2 = it;
    zt = s.z1; s.z1 = s.z2; s.z2 = zt;
  }
  if (s.y2 > s.y3) {
    it = s.y2; s.y2 = s.y3; s.y3 = it; it = s.x2; s.x2 = s.x3; s.x3 = it;
    zt = s.z2; s.z2 = s.z3; s.z3 = zt;
  }
  if (s.y1 > s.y2) {
    it = s.y1; s.y1 = s.y2; s.y2 = it; it = s.x1; s.x1 = s.x2; s.x2 = it;
    zt = s.z1; s.z1 = s.z2; s.z2 = zt;
  }
  const int x1 = s.x1, y1 = s.y1, x2 = s.x2, y2 = s.y2, x3 = s.x3, y3 = s.y3;
  const float z1 = s.z1, z2 = s.z2, z3 = s.z3;
  if (!(y1 < hF && y3 >= 0)) return;
  int x4 = f2i((float)x1 + ((float)(y2 - y1) / (float)(y3 - y1)) * (float)(x3 - x1));
  int y4 = y2;
  float z4 = z1 + (float)(y2 - y1) / (float)(y3 - y1) * (z3 - z1);
  float slope1 = (float)(x2 - x1) / (float)(y2 - y1);
  float slope2 = (float)(x4 - x1) / (float)(y4 - y1);
  float slopez1 = (z2 - z1) / (float)(y2 - y1);
  float slopez2 = (z4 - z1) / (float)(y4 - y1);
  float cx1 = (float)x1, cx2 = (float)x1, cz1 = z1, cz2 = z1;
  if (slope1 < slope2) {
    std::swap(slope1, slope2);
    std::swap(slopez1, slopez2);
  }
  for (int cy = y1; cy <= y2; cy++) {
    if (cy >= rlo && cy < rhi) {
      for (int ax = f2i(cx2); ax <= f2i(cx1); ax++) {
        if (cy >= 0 && cy < hF && ax >= 0 && ax < wF) {
          float t = ((float)ax - cx2) / (cx1 - cx2);
          t = omax(0.0f, omin(1.0f, t));
          float tz = 1 / ((1 - t) * cz2 + t * cz1);
          frag(ax, cy, tz, 0);
        }
      }
    }
    cx1 += slope1;
    cx2 += slope2;
    cz1 += slopez1;
    cz2 += slopez2;
  }
  slope1 = (float)(x3 - x2) / (float)(y3 - y2);
  slope2 = (float)(x3 - x4) / (float)(y3 - y4);
  slopez1 = (z3 - z2) / (float)(y3 - y2);
  slopez2 = (z3 - z4) / (float)(y3 - y4);
  cx1 = (float)x3;
  cx2 = (float)x3;
  cz1 = z3;
  cz2 = z3;
  if (slope1 < slope2) {
    std::swap(slope1, slope2);
    std::swap(slopez1, slopez2);
  }
  for (int cy = y3; cy > y2; cy--) {
    if (cy >= rlo && cy < rhi) {
      for (int ax = f2i(cx1); ax <= f2i(cx2); ax++) {
        if (cy >= 0 && cy < hF && ax >= 0 && ax < wF) {
          float t = ((float)ax - cx2) / (cx1 - cx2);
          t = omax(0.0f, omin(1.0f, t));
          float tz = 1 / ((1 - t) * cz2 + t * cz1);
          frag(ax, cy, tz, 1);
        }
      }
    }
    cx1 -= slope1;
    cx2 -= slope2;
    cz1 -= slopez1;
    cz2 -= slopez2;
  }
}

// ---------------------------------------------------------------- threads
template <class Fn>
void parallel_for(int n, int nthreads, Fn&& fn) {
  if (nthreads <= 1 || n <= 1) {
    for (int i = 0; i < n; i++) fn(i);
    return;
  }
  std::atomic<int> next{0};
  std::vector<std::thread> th;
  int nt = std::min(nthreads, n);
  for (int w = 0; w < nt; w++)
    th.emplace_back([&]() {
      for (;;) {
        int i = next.fetch_add(1);
        if (i >= n) break;
        fn(i);
      }
    });
  for (auto& t : th) t.join();
}

int resolve_threads(int n) {
  if (n > 0) return n;
  unsigned h = std::thread::hardware_concurrency();
  return h ? (int)h : 1;
}

constexpr int BAND = 16;

// Buckets item indices by the 16-row bands [ylo, yhi] touches, in order.
void bucket_rows(std::vector<std::vector<uint32_t>>& bands, uint32_t item, int ylo, int yhi,
                 int rows) {
  if (yhi < 0 || ylo >= rows) return;
  int b0 = std::max(ylo, 0) / BAND, b1 = std::min(yhi, rows - 1) / BAND;
  for (int b = b0; b <= b1; b++) bands[b].push_back(item);
}

int64_t render_shadow(Ctx& cx, int c, int nthreads) {
  const OrcShadowCfg& cfg = cx.fr->shadow[c];
  ShadowMap& sm = cx.sm[c];
  sm.dim = cfg.dim;
  sm.wS = cfg.dim / 2;
  sm.cam = cfg.cam;
  sm.F.assign((size_t)cfg.dim * cfg.dim, 65535.0f);  // Shaders/clearzb.c
  int64_t T = cx.base[cx.ngroups];
  std::vector<ShTri> tris(T);
  std::vector<int8_t> ok(T, 0);
  parallel_for((int)((T + 4095) / 4096), nthreads, [&](int blk) {
    int64_t lo = (int64_t)blk * 4096, hi = std::min<int64_t>(T, lo + 4096);
    for (int64_t gi = lo; gi < hi; gi++) {
      int g = (int)(std::upper_bound(cx.base.begin(), cx.base.end(), gi) - cx.base.begin()) - 1;
      const OrcGroup& G = cx.groups[g];
      if (!G.cast) continue;
      int ix[3], iy[3];
      float zz[3];
      if (shadow_setup_one(G.xyz, gi - cx.base[g], cfg.cam, cfg.dim, cfg.bias, ix, iy, zz)) {
        ok[gi] = 1;
        tris[gi] = {ix[0], iy[0], ix[1], iy[1], ix[2], iy[2], zz[0], zz[1], zz[2]};
      }
    }
  });
  int nb = (cfg.dim + BAND - 1) / BAND;
  std::vector<std::vector<uint32_t>> bands(nb);
  for (int64_t gi = 0; gi < T; gi++) {
    if (!ok[gi]) continue;
    const ShTri& s = tris[gi];
    bucket_rows(bands, (uint32_t)gi, std::min(s.y1, std::min(s.y2, s.y3)),
                std::max(s.y1, std::max(s.y2, s.y3)), cfg.dim);
  }
  float* F = sm.F.data();
  const int dim = cfg.dim;
  std::atomic<int64_t> nfrag{0};
  parallel_for(nb, nthreads, [&](int b) {
    int rlo = b * BAND, rhi = std::min(dim, rlo + BAND);
    int64_t nf = 0;
    for (uint32_t gi : bands[b])
      raster_shadow(tris[gi], dim, dim, rlo, rhi, [&](int ax, int cy, float tz, int) {
        nf++;
        float& f = F[(int64_t)dim * cy + ax];
        if (f > tz) f = tz;
      });
    nfrag += nf;
  });
  return nfrag.load();
}

}  // namespace

// =================================================================== C ABI
namespace {

// One drawAll pass over the groups of one kind (water = the SSR groups)
// in the reference's order: setup, small triangles per 16-row band in
// global order, then large triangles per group and tile.  on_frag(G, mi,
// gi, ax, cy, t, tz, c1, c2) sees every fragment in that order (per pixel).
template <class OnFrag>
void draw_pass(Ctx& cx, bool water, int nthreads, int64_t st[5], OnFrag&& on_frag) {
  const OrcGroup* groups = cx.groups;
  const int ngroups = cx.ngroups;
  const OrcFrame* fr = cx.fr;
  const int W = fr->W, H = fr->H;
  const int64_t T = cx.base[ngroups];
  // Triangle setup for every group drawn in this pass.
  cx.P.assign(3 * T, f2{0, 0});
  cx.Z.assign(3 * T, 0.0f);
  cx.cls.assign(T, 0);
  parallel_for((int)((T + 4095) / 4096), nthreads, [&](int blk) {
    int64_t lo = (int64_t)blk * 4096, hi = std::min<int64_t>(T, lo + 4096);
    for (int64_t gi = lo; gi < hi; gi++) {
      int g = (int)(std::upper_bound(cx.base.begin(), cx.base.end(), gi) - cx.base.begin()) - 1;
      const OrcGroup& G = groups[g];
      if (G.skip || (G.ssr >= 0) != water) continue;
      cx.cls[gi] = (int8_t)setup_one(G.xyz, G.vn, gi - cx.base[g], fr->view, W, H,
                                     (G.flags & ORC_CULL) ? 1 : 0, &cx.P[3 * gi], &cx.Z[3 * gi]);
    }
  });

  // Small triangles bucketed into 16-row bands in pass order.
  const int nb = (H + BAND - 1) / BAND;
  std::vector<std::vector<uint32_t>> sbands(nb);
  int64_t nsmall = 0, nlarge = 0;
  for (int64_t gi = 0; gi < T; gi++) {
    if (cx.cls[gi] == 1) {
      nsmall++;
      int y[3];
      for (int k = 0; k < 3; k++) y[k] = f2i(cx.P[3 * gi + k].y);
      bucket_rows(sbands, (uint32_t)gi, std::min(y[0], std::min(y[1], y[2])),
                  std::max(y[0], std::max(y[1], y[2])), H);
    } else if (cx.cls[gi] == 2) {
      nlarge++;
    }
  }
  st[0] = nsmall;
  st[1] = nlarge;

  // Coarse bins per group (Ops0.py:614-643: coarse then tiled, per group).
  const int WC = W / 16, HC = H / 16, NT = WC * HC;
  std::vector<std::vector<std::vector<uint32_t>>> bins(ngroups);
  int64_t maxbin = 0, entries = 0;
  for (int g = 0; g < ngroups; g++) {
    bool any = false;
    for (int64_t gi = cx.base[g]; gi < cx.base[g + 1]; gi++)
      if (cx.cls[gi] == 2) { any = true; break; }
    if (!any) continue;
    bins[g].assign(NT, {});
    for (int64_t gi = cx.base[g]; gi < cx.base[g + 1]; gi++) {
      if (cx.cls[gi] != 2) continue;
      coarse_walk(&cx.P[3 * gi], WC, HC, [&](int t) { bins[g][t].push_back((uint32_t)gi); });
    }
    for (auto& b : bins[g]) {
      maxbin = std::max<int64_t>(maxbin, (int64_t)b.size());
      entries += (int64_t)b.size();
    }
  }
  st[3] = maxbin;
  st[4] = entries;
  // The box each coarse tile's list is walked in: tile_box(bx, W), the tile
  // itself when W is a multiple of 16; for other widths some tiles fold onto
  // a box of a row above.  A box walked for several tiles (the reference
  // runs their work-groups concurrently on it) takes the union of their
  // lists in ascending global index: the large pass's order is the triangle
  // order, as for a box walked once.
  std::vector<std::vector<int>> src(NT);
  for (int bx = 0; bx < NT; bx++) {
    const TileBox tb = tile_box(bx, W);
    const int tx = f2i(tb.xMin / 16.0f), ty = f2i(tb.yMin / 16.0f);
    src[(size_t)ty * WC + tx].push_back(bx);
  }

  std::atomic<int64_t> frags{0};
  parallel_for(nb, nthreads, [&](int b) {
    const int rlo = b * BAND, rhi = std::min(H, rlo + BAND);
    int64_t nf = 0;
    // -- small pass, all groups in order
    for (uint32_t gi : sbands[b]) {
      int g = (int)(std::upper_bound(cx.base.begin(), cx.base.end(), (int64_t)gi) -
                    cx.base.begin()) - 1;
      const OrcGroup& G = groups[g];
      const int64_t li = gi - cx.base[g];
      SmallTri s;
      s.x1 = f2i(cx.P[3 * gi].x); s.y1 = f2i(cx.P[3 * gi].y);
      s.x2 = f2i(cx.P[3 * gi + 1].x); s.y2 = f2i(cx.P[3 * gi + 1].y);
      s.x3 = f2i(cx.P[3 * gi + 2].x); s.y3 = f2i(cx.P[3 * gi + 2].y);
      s.z1 = cx.Z[3 * gi]; s.z2 = cx.Z[3 * gi + 1]; s.z3 = cx.Z[3 * gi + 2];
      s.a1 = a_scale(raw_attr(G, 3 * li), s.z1);
      s.a2 = a_scale(raw_attr(G, 3 * li + 1), s.z2);
      s.a3 = a_scale(raw_attr(G, 3 * li + 2), s.z3);
      small_sort(s);
      small_fourth(s);
      MipInfo mi;
      if (G.flags & ORC_MIP) {
        Attr a[3] = {s.a1, s.a2, s.a3};
        float z[3] = {s.z1, s.z2, s.z3};
        mi = mip_setup<int>(a, z, s.x1, s.y1, s.x2, s.y2, s.x3, s.y3, G.tex_size, G.mip_bias);
        s.a1 = a[0]; s.a2 = a[1]; s.a3 = a[2];
      }
      raster_small(s, W, H, rlo, rhi,
                   [&](int ax, int cy, float t, float tz, const Attr& c1, const Attr& c2, int) {
                     nf++;
                     on_frag(G, mi, gi, ax, cy, t, tz, c1, c2);
                   });
    }
    // -- large pass, per group: tiles of this band row
    if (b < HC) {
      for (int g = 0; g < ngroups; g++) {
        if (bins[g].empty()) continue;
        const OrcGroup& G = groups[g];
        for (int tx = 0; tx < WC; tx++) {
          const std::vector<int>& from = src[(size_t)b * WC + tx];
          if (from.empty()) continue;
          const TileBox tb = tile_box(from[0], W);
          std::vector<uint32_t> merged;
          const std::vector<uint32_t>* list = &bins[g][from[0]];
          if (from.size() > 1) {
            for (int bx : from) merged.insert(merged.end(), bins[g][bx].begin(), bins[g][bx].end());
            std::stable_sort(merged.begin(), merged.end());
            list = &merged;
          }
          for (uint32_t gi : *list) {
            const int64_t li = gi - cx.base[g];
            TileTri s;
            s.x1 = cx.P[3 * gi].x; s.y1 = cx.P[3 * gi].y;
            s.x2 = cx.P[3 * gi + 1].x; s.y2 = cx.P[3 * gi + 1].y;
            s.x3 = cx.P[3 * gi + 2].x; s.y3 = cx.P[3 * gi + 2].y;
            s.z1 = cx.Z[3 * gi]; s.z2 = cx.Z[3 * gi + 1]; s.z3 = cx.Z[3 * gi + 2];
            s.a1 = a_scale(raw_attr(G, 3 * li), s.z1);
            s.a2 = a_scale(raw_attr(G, 3 * li + 1), s.z2);
            s.a3 = a_scale(raw_attr(G, 3 * li + 2), s.z3);
            tile_sort(s);
            tile_fourth(s);
            MipInfo mi;
            if (G.flags & ORC_MIP) {
              Attr a[3] = {s.a1, s.a2, s.a3};
              float z[3] = {s.z1, s.z2, s.z3};
              mi = mip_setup<float>(a, z, s.x1, s.y1, s.x2, s.y2, s.x3, s.y3, G.tex_size,
                                    G.mip_bias);
              s.a1 = a[0]; s.a2 = a[1]; s.a3 = a[2];
            }
            raster_tile(s, tb,
                        [&](int ax, int cy, float t, float tz, const Attr& c1, const Attr& c2,
                            int) {
                          nf++;
                          on_frag(G, mi, gi, ax, cy, t, tz, c1, c2);
                        });
          }
        }
      }
    }
    frags += nf;
  });
  st[2] = frags.load();
}

}  // namespace

extern "C" {

int orc_trisetup(const float* xyz, const float* vn, int ntri, const OrcCam* cam, int W, int H,
                 int cull, float* P, float* Z, int8_t* cls) {
  for (int64_t i = 0; i < ntri; i++) {
    f2 p[3];
    float z[3];
    cls[i] = (int8_t)setup_one(xyz, vn, i, *cam, W, H, cull, p, z);
    for (int k = 0; k < 3; k++) {
      P[6 * i + 2 * k] = p[k].x;
      P[6 * i + 2 * k + 1] = p[k].y;
      Z[3 * i + k] = z[k];
    }
  }
  return 0;
}

int orc_shadow_setup(const float* xyz, int ntri, const OrcCam* cam, int dim, float bias,
                     int32_t* XY, float* Z, int8_t* ok) {
  for (int64_t i = 0; i < ntri; i++) {
    int ix[3], iy[3];
    float zz[3];
    ok[i] = shadow_setup_one(xyz, i, *cam, dim, bias, ix, iy, zz) ? 1 : 0;
    for (int k = 0; k < 3; k++) {
      XY[6 * i + 2 * k] = ix[k];
      XY[6 * i + 2 * k + 1] = iy[k];
      Z[3 * i + k] = zz[k];
    }
  }
  return 0;
}

int orc_coarse_tri(const float* P6, int WC, int HC, int32_t* tiles, int cap) {
  f2 P[3] = {{P6[0], P6[1]}, {P6[2], P6[3]}, {P6[4], P6[5]}};
  int n = 0;
  coarse_walk(P, WC, HC, [&](int t) {
    if (n < cap) tiles[n] = t;
    n++;
  });
  return n <= cap ? n : -n;
}

int orc_fragments(int mode, const float* P6, const float* Z3, int W, int H, int tile, int32_t* xs,
                  int32_t* ys, float* tz, int8_t* half, int cap) {
  int n = 0;
  auto put = [&](int ax, int cy, float z, int h) {
    if (n < cap) {
      xs[n] = ax;
      ys[n] = cy;
      tz[n] = z;
      half[n] = (int8_t)h;
    }
    n++;
  };
  Attr zero;
  for (int i = 0; i < NA; i++) zero.v[i] = 0.0f;
  if (mode == 0) {
    SmallTri s;
    s.x1 = f2i(P6[0]); s.y1 = f2i(P6[1]);
    s.x2 = f2i(P6[2]); s.y2 = f2i(P6[3]);
    s.x3 = f2i(P6[4]); s.y3 = f2i(P6[5]);
    s.z1 = Z3[0]; s.z2 = Z3[1]; s.z3 = Z3[2];
    s.a1 = s.a2 = s.a3 = zero;
    small_sort(s);
    small_fourth(s);
    raster_small(s, W, H, -2147483647, 2147483647,
                 [&](int ax, int cy, float, float z, const Attr&, const Attr&, int h) {
                   put(ax, cy, z, h);
                 });
  } else if (mode == 1) {
    TileTri s;
    s.x1 = P6[0]; s.y1 = P6[1]; s.x2 = P6[2]; s.y2 = P6[3]; s.x3 = P6[4]; s.y3 = P6[5];
    s.z1 = Z3[0]; s.z2 = Z3[1]; s.z3 = Z3[2];
    s.a1 = s.a2 = s.a3 = zero;
    tile_sort(s);
    tile_fourth(s);
    raster_tile(s, tile_box(tile, W),
                [&](int ax, int cy, float, float z, const Attr&, const Attr&, int h) {
                  put(ax, cy, z, h);
                });
  } else {
    ShTri s = {f2i(P6[0]), f2i(P6[1]), f2i(P6[2]), f2i(P6[3]), f2i(P6[4]), f2i(P6[5]),
               Z3[0], Z3[1], Z3[2]};
    raster_shadow(s, W, H, -2147483647, 2147483647,
                  [&](int ax, int cy, float z, int h) { put(ax, cy, z, h); });
  }
  return n <= cap ? n : -n;
}

int orc_gamma(uint16_t* r, uint16_t* g, uint16_t* b, int64_t n, float ex) {
  // Shaders/Post/gamma.c:13-29
  const float k = 8 * ex;
  for (int64_t i = 0; i < n; i++) {
    float j[3] = {std::sqrt(k * (float)r[i]) / 1.1f, std::sqrt(k * (float)g[i]) / 1.1f,
                  std::sqrt(k * (float)b[i]) / 1.1f};
    float m = omax(j[0], omax(j[1], j[2])) / 255.0f;
    if (m > 1.0f) {
      float s = std::sqrt(m);
      for (int c = 0; c < 3; c++) j[c] = j[c] / s;
    }
    r[i] = (uint16_t)omin(255.0f, j[0]);
    g[i] = (uint16_t)omin(255.0f, j[1]);
    b[i] = (uint16_t)omin(255.0f, j[2]);
  }
  return 0;
}

// ------------------------------------------------------- a13+: blur
// CLDraw.blur (Ops.py:1009-1020): Shaders/Post/blur.c blurH, then
// Shaders/Post/blur2.c blurV, as serial phases (the kernels' global
// barriers do not order work-groups).  Decrees where the text is undefined:
// literals are float; float->ushort stores saturate toward zero; the
// half-resolution buffers start at zero each call (rows the kernels never
// write stay zero); the vertical blur skips taps outside the half-res rows
// (blur2.c:47-56 bounds them by the width, reading past the buffer).
namespace {
inline uint16_t sat_rtz(float v) { return f2u16(v); }
const float kTap[10] = {0.10125f, 0.09815f, 0.08941f, 0.07654f, 0.06157f,
                        0.04654f, 0.03306f, 0.02207f, 0.01384f, 0.00816f};
}  // namespace

int orc_blur(uint16_t* R, uint16_t* G, uint16_t* B, int W, int H) {
  if (W < 4 || H < 8 || (W & 1) || (H & 1)) return -1;
  const int wF = W / 2, hF = H / 2;
  const size_t n = (size_t)wF * hF;
  std::vector<uint16_t> D(3 * n, 0), T(3 * n, 0), U(3 * n, 0);
  uint16_t* P[3] = {R, G, B};
  // blurH phase 1 (blur.c:23-36): brightness of a 2-px wide sample, as indexed
  for (int cy = 0; cy < hF - 2; cy++)
    for (int cx = 0; cx < wF; cx++) {
      const int y1 = imin(wF * (hF - 1) * 4, wF * cy * 4 + 1);
      const int x1 = imin(wF * 2 - 1, cx * 2 + 1);
      float c[3];
      for (int k = 0; k < 3; k++) {
        const uint16_t* S = P[k];
        c[k] = (float)(S[wF * cy * 4 + cx * 2] / 4 + S[y1 + cx * 2] / 4 + S[wF * cy * 4 + x1] / 4 +
                       S[y1 + x1] / 4);
      }
      const float lum = (0.3626f * c[0] + 0.5152f * c[1]) + 0.1222f * c[2];
      for (int k = 0; k < 3; k++) D[k * n + (size_t)wF * cy + cx] = sat_rtz(c[k] * lum / 256 / 256);
    }
  // blurH phase 2 (blur.c:40-64): 19 taps along x, /0.9
  for (int cy = 0; cy < hF - 1; cy++)
    for (int cx = 0; cx < wF; cx++)
      for (int k = 0; k < 3; k++) {
        const uint16_t* S = D.data() + k * n + (size_t)wF * cy;
        float a = 0.0f;
        for (int o = -9; o <= 9; o++) {
          const int ao = o < 0 ? -o : o;
          if (o < 0 && cx <= -o - 1) continue;
          if (o > 0 && cx >= wF - o) continue;
          a = a + kTap[ao] * (float)S[cx + o];
        }
        a /= 0.9f;
        T[k * n + (size_t)wF * cy + cx] = sat_rtz(a);
      }
  // blurV phase 1 (blur2.c:21-60): 19 taps along y, /0.9
  for (int cy = 0; cy < hF - 1; cy++)
    for (int cx = 0; cx < wF; cx++)
      for (int k = 0; k < 3; k++) {
        const uint16_t* S = T.data() + k * n;
        float a = 0.0f;
        for (int o = -9; o <= 9; o++) {
          const int ao = o < 0 ? -o : o;
          if (o < 0 && cy <= -o - 1) continue;
          if (o > 0 && (cy >= wF - o || cy + o >= hF)) continue;
          a = a + kTap[ao] * (float)S[(size_t)wF * (cy + o) + cx];
        }
        a /= 0.9f;
        U[k * n + (size_t)wF * cy + cx] = sat_rtz(a);
      }
  // blurV phase 2 (blur2.c:62-88): bilinear upsample added to the frame
  const int lenT = wF;
  for (int cy = 0; cy < H; cy++)
    for (int cx = 0; cx < W; cx++) {
      float texr1 = (float)cx / 2.f;
      const int tex1 = (int)texr1;
      texr1 -= (float)tex1;
      float texr2 = (float)cy / 2.f;
      const int tex2 = (int)texr2;
      texr2 -= (float)tex2;
      const int t00 = tex1 + lenT * tex2;
      const int t10 = imin(tex1 + 1, lenT - 1) + lenT * tex2;
      const int t01 = tex1 + lenT * imin(tex2 + 1, H / 2 - 2);
      const int t11 = imin(tex1 + 1, lenT - 1) + lenT * imin(tex2 + 1, H / 2 - 2);
      const float texi1 = 1 - texr1, texi2 = 1 - texr2;
      for (int k = 0; k < 3; k++) {
        const uint16_t* S = U.data() + k * n;
        const float up = ((texi1 * texi2 * (float)S[t00] + texr1 * texi2 * (float)S[t10]) +
                          texi1 * texr2 * (float)S[t01]) +
                         texr1 * texr2 * (float)S[t11];
        uint16_t& px = P[k][(size_t)W * cy + cx];
        px = sat_rtz((float)px + up);
      }
    }
  return 0;
}

int orc_math(int which, const float* x, float* y, int64_t n) {
  if (which != 0 && which != 1) return -1;
  for (int64_t i = 0; i < n; i++) y[i] = which == 0 ? log2_cw(x[i]) : exp_cw(x[i]);
  return 0;
}

int orc_create_mips(const uint16_t* plane, int N, uint16_t* out) {
  // Utils.py:79-93: 2x2 box sum >> 2 per level; levels concatenated
  // smallest first.
  if (N <= 0 || (N & (N - 1))) return -1;
  std::vector<std::vector<int64_t>> lv;
  std::vector<int64_t> a((size_t)N * N);
  for (size_t i = 0; i < a.size(); i++) a[i] = plane[i];
  lv.push_back(a);
  int n = N;
  while (n > 1) {
    int h = n / 2;
    std::vector<int64_t> c((size_t)h * h);
    for (int i = 0; i < h; i++)
      for (int j = 0; j < h; j++) {
        int64_t b0 = a[(size_t)(2 * i) * n + 2 * j] + a[(size_t)(2 * i + 1) * n + 2 * j];
        int64_t b1 = a[(size_t)(2 * i) * n + 2 * j + 1] + a[(size_t)(2 * i + 1) * n + 2 * j + 1];
        c[(size_t)i * h + j] = (b0 + b1) >> 2;
      }
    a.swap(c);
    n = h;
    lv.push_back(a);
  }
  size_t o = 0;
  for (auto it = lv.rbegin(); it != lv.rend(); ++it)
    for (int64_t v : *it) out[o++] = (uint16_t)v;
  return (int)o;
}

int orc_vertlight(const float* xyz, const float* vn, int nvert, const float* dirI,
                  const float* dirD, int ndir, const float* pointI, const float* pointP,
                  int npoint, const float* spotI, const float* spotD, const float* spotP,
                  int nspot, float amb, float* out) {
  // Pipe/vert.c:52-86
  for (int64_t v = 0; v < nvert; v++) {
    f3 x = ld3(xyz, v), n = ld3(vn, v);
    f3 light = {0.0f, 0.0f, 0.0f};
    for (int i = 1; i < ndir; i++) {
      float s = omax(0.0f, dot3(n, ld3(dirD, i)) + 0.5f) * 0.66f;
      light = add3(light, scl3(ld3(dirI, i), s));
    }
    for (int i = 0; i < npoint; i++) {
      f3 pl = sub3(x, ld3(pointP, i));
      if (dot3(n, pl) > 0.0f) {
        float len = std::sqrt(dot3(pl, pl));
        float s = dot3(n, normalize3(pl)) / (1.0f + len * len);
        light = add3(light, scl3(ld3(pointI, i), s));
      }
    }
    for (int i = 0; i < nspot; i++) {
      f3 pl = sub3(x, ld3(spotP, i));
      f3 sd = ld3(spotD, i);
      if (dot3(sd, pl) > 0.0f && dot3(n, pl) > 0.0f) {
        float len = std::sqrt(dot3(pl, pl));
        f3 npl = normalize3(pl);
        float s = dot3(n, npl) * dot3(npl, sd) / (1.5f + len * len);
        light = add3(light, scl3(ld3(spotI, i), s));
      }
    }
    out[3 * v] = omax(light.x, amb);
    out[3 * v + 1] = omax(light.y, amb);
    out[3 * v + 2] = omax(light.z, amb);
  }
  return 0;
}

int orc_render(const OrcGroup* groups, int ngroups, const OrcFrame* fr, OrcOut* out) {
  if (!groups || ngroups <= 0 || !fr || !out) return -1;
  const int W = fr->W, H = fr->H;
  if (W <= 0 || H <= 0) return -2;
  const int nthreads = resolve_threads(fr->nthreads);
  Ctx cx;
  cx.groups = groups;
  cx.ngroups = ngroups;
  cx.fr = fr;
  cx.base.assign(ngroups + 1, 0);
  for (int g = 0; g < ngroups; g++) cx.base[g + 1] = cx.base[g] + groups[g].ntri;
  for (int g = 0; g < ngroups; g++) {
    const OrcGroup& G = groups[g];
    if (G.ssr < 0 || G.skip) continue;
    if (G.ssr >= 4 || !fr->refl_r[G.ssr] || !fr->refl_g[G.ssr] || !fr->refl_b[G.ssr] ||
        fr->refl_size[G.ssr] < 2 || (G.flags & ORC_MIP) || fr->ssr_steps < 1 || fr->ssr_step < 1)
      return -3;
  }
  for (int i = 0; i < 8; i++) out->stats[i] = 0;  // 5, 6: shadow fragments per cascade

  // Shadow cascades (Ops.py:1089-1141 per map), before the main pass.
  cx.nsm = std::min(std::max(fr->nshadow, 0), 2);
  for (int c = 0; c < cx.nsm; c++) {
    out->stats[5 + c] = render_shadow(cx, c, nthreads);
    if (out->shadow[c])
      std::memcpy(out->shadow[c], cx.sm[c].F.data(), cx.sm[c].F.size() * sizeof(float));
  }

  // clearZBuffer: depth 65535 (clearzb.c); colour to the clear colour.  A
  // drawAll with no clear before it draws onto the frame it is given.
  const int64_t NP = (int64_t)W * H;
  if (fr->init_depth && fr->init_id && fr->init_r && fr->init_g && fr->init_b) {
    cx.F.assign(fr->init_depth, fr->init_depth + NP);
    cx.id.assign(fr->init_id, fr->init_id + NP);
    cx.R.assign(fr->init_r, fr->init_r + NP);
    cx.G.assign(fr->init_g, fr->init_g + NP);
    cx.B.assign(fr->init_b, fr->init_b + NP);
  } else {
    cx.F.assign(NP, 65535.0f);
    cx.id.assign(NP, 0xFFFFFFFFu);
    cx.R.assign(NP, fr->clear_rgb[0]);
    cx.G.assign(NP, fr->clear_rgb[1]);
    cx.B.assign(NP, fr->clear_rgb[2]);
  }

  int64_t st[5];
  draw_pass(cx, false, nthreads, st, [&](const OrcGroup& G, const MipInfo& mi, uint32_t gi,
                                          int ax, int cy, float t, float tz, const Attr& c1,
                                          const Attr& c2) {
    const int64_t p = (int64_t)W * cy + ax;
    if (cx.F[p] >= tz) {
      cx.F[p] = tz;
      cx.id[p] = gi;
      uint16_t rgb[3];
      shade(cx, G, mi, t, tz, c1, c2, rgb);
      cx.R[p] = rgb[0];
      cx.G[p] = rgb[1];
      cx.B[p] = rgb[2];
    }
  });
  for (int k = 0; k < 5; k++) out->stats[k] = st[k];

  // SSR water groups over the opaque frame (av_SSR.c): the shader reads a
  // copy of the opaque frame; the strict test keeps the nearest water
  // fragment, the first in draw order on a tie.
  bool any_water = false;
  for (int g = 0; g < ngroups; g++) any_water |= !groups[g].skip && groups[g].ssr >= 0;
  if (any_water) {
    Snap S;
    S.F = cx.F;
    S.R = cx.R;
    S.G = cx.G;
    S.B = cx.B;
    int64_t wst[5];
    draw_pass(cx, true, nthreads, wst, [&](const OrcGroup& G, const MipInfo&, uint32_t,
                                           int ax, int cy, float t, float tz, const Attr& c1,
                                           const Attr& c2) {
      const int64_t p = (int64_t)W * cy + ax;
      if (cx.F[p] > tz) {
        cx.F[p] = tz;
        uint16_t rgb[3];
        shade_ssr(cx, G, S, ax, cy, t, tz, c1, c2, rgb);
        cx.R[p] = rgb[0];
        cx.G[p] = rgb[1];
        cx.B[p] = rgb[2];
      }
    });
  }

  if (fr->exposure > 0)
    orc_gamma(cx.R.data(), cx.G.data(), cx.B.data(), NP, fr->exposure);

  if (out->depth) std::memcpy(out->depth, cx.F.data(), NP * sizeof(float));
  if (out->id) std::memcpy(out->id, cx.id.data(), NP * sizeof(uint32_t));
  if (out->r) std::memcpy(out->r, cx.R.data(), NP * sizeof(uint16_t));
  if (out->g) std::memcpy(out->g, cx.G.data(), NP * sizeof(uint16_t));
  if (out->b) std::memcpy(out->b, cx.B.data(), NP * sizeof(uint16_t));
  return 0;
}

}  // extern "C"
