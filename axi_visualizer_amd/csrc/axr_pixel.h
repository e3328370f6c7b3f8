// axr_pixel.h -- per-pixel resolve of a winning depth key and the pixel
// store, shared by the single-frame and the batched raster kernels.
#pragma once
#include "axr_launch.h"
#include "axr_tri.h"

namespace axr {

__device__ __forceinline__ int find_group(const AxrGroupDev* __restrict__ groups, int ng,
                                          uint32_t gi) {
  int lo = 0, hi = ng - 1;
  while (lo < hi) {
    int mid = (lo + hi + 1) >> 1;
    if (groups[mid].base <= gi)
      lo = mid;
    else
      hi = mid - 1;
  }
  return lo;
}

// A rebuilt fragment: its group, mip level and the span interpolants.
struct FragState {
  AxrGroupDev grp;
  MipInfo mi;
  float t, tz;
  Attr c1, c2;
};

// Rebuilds fragment seq (frag_seq of its triangle's order and scanline
// half) at pixel (x, y): the winner's walk replayed up to the pixel's row.
__device__ __forceinline__ void replay_frag(uint32_t seq, int x, int y,
                                            const AxrTriRec* __restrict__ rec,
                                            const float4* __restrict__ vpos,
                                            const float4* __restrict__ vattr,
                                            const AxrGroupDev* __restrict__ groups,
                                            const AxrDrawConst& dc, FragState& f) {
  const uint32_t order = (seq - 1u) >> 1;
  const int half = (int)((seq - 1u) & 1u);
  const bool large = order >= dc.T;
  const uint32_t gi = large ? order - dc.T : order;
  AxrGroupDev grp = groups[find_group(groups, (int)dc.ngroups, gi)];
  grp.flags &= dc.flag_mask;

  const AxrTriRec tr = rec[gi];
  const float* px = tr.px;
  const float* py = tr.py;
  const float* z = tr.z;
  // the vertices are loaded in the walk's sorted order (k1, k2, k3 with
  // their sorted reciprocal depths): no per-attribute selects

  float cx1, cx2, cz1, cz2;
  Attr c1, c2;
  MipInfo mi = {0, 0.0f, 0, 0, 0, 0};
  if (!large) {
    const Geom<int> g = small_geom(px, py, z);
    Attr a1 = load_attr(vpos, vattr, 3 * gi + (uint32_t)g.k1, g.z1);
    Attr a2 = load_attr(vpos, vattr, 3 * gi + (uint32_t)g.k2, g.z2);
    Attr a3 = load_attr(vpos, vattr, 3 * gi + (uint32_t)g.k3, g.z3);
    const Attr a4 = a_lerp4(a1, a3, g.yd);
    if (grp.flags & AXR_SH_MIP) mi = mip_setup(a1, a2, a3, g, grp.tex_size, grp.mip_bias);
    const Edges e = small_edges(g, half);
    if (half == 0) {
      Attr s1 = a_slope(a2, a1, e.inv), s2 = a_slope(a4, a1, e.inv);
      if (e.swapped) swp(s1, s2);
      cx1 = cx2 = (float)g.x1;
      cz1 = cz2 = g.z1;
      c1 = c2 = a1;
      for (int r = g.y1; r < y; r++) {
        cx1 += e.s1; cx2 += e.s2; cz1 += e.sz1; cz2 += e.sz2;
        a_inc(c1, s1);
        a_inc(c2, s2);
      }
    } else {
      Attr s1 = a_nslope(a2, a3, e.inv), s2 = a_nslope(a4, a3, e.inv);
      if (e.swapped) swp(s1, s2);
      cx1 = cx2 = (float)g.x3;
      cz1 = cz2 = g.z3;
      c1 = c2 = a3;
      for (int r = g.y3; r > y; r--) {
        cx1 -= e.s1; cx2 -= e.s2; cz1 -= e.sz1; cz2 -= e.sz2;
        a_dec(c1, s1);
        a_dec(c2, s2);
      }
    }
  } else {
    const Geom<float> g = tile_geom(px, py, z);
    Attr a1 = load_attr(vpos, vattr, 3 * gi + (uint32_t)g.k1, g.z1);
    Attr a2 = load_attr(vpos, vattr, 3 * gi + (uint32_t)g.k2, g.z2);
    Attr a3 = load_attr(vpos, vattr, 3 * gi + (uint32_t)g.k3, g.z3);
    const Attr a4 = a_lerp4(a1, a3, g.yd);
    if (grp.flags & AXR_SH_MIP) mi = mip_setup(a1, a2, a3, g, grp.tex_size, grp.mip_bias);
    const TileBox box = tile_box_xy(x >> 4, y >> 4);  // the walked tile holding the pixel
    const Edges e = tile_edges(g, half);
    if (half == 0) {
      Attr s1 = a_slope(a2, a1, e.inv), s2 = a_slope(a4, a1, e.inv);
      if (e.swapped) swp(s1, s2);
      int cy = f2i(oclamp(g.y1, box.yMin, box.yMax));
      const float dy = (float)cy - g.y1;
      cx1 = g.x1 + dy * e.s1; cx2 = g.x1 + dy * e.s2;
      cz1 = g.z1 + dy * e.sz1; cz2 = g.z1 + dy * e.sz2;
      c1 = a_mad(a1, dy, s1);
      c2 = a_mad(a1, dy, s2);
      for (; cy < y; cy++) {
        cx1 += e.s1; cx2 += e.s2; cz1 += e.sz1; cz2 += e.sz2;
        a_inc(c1, s1);
        a_inc(c2, s2);
      }
    } else {
      Attr s1 = a_nslope(a2, a3, e.inv), s2 = a_nslope(a4, a3, e.inv);
      if (e.swapped) swp(s1, s2);
      int cy = f2i(oclamp(g.y3, box.yMin, box.yMax));
      const float dy = (float)cy - g.y3;
      cx1 = g.x3 + dy * e.s1; cx2 = g.x3 + dy * e.s2;
      cz1 = g.z3 + dy * e.sz1; cz2 = g.z3 + dy * e.sz2;
      c1 = a_mad(a3, dy, s1);
      c2 = a_mad(a3, dy, s2);
      for (; cy > y; cy--) {
        cx1 -= e.s1; cx2 -= e.s2; cz1 -= e.sz1; cz2 -= e.sz2;
        a_dec(c1, s1);
        a_dec(c2, s2);
      }
    }
  }
  f.t = span_t(x, cx1, cx2);
  f.tz = span_tz(f.t, cz1, cz2);
  f.grp = grp;
  f.mi = mi;
  f.c1 = c1;
  f.c2 = c2;
}

// Rebuilds the winning fragment of a depth key at (x, y) and shades it.
__device__ __forceinline__ void resolve_pixel(uint64_t key, int x, int y,
                                              const AxrTriRec* __restrict__ rec,
                                              const float4* __restrict__ vpos,
                                              const float4* __restrict__ vattr,
                                              const AxrGroupDev* __restrict__ groups,
                                              const AxrDrawConst& dc, uint16_t rgb[3]) {
  FragState f;
  replay_frag(~(uint32_t)key, x, y, rec, vpos, vattr, groups, dc, f);
  shade(dc, f.grp, f.mi, f.t, f.tz, f.c1, f.c2, rgb);
}

template <int MODE>
__device__ __forceinline__ void put_px(size_t p, float d, uint16_t rgb[3], float exposure,
                                       float* __restrict__ depth, uint16_t* __restrict__ ro,
                                       uint16_t* __restrict__ go, uint16_t* __restrict__ bo,
                                       uint32_t* __restrict__ rgba) {
  depth[p] = d;
  if (MODE == SHADE_LDR8) {
    gamma_px(exposure, rgb);
    rgba[p] = (uint32_t)rgb[0] | ((uint32_t)rgb[1] << 8) | ((uint32_t)rgb[2] << 16) | 0xFF000000u;
  } else {
    ro[p] = rgb[0];
    go[p] = rgb[1];
    bo[p] = rgb[2];
  }
}

}  // namespace axr
