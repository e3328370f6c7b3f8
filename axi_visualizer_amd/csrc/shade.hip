// shade.hip -- per-bin depth walk and shading of the gfx950 raster path,
// and the small per-pixel / per-vertex kernels around it (gamma, readback
// helpers, vertL).
//
// One block per 32x32 screen bin: the bin's 64-bit depth keys start at the
// clear key in LDS, every small triangle binned there walks its scanlines
// clipped to the bin (a7) and every large-triangle tile entry walks the
// reference's 16-px tile (a9); each fragment issues an LDS min of
// (orderable tz << 32) | ~seq.  The surviving key names the winning fragment
// (triangle, pass, scanline half).  Each pixel is then shaded exactly once:
// the winner's scanline walk is replayed up to the pixel's row -- the same
// accumulation sequence, so the interpolants are bit-identical to the
// reference's at that fragment -- and the colour is stored coalesced (RGBA8
// with gamma fused, or HDR u16 planes).  Keys never leave LDS and overdraw
// is never shaded.
#include "axr_launch.h"
#include "axr_tri.h"

using namespace axr;

namespace {

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

// Vertex k of three, by selects (a dynamic index would place the array in
// scratch memory).
__device__ __forceinline__ Attr pick3(const Attr& a0, const Attr& a1, const Attr& a2, int k) {
  Attr r;
#define M(f) r.f = k == 0 ? a0.f : (k == 1 ? a1.f : a2.f);
  AXR_ATTR_FIELDS(M)
#undef M
  return r;
}

// Rebuilds the winning fragment at (x, y) and shades it.
__device__ __forceinline__ void resolve_pixel(uint64_t key, int x, int y,
                                              const AxrTriRec* __restrict__ rec,
                                              const float4* __restrict__ vpos,
                                           const float4* __restrict__ vattr,
                                           const AxrGroupDev* __restrict__ groups,
                                           const AxrDrawConst& dc, uint16_t rgb[3]) {
  const uint32_t seq = ~(uint32_t)key;
  const uint32_t order = (seq - 1u) >> 1;
  const int half = (int)((seq - 1u) & 1u);
  const bool large = order >= dc.T;
  const uint32_t gi = large ? order - dc.T : order;
  const AxrGroupDev grp = groups[find_group(groups, (int)dc.ngroups, gi)];

  const AxrTriRec tr = rec[gi];
  const float* px = tr.px;
  const float* py = tr.py;
  const float* z = tr.z;
  const Attr A0 = load_attr(vpos, vattr, 3 * gi, z[0]);
  const Attr A1 = load_attr(vpos, vattr, 3 * gi + 1, z[1]);
  const Attr A2 = load_attr(vpos, vattr, 3 * gi + 2, z[2]);

  float cx1, cx2, cz1, cz2;
  Attr c1, c2;
  MipInfo mi = {0, 0.0f, 0, 0, 0, 0};
  if (!large) {
    const Geom<int> g = small_geom(px, py, z);
    Attr a1 = pick3(A0, A1, A2, g.k1), a2 = pick3(A0, A1, A2, g.k2),
         a3 = pick3(A0, A1, A2, g.k3);
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
    Attr a1 = pick3(A0, A1, A2, g.k1), a2 = pick3(A0, A1, A2, g.k2),
         a3 = pick3(A0, A1, A2, g.k3);
    const Attr a4 = a_lerp4(a1, a3, g.yd);
    if (grp.flags & AXR_SH_MIP) mi = mip_setup(a1, a2, a3, g, grp.tex_size, grp.mip_bias);
    const TileBox box = tile_box((y >> 4) * dc.WC + (x >> 4), dc.W);
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
  const float t = span_t(x, cx1, cx2);
  const float tz = span_tz(t, cz1, cz2);
  shade(dc, grp, mi, t, tz, c1, c2, rgb);
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

constexpr int kBinPx = AXR_BIN * AXR_BIN;

template <int MODE>
__global__ __launch_bounds__(AXR_RASTER_THREADS) void k_bin_raster(
    const AxrTriRec* __restrict__ rec, AxrBins B, int empty, const float4* __restrict__ vpos,
    const float4* __restrict__ vattr, const AxrGroupDev* __restrict__ groups, AxrDrawConst dc,
    int NBX, float exposure, float* __restrict__ depth, uint16_t* __restrict__ ro,
    uint16_t* __restrict__ go, uint16_t* __restrict__ bo, uint32_t* __restrict__ rgba,
    uint32_t* __restrict__ ids) {
  __shared__ unsigned long long lk[kBinPx];
  __shared__ uint16_t list[kBinPx];
  __shared__ uint32_t ncov;
  // Heavy bins first (order from the order pass); an empty draw has none.
  const int b = empty ? (int)blockIdx.x : (int)B.order[blockIdx.x];
  const int bx = b % NBX, by = b / NBX;
  const int x0 = bx << AXR_BIN_SHIFT, y0 = by << AXR_BIN_SHIFT;
  const int xe = imin(x0 + AXR_BIN, dc.W), ye = imin(y0 + AXR_BIN, dc.H);
  const int tid = threadIdx.x;
  for (int i = tid; i < kBinPx; i += blockDim.x) lk[i] = key_clear();
  if (tid == 0) ncov = 0;
  __syncthreads();
  if (!empty) {
    // a7: small triangles, walk clipped to the bin
    for_bin_entries(B, (uint32_t)b, B.num[b], tid, (int)blockDim.x, [&](uint32_t gi) {
      const AxrTriRec r = rec[gi];
      const Geom<int> g = small_geom(r.px, r.py, r.z);
      small_walk(g, x0, y0, xe, ye, [&](int ax, int cy, float tz, int half) {
        if (!(tz <= AXR_DEPTH_CLEAR)) return;
        const unsigned lc = (unsigned)(((cy - y0) << AXR_BIN_SHIFT) + (ax - x0));
        if (lc < (unsigned)kBinPx) atomicMin(&lk[lc], (unsigned long long)make_key(tz, frag_seq(gi, half)));
      });
    });
    // a9: large triangles, one reference 16-px tile per entry
    const uint32_t ls = (uint32_t)B.nb + (uint32_t)b;
    for_bin_entries(B, ls, B.num[ls], tid, (int)blockDim.x, [&](uint32_t v) {
      const uint32_t gi = v >> 2, sub = v & 3u;
      const int t16 = ((by << 1) + (int)(sub >> 1)) * dc.WC + (bx << 1) + (int)(sub & 1u);
      const TileBox box = tile_box(t16, dc.W);
      const AxrTriRec r = rec[gi];
      const Geom<float> g = tile_geom(r.px, r.py, r.z);
      const uint32_t order = dc.T + gi;
      tile_walk(g, box, [&](int ax, int cy, float tz, int half) {
        if (!(tz <= AXR_DEPTH_CLEAR)) return;
        const unsigned lc = (unsigned)(((cy - y0) << AXR_BIN_SHIFT) + (ax - x0));
        if (lc < (unsigned)kBinPx) atomicMin(&lk[lc], (unsigned long long)make_key(tz, frag_seq(order, half)));
      });
    });
  }
  __syncthreads();
  for (int i = tid; i < kBinPx; i += blockDim.x) {
    const int x = x0 + (i & (AXR_BIN - 1)), y = y0 + (i >> AXR_BIN_SHIFT);
    if (x >= xe || y >= ye) continue;
    const size_t p = (size_t)y * dc.W + x;
    const uint64_t key = lk[i];
    const uint32_t low = (uint32_t)key;
    if (MODE == SHADE_IDS) {
      uint32_t id = AXR_NONE;
      if (low != 0xFFFFFFFFu) {
        const uint32_t order = ((~low) - 1u) >> 1;
        id = order >= dc.T ? order - dc.T : order;
      }
      ids[p] = id;
      depth[p] = unordf((uint32_t)(key >> 32));
    } else if (low == 0xFFFFFFFFu) {
      uint16_t rgb[3] = {dc.clear_rgb[0], dc.clear_rgb[1], dc.clear_rgb[2]};
      put_px<MODE>(p, AXR_DEPTH_CLEAR, rgb, exposure, depth, ro, go, bo, rgba);
    } else {
      list[atomicAdd(&ncov, 1u)] = (uint16_t)i;
    }
  }
  if (MODE == SHADE_IDS) return;
  __syncthreads();
  const uint32_t n = ncov;
  for (uint32_t j = tid; j < n; j += blockDim.x) {
    const int i = list[j];
    const int x = x0 + (i & (AXR_BIN - 1)), y = y0 + (i >> AXR_BIN_SHIFT);
    const uint64_t key = lk[i];
    uint16_t rgb[3];
    resolve_pixel(key, x, y, rec, vpos, vattr, groups, dc, rgb);
    put_px<MODE>((size_t)y * dc.W + x, unordf((uint32_t)(key >> 32)), rgb, exposure, depth, ro,
                 go, bo, rgba);
  }
}

__global__ void k_gamma_planes(uint16_t* __restrict__ r, uint16_t* __restrict__ g,
                               uint16_t* __restrict__ b, size_t n, float ex) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x) {
    uint16_t v[3] = {r[i], g[i], b[i]};
    gamma_px(ex, v);
    r[i] = v[0];
    g[i] = v[1];
    b[i] = v[2];
  }
}

__global__ void k_expand_rgba8(const uint32_t* __restrict__ rgba, uint16_t* __restrict__ r,
                               uint16_t* __restrict__ g, uint16_t* __restrict__ b, size_t n) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x) {
    const uint32_t v = rgba[i];
    r[i] = (uint16_t)(v & 0xFF);
    g[i] = (uint16_t)((v >> 8) & 0xFF);
    b[i] = (uint16_t)((v >> 16) & 0xFF);
  }
}

__global__ void k_decode_shadow(const uint32_t* __restrict__ map, float* __restrict__ out,
                                size_t n) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x)
    out[i] = unordf(map[i]);
}

// Pipe/vert.c:52-86: per-vertex light from directional lights 1.., point
// and spot lights, floored at the ambient level.
__global__ void k_vertlight(const float4* __restrict__ vpos, float4* __restrict__ vattr,
                            uint32_t v0, uint32_t nv, AxrLights L) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nv) return;
  const uint32_t v = v0 + i;
  const float4 p4 = vpos[v];
  const float4 a0 = vattr[2 * (size_t)v];
  const V3 x = v3(p4.x, p4.y, p4.z), n = v3(a0.x, a0.y, a0.z);
  V3 light = v3(0.0f, 0.0f, 0.0f);
  for (int k = 1; k < L.ndir; k++) {
    const V3 d = v3(L.dirD[3 * k], L.dirD[3 * k + 1], L.dirD[3 * k + 2]);
    const V3 c = v3(L.dirI[3 * k], L.dirI[3 * k + 1], L.dirI[3 * k + 2]);
    const float s = omax(0.0f, dot(n, d) + 0.5f) * 0.66f;
    light = add(light, scl(c, s));
  }
  for (int k = 0; k < L.npoint; k++) {
    const V3 pl = sub(x, v3(L.pointP[3 * k], L.pointP[3 * k + 1], L.pointP[3 * k + 2]));
    if (dot(n, pl) > 0.0f) {
      const float len = sqrtf(dot(pl, pl));
      const float s = dot(n, normalize(pl)) / (1.0f + len * len);
      light = add(light, scl(v3(L.pointI[3 * k], L.pointI[3 * k + 1], L.pointI[3 * k + 2]), s));
    }
  }
  for (int k = 0; k < L.nspot; k++) {
    const V3 pl = sub(x, v3(L.spotP[3 * k], L.spotP[3 * k + 1], L.spotP[3 * k + 2]));
    const V3 sd = v3(L.spotD[3 * k], L.spotD[3 * k + 1], L.spotD[3 * k + 2]);
    if (dot(sd, pl) > 0.0f && dot(n, pl) > 0.0f) {
      const float len = sqrtf(dot(pl, pl));
      const V3 npl = normalize(pl);
      const float s = dot(n, npl) * dot(npl, sd) / (1.5f + len * len);
      light = add(light, scl(v3(L.spotI[3 * k], L.spotI[3 * k + 1], L.spotI[3 * k + 2]), s));
    }
  }
  float4 a1 = vattr[2 * (size_t)v + 1];
  a1.y = omax(light.x, L.amb);
  a1.z = omax(light.y, L.amb);
  a1.w = omax(light.z, L.amb);
  vattr[2 * (size_t)v + 1] = a1;
}

inline unsigned grid_for(size_t n, unsigned block, unsigned cap) {
  size_t g = (n + block - 1) / block;
  if (g < 1) g = 1;
  if (g > cap) g = cap;
  return (unsigned)g;
}

}  // namespace

namespace axr {

hipError_t launch_bin_raster(int mode, const AxrTriRec* rec, const AxrBins& B, int empty,
                             const float4* vpos, const float4* vattr, const AxrGroupDev* groups,
                             const AxrDrawConst& dc, int NBX, float exposure, float* depth,
                             uint16_t* r, uint16_t* g, uint16_t* b, uint32_t* rgba,
                             uint32_t* ids, hipStream_t s) {
  if (B.nb == 0) return hipSuccess;
#define AXR_LAUNCH_BIN(M)                                                                     \
  hipLaunchKernelGGL(k_bin_raster<M>, dim3(B.nb), dim3(AXR_RASTER_THREADS), 0, s, rec, B, empty, \
                     vpos, vattr, groups, dc, NBX, exposure, depth, r, g, b, rgba, ids)
  if (mode == SHADE_LDR8)
    AXR_LAUNCH_BIN(SHADE_LDR8);
  else if (mode == SHADE_HDR)
    AXR_LAUNCH_BIN(SHADE_HDR);
  else
    AXR_LAUNCH_BIN(SHADE_IDS);
#undef AXR_LAUNCH_BIN
  return hipGetLastError();
}

hipError_t launch_gamma_planes(uint16_t* r, uint16_t* g, uint16_t* b, size_t n, float ex,
                               hipStream_t s) {
  hipLaunchKernelGGL(k_gamma_planes, dim3(grid_for(n, 256, 4096)), dim3(256), 0, s, r, g, b, n,
                     ex);
  return hipGetLastError();
}

hipError_t launch_expand_rgba8(const uint32_t* rgba, uint16_t* r, uint16_t* g, uint16_t* b,
                               size_t n, hipStream_t s) {
  hipLaunchKernelGGL(k_expand_rgba8, dim3(grid_for(n, 256, 4096)), dim3(256), 0, s, rgba, r, g,
                     b, n);
  return hipGetLastError();
}

hipError_t launch_decode_shadow(const uint32_t* map, float* out, size_t n, hipStream_t s) {
  hipLaunchKernelGGL(k_decode_shadow, dim3(grid_for(n, 256, 4096)), dim3(256), 0, s, map, out, n);
  return hipGetLastError();
}

hipError_t launch_vertlight(const float4* vpos, float4* vattr, uint32_t v0, uint32_t nv,
                            const AxrLights& L, hipStream_t s) {
  if (nv == 0) return hipSuccess;
  hipLaunchKernelGGL(k_vertlight, dim3((nv + 255) / 256), dim3(256), 0, s, vpos, vattr, v0, nv, L);
  return hipGetLastError();
}

}  // namespace axr
