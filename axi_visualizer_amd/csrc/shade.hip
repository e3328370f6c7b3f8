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
#include "axr_bin.h"

using namespace axr;

namespace {

// A block takes one heavy bin with all its threads, two medium bins with
// half of them each, or four light bins with a quarter each: heavy = weight
// >= kHeavyWeight, medium >= kMediumWeight (the order pass counts both into
// tiers; heavier bins lead the order).  Dense bins keep their parallelism
// while sparse ones keep most lanes busy and more bins in flight.  Without
// tiers (a frame of fewer bins than the device holds blocks) each bin gets
// a whole block.
template <int MODE>
__global__ __launch_bounds__(AXR_RASTER_THREADS) void k_bin_raster(
    const AxrTriRec* __restrict__ rec, AxrBins B, int empty, const float4* __restrict__ vpos,
    const float4* __restrict__ vattr, const AxrGroupDev* __restrict__ groups, AxrDrawConst dc,
    int NBX, float exposure, float* __restrict__ depth, uint16_t* __restrict__ ro,
    uint16_t* __restrict__ go, uint16_t* __restrict__ bo, uint32_t* __restrict__ rgba,
    uint32_t* __restrict__ ids, const uint32_t* __restrict__ tiers) {
  constexpr int NPX = AXR_BIN * AXR_BIN, HALF = AXR_RASTER_THREADS / 2,
                QUARTER = AXR_RASTER_THREADS / 4;
  static_assert(QUARTER % 64 == 0, "a quarter block is whole waves");
  __shared__ unsigned long long lk[4][NPX];
  __shared__ uint16_t list[4][NPX];
  __shared__ uint32_t ncov[4];
  const int tid = (int)threadIdx.x;
  // one inlined walk for every block kind (no order pass: a bin per block;
  // a heavy bin; a pair of medium bins; four light bins), so the kernel's
  // code stays small enough for the instruction cache
  int h = 0, t = tid, nthr = (int)blockDim.x, b = (int)blockIdx.x;
  bool active = true;
  if (!empty) {
    // no tiers: every bin is given a whole block, in the order pass's order
    const uint32_t nb = (uint32_t)B.nb, H = tiers ? tiers[0] : nb, M = tiers ? tiers[1] : nb;
    const uint32_t blk = blockIdx.x, NM = (M - H + 1) / 2;  // blocks of medium pairs
    if (blk < H) {
      b = (int)B.order[blk];
    } else if (blk < H + NM) {
      h = tid >= HALF ? 1 : 0;
      const uint32_t j = H + 2 * (blk - H) + (uint32_t)h;
      active = j < M;
      b = active ? (int)B.order[j] : 0;
      t = tid - h * HALF;
      nthr = HALF;
    } else {
      // one block per bin is launched; pairs and quads need fewer
      const uint32_t j0 = M + 4 * (blk - H - NM);
      if (j0 >= nb) return;
      h = tid / QUARTER;
      const uint32_t j = j0 + (uint32_t)h;
      active = j < nb;
      b = active ? (int)B.order[j] : 0;
      t = tid - h * QUARTER;
      nthr = QUARTER;
    }
  }
  raster_bin<AXR_BIN_SHIFT, MODE, true>(lk[h], list[h], &ncov[h], rec, B, (uint32_t)b,
                                  (uint32_t)(B.nb + b), empty, b % NBX, b / NBX, vpos, vattr,
                                  groups, dc, exposure, depth, ro, go, bo, rgba, ids, t, nthr,
                                  active);
}

// A frame of few bins (no more than two per CU) is rastered a 16-px
// quadrant per block: four blocks share each 32-px bin, each walking the
// bin's small triangles clipped to its quadrant and the large-triangle tiles
// of its quadrant, then shading its 256 pixels.  The pixels and their
// winners are those of the whole-bin raster; four times the blocks run.
constexpr int kQuadThreads = 256;
template <int MODE>
__global__ __launch_bounds__(kQuadThreads) void k_bin_raster_quads(
    const AxrTriRec* __restrict__ rec, AxrBins B, int empty, const float4* __restrict__ vpos,
    const float4* __restrict__ vattr, const AxrGroupDev* __restrict__ groups, AxrDrawConst dc,
    int NBX, float exposure, float* __restrict__ depth, uint16_t* __restrict__ ro,
    uint16_t* __restrict__ go, uint16_t* __restrict__ bo, uint32_t* __restrict__ rgba,
    uint32_t* __restrict__ ids) {
  constexpr int NPX = (AXR_BIN / 2) * (AXR_BIN / 2);
  __shared__ unsigned long long lk[NPX];
  __shared__ uint16_t list[NPX];
  __shared__ uint32_t ncov;
  const int q = (int)(blockIdx.x & 3u);
  const uint32_t j = blockIdx.x >> 2;
  const int b = empty ? (int)j : (int)B.order[j];  // heavy bins first
  const int bx = 2 * (b % NBX) + (q & 1), by = 2 * (b / NBX) + (q >> 1);
  raster_bin<AXR_BIN_SHIFT - 1, MODE, true>(lk, list, &ncov, rec, B, (uint32_t)b,
                                            (uint32_t)(B.nb + b), empty, bx, by, vpos, vattr,
                                            groups, dc, exposure, depth, ro, go, bo, rgba, ids,
                                            (int)threadIdx.x, kQuadThreads, true, q);
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

__global__ void k_gamma_pack(const uint16_t* __restrict__ r, const uint16_t* __restrict__ g,
                             const uint16_t* __restrict__ b, uint32_t* __restrict__ rgba,
                             size_t n, float ex) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x) {
    uint16_t v[3] = {r[i], g[i], b[i]};
    gamma_px(ex, v);
    rgba[i] = (uint32_t)v[0] | ((uint32_t)v[1] << 8) | ((uint32_t)v[2] << 16) | 0xFF000000u;
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

// Gamma'd u16 planes (values 0..255) packed into RGBA8 for readback.
__global__ void k_pack_rgba8(const uint16_t* __restrict__ r, const uint16_t* __restrict__ g,
                             const uint16_t* __restrict__ b, uint32_t* __restrict__ rgba,
                             size_t n) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x)
    rgba[i] = (uint32_t)(r[i] & 0xFF) | ((uint32_t)(g[i] & 0xFF) << 8) |
              ((uint32_t)(b[i] & 0xFF) << 16) | 0xFF000000u;
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

// Utils.createMips (Utils.py:79-93) one level at a time: level k (side
// n = 2^k, at offset (4^k - 1)/3 of the smallest-first chain) from level
// k + 1 as the 2x2 sums of each channel shifted right by 2.
__global__ void k_mip_level(ushort4* __restrict__ chain, int k) {
  const int n = 1 << k;
  const size_t dst0 = (((size_t)1 << (2 * k)) - 1) / 3;
  const size_t src0 = dst0 + (size_t)n * n;
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= n * n) return;
  const int x = idx % n, y = idx / n, m = 2 * n;
  const ushort4 a = chain[src0 + (size_t)(2 * y) * m + 2 * x];
  const ushort4 b = chain[src0 + (size_t)(2 * y) * m + 2 * x + 1];
  const ushort4 c = chain[src0 + (size_t)(2 * y + 1) * m + 2 * x];
  const ushort4 d = chain[src0 + (size_t)(2 * y + 1) * m + 2 * x + 1];
  chain[dst0 + (size_t)y * n + x] =
      make_ushort4((uint16_t)(((uint32_t)a.x + b.x + c.x + d.x) >> 2),
                   (uint16_t)(((uint32_t)a.y + b.y + c.y + d.y) >> 2),
                   (uint16_t)(((uint32_t)a.z + b.z + c.z + d.z) >> 2), 0);
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
                             uint32_t* ids, const uint32_t* tiers, hipStream_t s) {
  if (B.nb == 0) return hipSuccess;
  // no tiers: a frame of few bins, rastered a quadrant per block
#define AXR_LAUNCH_BIN(M)                                                                     \
  do {                                                                                        \
    if (tiers)                                                                                \
      hipLaunchKernelGGL(k_bin_raster<M>, dim3(B.nb), dim3(AXR_RASTER_THREADS), 0, s, rec, B,  \
                         empty, vpos, vattr, groups, dc, NBX, exposure, depth, r, g, b, rgba,   \
                         ids, tiers);                                                         \
    else                                                                                      \
      hipLaunchKernelGGL(k_bin_raster_quads<M>, dim3(4 * B.nb), dim3(kQuadThreads), 0, s, rec,  \
                         B, empty, vpos, vattr, groups, dc, NBX, exposure, depth, r, g, b,      \
                         rgba, ids);                                                          \
  } while (0)
  if (mode == SHADE_LDR8)
    AXR_LAUNCH_BIN(SHADE_LDR8);
  else if (mode == SHADE_HDR)
    AXR_LAUNCH_BIN(SHADE_HDR);
  else if (mode == SHADE_ACCUM)
    AXR_LAUNCH_BIN(SHADE_ACCUM);
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

hipError_t launch_gamma_pack(const uint16_t* r, const uint16_t* g, const uint16_t* b,
                             uint32_t* rgba, size_t n, float ex, hipStream_t s) {
  hipLaunchKernelGGL(k_gamma_pack, dim3(grid_for(n, 256, 4096)), dim3(256), 0, s, r, g, b, rgba, n,
                     ex);
  return hipGetLastError();
}

hipError_t launch_expand_rgba8(const uint32_t* rgba, uint16_t* r, uint16_t* g, uint16_t* b,
                               size_t n, hipStream_t s) {
  hipLaunchKernelGGL(k_expand_rgba8, dim3(grid_for(n, 256, 4096)), dim3(256), 0, s, rgba, r, g,
                     b, n);
  return hipGetLastError();
}

hipError_t launch_pack_rgba8(const uint16_t* r, const uint16_t* g, const uint16_t* b,
                             uint32_t* rgba, size_t n, hipStream_t s) {
  hipLaunchKernelGGL(k_pack_rgba8, dim3(grid_for(n, 256, 4096)), dim3(256), 0, s, r, g, b, rgba,
                     n);
  return hipGetLastError();
}

hipError_t launch_decode_shadow(const uint32_t* map, float* out, size_t n, hipStream_t s) {
  hipLaunchKernelGGL(k_decode_shadow, dim3(grid_for(n, 256, 4096)), dim3(256), 0, s, map, out, n);
  return hipGetLastError();
}

hipError_t launch_create_mips(ushort4* chain, int levels, hipStream_t s) {
  for (int k = levels - 2; k >= 0; k--) {
    const int n = 1 << k;
    hipLaunchKernelGGL(k_mip_level, dim3((n * n + 255) / 256), dim3(256), 0, s, chain, k);
  }
  return hipGetLastError();
}

hipError_t launch_vertlight(const float4* vpos, float4* vattr, uint32_t v0, uint32_t nv,
                            const AxrLights& L, hipStream_t s) {
  if (nv == 0) return hipSuccess;
  hipLaunchKernelGGL(k_vertlight, dim3((nv + 255) / 256), dim3(256), 0, s, vpos, vattr, v0, nv, L);
  return hipGetLastError();
}

}  // namespace axr
