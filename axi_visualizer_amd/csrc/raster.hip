// raster.hip -- setup and binning of the gfx950 raster path, and the
// shadow-map depth pass.
//
// Main pass (drawAll), three launches:
//   k_setup_bin  triangle setup (a5) of every drawn triangle; the setup
//                record is stored by global index; small triangles count the
//                32x32 bins their pixel rectangle touches, large triangles the
//                bins of their coarse 16-px tiles (a8, Pipe/coarse.c)
//   k_scan       bin offsets; zeroes the counts for the next frame
//   k_bin_fill   writes bin entries (same bins, same order of rounds)
// The per-bin depth walk and shading run later in one launch (shade.hip).
//
// Shadow pass (shadowMap), four launches: k_shadow_setup_bin, k_scan,
// k_shadow_fill and k_shadow_bins, which walks each bin's casters (a11,
// Shaders/drawmin.c) with the bin's texels in LDS and writes them once.
//
// Bin counting aggregates within a wave: lanes of one wave come from one
// mesh region and mostly share a bin, so one atomic serves all of them.
// Depth minima live in LDS; no global atomics touch pixels or texels.
#include "axr_launch.h"
#include "axr_tri.h"

using namespace axr;

namespace {

constexpr int kBlock = 256;

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

__device__ __forceinline__ uint64_t lanemask_lt() {
  const unsigned lane = __lane_id();
  return lane ? (~0ull >> (64 - lane)) : 0ull;
}

// Adds one per active lane to ctr[bin] with one atomic per distinct bin in
// the wave; returns the counter's previous value plus the lane's rank among
// the lanes that share its bin.  All lanes of the wave must call it.
__device__ __forceinline__ uint32_t wave_bin_add(uint32_t* __restrict__ ctr, bool active,
                                                 uint32_t bin) {
  uint64_t pending = __ballot(active);
  uint32_t result = 0;
  while (pending) {
    const int leader = __ffsll((unsigned long long)pending) - 1;
    const uint32_t b = __shfl(bin, leader);
    const uint64_t grp = __ballot(active && bin == b) & pending;
    uint32_t base = 0;
    if ((int)__lane_id() == leader) base = atomicAdd(&ctr[b], (uint32_t)__popcll(grp));
    base = __shfl(base, leader);
    if ((grp >> __lane_id()) & 1ull) result = base + (uint32_t)__popcll(grp & lanemask_lt());
    pending &= ~grp;
  }
  return result;
}

// Bin rectangle of a pixel rectangle clipped to [0,w) x [0,h).
struct BinRect {
  int bx0, by0, nbx, n;
};
__device__ __forceinline__ BinRect bin_rect(int x0, int y0, int x1, int y1, int w, int h) {
  BinRect r = {0, 0, 1, 0};
  x0 = imax(x0, 0);
  y0 = imax(y0, 0);
  x1 = imin(x1, w - 1);
  y1 = imin(y1, h - 1);
  if (x0 <= x1 && y0 <= y1) {
    r.bx0 = x0 >> AXR_BIN_SHIFT;
    r.by0 = y0 >> AXR_BIN_SHIFT;
    r.nbx = (x1 >> AXR_BIN_SHIFT) - r.bx0 + 1;
    r.n = r.nbx * ((y1 >> AXR_BIN_SHIFT) - r.by0 + 1);
  }
  return r;
}
__device__ __forceinline__ uint32_t rect_bin(const BinRect& r, int k, int nbx_total) {
  return (uint32_t)((r.by0 + k / r.nbx) * nbx_total + r.bx0 + k % r.nbx);
}

// 32-px bin and quadrant of a reference 16-px tile index.
__device__ __forceinline__ void tile_to_bin(int t, int WC, int NBX, uint32_t& bin, uint32_t& sub) {
  const int tx = t % WC, ty = t / WC;
  bin = (uint32_t)((ty >> 1) * NBX + (tx >> 1));
  sub = (uint32_t)(((ty & 1) << 1) | (tx & 1));
}

__global__ void k_fill_u32(uint32_t* __restrict__ p, size_t n, uint32_t v) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x)
    p[i] = v;
}

// K1: setup + bin counts.  cnt[0, NB): small entries per bin; cnt[NB, 2NB):
// large (coarse tile) entries per bin.
__global__ __launch_bounds__(kBlock) void k_setup_bin(const float4* __restrict__ vpos,
                                                      const float4* __restrict__ vattr,
                                                      const AxrGroupDev* __restrict__ groups,
                                                      AxrDrawConst dc, AxrTriRec* __restrict__ rec,
                                                      uint8_t* __restrict__ cls_out,
                                                      uint32_t* __restrict__ cnt, int NBX,
                                                      int NB) {
  const uint32_t gi = blockIdx.x * blockDim.x + threadIdx.x;
  int cls = 0;
  AxrTriRec r;
  if (gi < dc.T) {
    const uint32_t f = groups[find_group(groups, (int)dc.ngroups, gi)].flags;
    if (f & AXR_GRP_DRAW)
      cls = tri_setup(vpos, vattr, gi, dc.view, dc.W, dc.H, (f & AXR_SH_CULL) != 0, r.px, r.py,
                      r.z);
    cls_out[gi] = (uint8_t)cls;
    if (cls) rec[gi] = r;
  }
  BinRect br = {0, 0, 1, 0};
  if (cls == 1) {
    int x0, y0, x1, y1;
    small_bounds(r.px, r.py, x0, y0, x1, y1);
    br = bin_rect(x0, y0, x1, y1, dc.W, dc.H);
  }
  for (int k = 0;; k++) {
    const bool act = k < br.n;
    if (__ballot(act) == 0) break;
    wave_bin_add(cnt, act, act ? rect_bin(br, k, NBX) : 0u);
  }
  if (cls == 2) {
    coarse_walk(r.px, r.py, dc.WC, dc.HC, [&](int t) {
      uint32_t bin, sub;
      tile_to_bin(t, dc.WC, NBX, bin, sub);
      atomicAdd(&cnt[NB + bin], 1u);
    });
  }
}

// K3: bin entries.  Small: global index; large: (global index << 2) | the
// 16-px tile's quadrant in its bin.
__global__ __launch_bounds__(kBlock) void k_bin_fill(const AxrTriRec* __restrict__ rec,
                                                     const uint8_t* __restrict__ cls_in,
                                                     uint32_t T, int W, int H, int WC, int HC,
                                                     int NBX, int NB, uint32_t* __restrict__ cur,
                                                     uint32_t* __restrict__ ent, uint32_t cap) {
  const uint32_t gi = blockIdx.x * blockDim.x + threadIdx.x;
  const int cls = gi < T ? cls_in[gi] : 0;
  AxrTriRec r;
  if (cls) r = rec[gi];
  BinRect br = {0, 0, 1, 0};
  if (cls == 1) {
    int x0, y0, x1, y1;
    small_bounds(r.px, r.py, x0, y0, x1, y1);
    br = bin_rect(x0, y0, x1, y1, W, H);
  }
  for (int k = 0;; k++) {
    const bool act = k < br.n;
    if (__ballot(act) == 0) break;
    const uint32_t pos = wave_bin_add(cur, act, act ? rect_bin(br, k, NBX) : 0u);
    if (act && pos < cap) ent[pos] = gi;
  }
  if (cls == 2) {
    coarse_walk(r.px, r.py, WC, HC, [&](int t) {
      uint32_t bin, sub;
      tile_to_bin(t, WC, NBX, bin, sub);
      const uint32_t pos = atomicAdd(&cur[NB + bin], 1u);
      if (pos < cap) ent[pos] = (gi << 2) | sub;
    });
  }
}

constexpr int kOrderClasses = 66;
// 0 for an empty bin, else 2*floor(log2 w) + (next bit) + 2.
__device__ __forceinline__ int order_class(uint32_t w) {
  if (w == 0) return 0;
  const int l = 31 - __clz(w);
  const int half = l > 0 ? (int)((w >> (l - 1)) & 1u) : 0;
  return 2 * l + half + 2;
}

// K2: exclusive scan of n = nbins * nseg bin counts (one block).  off[0..n],
// cur = off, counts zeroed for the next frame; total and overflow into ctr;
// order[0..nbins) = bins by descending weight.
__global__ __launch_bounds__(1024) void k_scan(uint32_t* __restrict__ cnt,
                                               uint32_t* __restrict__ off,
                                               uint32_t* __restrict__ cur, int nbins, int nseg,
                                               uint32_t cap, uint32_t* __restrict__ ctr_total,
                                               uint32_t* __restrict__ ctr_ovf,
                                               uint32_t* __restrict__ order) {
  const int n = nbins * nseg;
  __shared__ uint32_t wsum[16];
  const int tid = threadIdx.x;
  const int lane = tid & 63, w = tid >> 6;
  const int per = (n + 1023) / 1024;
  const int b0 = imin(n, tid * per), b1 = imin(n, b0 + per);
  uint32_t s = 0;
  for (int i = b0; i < b1; i++) s += cnt[i];
  uint32_t v = s;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const uint32_t t = __shfl_up(v, o);
    if (lane >= o) v += t;
  }
  if (lane == 63) wsum[w] = v;
  __syncthreads();
  if (w == 0) {
    uint32_t x = lane < 16 ? wsum[lane] : 0u;
#pragma unroll
    for (int o = 1; o < 16; o <<= 1) {
      const uint32_t t = __shfl_up(x, o);
      if (lane >= o) x += t;
    }
    if (lane < 16) wsum[lane] = x;
  }
  __syncthreads();
  uint32_t run = v - s + (w > 0 ? wsum[w - 1] : 0u);
  for (int i = b0; i < b1; i++) {
    const uint32_t c = cnt[i];
    off[i] = run;
    cur[i] = run;
    run += c;
    cnt[i] = 0;
  }
  if (tid == 1023) {
    const uint32_t total = wsum[15];
    off[n] = total;
    *ctr_total = total;
    *ctr_ovf = total > cap ? total : 0u;
  }
  // Heavy-first launch order of the nbins bins (weight = entries over the
  // nseg segments), so the longest bins start first and the short ones
  // fill the tail.  Classes are half-octaves of the weight; order within a
  // class is arbitrary (every bin is computed by one block, so the result
  // does not depend on it).
  __shared__ uint32_t hist[kOrderClasses];
  for (int i = tid; i < kOrderClasses; i += 1024) hist[i] = 0;
  __syncthreads();  // also publishes off[] to the whole block
  for (int i = tid; i < nbins; i += 1024) {
    uint32_t wgt = 0;
    for (int sg = 0; sg < nseg; sg++) wgt += off[sg * nbins + i + 1] - off[sg * nbins + i];
    atomicAdd(&hist[order_class(wgt)], 1u);
  }
  __syncthreads();
  if (tid == 0) {
    uint32_t run = 0;
    for (int k = kOrderClasses - 1; k >= 0; k--) {
      const uint32_t c = hist[k];
      hist[k] = run;
      run += c;
    }
  }
  __syncthreads();
  for (int i = tid; i < nbins; i += 1024) {
    uint32_t wgt = 0;
    for (int sg = 0; sg < nseg; sg++) wgt += off[sg * nbins + i + 1] - off[sg * nbins + i];
    order[atomicAdd(&hist[order_class(wgt)], 1u)] = (uint32_t)i;
  }
}

// S1: ortho setup of every casting group + bin counts.
__global__ __launch_bounds__(kBlock) void k_shadow_setup_bin(
    const float4* __restrict__ vpos, const AxrGroupDev* __restrict__ groups, uint32_t ngroups,
    uint32_t T, AxrCastMask cast, AxrCam cam, int dim, float bias, AxrShRec* __restrict__ rec,
    uint8_t* __restrict__ ok_out, uint32_t* __restrict__ cnt, int NBX) {
  const uint32_t gi = blockIdx.x * blockDim.x + threadIdx.x;
  bool ok = false;
  AxrShRec r;
  if (gi < T) {
    if (cast_bit(cast, find_group(groups, (int)ngroups, gi)))
      ok = shadow_setup(vpos, gi, cam, dim, bias, r);
    ok_out[gi] = ok ? 1 : 0;
    if (ok) rec[gi] = r;
  }
  BinRect br = {0, 0, 1, 0};
  if (ok) {
    int x0, y0, x1, y1;
    shadow_bounds(r, x0, y0, x1, y1);
    br = bin_rect(x0, y0, x1, y1, dim, dim);
  }
  for (int k = 0;; k++) {
    const bool act = k < br.n;
    if (__ballot(act) == 0) break;
    wave_bin_add(cnt, act, act ? rect_bin(br, k, NBX) : 0u);
  }
}

// S3: shadow bin entries (global index of the caster).
__global__ __launch_bounds__(kBlock) void k_shadow_fill(const AxrShRec* __restrict__ rec,
                                                        const uint8_t* __restrict__ ok_in,
                                                        uint32_t T, int dim, int NBX,
                                                        uint32_t* __restrict__ cur,
                                                        uint32_t* __restrict__ ent, uint32_t cap) {
  const uint32_t gi = blockIdx.x * blockDim.x + threadIdx.x;
  const bool ok = gi < T && ok_in[gi];
  BinRect br = {0, 0, 1, 0};
  if (ok) {
    const AxrShRec r = rec[gi];
    int x0, y0, x1, y1;
    shadow_bounds(r, x0, y0, x1, y1);
    br = bin_rect(x0, y0, x1, y1, dim, dim);
  }
  for (int k = 0;; k++) {
    const bool act = k < br.n;
    if (__ballot(act) == 0) break;
    const uint32_t pos = wave_bin_add(cur, act, act ? rect_bin(br, k, NBX) : 0u);
    if (act && pos < cap) ent[pos] = gi;
  }
}

// S4: one block per 32x32 texel bin; minimum depth keys in LDS, written
// once.  The bin starts from the map (from_map) or from the clear depth.
__global__ __launch_bounds__(AXR_SHADOW_THREADS) void k_shadow_bins(const AxrShRec* __restrict__ rec,
                                                        const uint32_t* __restrict__ off,
                                                        const uint32_t* __restrict__ ent,
                                                        uint32_t cap, uint32_t* __restrict__ map,
                                                        int dim, int NBX, int from_map,
                                                        const uint32_t* __restrict__ order) {
  __shared__ uint32_t lk[AXR_BIN * AXR_BIN];
  const int b = (int)order[blockIdx.x];
  const uint32_t s0 = imin(off[b], cap), s1 = imin(off[b + 1], cap);
  if (from_map && s0 >= s1) return;
  const int x0 = (b % NBX) << AXR_BIN_SHIFT, y0 = (b / NBX) << AXR_BIN_SHIFT;
  const int xe = imin(x0 + AXR_BIN, dim), ye = imin(y0 + AXR_BIN, dim);
  const uint32_t clear = ordf(AXR_DEPTH_CLEAR);
  for (int i = threadIdx.x; i < AXR_BIN * AXR_BIN; i += blockDim.x) {
    const int x = x0 + (i & (AXR_BIN - 1)), y = y0 + (i >> AXR_BIN_SHIFT);
    lk[i] = (from_map && x < xe && y < ye) ? map[(size_t)y * dim + x] : clear;
  }
  __syncthreads();
  for (uint32_t e = s0 + threadIdx.x; e < s1; e += blockDim.x) {
    const AxrShRec r = rec[ent[e]];
    shadow_walk(r, dim, x0, y0, xe, ye, [&](int ax, int cy, float tz) {
      if (tz != tz) return;
      const unsigned lc = (unsigned)(((cy - y0) << AXR_BIN_SHIFT) + (ax - x0));
      if (lc < (unsigned)(AXR_BIN * AXR_BIN)) atomicMin(&lk[lc], ordf(tz == 0.0f ? 0.0f : tz));
    });
  }
  __syncthreads();
  for (int i = threadIdx.x; i < AXR_BIN * AXR_BIN; i += blockDim.x) {
    const int x = x0 + (i & (AXR_BIN - 1)), y = y0 + (i >> AXR_BIN_SHIFT);
    if (x < xe && y < ye) map[(size_t)y * dim + x] = lk[i];
  }
}

// Counts small / large / casting triangles on request (statistics only).
__global__ void k_stats(const uint8_t* __restrict__ cls, const uint8_t* __restrict__ ok,
                        uint32_t T, uint32_t* __restrict__ out) {
  uint32_t s = 0, l = 0, c = 0;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < T; i += gridDim.x * blockDim.x) {
    s += cls[i] == 1;
    l += cls[i] == 2;
    c += ok[i] != 0;
  }
  atomicAdd(&out[0], s);
  atomicAdd(&out[1], l);
  atomicAdd(&out[2], c);
}

inline unsigned grid_for(size_t n, unsigned block, unsigned cap) {
  size_t g = (n + block - 1) / block;
  if (g < 1) g = 1;
  if (g > cap) g = cap;
  return (unsigned)g;
}

}  // namespace

namespace axr {

hipError_t launch_fill_u32(uint32_t* p, size_t n, uint32_t v, hipStream_t s) {
  hipLaunchKernelGGL(k_fill_u32, dim3(grid_for(n, kBlock, 4096)), dim3(kBlock), 0, s, p, n, v);
  return hipGetLastError();
}

hipError_t launch_setup_bin(const float4* vpos, const float4* vattr, const AxrGroupDev* groups,
                            const AxrDrawConst& dc, AxrTriRec* rec, uint8_t* cls,
                            uint32_t* cnt, int NBX, int NB, hipStream_t s) {
  if (dc.T == 0) return hipSuccess;
  hipLaunchKernelGGL(k_setup_bin, dim3((dc.T + kBlock - 1) / kBlock), dim3(kBlock), 0, s, vpos,
                     vattr, groups, dc, rec, cls, cnt, NBX, NB);
  return hipGetLastError();
}

hipError_t launch_scan(uint32_t* cnt, uint32_t* off, uint32_t* cur, int nbins, int nseg,
                       uint32_t cap, uint32_t* ctr_total, uint32_t* ctr_ovf, uint32_t* order,
                       hipStream_t s) {
  hipLaunchKernelGGL(k_scan, dim3(1), dim3(1024), 0, s, cnt, off, cur, nbins, nseg, cap,
                     ctr_total, ctr_ovf, order);
  return hipGetLastError();
}

hipError_t launch_bin_fill(const AxrTriRec* rec, const uint8_t* cls, uint32_t T, int W, int H,
                           int WC, int HC, int NBX, int NB, uint32_t* cur, uint32_t* ent,
                           uint32_t cap, hipStream_t s) {
  if (T == 0) return hipSuccess;
  hipLaunchKernelGGL(k_bin_fill, dim3((T + kBlock - 1) / kBlock), dim3(kBlock), 0, s, rec, cls, T,
                     W, H, WC, HC, NBX, NB, cur, ent, cap);
  return hipGetLastError();
}

hipError_t launch_shadow_setup_bin(const float4* vpos, const AxrGroupDev* groups,
                                   uint32_t ngroups, uint32_t T, const AxrCastMask& cast,
                                   const AxrCam& cam, int dim, float bias, AxrShRec* rec,
                                   uint8_t* ok, uint32_t* cnt, int NBX, hipStream_t s) {
  if (T == 0) return hipSuccess;
  hipLaunchKernelGGL(k_shadow_setup_bin, dim3((T + kBlock - 1) / kBlock), dim3(kBlock), 0, s,
                     vpos, groups, ngroups, T, cast, cam, dim, bias, rec, ok, cnt, NBX);
  return hipGetLastError();
}

hipError_t launch_shadow_fill(const AxrShRec* rec, const uint8_t* ok, uint32_t T, int dim,
                              int NBX, uint32_t* cur, uint32_t* ent, uint32_t cap,
                              hipStream_t s) {
  if (T == 0) return hipSuccess;
  hipLaunchKernelGGL(k_shadow_fill, dim3((T + kBlock - 1) / kBlock), dim3(kBlock), 0, s, rec, ok,
                     T, dim, NBX, cur, ent, cap);
  return hipGetLastError();
}

hipError_t launch_shadow_bins(const AxrShRec* rec, const uint32_t* off, const uint32_t* ent,
                              uint32_t cap, uint32_t* map, int dim, int NBX, int NB, int from_map,
                              const uint32_t* order, hipStream_t s) {
  hipLaunchKernelGGL(k_shadow_bins, dim3(NB), dim3(AXR_SHADOW_THREADS), 0, s, rec, off, ent, cap, map, dim,
                     NBX, from_map, order);
  return hipGetLastError();
}

hipError_t launch_stats(const uint8_t* cls, const uint8_t* ok, uint32_t T, uint32_t* out,
                        hipStream_t s) {
  hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(kBlock), 0, s, out, (size_t)4, 0u);
  if (T)
    hipLaunchKernelGGL(k_stats, dim3(grid_for(T, kBlock, 1024)), dim3(kBlock), 0, s, cls, ok, T,
                       out);
  return hipGetLastError();
}

}  // namespace axr
