// raster.hip -- setup and binning of the gfx950 raster path, and the
// shadow-map depth pass.
//
// Main pass (drawAll), two launches:
//   k_setup_bin  triangle setup (a5) of every drawn triangle; the setup
//                record is stored by global index; small triangles are
//                appended to the 32x32 bins their pixel rectangle touches,
//                large triangles to the bins of their coarse 16-px tiles
//                (a8, Pipe/coarse.c)
//   k_bin_order  entry counts of the pass, heavy-first bin order; resets
//                the append cursors for the next pass
// The per-bin depth walk and shading run later in one launch (shade.hip).
//
// Shadow pass (shadowMap), three launches: k_shadow_setup_bin, k_bin_order
// and k_shadow_bins, which walks each bin's casters (a11, Shaders/drawmin.c)
// with the bin's texels in LDS and writes them once.
//
// Appends aggregate within a wave: lanes of one wave come from one mesh
// region and mostly share a bin, so one atomic serves all of them.  Depth
// minima live in LDS; no global atomics touch pixels or texels.
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

// K1: setup + bin appends.  Slot b: small triangles (global index); slot
// NB + b: large triangles' coarse tiles, (global index << 2) | the 16-px
// tile's quadrant in bin b.
__global__ __launch_bounds__(kBlock) void k_setup_bin(const float4* __restrict__ vpos,
                                                      const float4* __restrict__ vattr,
                                                      const AxrGroupDev* __restrict__ groups,
                                                      AxrDrawConst dc, AxrTriRec* __restrict__ rec,
                                                      uint8_t* __restrict__ cls_out, AxrBins B,
                                                      int NBX) {
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
    const uint32_t slot = act ? rect_bin(br, k, NBX) : 0u;
    const uint32_t pos = wave_bin_add(B.cnt, act, slot);
    if (act) bin_store(B, slot, pos, gi);
  }
  if (cls == 2) {
    coarse_walk(r.px, r.py, dc.WC, dc.HC, [&](int t) {
      uint32_t bin, sub;
      tile_to_bin(t, dc.WC, NBX, bin, sub);
      const uint32_t slot = (uint32_t)B.nb + bin;
      bin_store(B, slot, atomicAdd(&B.cnt[slot], 1u), (gi << 2) | sub);
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

// K2 (one block): num = cnt, cnt = 0 for the next pass; the pass's entry
// total, spill and loss into the counter block; bins in heavy-first launch
// order (weight = small entries + 16 x large-tile entries), so the longest
// bins start first and short ones fill the tail.  Order within a weight
// class is arbitrary: each bin is computed by one block, so results do not
// depend on it.
__global__ __launch_bounds__(1024) void k_bin_order(AxrBins B, int nseg,
                                                    uint32_t* __restrict__ ctr_total,
                                                    uint32_t* __restrict__ ctr_spill,
                                                    uint32_t* __restrict__ ctr_lost) {
  __shared__ uint32_t hist[kOrderClasses];
  __shared__ uint32_t total;
  const int tid = threadIdx.x;
  for (int i = tid; i < kOrderClasses; i += 1024) hist[i] = 0;
  if (tid == 0) total = 0;
  __syncthreads();
  uint32_t tot = 0;
  for (int b = tid; b < B.nb; b += 1024) {
    uint32_t w = 0;
    for (int sg = 0; sg < nseg; sg++) {
      const uint32_t slot = (uint32_t)(sg * B.nb + b);
      const uint32_t c = B.cnt[slot];
      B.num[slot] = c;
      B.cnt[slot] = 0;
      tot += c;
      w += sg ? 16u * c : c;
    }
    atomicAdd(&hist[order_class(w)], 1u);
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) tot += __shfl_xor(tot, o);
  if ((tid & 63) == 0) atomicAdd(&total, tot);
  __syncthreads();
  if (tid == 0) {
    uint32_t run = 0;
    for (int k = kOrderClasses - 1; k >= 0; k--) {
      const uint32_t c = hist[k];
      hist[k] = run;
      run += c;
    }
    *ctr_total = total;
    const uint32_t used = B.spill_ctr[0];
    B.spill_ctr[1] = used < B.spill_cap ? used : B.spill_cap;
    B.spill_ctr[0] = 0;
    if (used) atomicMax(ctr_spill, used);
    if (used > B.spill_cap) atomicAdd(ctr_lost, used - B.spill_cap);
  }
  __syncthreads();
  for (int b = tid; b < B.nb; b += 1024) {
    uint32_t w = 0;
    for (int sg = 0; sg < nseg; sg++) {
      const uint32_t c = B.num[sg * B.nb + b];  // written by this thread above
      w += sg ? 16u * c : c;
    }
    B.order[atomicAdd(&hist[order_class(w)], 1u)] = (uint32_t)b;
  }
}

// S1: ortho setup of every casting group + bin appends (global index).
__global__ __launch_bounds__(kBlock) void k_shadow_setup_bin(
    const float4* __restrict__ vpos, const AxrGroupDev* __restrict__ groups, uint32_t ngroups,
    uint32_t T, AxrCastMask cast, AxrCam cam, int dim, float bias, AxrShRec* __restrict__ rec,
    uint8_t* __restrict__ ok_out, AxrBins B, int NBX) {
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
    const uint32_t slot = act ? rect_bin(br, k, NBX) : 0u;
    const uint32_t pos = wave_bin_add(B.cnt, act, slot);
    if (act) bin_store(B, slot, pos, gi);
  }
}

// S3: one block per 32x32 texel bin, heavy bins first; minimum depth keys
// in LDS, written once.  The bin starts from the map (from_map) or from the
// clear depth.
__global__ __launch_bounds__(AXR_SHADOW_THREADS) void k_shadow_bins(
    const AxrShRec* __restrict__ rec, AxrBins B, uint32_t* __restrict__ map, int dim, int NBX,
    int from_map) {
  __shared__ uint32_t lk[AXR_BIN * AXR_BIN];
  const int b = (int)B.order[blockIdx.x];
  const uint32_t n = B.num[b];
  if (from_map && n == 0) return;
  const int x0 = (b % NBX) << AXR_BIN_SHIFT, y0 = (b / NBX) << AXR_BIN_SHIFT;
  const int xe = imin(x0 + AXR_BIN, dim), ye = imin(y0 + AXR_BIN, dim);
  const uint32_t clear = ordf(AXR_DEPTH_CLEAR);
  for (int i = threadIdx.x; i < AXR_BIN * AXR_BIN; i += blockDim.x) {
    const int x = x0 + (i & (AXR_BIN - 1)), y = y0 + (i >> AXR_BIN_SHIFT);
    lk[i] = (from_map && x < xe && y < ye) ? map[(size_t)y * dim + x] : clear;
  }
  __syncthreads();
  for_bin_entries(B, (uint32_t)b, n, (int)threadIdx.x, (int)blockDim.x, [&](uint32_t gi) {
    const AxrShRec r = rec[gi];
    shadow_walk(r, dim, x0, y0, xe, ye, [&](int ax, int cy, float tz) {
      if (tz != tz) return;
      const unsigned lc = (unsigned)(((cy - y0) << AXR_BIN_SHIFT) + (ax - x0));
      if (lc < (unsigned)(AXR_BIN * AXR_BIN)) atomicMin(&lk[lc], ordf(tz == 0.0f ? 0.0f : tz));
    });
  });
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
                            const AxrBins& B, int NBX, hipStream_t s) {
  if (dc.T == 0) return hipSuccess;
  hipLaunchKernelGGL(k_setup_bin, dim3((dc.T + kBlock - 1) / kBlock), dim3(kBlock), 0, s, vpos,
                     vattr, groups, dc, rec, cls, B, NBX);
  return hipGetLastError();
}

hipError_t launch_bin_order(const AxrBins& B, int nseg, uint32_t* ctr_total,
                            uint32_t* ctr_spill, uint32_t* ctr_lost, hipStream_t s) {
  hipLaunchKernelGGL(k_bin_order, dim3(1), dim3(1024), 0, s, B, nseg, ctr_total, ctr_spill,
                     ctr_lost);
  return hipGetLastError();
}

hipError_t launch_shadow_setup_bin(const float4* vpos, const AxrGroupDev* groups,
                                   uint32_t ngroups, uint32_t T, const AxrCastMask& cast,
                                   const AxrCam& cam, int dim, float bias, AxrShRec* rec,
                                   uint8_t* ok, const AxrBins& B, int NBX, hipStream_t s) {
  if (T == 0) return hipSuccess;
  hipLaunchKernelGGL(k_shadow_setup_bin, dim3((T + kBlock - 1) / kBlock), dim3(kBlock), 0, s,
                     vpos, groups, ngroups, T, cast, cam, dim, bias, rec, ok, B, NBX);
  return hipGetLastError();
}

hipError_t launch_shadow_bins(const AxrShRec* rec, const AxrBins& B, uint32_t* map, int dim,
                              int NBX, int from_map, hipStream_t s) {
  hipLaunchKernelGGL(k_shadow_bins, dim3(B.nb), dim3(AXR_SHADOW_THREADS), 0, s, rec, B, map, dim,
                     NBX, from_map);
  return hipGetLastError();
}

hipError_t launch_stats(const uint8_t* cls, const uint8_t* ok, uint32_t T, uint32_t* out,
                        hipStream_t s) {
  hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(kBlock), 0, s, out, (size_t)3, 0u);
  if (T)
    hipLaunchKernelGGL(k_stats, dim3(grid_for(T, kBlock, 1024)), dim3(kBlock), 0, s, cls, ok, T,
                       out);
  return hipGetLastError();
}

}  // namespace axr
