// raster.hip -- depth pass of the gfx950 raster path.
//
//   k_clear_frame    depth keys to the clear key, per-frame counters to 0
//   k_setup_raster   triangle setup of every drawn triangle (a5); a small
//                    triangle is walked by the same lane at once (a7), a
//                    large one is appended to the large list
//   k_bin_count/scan/fill   coarse 16x16 tile binning of large triangles (a8)
//   k_raster_tiles   one block per tile, keys resident in LDS (a9)
//   k_clear_shadow / k_shadow   ortho setup + drawmin walk per caster (a11)
//
// Every fragment issues min(key) with key = (orderable tz << 32) | ~seq
// (axr_common.h), so the result does not depend on launch order or on which
// lane visits a fragment first: bins need no sorting, small triangles need
// no compaction, and the reference's per-group launch sequence collapses
// into single launches.
#include "axr_launch.h"
#include "axr_tri.h"

using namespace axr;

namespace {

constexpr int kBlock = 256;

// Posted (no-return) device-scope minimum: the lane never waits on it.
// Measured faster than reading the current value first to skip occluded
// fragments, whose dependent load stalls every fragment of the walk.
template <class T>
__device__ __forceinline__ void atomic_min_posted(T* p, T k) {
  __hip_atomic_fetch_min(p, k, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

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

// Appends `item` for every lane with `take`, one atomic per wave that has
// any taker.
template <class T>
__device__ __forceinline__ void wave_append(bool take, const T& item, T* __restrict__ out,
                                            uint32_t* __restrict__ counter, uint32_t cap) {
  const uint64_t m = __ballot(take);
  if (m == 0) return;
  const int leader = __ffsll((unsigned long long)m) - 1;
  uint32_t base = 0;
  if ((int)__lane_id() == leader) base = atomicAdd(counter, (uint32_t)__popcll(m));
  base = __shfl(base, leader);
  if (take) {
    const uint32_t slot = base + (uint32_t)__popcll(m & lanemask_lt());
    if (slot < cap) out[slot] = item;
  }
}

// Adds the wave's count of `flag` lanes to one of 64 spread stat slots.
__device__ __forceinline__ void wave_count(bool flag, uint32_t* __restrict__ slots) {
  const uint64_t m = __ballot(flag);
  if (m != 0 && __lane_id() == 0)
    atomicAdd(&slots[(blockIdx.x & 63) * AXR_CTR_STRIDE], (uint32_t)__popcll(m));
}

__global__ void k_fill_u64(uint64_t* __restrict__ p, size_t n, uint64_t v) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x)
    p[i] = v;
}

__global__ void k_fill_u32(uint32_t* __restrict__ p, size_t n, uint32_t v) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x)
    p[i] = v;
}

// clearZBuffer: keys to the clear key; draw counters and tile counts to 0.
__global__ void k_clear_frame(uint64_t* __restrict__ keys, size_t n, uint64_t v,
                              uint32_t* __restrict__ counters, uint32_t* __restrict__ tile_cnt,
                              int ntile) {
  const size_t t0 = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const size_t st = (size_t)gridDim.x * blockDim.x;
  for (size_t i = t0; i < n; i += st) keys[i] = v;
  for (size_t i = t0; i < (size_t)AXR_CTR_CAST; i += st) counters[i] = 0;
  for (size_t i = t0; i < (size_t)ntile; i += st) tile_cnt[i] = 0;
}

// clearShadowMap: map to the clear depth key; caster counts to 0.
__global__ void k_clear_shadow(uint32_t* __restrict__ map, size_t n, uint32_t v,
                               uint32_t* __restrict__ counters) {
  const size_t t0 = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const size_t st = (size_t)gridDim.x * blockDim.x;
  for (size_t i = t0; i < n; i += st) map[i] = v;
  for (size_t i = AXR_CTR_CAST + t0; i < (size_t)AXR_CTR_WORDS; i += st) counters[i] = 0;
}

// a5 + a7: one lane per triangle of every drawn group.
__global__ __launch_bounds__(kBlock) void k_setup_raster(const float4* __restrict__ vpos,
                                                         const float4* __restrict__ vattr,
                                                         const AxrGroupDev* __restrict__ groups,
                                                         AxrDrawConst dc,
                                                         AxrTriRec* __restrict__ large,
                                                         uint32_t* __restrict__ counters,
                                                         unsigned long long* __restrict__ keys) {
  const uint32_t gi = blockIdx.x * blockDim.x + threadIdx.x;
  int cls = 0;
  AxrTriRec r;
  r.gi = gi;
  if (gi < dc.T) {
    const uint32_t f = groups[find_group(groups, (int)dc.ngroups, gi)].flags;
    if (f & AXR_GRP_DRAW)
      cls = tri_setup(vpos, vattr, gi, dc.view, dc.W, dc.H, (f & AXR_SH_CULL) != 0, r.px, r.py,
                      r.z);
  }
  wave_count(cls == 1, counters + AXR_CTR_SMALL);
  wave_append(cls == 2, r, large, &counters[AXR_CTR_LARGE], dc.T);
  if (cls != 1) return;
  const Geom<int> g = small_geom(r.px, r.py, r.z);
  const int W = dc.W;
  small_walk(g, W, dc.H, [&](int ax, int cy, float tz, int half) {
    if (!(tz <= AXR_DEPTH_CLEAR)) return;
    const unsigned long long k = make_key(tz, frag_seq(gi, half));
    atomic_min_posted(&keys[(size_t)cy * W + ax], k);
  });
}

// a8: coarse binning, pass 1 -- entries per tile.
__global__ __launch_bounds__(kBlock) void k_bin_count(const AxrTriRec* __restrict__ rec,
                                                      const uint32_t* __restrict__ counters,
                                                      uint32_t* __restrict__ tile_cnt, int WC,
                                                      int HC) {
  const uint32_t n = counters[AXR_CTR_LARGE];
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const AxrTriRec r = rec[i];
    coarse_walk(r.px, r.py, WC, HC, [&](int t) { atomicAdd(&tile_cnt[t], 1u); });
  }
}

// Exclusive scan of the tile counts (single block); zeroes the counts so
// they serve as fill cursors; flags capacity overflow.
__global__ __launch_bounds__(1024) void k_bin_scan(uint32_t* __restrict__ tile_cnt,
                                                   uint32_t* __restrict__ tile_off, int NT,
                                                   uint32_t* __restrict__ counters,
                                                   uint32_t cap) {
  __shared__ uint32_t part[1024];
  const int tid = threadIdx.x;
  const int per = (NT + 1023) / 1024;
  const int b0 = tid * per, b1 = imin(NT, b0 + per);
  uint32_t s = 0;
  for (int i = b0; i < b1; i++) s += tile_cnt[i];
  part[tid] = s;
  __syncthreads();
  for (int off = 1; off < 1024; off <<= 1) {
    uint32_t v = tid >= off ? part[tid - off] : 0u;
    __syncthreads();
    part[tid] += v;
    __syncthreads();
  }
  uint32_t run = part[tid] - s;
  for (int i = b0; i < b1; i++) {
    const uint32_t c = tile_cnt[i];
    tile_off[i] = run;
    run += c;
    tile_cnt[i] = 0;
  }
  if (tid == 1023) {
    tile_off[NT] = part[1023];
    counters[AXR_CTR_BIN_TOTAL] = part[1023];
    counters[AXR_CTR_BIN_OVERFLOW] = part[1023] > cap ? part[1023] : 0u;
  }
}

// a8: coarse binning, pass 2 -- scatter record indices into the bins.
__global__ __launch_bounds__(kBlock) void k_bin_fill(const AxrTriRec* __restrict__ rec,
                                                     const uint32_t* __restrict__ counters,
                                                     uint32_t* __restrict__ cursor,
                                                     const uint32_t* __restrict__ tile_off,
                                                     uint32_t* __restrict__ bins, int WC, int HC,
                                                     uint32_t cap) {
  const uint32_t n = counters[AXR_CTR_LARGE];
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const AxrTriRec r = rec[i];
    coarse_walk(r.px, r.py, WC, HC, [&](int t) {
      const uint32_t pos = tile_off[t] + atomicAdd(&cursor[t], 1u);
      if (pos < cap) bins[pos] = i;
    });
  }
}

// a9: one block per 16x16 tile; the tile's keys live in LDS for the pass.
__global__ __launch_bounds__(kBlock) void k_raster_tiles(const AxrTriRec* __restrict__ rec,
                                                         const uint32_t* __restrict__ tile_off,
                                                         const uint32_t* __restrict__ bins,
                                                         unsigned long long* __restrict__ keys,
                                                         int W, uint32_t T, uint32_t cap) {
  __shared__ unsigned long long lk[256];
  const int bx = blockIdx.x;
  const uint32_t b0 = imin(tile_off[bx], cap), b1 = imin(tile_off[bx + 1], cap);
  if (b0 >= b1) return;
  const TileBox box = tile_box(bx, W);
  const int x0 = (int)box.xMin, y0 = (int)box.yMin;
  const int tid = threadIdx.x;
  unsigned long long* gk = &keys[(size_t)(y0 + (tid >> 4)) * W + x0 + (tid & 15)];
  lk[tid] = *gk;
  __syncthreads();
  for (uint32_t e = b0 + tid; e < b1; e += blockDim.x) {
    const AxrTriRec r = rec[bins[e]];
    const Geom<float> g = tile_geom(r.px, r.py, r.z);
    const uint32_t order = T + r.gi;
    tile_walk(g, box, [&](int ax, int cy, float tz, int half) {
      if (!(tz <= AXR_DEPTH_CLEAR)) return;
      const unsigned lc = (unsigned)((cy - y0) * 16 + (ax - x0));
      if (lc < 256u) atomicMin(&lk[lc], (unsigned long long)make_key(tz, frag_seq(order, half)));
    });
  }
  __syncthreads();
  *gk = lk[tid];
}

// a11: ortho setup + drawmin walk per caster; the map holds orderable keys
// of the minimum depth.
__global__ __launch_bounds__(kBlock) void k_shadow(const float4* __restrict__ vpos,
                                                   const AxrGroupDev* __restrict__ groups,
                                                   uint32_t ngroups, uint32_t T, AxrCastMask cast,
                                                   AxrCam cam, int dim, float bias,
                                                   uint32_t* __restrict__ map,
                                                   uint32_t* __restrict__ counters) {
  const uint32_t gi = blockIdx.x * blockDim.x + threadIdx.x;
  bool ok = false;
  AxrShRec r;
  if (gi < T) {
    const int g = find_group(groups, (int)ngroups, gi);
    if (cast_bit(cast, g)) ok = shadow_setup(vpos, gi, cam, dim, bias, r);
  }
  wave_count(ok, counters + AXR_CTR_CAST);
  if (!ok) return;
  shadow_walk(r, dim, dim, [&](int ax, int cy, float tz) {
    if (tz != tz) return;
    const uint32_t k = ordf(tz == 0.0f ? 0.0f : tz);
    atomic_min_posted(&map[(size_t)cy * dim + ax], k);
  });
}

inline unsigned grid_for(size_t n, unsigned block, unsigned cap) {
  size_t g = (n + block - 1) / block;
  if (g < 1) g = 1;
  if (g > cap) g = cap;
  return (unsigned)g;
}

}  // namespace

namespace axr {

hipError_t launch_fill_u64(uint64_t* p, size_t n, uint64_t v, hipStream_t s) {
  hipLaunchKernelGGL(k_fill_u64, dim3(grid_for(n, kBlock, 4096)), dim3(kBlock), 0, s, p, n, v);
  return hipGetLastError();
}

hipError_t launch_fill_u32(uint32_t* p, size_t n, uint32_t v, hipStream_t s) {
  hipLaunchKernelGGL(k_fill_u32, dim3(grid_for(n, kBlock, 4096)), dim3(kBlock), 0, s, p, n, v);
  return hipGetLastError();
}

hipError_t launch_clear_frame(uint64_t* keys, size_t n, uint32_t* counters, uint32_t* tile_cnt,
                              int ntile, hipStream_t s) {
  hipLaunchKernelGGL(k_clear_frame, dim3(grid_for(n, kBlock, 4096)), dim3(kBlock), 0, s, keys, n,
                     (uint64_t)key_clear(), counters, tile_cnt, ntile);
  return hipGetLastError();
}

hipError_t launch_clear_shadow(uint32_t* map, size_t n, uint32_t* counters, hipStream_t s) {
  hipLaunchKernelGGL(k_clear_shadow, dim3(grid_for(n, kBlock, 4096)), dim3(kBlock), 0, s, map, n,
                     ordf(AXR_DEPTH_CLEAR), counters);
  return hipGetLastError();
}

hipError_t launch_setup_raster(const float4* vpos, const float4* vattr, const AxrGroupDev* groups,
                               const AxrDrawConst& dc, AxrTriRec* large, uint32_t* counters,
                               uint64_t* keys, hipStream_t s) {
  if (dc.T == 0) return hipSuccess;
  hipLaunchKernelGGL(k_setup_raster, dim3((dc.T + kBlock - 1) / kBlock), dim3(kBlock), 0, s, vpos,
                     vattr, groups, dc, large, counters, (unsigned long long*)keys);
  return hipGetLastError();
}

hipError_t launch_bin_count(const AxrTriRec* large, const uint32_t* counters, uint32_t* tile_cnt,
                            int WC, int HC, uint32_t max_n, hipStream_t s) {
  hipLaunchKernelGGL(k_bin_count, dim3(grid_for(max_n, kBlock, 2048)), dim3(kBlock), 0, s, large,
                     counters, tile_cnt, WC, HC);
  return hipGetLastError();
}

hipError_t launch_bin_scan(uint32_t* tile_cnt, uint32_t* tile_off, int NT, uint32_t* counters,
                           uint32_t cap, hipStream_t s) {
  hipLaunchKernelGGL(k_bin_scan, dim3(1), dim3(1024), 0, s, tile_cnt, tile_off, NT, counters, cap);
  return hipGetLastError();
}

hipError_t launch_bin_fill(const AxrTriRec* large, const uint32_t* counters, uint32_t* tile_cnt,
                           const uint32_t* tile_off, uint32_t* bins, int WC, int HC, uint32_t cap,
                           uint32_t max_n, hipStream_t s) {
  hipLaunchKernelGGL(k_bin_fill, dim3(grid_for(max_n, kBlock, 2048)), dim3(kBlock), 0, s, large,
                     counters, tile_cnt, tile_off, bins, WC, HC, cap);
  return hipGetLastError();
}

hipError_t launch_raster_tiles(const AxrTriRec* large, const uint32_t* tile_off,
                               const uint32_t* bins, uint64_t* keys, int W, int WC, int HC,
                               uint32_t T, uint32_t cap, hipStream_t s) {
  const int NT = WC * HC;
  if (NT == 0) return hipSuccess;
  hipLaunchKernelGGL(k_raster_tiles, dim3(NT), dim3(kBlock), 0, s, large, tile_off, bins,
                     (unsigned long long*)keys, W, T, cap);
  return hipGetLastError();
}

hipError_t launch_shadow(const float4* vpos, const AxrGroupDev* groups, uint32_t ngroups,
                         uint32_t T, const AxrCastMask& cast, const AxrCam& cam, int dim,
                         float bias, uint32_t* map, uint32_t* counters, hipStream_t s) {
  if (T == 0) return hipSuccess;
  hipLaunchKernelGGL(k_shadow, dim3((T + kBlock - 1) / kBlock), dim3(kBlock), 0, s, vpos, groups,
                     ngroups, T, cast, cam, dim, bias, map, counters);
  return hipGetLastError();
}

}  // namespace axr
