// batch.hip -- many frames (jobs) in one launch sequence.
//
// A job is a resident scene seen by one camera at one resolution.  A batch
// of jobs renders with three launches for all of them:
//   k_batch_setup  one thread per triangle of each scene in the batch: the
//                  triangle is loaded once and set up (a5) with the camera of
//                  every job on that scene in turn -- record and class in
//                  the job's arrays, appends to the job's screen bins (slots
//                  slot0.. of one shared store; large triangles' coarse
//                  tiles in the second segment)
//   k_bin_order    the shared store's counts and heavy-first order across
//                  every job's bins (raster.hip); it also resets the queue
//   k_batch_raster a persistent kernel: each block takes the next bin from
//                  that order (an atomic queue head) until none is left, and
//                  walks and shades it exactly as the single-frame raster does
// Bins are 8, 16 or 32 pixels on a side -- a scheduling unit only: small
// triangles walk clipped to any bin, large ones keep the reference's 16-px
// tiles.  The result of a job is bit-identical whatever the bin size and
// whatever else is in the batch.
#include "axr_bin.h"

using namespace axr;

namespace {

constexpr int kSetupBlock = 256;
#ifndef AXR_BATCH_THREADS
#define AXR_BATCH_THREADS 256  // threads of a raster block (a bin of up to 32x32)
#endif

// Scene group owning setup block blk: groups (blk0, first job index, jobs,
// triangles) ascend in blk0.
__device__ __forceinline__ int group_of_block(const uint4* __restrict__ sg, int n, uint32_t blk) {
  int lo = 0, hi = n - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (sg[mid].x <= blk)
      lo = mid;
    else
      hi = mid - 1;
  }
  return lo;
}

// Job owning bin slot s (jobs ascend in slot0).
__device__ __forceinline__ int job_of_slot(const AxrJobDev* __restrict__ jobs, int n, uint32_t s) {
  int lo = 0, hi = n - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (jobs[mid].slot0 <= s)
      lo = mid;
    else
      hi = mid - 1;
  }
  return lo;
}

// The jobs of one scene group share geometry, groups and chunk spheres, so
// a thread loads its triangle once and runs every job's setup on it; the
// job loop is uniform over the block, so the wave-wide appends work as in
// a single frame.
template <int SHIFT>
__global__ __launch_bounds__(kSetupBlock) void k_batch_setup(const AxrJobDev* __restrict__ jobs,
                                                             const uint4* __restrict__ sg, int nsg,
                                                             const uint32_t* __restrict__ jidx,
                                                             AxrBins B, uint2* __restrict__ large,
                                                             uint32_t* __restrict__ large_ctl,
                                                             uint32_t large_cap,
                                                             uint32_t* __restrict__ ctr,
                                                             uint32_t* __restrict__ visible) {
  const uint4 G = sg[group_of_block(sg, nsg, blockIdx.x)];
  const AxrJobDev& J0 = jobs[jidx[G.y]];
  const uint32_t T = G.w;
  const uint32_t gi = (blockIdx.x - G.x) * kSetupBlock + threadIdx.x;
  const bool live = gi < T;
  // the wave's 64 triangles are one chunk: a job skips it when out of view.
  // The first wave tests the block's chunks against every job of the group
  // (a lane per (job, chunk) pair) so the other waves issue none of the
  // tests; groups of more jobs than the table holds test per wave.
  constexpr int kWaves = kSetupBlock / 64;
  constexpr uint32_t kVisJobs = 512;
  __shared__ uint8_t jvis[kVisJobs * kWaves];
  const int last_chunk = (int)((T - 1) >> 6);
  const float4 sph = J0.chunks[imin((int)(gi >> 6), last_chunk)];
  const bool pre = G.z <= kVisJobs;
  if (pre) {
    if (threadIdx.x < 64) {
      const int c0 = (int)((blockIdx.x - G.x) * kWaves);
      for (uint32_t k = threadIdx.x; k < kWaves * G.z; k += 64) {
        const uint32_t q = k / kWaves;
        const float4 cs = J0.chunks[imin(c0 + (int)(k % kWaves), last_chunk)];
        jvis[k] = chunk_outside_view(cs, jobs[jidx[G.y + q]].dc.view, jobs[jidx[G.y + q]].dc.cull) ? 0 : 1;
      }
    }
    __syncthreads();
  }
  float4 vp[3];
  uint32_t f = 0;
  if (live) {
#pragma unroll
    for (int k = 0; k < 3; k++) vp[k] = J0.vpos[3 * (size_t)gi + k];
    f = J0.groups[find_group(J0.groups, (int)J0.dc.ngroups, gi)].flags;
  } else {
#pragma unroll
    for (int k = 0; k < 3; k++) vp[k] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  }
  for (uint32_t q = 0; q < G.z; q++) {
    const AxrJobDev& J = jobs[jidx[G.y + q]];
    const AxrDrawConst& dc = J.dc;
    int cls = 0;
    AxrTriRec r;
    const bool in_view = pre ? jvis[q * kWaves + (threadIdx.x >> 6)] != 0
                             : !chunk_outside_view(sph, dc.view, dc.cull);
    if (live && (f & AXR_GRP_DRAW) && in_view) {
      cls = tri_setup(vp, J.vattr, gi, dc.view, dc.W, dc.H, (f & AXR_SH_CULL) != 0, r.px, r.py,
                      r.z);
      if (cls) J.rec[gi] = r;  // read by the raster through the bin entries only
    }
    BinRect br = {0, 0, 1, 0};
    if (cls == 1) {
      int x0, y0, x1, y1;
      small_bounds(r.px, r.py, x0, y0, x1, y1);
      br = bin_rect(x0, y0, x1, y1, dc.W, dc.H, SHIFT);
    }
    BinCursor bc;
    for (int k = 0;; k++) {
      const bool act = k < br.n;
      if (__ballot(act) == 0) break;
      const uint32_t slot = act ? J.slot0 + bc.bin(br, J.nbx) : 0u;
      const uint32_t pos = wave_bin_add(B.cnt, act, slot);
      if (act) bin_store(B, slot, pos, gi);
      bc.next(br);
    }
    // large triangles are binned by k_batch_coarse, a block per (job,
    // triangle); past the list's capacity they count as lost entries (the
    // wait reports them and the list grows)
    if (visible) {  // statistics renders only: visible triangles per job
      const uint64_t vm = __ballot(cls != 0);
      if (vm && __lane_id() == 0) atomicAdd(&visible[jidx[G.y + q]], (uint32_t)__popcll(vm));
    }
    const uint32_t at = wave_bin_add(large_ctl, cls == 2, 0u);
    if (cls == 2) {
      if (at < large_cap)
        large[at] = make_uint2(jidx[G.y + q], gi);
      else
        atomicAdd(&ctr[AXR_CTR_LOST], 1u);
    }
  }
}

// Coarse binning of the batch's large triangles (Pipe/coarse.c), a block
// per (job, triangle) of the list, each tile row to a group of threads
// (coarse_block_tiles), as the single-frame k_coarse does; the tiles go to
// the job's large-tile bins at this batch's bin size (append_tile).  The
// order pass that follows resets the list.
constexpr int kCoarseThreads = 256;
template <int SHIFT>
__global__ __launch_bounds__(kCoarseThreads) void k_batch_coarse(
    const AxrJobDev* __restrict__ jobs, const uint2* __restrict__ large,
    const uint32_t* __restrict__ large_ctl, uint32_t large_cap, AxrBins B) {
  const uint32_t n = umin(large_ctl[0], large_cap);
  for (uint32_t i = blockIdx.x; i < n; i += gridDim.x) {
    const uint2 e = large[i];
    const AxrJobDev& J = jobs[e.x];
    const uint32_t gi = e.y;
    const AxrTriRec r = J.rec[gi];
    const int WC = J.dc.WC;
    coarse_block_tiles(r.px, r.py, WC, J.dc.HC, (int)threadIdx.x, kCoarseThreads,
                       [&](int cy, int ax, int G, float xe) {
                         // columns in [0, WC - 1], cy in [0, HC - 1]
#pragma unroll
                         for (int j = 0; j < 8; j++) {
                           const int c = ax + j * G;
                           if ((float)c <= xe) {
                             int tx, ty;
                             walked_tile(cy, c, WC, J.dc.W, tx, ty);
                             append_tile<SHIFT>(B, (uint32_t)B.nb + J.slot0, tx, ty, J.nbx, gi);
                           }
                         }
                       });
  }
}

template <int SHIFT, int MODE>
__global__ __launch_bounds__(1 << (2 * SHIFT) > AXR_BATCH_THREADS ? AXR_BATCH_THREADS
                                                                  : 1 << (2 * SHIFT)) void k_batch_raster(
    const AxrJobDev* __restrict__ jobs, int njobs, AxrBins B) {
  constexpr int NPX = 1 << (2 * SHIFT);
  __shared__ unsigned long long lk[NPX];
  __shared__ uint16_t list[NPX];
  __shared__ uint32_t ncov, item;
  for (;;) {
    if (threadIdx.x == 0) item = atomicAdd(B.head, 1u);
    __syncthreads();
    const uint32_t it = item;
    if (it >= (uint32_t)B.nb) break;
    const uint32_t slot = B.order[it];
    const AxrJobDev& J = jobs[job_of_slot(jobs, njobs, slot)];
    const int b = (int)(slot - J.slot0);
    raster_bin<SHIFT, MODE>(lk, list, &ncov, J.rec, B, slot, (uint32_t)B.nb + slot, 0,
                            b % J.nbx, b / J.nbx, J.vpos, J.vattr, J.groups, J.dc, J.exposure,
                            J.depth, J.r, J.g, J.b, J.rgba, J.ids, (int)threadIdx.x,
                            (int)blockDim.x);
    __syncthreads();  // the block's LDS is reused by the next bin
  }
}

template <int SHIFT, int MODE>
hipError_t raster_launch(const AxrJobDev* jobs, int njobs, const AxrBins& B, int grid,
                         hipStream_t s) {
  constexpr int threads = (1 << (2 * SHIFT)) > AXR_BATCH_THREADS ? AXR_BATCH_THREADS : (1 << (2 * SHIFT));
  hipLaunchKernelGGL((k_batch_raster<SHIFT, MODE>), dim3(grid), dim3(threads), 0, s, jobs, njobs,
                     B);
  return hipGetLastError();
}

template <int SHIFT, int MODE>
int raster_occupancy() {
  constexpr int threads = (1 << (2 * SHIFT)) > AXR_BATCH_THREADS ? AXR_BATCH_THREADS : (1 << (2 * SHIFT));
  int blocks = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, k_batch_raster<SHIFT, MODE>, threads,
                                                   0) != hipSuccess)
    return 1;
  return blocks > 0 ? blocks : 1;
}

}  // namespace

namespace axr {

hipError_t launch_batch_setup(int bin_shift, const AxrJobDev* jobs, const uint4* groups,
                              int ngroups, const uint32_t* jidx, uint32_t blocks,
                              const AxrBins& B, uint2* large, uint32_t* large_ctl,
                              uint32_t large_cap, uint32_t* ctr, uint32_t* visible,
                              hipStream_t s) {
  if (ngroups == 0 || blocks == 0) return hipSuccess;
  switch (bin_shift) {
    case 3:
      hipLaunchKernelGGL(k_batch_setup<3>, dim3(blocks), dim3(kSetupBlock), 0, s, jobs, groups,
                         ngroups, jidx, B, large, large_ctl, large_cap, ctr, visible);
      break;
    case 4:
      hipLaunchKernelGGL(k_batch_setup<4>, dim3(blocks), dim3(kSetupBlock), 0, s, jobs, groups,
                         ngroups, jidx, B, large, large_ctl, large_cap, ctr, visible);
      break;
    default:
      hipLaunchKernelGGL(k_batch_setup<5>, dim3(blocks), dim3(kSetupBlock), 0, s, jobs, groups,
                         ngroups, jidx, B, large, large_ctl, large_cap, ctr, visible);
  }
  return hipGetLastError();
}

hipError_t launch_batch_coarse(int bin_shift, const AxrJobDev* jobs, const uint2* large,
                               const uint32_t* large_ctl, uint32_t large_cap, const AxrBins& B,
                               int grid, hipStream_t s) {
  switch (bin_shift) {
    case 3:
      hipLaunchKernelGGL(k_batch_coarse<3>, dim3(grid), dim3(kCoarseThreads), 0, s, jobs, large,
                         large_ctl, large_cap, B);
      break;
    case 4:
      hipLaunchKernelGGL(k_batch_coarse<4>, dim3(grid), dim3(kCoarseThreads), 0, s, jobs, large,
                         large_ctl, large_cap, B);
      break;
    default:
      hipLaunchKernelGGL(k_batch_coarse<5>, dim3(grid), dim3(kCoarseThreads), 0, s, jobs, large,
                         large_ctl, large_cap, B);
  }
  return hipGetLastError();
}

hipError_t launch_batch_raster(int bin_shift, int mode, const AxrJobDev* jobs, int njobs,
                               const AxrBins& B, int grid, hipStream_t s) {
  if (njobs == 0 || B.nb == 0) return hipSuccess;
#define AXR_RL(SH)                                                       \
  return mode == SHADE_IDS   ? raster_launch<SH, SHADE_IDS>(jobs, njobs, B, grid, s) \
         : mode == SHADE_HDR ? raster_launch<SH, SHADE_HDR>(jobs, njobs, B, grid, s) \
                             : raster_launch<SH, SHADE_LDR8>(jobs, njobs, B, grid, s);
  if (bin_shift == 3) AXR_RL(3)
  if (bin_shift == 4) AXR_RL(4)
  AXR_RL(5)
#undef AXR_RL
}

int batch_raster_blocks_per_cu(int bin_shift, int mode) {
#define AXR_OCC(SH)                                                 \
  return mode == SHADE_IDS   ? raster_occupancy<SH, SHADE_IDS>()    \
         : mode == SHADE_HDR ? raster_occupancy<SH, SHADE_HDR>()    \
                             : raster_occupancy<SH, SHADE_LDR8>();
  if (bin_shift == 3) AXR_OCC(3)
  if (bin_shift == 4) AXR_OCC(4)
  AXR_OCC(5)
#undef AXR_OCC
}

}  // namespace axr
