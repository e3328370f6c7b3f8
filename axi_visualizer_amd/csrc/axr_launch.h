// axr_launch.h -- host-side launchers of the gfx950 kernels (one stream each).
#pragma once
#include "axr_common.h"

namespace axr {

hipError_t launch_fill_u32(uint32_t* p, size_t n, uint32_t v, hipStream_t s);

// Setup + bin appends of the drawAll (do_main) and of the pending shadow
// passes in ps, one launch over global triangles [t0, t1) (t0 is rounded
// down to a 64-triangle chunk; the other triangles' records and classes are
// left as they are).  The draw's large triangles go to the list `large`
// (count in large_ctl[0]) and a second launch bins their tiles; the order
// pass zeroes the count.
hipError_t launch_setup(const float4* vpos, const float4* vattr, const AxrGroupDev* groups,
                        const AxrDrawConst& dc, int do_main, AxrTriRec* rec, uint8_t* cls,
                        const AxrBins& MB, int NBX, const AxrPassSet& ps, const AxrBins& SB,
                        const float4* chunks, uint32_t t0, uint32_t t1, uint32_t* large,
                        uint32_t* large_ctl, hipStream_t s);
// HDR planes -> gamma (Shaders/Post/gamma.c) packed as RGBA8.
hipError_t launch_gamma_pack(const uint16_t* r, const uint16_t* g, const uint16_t* b,
                             uint32_t* rgba, size_t n, float ex, hipStream_t s);
// Bounding sphere per 64-triangle chunk (setup skips chunks out of view).
hipError_t launch_chunk_bounds(const float4* vpos, uint32_t T, float4* chunks, hipStream_t s);
// Counts of the pass, heavy-first order, cursor reset for the screen bins
// and/or the shadow bins (either may be null), one launch; with the screen
// bins it also zeroes *clear (the setup's large-triangle count) and, if
// tiers is given, writes the number of heavy bins to tiers[0] and of bins
// at least medium to tiers[1].
// A screen bin of weight (small entries + 16 x large-tile entries) of at
// least kHeavyWeight is heavy: the single-frame raster gives it a whole
// block; of at least kMediumWeight, medium: half a block; lighter bins a
// quarter.  Both are powers of two.
// Measured on MI355X over the thresholds 256..4096 and 16..512: these keep
// the 1080p/1M-triangle frame's dense bins fast and take the 1440p full
// frame's opaque raster from 351 to 300 us.
#ifndef AXR_HEAVY_WEIGHT
#define AXR_HEAVY_WEIGHT 1024
#endif
#ifndef AXR_MEDIUM_WEIGHT
#define AXR_MEDIUM_WEIGHT 256
#endif
// Frames of at most this many 32-px bins per CU are rastered a 16-px
// quadrant per block; larger ones a bin (or two, or four) per block.
#ifndef AXR_QUAD_BINS_PER_CU
#define AXR_QUAD_BINS_PER_CU 2
#endif
constexpr uint32_t kHeavyWeight = AXR_HEAVY_WEIGHT;
constexpr uint32_t kMediumWeight = AXR_MEDIUM_WEIGHT;
hipError_t launch_bin_order(const AxrBins* main_bins, const AxrBins* shadow_bins,
                            uint32_t* ctr, uint32_t* clear, uint32_t* tiers, hipStream_t s);
// Device scratch of the many-block order pass (zeroed once at allocation;
// each pass leaves it zeroed again).
struct OrderScratch {
  uint32_t hist[64];  // bins per weight class
  uint32_t cur[64];   // positions taken per class
  uint32_t total, ticket;
};
// The same order pass for one large two-segment screen store (a batch's
// bins, no tiers) over many blocks in two launches: the same counts,
// counters and class order.
hipError_t launch_bin_order_wide(const AxrBins& B, OrderScratch* sc, uint32_t* ctr,
                                 uint32_t* clear, hipStream_t s);
// Depth walk of every pending map's bins, one launch.
hipError_t launch_shadow_bins(const AxrBins& B, const AxrPassSet& ps, hipStream_t s);
hipError_t launch_stats(const uint8_t* cls, const uint8_t* ok, uint32_t T, uint32_t* out,
                        hipStream_t s);

// Bin raster modes: HDR u16 planes, gamma fused into RGBA8, or winner ids;
// SHADE_ACCUM draws onto the frame already there (a drawAll with no clear
// before it): each pixel's key starts from its depth, and only pixels a new
// fragment wins are written (HDR planes, depth and winner id).
enum ShadeMode { SHADE_HDR = 0, SHADE_LDR8 = 1, SHADE_IDS = 2, SHADE_ACCUM = 3 };
hipError_t launch_bin_raster(int mode, const AxrTriRec* rec, const AxrBins& B, int empty,
                             const float4* vpos, const float4* vattr, const AxrGroupDev* groups,
                             const AxrDrawConst& dc, int NBX, float exposure, float* depth,
                             uint16_t* r, uint16_t* g, uint16_t* b, uint32_t* rgba,
                             uint32_t* ids, const uint32_t* tiers, hipStream_t s);
hipError_t launch_gamma_planes(uint16_t* r, uint16_t* g, uint16_t* b, size_t n, float ex,
                               hipStream_t s);
hipError_t launch_expand_rgba8(const uint32_t* rgba, uint16_t* r, uint16_t* g, uint16_t* b,
                               size_t n, hipStream_t s);
hipError_t launch_pack_rgba8(const uint16_t* r, const uint16_t* g, const uint16_t* b,
                             uint32_t* rgba, size_t n, hipStream_t s);
// Smallest-first mip chain of `levels` levels whose last (largest) level
// is filled: builds the others on the device.
hipError_t launch_create_mips(ushort4* chain, int levels, hipStream_t s);
// SSR water pass (ssr.hip): F0/R0/G0/B0 are a copy of the opaque frame.
struct AxrSsrConst {
  const float* F0;
  const uint16_t *R0, *G0, *B0;
  int steps, step;  // REFL_LENGTH, REFL_STEP
  float absorb;     // ABSORB
  int fresnel;      // 0: fr = 1 (the metal variant)
};
hipError_t launch_ssr(const AxrTriRec* rec, const AxrBins& B, const float4* vpos,
                      const float4* vattr, const AxrGroupDev* groups, const AxrDrawConst& dc,
                      const AxrSsrConst& S, int NBX, float* depth, uint16_t* r, uint16_t* g,
                      uint16_t* b, hipStream_t s);
// Copies depth and the three HDR planes (np pixels, every pointer 16-byte
// aligned) into the snapshot the SSR pass reads.
hipError_t launch_snap_frame(const float* depth, const uint16_t* r, const uint16_t* g,
                             const uint16_t* b, size_t np, float* F0, uint16_t* R0,
                             uint16_t* G0, uint16_t* B0, hipStream_t s);
// Six faces' planes [6][S][S] into the cube strip [S][6S] of (R,G,B,0).
hipError_t launch_pack_cube(const uint16_t* r, const uint16_t* g, const uint16_t* b, int S,
                            ushort4* strip, hipStream_t s);

// Post passes on the u16 planes (post.hip).  blur: tmp holds
// 6 * (W/2) * (H/2) u16.  bloom: tmp holds bloom_halves(W, H, levels) fp16.
hipError_t launch_blur(uint16_t* r, uint16_t* g, uint16_t* b, int W, int H, uint16_t* tmp,
                       hipStream_t s);
size_t bloom_halves(int W, int H, int levels);
hipError_t launch_bloom(uint16_t* r, uint16_t* g, uint16_t* b, int W, int H, int levels,
                        float threshold, float white, float strength, void* tmp, hipStream_t s);
hipError_t launch_decode_shadow(const uint32_t* map, float* out, size_t n, hipStream_t s);

// Batched jobs (batch.hip).  bin_shift 3/4/5 = 8/16/32-px bins.
// groups: per scene of the batch (first setup block, offset of its job
// indices in jidx, job count, triangles), ascending in first block.
// large: (job, triangle) pairs of the large triangles, counted in
// large_ctl[0] (reset by the order pass); pairs past large_cap add to
// ctr[AXR_CTR_LOST].  visible (may be null): per-job counts of the
// triangles the setup accepts, added to.
hipError_t launch_batch_setup(int bin_shift, const AxrJobDev* jobs, const uint4* groups,
                              int ngroups, const uint32_t* jidx, uint32_t blocks,
                              const AxrBins& B, uint2* large, uint32_t* large_ctl,
                              uint32_t large_cap, uint32_t* ctr, uint32_t* visible,
                              hipStream_t s);
hipError_t launch_batch_coarse(int bin_shift, const AxrJobDev* jobs, const uint2* large,
                               const uint32_t* large_ctl, uint32_t large_cap, const AxrBins& B,
                               int grid, hipStream_t s);
// mode: a ShadeMode (RGBA8 with gamma, winner ids, or HDR planes).
hipError_t launch_batch_raster(int bin_shift, int mode, const AxrJobDev* jobs, int njobs,
                               const AxrBins& B, int grid, hipStream_t s);
int batch_raster_blocks_per_cu(int bin_shift, int mode);

struct AxrLights {
  int ndir, npoint, nspot;
  float amb;
  const float* dirI;  // [ndir][3] on device
  const float* dirD;
  const float* pointI;
  const float* pointP;
  const float* spotI;
  const float* spotD;
  const float* spotP;
};
hipError_t launch_vertlight(const float4* vpos, float4* vattr, uint32_t v0, uint32_t nv,
                            const AxrLights& L, hipStream_t s);

}  // namespace axr
