// axr_launch.h -- host-side launchers of the gfx950 kernels (one stream each).
#pragma once
#include "axr_common.h"

namespace axr {

hipError_t launch_fill_u64(uint64_t* p, size_t n, uint64_t v, hipStream_t s);
hipError_t launch_fill_u32(uint32_t* p, size_t n, uint32_t v, hipStream_t s);

hipError_t launch_clear_frame(uint64_t* keys, size_t n, uint32_t* counters, uint32_t* tile_cnt,
                              int ntile, hipStream_t s);
hipError_t launch_clear_shadow(uint32_t* map, size_t n, uint32_t* counters, hipStream_t s);
hipError_t launch_setup_raster(const float4* vpos, const float4* vattr, const AxrGroupDev* groups,
                               const AxrDrawConst& dc, AxrTriRec* large, uint32_t* counters,
                               uint64_t* keys, hipStream_t s);
hipError_t launch_bin_count(const AxrTriRec* large, const uint32_t* counters, uint32_t* tile_cnt,
                            int WC, int HC, uint32_t max_n, hipStream_t s);
hipError_t launch_bin_scan(uint32_t* tile_cnt, uint32_t* tile_off, int NT, uint32_t* counters,
                           uint32_t cap, hipStream_t s);
hipError_t launch_bin_fill(const AxrTriRec* large, const uint32_t* counters, uint32_t* tile_cnt,
                           const uint32_t* tile_off, uint32_t* bins, int WC, int HC, uint32_t cap,
                           uint32_t max_n, hipStream_t s);
hipError_t launch_raster_tiles(const AxrTriRec* large, const uint32_t* tile_off,
                               const uint32_t* bins, uint64_t* keys, int W, int WC, int HC,
                               uint32_t T, uint32_t cap, hipStream_t s);
hipError_t launch_shadow(const float4* vpos, const AxrGroupDev* groups, uint32_t ngroups,
                         uint32_t T, const AxrCastMask& cast, const AxrCam& cam, int dim,
                         float bias, uint32_t* map, uint32_t* counters, hipStream_t s);

// Shade modes: HDR u16 planes, or gamma fused into RGBA8.
enum ShadeMode { SHADE_HDR = 0, SHADE_LDR8 = 1 };
hipError_t launch_shade(int mode, const uint64_t* keys, const float4* vpos, const float4* vattr,
                        const AxrGroupDev* groups, const AxrDrawConst& dc, float exposure,
                        float* depth, uint16_t* r, uint16_t* g, uint16_t* b, uint32_t* rgba,
                        hipStream_t s);
hipError_t launch_gamma_planes(uint16_t* r, uint16_t* g, uint16_t* b, size_t n, float ex,
                               hipStream_t s);
hipError_t launch_expand_rgba8(const uint32_t* rgba, uint16_t* r, uint16_t* g, uint16_t* b,
                               size_t n, hipStream_t s);
hipError_t launch_ids(const uint64_t* keys, uint32_t T, uint32_t* ids, size_t n, hipStream_t s);
hipError_t launch_decode_shadow(const uint32_t* map, float* out, size_t n, hipStream_t s);

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
