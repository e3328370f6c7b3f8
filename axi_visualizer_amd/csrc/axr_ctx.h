// axr_ctx.h -- host-side state of a libaxr context, shared by the
// single-frame API (axr_api.cpp) and the batch API (batch_api.cpp).
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <map>
#include <string>
#include <vector>

#include "../../include/axr.h"
#include "axr_launch.h"

using namespace axr;

namespace axri {

int fail(int code, const char* fmt, ...);

#define HIPCHK(expr)                                                                 \
  do {                                                                               \
    hipError_t e_ = (expr);                                                          \
    if (e_ != hipSuccess) return fail(AXR_EHIP, "%s: %s", #expr, hipGetErrorString(e_)); \
  } while (0)

struct HostGroup {
  uint32_t base = 0, ntri = 0;
  int tex_size = 0;
  int is_mip = 0;
  int64_t ntexels = 0;
  float mip_bias = 0.0f;
  int spec_pow = 16;
  uint32_t flags = 0;  // as last uploaded to the device table
  ushort4* d_tex = nullptr;
  const ushort4* d_refl = nullptr;  // SSR draws: the group's reflection cube
  int refl_dim = 0;
};

// A reflection cube strip (setReflTex, Ops.py:458-467): faces of side S
// side by side, S rows of 6S texels, stored as (R,G,B,0).
struct ReflTex {
  ushort4* d = nullptr;
  int size = 0;
};

struct HostShadow {
  bool exists = false;
  bool placed = false;
  bool clear_pending = false;  // cleared since its last pass; map not yet written
  // a pass ran on the map since the bin counters were last read back
  bool unchecked = false;
  // a pass on the map lost bin entries: incomplete until cleared
  bool incomplete = false;
  int dim = 0;
  int nbx = 0, nb = 0;  // 64x64 texel bins
  float scale = 0.0f;
  AxrCam cam{};
  uint32_t* d_map = nullptr;
};

enum State {
  ST_UNCLEARED,  // nothing drawn since the context was made
  ST_CLEARED,    // cleared, nothing drawn
  ST_PENDING,    // triangles binned; colour not materialised
  ST_HDR,        // u16 planes hold HDR colour
  ST_LDR8,       // rgba8 holds gamma'd colour
  ST_LDR16       // u16 planes hold gamma'd colour
};

// A shadowMap call recorded for the next batch of shadow passes.
struct PendingPass {
  int map;
  AxrCastMask cast;
  float bias;
  AxrCam cam;    // placement at the time of the call
  int from_map;  // 0 when the map was cleared before the pass
};

struct TimingRec {
  const char* name;
  hipEvent_t a, b;
};

constexpr int kCounters = AXR_CTR_WORDS;

// Bucket capacities (entries per bin slot).  A pass appends into fixed
// buckets; what does not fit goes to a spill list that the raster scans for
// the affected bins only, and a pass that spilled grows the buckets before
// the next pass of its kind.
constexpr uint32_t kMainCap0 = 4096;   // small triangles per screen bin
constexpr uint32_t kMainCap1 = 1024;   // large-triangle tiles per screen bin
constexpr uint32_t kShadowCap = 8192;  // casters per shadow bin
constexpr uint32_t kMinCap = 256;
constexpr size_t kBucketBudget = (size_t)256 << 20;      // bytes for the initial buckets
constexpr size_t kBucketGrowLimit = (size_t)4 << 30;     // bytes a growth may reach
constexpr uint32_t kSpillCap = 1u << 20;                 // spill entries

// Device bin storage of one pass kind.
struct BinStore {
  AxrBins B{};            // pointers and capacities as kernels see them
  int nseg = 1;
  int nb_alloc = 0;       // bins the cursor/count/order arrays hold
  size_t ent_alloc = 0;   // bucket entries allocated
  uint32_t want0 = 0, want1 = 0, want_spill = 0;  // capacities for the next pass
};

}  // namespace axri

using namespace axri;

struct axr_ctx {
  int device = 0;
  int cu_count = 0;  // compute units of the device
  int W = 0, H = 0, WC = 0, HC = 0;
  int NBX = 0, NB = 0;  // 32x32 screen bins
  hipStream_t stream = nullptr;

  std::vector<HostGroup> groups;
  uint32_t T = 0;
  size_t tri_cap = 0;
  float4* d_vpos = nullptr;   // [3*tri_cap] (x,y,z,0)
  float4* d_vattr = nullptr;  // [6*tri_cap] (nx,ny,nz,u),(v,Ix,Iy,Iz)
  float4* d_chunks = nullptr;  // [tri_cap/64] bounding sphere per 64-triangle chunk
  AxrGroupDev* d_groups = nullptr;
  AxrTriRec* d_rec = nullptr;   // [tri_cap] setup records of the last drawAll
  uint8_t* d_cls = nullptr;     // [tri_cap] 0 culled, 1 small, 2 large
  uint32_t* d_large = nullptr;  // [tri_cap] large triangles of a setup, for coarse binning
  // {count of d_large (zeroed by the order pass), heavy and at least
  // medium screen bins of the last order pass}
  uint32_t* d_large_ctl = nullptr;
  // [tri_cap] ortho setup and caster flags, one array per pass of a batch
  AxrShRec* d_shrec[AXR_MAX_SHADOWS] = {};
  uint8_t* d_ok[AXR_MAX_SHADOWS] = {};
  const uint8_t* last_ok = nullptr;  // caster flags of the last shadow pass run

  BinStore mbins;  // screen bins: slots [0, NB) small, [NB, 2NB) large tiles
  BinStore sbins;  // shadow bins, sized for the largest map

  float* d_depth = nullptr;
  uint16_t *d_r = nullptr, *d_g = nullptr, *d_b = nullptr;
  uint32_t* d_rgba = nullptr;
  uint32_t* d_ids = nullptr;
  uint32_t* d_counters = nullptr;
  void* d_post = nullptr;     // blur / bloom intermediates
  size_t post_bytes = 0;
  void* d_snap = nullptr;     // SSR: copy of the opaque frame (depth, R, G, B)
  size_t snap_bytes = 0;
  AxrGroupDev* d_groups_ssr = nullptr;   // group table of SSR draws
  std::vector<AxrGroupDev> h_groups_ssr;  // as last uploaded
  std::vector<ReflTex> refl;
  // A cube pass (axr_render_cube) renders this scene on a batch's stream:
  // its reads of the geometry, group table, textures and shadow maps and its
  // write of the reflection strip end at ev_cube.  The scene's stream waits
  // on it (wait_cube) only before it writes one of those or reads a strip,
  // so the scene's own drawAll overlaps the cube pass.
  hipEvent_t ev_cube = nullptr;
  bool cube_pending = false;
  float* d_light = nullptr;  // vertLight light arrays
  size_t light_cap = 0;

  HostShadow sm[AXR_MAX_SHADOWS];
  std::vector<PendingPass> pend;  // shadow passes not yet launched (one per map)
  AxrCam view{};
  float lint[3] = {1.0f, 1.0f, 1.0f};
  float ldir[3] = {0.0f, -1.0f, 0.0f};
  float amb = 0.2f;
  uint16_t clear_rgb[3] = {1200, 4033, 21675};  // (60,110,255)^2/3, Ops.py:941

  int state = ST_UNCLEARED;
  AxrDrawConst last{};
  bool records_valid = false;  // setup records and bins belong to `last`
  // the frame since the last clearZBuffer lost bin entries, or samples an
  // incomplete shadow map: every frame readback fails until the next clear
  bool frame_incomplete = false;
  // More than one drawAll since the clear: the frame's winner ids are kept
  // in d_ids as each draw is composed (composed), and the pending draw is to
  // be rastered onto the frame rather than from the clear (accum).
  bool composed = false, accum = false;

  bool timing = false;
  std::vector<TimingRec> trecs;
  std::vector<hipEvent_t> free_events;
  std::map<std::string, std::pair<double, int64_t>> tacc;
};

namespace axri {

// Scoped event pair around a launch when timing is enabled.
struct Timed {
  axr_ctx* c;
  const char* name;
  hipEvent_t a = nullptr, b = nullptr;
  hipEvent_t take() {
    if (!c->free_events.empty()) {
      hipEvent_t e = c->free_events.back();
      c->free_events.pop_back();
      return e;
    }
    hipEvent_t e = nullptr;
    (void)hipEventCreate(&e);
    return e;
  }
  Timed(axr_ctx* ctx, const char* n) : c(ctx), name(n) {
    if (!c->timing) return;
    a = take();
    b = take();
    (void)hipEventRecord(a, c->stream);
  }
  ~Timed() {
    if (!c->timing) return;
    (void)hipEventRecord(b, c->stream);
    c->trecs.push_back({name, a, b});
  }
};

template <class T>
int dalloc(T** p, size_t n) {
  *p = nullptr;
  if (n == 0) n = 1;
  HIPCHK(hipMalloc((void**)p, n * sizeof(T)));
  return 0;
}

template <class T>
void dfree(T*& p) {
  if (p) (void)hipFree((void*)p);
  p = nullptr;
}

void bins_free(BinStore& S);
uint32_t fit_cap(uint32_t cap, size_t nb, int nseg);
int bins_prepare(BinStore& S, int nb, hipStream_t st);
void bins_grow(BinStore& S, uint32_t spilled, uint32_t lost);
int check_ctx(axr_ctx* c);
int apply_shadow_clear(axr_ctx* c, int i);
int flush_shadows(axr_ctx* c);
int wait_cube(axr_ctx* c);

}  // namespace axri
