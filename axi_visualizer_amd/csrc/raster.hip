// raster.hip -- setup and binning of the gfx950 raster path, and the
// shadow-map depth pass.
//
// A drawAll and the shadow passes recorded before it run as three launches:
//   k_setup      per triangle, positions read once: the drawAll's setup (a5)
//                with its record stored by global index, small triangles
//                appended to the 32x32 screen bins their pixel rectangle
//                touches and large triangles to the bins of their coarse
//                16-px tiles (a8, Pipe/coarse.c); then the ortho setup of each
//                pending shadow pass (a11, Pipe/trisetupOrtho.c) appended to
//                that map's 64x64 texel bins
//   k_bin_order  entry counts of the pass, heavy-first bin order; resets the
//                append cursors for the next pass (one block per bin store;
//                a batch's store of every job's bins takes k_order_count and
//                k_order_scatter over many blocks instead)
//   k_shadow_bins  each pending map's bins walked with the bin's texels in
//                LDS (Shaders/drawmin.c) and written once
// The screen bins' depth walk and shading run later in one launch
// (shade.hip).  Shadow passes with no drawAll after them run the same
// launches without the screen part.
//
// Appends aggregate within a wave: lanes of one wave come from one mesh
// region and mostly share a bin, so one atomic serves all of them.  Depth
// minima live in LDS; no global atomics touch pixels or texels.
#include "axr_bin.h"

using namespace axr;

namespace {

constexpr int kBlock = 256;

__global__ void k_fill_u32(uint32_t* __restrict__ p, size_t n, uint32_t v) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x)
    p[i] = v;
}

// Bounding sphere of each 64-triangle chunk (one wave per chunk): the box
// centre, and the largest vertex distance from it, widened by 1e-5 relative
// and by 1e-6 of the coordinate magnitude so it holds every vertex.
__global__ __launch_bounds__(256) void k_chunk_bounds(const float4* __restrict__ vpos, uint32_t T,
                                                      float4* __restrict__ chunks) {
  const uint32_t gi = blockIdx.x * blockDim.x + threadIdx.x;
  const uint32_t chunk = gi >> 6, nch = (T + 63) >> 6;
  const bool live = gi < T;
  float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
  float4 v[3];
  bool finite = true;
#pragma unroll
  for (int k = 0; k < 3; k++) {
    v[k] = live ? vpos[3 * (size_t)gi + k] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    if (live) {
      const float c[3] = {v[k].x, v[k].y, v[k].z};
#pragma unroll
      for (int a = 0; a < 3; a++) {
        finite = finite && isfinite(c[a]);
        lo[a] = fminf(lo[a], c[a]);
        hi[a] = fmaxf(hi[a], c[a]);
      }
    }
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
#pragma unroll
    for (int a = 0; a < 3; a++) {
      lo[a] = fminf(lo[a], __shfl_xor(lo[a], o));
      hi[a] = fmaxf(hi[a], __shfl_xor(hi[a], o));
    }
  }
  const float cx = 0.5f * (lo[0] + hi[0]), cy = 0.5f * (lo[1] + hi[1]), cz = 0.5f * (lo[2] + hi[2]);
  float r = 0.0f;
  if (live) {
#pragma unroll
    for (int k = 0; k < 3; k++) {
      const float dx = v[k].x - cx, dy = v[k].y - cy, dz = v[k].z - cz;
      r = fmaxf(r, sqrtf(dx * dx + dy * dy + dz * dz));
    }
  }
  const bool all_finite = __ballot(live && !finite) == 0;
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) r = fmaxf(r, __shfl_xor(r, o));
  if (__lane_id() == 0 && chunk < nch) {
    const float mag = fabsf(cx) + fabsf(cy) + fabsf(cz) + r;
    chunks[chunk] = all_finite ? make_float4(cx, cy, cz, r * 1.00001f + 1e-6f * mag + 1e-30f)
                               : make_float4(0.0f, 0.0f, 0.0f, NAN);
  }
}

// K1: setup + bin appends.  Screen slot b: small triangles (global index);
// slot NB + b: large triangles' coarse tiles, (global index << 2) | the
// 16-px tile's quadrant in bin b.  Shadow slot slot0 + b of a pass: casters
// (global index).
__global__ __launch_bounds__(kBlock) void k_setup(const float4* __restrict__ vpos,
                                                  const float4* __restrict__ vattr,
                                                  const AxrGroupDev* __restrict__ groups,
                                                  AxrDrawConst dc, int do_main,
                                                  AxrTriRec* __restrict__ rec,
                                                  uint8_t* __restrict__ cls_out, AxrBins MB,
                                                  int NBX, AxrPassSet ps, AxrBins SB,
                                                  const float4* __restrict__ chunks,
                                                  uint32_t t0, uint32_t t1,
                                                  uint32_t* __restrict__ large,
                                                  uint32_t* __restrict__ large_ctl) {
  const uint32_t gi = t0 + blockIdx.x * blockDim.x + threadIdx.x;
  const bool live = gi < t1;
  // The wave's 64 triangles are one chunk: skip the passes it cannot reach.
  const float4 sph = chunks[imin((int)(gi >> 6), (int)((dc.T - 1) >> 6))];
  const bool main_on = do_main && !chunk_outside_view(sph, dc.view, dc.cull);
  bool any_on = main_on;
  uint32_t pass_on = 0;
#pragma unroll
  for (int p = 0; p < AXR_MAX_SHADOWS; p++) {
    if (p >= ps.n) break;
    if (!chunk_outside_ortho(sph, ps.p[p].cam, ps.p[p].cull, ps.p[p].dim, ps.p[p].bias)) pass_on |= 1u << p;
  }
  any_on |= pass_on != 0;
  float4 vp[3];
  int grp = 0;
  if (any_on) {
    // The wave's triangles are consecutive: one wave-wide search (scalar
    // loads) finds the group of its first one, and only lanes past that
    // group's end search on their own.
    const int ng = (int)dc.ngroups;
    const int gw = find_group(groups, ng, __builtin_amdgcn_readfirstlane(gi));
    const uint32_t gend = gw + 1 < ng ? groups[gw + 1].base : 0xFFFFFFFFu;
    grp = gw;
    if (live && gi >= gend) grp = find_group(groups, ng, gi);
  }
  if (live && any_on) {
#pragma unroll
    for (int k = 0; k < 3; k++) vp[k] = vpos[3 * (size_t)gi + k];
  } else {
#pragma unroll
    for (int k = 0; k < 3; k++) vp[k] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  }
  // The screen setup and the first two shadow passes' setups run first;
  // their bin appends then share one loop, so that a wave's three streams
  // of atomics are in flight together instead of one round trip each.
  BinRect brm = {0, 0, 1, 0};
  if (do_main) {
    int cls = 0;
    AxrTriRec r;
    if (live && !main_on) cls_out[gi] = 0;
    if (live && main_on) {
      const uint32_t f = groups[grp].flags;
      if (f & AXR_GRP_DRAW)
        cls = tri_setup(vp, vattr, gi, dc.view, dc.W, dc.H, (f & AXR_SH_CULL) != 0, r.px, r.py,
                        r.z);
      cls_out[gi] = (uint8_t)cls;
      if (cls) rec[gi] = r;
    }
    if (cls == 1) {
      int x0, y0, x1, y1;
      small_bounds(r.px, r.py, x0, y0, x1, y1);
      brm = bin_rect(x0, y0, x1, y1, dc.W, dc.H, AXR_BIN_SHIFT);
    }
    // large triangles are binned by k_coarse, a block per triangle
    const uint32_t at = wave_bin_add(large_ctl, cls == 2, 0u);
    if (cls == 2) large[at] = gi;
  }
  // shadow pass p: ortho setup, its record, and the bin rectangle and entry
  // of its caster
  auto shadow_pass = [&](const AxrShadowPass& P, int p, BinRect& br, uint32_t& entry) {
    bool ok = false;
    AxrShRec r;
    if (live) {
      if (((pass_on >> p) & 1u) && cast_bit(P.cast, grp))
        ok = shadow_setup(vp, P.cam, P.dim, P.bias, r);
      P.ok[gi] = ok ? 1 : 0;
      if (ok) P.rec[gi] = r;
    }
    br = BinRect{0, 0, 1, 0};
    entry = gi;
    if (ok) {
      int x0, y0, x1, y1;
      shadow_bounds(r, x0, y0, x1, y1);
      br = bin_rect(x0, y0, x1, y1, P.dim, P.dim, AXR_SHADOW_SHIFT);
      entry |= shadow_class(r) << 30;
    }
  };
  BinRect br0 = {0, 0, 1, 0}, br1 = {0, 0, 1, 0};
  uint32_t e0 = gi, e1 = gi;
  if (ps.n > 0) shadow_pass(ps.p[0], 0, br0, e0);
  if (ps.n > 1) shadow_pass(ps.p[1], 1, br1, e1);
  {
    BinCursor cm, c0, c1;
    for (int k = 0;; k++) {
      const bool am = k < brm.n, a0 = k < br0.n, a1 = k < br1.n;
      if (__ballot(am || a0 || a1) == 0) break;
      const uint32_t sm = am ? cm.bin(brm, NBX) : 0u;
      const uint32_t s0 = a0 ? ps.p[0].slot0 + c0.bin(br0, ps.p[0].nbx) : 0u;
      const uint32_t s1 = a1 ? ps.p[1].slot0 + c1.bin(br1, ps.p[1].nbx) : 0u;
      const WaveAdd wm = wave_bin_add_issue(MB.cnt, am, sm);
      const WaveAdd w0 = wave_bin_add_issue(SB.cnt, a0, s0);
      const WaveAdd w1 = wave_bin_add_issue(SB.cnt, a1, s1);
      const uint32_t qm = wm.finish(), q0 = w0.finish(), q1 = w1.finish();
      if (am) bin_store(MB, sm, qm, gi);
      if (a0) bin_store(SB, s0, q0, e0);
      if (a1) bin_store(SB, s1, q1, e1);
      cm.next(brm);
      c0.next(br0);
      c1.next(br1);
    }
  }
  // further pending passes (more than two maps rendered before a draw)
#pragma unroll
  for (int p = 2; p < AXR_MAX_SHADOWS; p++) {
    if (p >= ps.n) break;
    const AxrShadowPass& P = ps.p[p];
    BinRect br;
    uint32_t entry;
    shadow_pass(P, p, br, entry);
    BinCursor bc;
    for (int k = 0;; k++) {
      const bool act = k < br.n;
      if (__ballot(act) == 0) break;
      const uint32_t slot = act ? P.slot0 + bc.bin(br, P.nbx) : 0u;
      const uint32_t pos = wave_bin_add(SB.cnt, act, slot);
      if (act) bin_store(SB, slot, pos, entry);
      bc.next(br);
    }
  }
}

// Coarse binning of the setup's large triangles (a8, Pipe/coarse.c:26-236):
// a block per triangle of the list; each tile row (coarse_rows) goes to a
// group of threads which share the row's tiles (coarse_block_tiles); the
// appends are issued eight at a time so their atomics overlap.  A wide flat
// triangle (few rows, spans of a hundred tiles) thus keeps the block busy
// instead of a handful of threads.  Every thread of a group walks the rows
// with the same additions, so each tile is emitted once, as the serial walk
// emits it.  The order pass that follows resets the list for the next
// setup.
constexpr int kCoarseBlocks = 256;
constexpr int kCoarseThreads = 256;
__global__ __launch_bounds__(kCoarseThreads) void k_coarse(const AxrTriRec* __restrict__ rec,
                                                           const uint32_t* __restrict__ large,
                                                           const uint32_t* __restrict__ ctl,
                                                           AxrBins MB,
                                                           int W, int WC, int HC, int NBX) {
  const uint32_t n = ctl[0];
  for (uint32_t i = blockIdx.x; i < n; i += gridDim.x) {
    const uint32_t gi = large[i];
    const AxrTriRec r = rec[gi];
    coarse_block_tiles(r.px, r.py, WC, HC, (int)threadIdx.x, kCoarseThreads,
                       [&](int cy, int ax, int G, float xe) {
                         // columns in [0, WC - 1] and cy in [0, HC - 1]
                         // (coarse_block_tiles, coarse_rows)
                         // each to the bin and quadrant of the tile the
                         // tiled draw walks for it (walked_tile)
                         uint32_t slot[8], pos[8], sb[8];
                         bool on[8];
#pragma unroll
                         for (int j = 0; j < 8; j++) {
                           int tx, ty;
                           walked_tile(cy, ax + j * G, WC, W, tx, ty);
                           on[j] = (float)(ax + j * G) <= xe;
                           slot[j] = (uint32_t)MB.nb + (uint32_t)((ty >> 1) * NBX + (tx >> 1));
                           sb[j] = (uint32_t)(((ty & 1) << 1) | (tx & 1));
                           pos[j] = on[j] ? atomicAdd(&MB.cnt[slot[j]], 1u) : 0u;
                         }
#pragma unroll
                         for (int j = 0; j < 8; j++)
                           if (on[j]) bin_store(MB, slot[j], pos[j], (gi << 2) | sb[j]);
                       });
  }
}

constexpr int kOrderClasses = 64;
// 0 for an empty bin, else 2*floor(log2 w) + (next bit) + 2, capped at 63.
__device__ __forceinline__ uint32_t order_class(uint32_t w) {
  if (w == 0) return 0;
  const int l = imin(30, 31 - __clz(w));
  const int half = l > 0 ? (int)((w >> (l - 1)) & 1u) : 0;
  return (uint32_t)(2 * l + half + 2);
}

// One bin store to order, with the counter words it reports into.
struct OrderJob {
  AxrBins B;
  int nseg;
  int ctr_total, ctr_spill;
  uint32_t* heavy;  // if set: bins of weight >= kHeavyWeight, then >= kMediumWeight
};

// K2, one block per bin store: num = cnt and cnt = 0 for the next pass; the
// pass's entry total, spill and loss into the counter block; bins in
// heavy-first launch order (weight = small entries + 16 x large-tile
// entries), so the longest bins start first and short ones fill the tail.
// Order within a weight class is arbitrary: each bin is computed by one
// block, so results do not depend on it.
constexpr int kOrderPer = 8;               // bins per thread per round
constexpr int kOrderLds = 1024 * kOrderPer * 2;  // bins whose class is kept in LDS

__global__ __launch_bounds__(1024) void k_bin_order(OrderJob j0, OrderJob j1,
                                                    uint32_t* __restrict__ ctr,
                                                    uint32_t* __restrict__ clear) {
  __shared__ uint32_t hist[kOrderClasses];
  __shared__ uint32_t total;
  __shared__ uint8_t klass[kOrderLds];
  const bool second = blockIdx.x != 0;
  const AxrBins B = second ? j1.B : j0.B;
  const int nseg = second ? j1.nseg : j0.nseg;
  const int nb = B.nb;
  const int tid = threadIdx.x, lane = tid & 63;
  if (tid < kOrderClasses) hist[tid] = 0;
  if (tid == 0) total = 0;
  if (clear && tid == 0 && blockIdx.x == 0) *clear = 0;  // the large list, read by now
  __syncthreads();
  uint32_t tot = 0;
  // Counts are loaded a round at a time so the loads overlap.
  for (int base = 0; base < nb; base += 1024 * kOrderPer) {
    uint32_t c0[kOrderPer], c1[kOrderPer];
#pragma unroll
    for (int u = 0; u < kOrderPer; u++) {
      const int b = base + u * 1024 + tid;
      c0[u] = b < nb ? B.cnt[b] : 0u;
      c1[u] = (nseg > 1 && b < nb) ? B.cnt[nb + b] : 0u;
    }
#pragma unroll
    for (int u = 0; u < kOrderPer; u++) {
      const int b = base + u * 1024 + tid;
      const bool valid = b < nb;
      uint32_t k = 0;
      if (valid) {
        B.num[b] = c0[u];
        B.cnt[b] = 0;
        if (nseg > 1) {
          B.num[nb + b] = c1[u];
          B.cnt[nb + b] = 0;
        }
        tot += c0[u] + c1[u];
        k = order_class(c0[u] + 16u * c1[u]);
        if (b < kOrderLds) klass[b] = (uint8_t)k;
      }
      // empty bins (most of a sparse frame) count with one atomic per wave
      const uint64_t em = __ballot(valid && k == 0);
      if (em && lane == __ffsll((unsigned long long)em) - 1) atomicAdd(&hist[0], (uint32_t)__popcll(em));
      if (valid && k != 0) atomicAdd(&hist[k], 1u);
    }
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) tot += __shfl_xor(tot, o);
  if (lane == 0) atomicAdd(&total, tot);
  __syncthreads();
  if (tid < 64) {
    // descending exclusive offsets: classes above this one come first
    const uint32_t h = hist[tid];
    uint32_t sfx = h;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
      const uint32_t t = __shfl_down(sfx, o);
      if (lane + o < 64) sfx += t;
    }
    hist[tid] = sfx - h;
  }
  if (tid == 0) {
    ctr[second ? j1.ctr_total : j0.ctr_total] = total;
    const uint32_t used = B.spill_ctr[0];
    B.spill_ctr[1] = used < B.spill_cap ? used : B.spill_cap;
    B.spill_ctr[0] = 0;
    if (B.head) *B.head = 0;
    if (used) atomicMax(&ctr[second ? j1.ctr_spill : j0.ctr_spill], used);
    if (used > B.spill_cap) atomicAdd(&ctr[AXR_CTR_LOST], used - B.spill_cap);
  }
  __syncthreads();
  uint32_t* heavy = second ? j1.heavy : j0.heavy;
  if (heavy) {
    // hist[c] now counts the bins of classes above c; weight 2^k starts
    // class 2k + 2
    constexpr int kHeavyClass = 2 * (31 - __builtin_clz(kHeavyWeight)) + 2;
    constexpr int kMediumClass = 2 * (31 - __builtin_clz(kMediumWeight)) + 2;
    static_assert((kHeavyWeight & (kHeavyWeight - 1)) == 0, "a power of two");
    static_assert((kMediumWeight & (kMediumWeight - 1)) == 0, "a power of two");
    static_assert(kMediumWeight <= kHeavyWeight, "medium below heavy");
    if (tid == 0) {
      heavy[0] = hist[kHeavyClass - 1];
      heavy[1] = hist[kMediumClass - 1];
    }
    __syncthreads();
  }
  for (int b0 = 0; b0 < nb; b0 += 1024) {
    const int b = b0 + tid;
    const bool valid = b < nb;
    uint32_t k = 0;
    if (valid) {
      if (b < kOrderLds) {
        k = klass[b];
      } else {  // written by this thread above
        k = order_class(B.num[b] + (nseg > 1 ? 16u * B.num[nb + b] : 0u));
      }
    }
    const uint64_t em = __ballot(valid && k == 0);
    uint32_t base = 0;
    const int leader = em ? __ffsll((unsigned long long)em) - 1 : 0;
    if (em && lane == leader) base = atomicAdd(&hist[0], (uint32_t)__popcll(em));
    base = __shfl(base, leader);
    if (valid) {
      const uint32_t pos = k == 0 ? base + (uint32_t)__popcll(em & lanemask_lt())
                                  : atomicAdd(&hist[k], 1u);
      B.order[pos] = (uint32_t)b;
    }
  }
}

// The order pass of a large store (a batch: every job's bins) over many
// blocks, in two launches.  k_order_count: num = cnt, cnt = 0, each block's
// weight-class histogram and entry total added into OrderScratch; block 0
// also settles the spill list, resets the work-queue head and zeroes
// *clear.  k_order_scatter: every block derives the heavy-first class
// offsets from the histogram, reserves each class's positions for its bins
// with one atomic per class and writes them; the last block to read the
// histogram clears it for the next pass (the count launch clears the
// cursors), so the scratch needs no memset between passes.
constexpr int kWideThreads = 256;
constexpr int kWidePer = 4;  // bins per thread
constexpr int kWideBins = kWideThreads * kWidePer;

__global__ __launch_bounds__(kWideThreads) void k_order_count(AxrBins B, OrderScratch* sc,
                                                              uint32_t* __restrict__ ctr,
                                                              int ctr_spill,
                                                              uint32_t* __restrict__ clear) {
  __shared__ uint32_t hist[kOrderClasses];
  const int tid = threadIdx.x, lane = tid & 63;
  const int nb = B.nb;
  if (tid < kOrderClasses) hist[tid] = 0;
  if (blockIdx.x == 0) {
    if (tid < kOrderClasses) sc->cur[tid] = 0;
    if (tid == 0) {
      if (clear) *clear = 0;  // the large list, read by now
      const uint32_t used = B.spill_ctr[0];
      B.spill_ctr[1] = used < B.spill_cap ? used : B.spill_cap;
      B.spill_ctr[0] = 0;
      if (B.head) *B.head = 0;
      if (used) atomicMax(&ctr[ctr_spill], used);
      if (used > B.spill_cap) atomicAdd(&ctr[AXR_CTR_LOST], used - B.spill_cap);
    }
  }
  __syncthreads();
  uint32_t tot = 0;
  const int b0 = blockIdx.x * kWideBins;
  uint32_t c0[kWidePer], c1[kWidePer];
#pragma unroll
  for (int u = 0; u < kWidePer; u++) {
    const int b = b0 + u * kWideThreads + tid;
    c0[u] = b < nb ? B.cnt[b] : 0u;
    c1[u] = b < nb ? B.cnt[nb + b] : 0u;
  }
#pragma unroll
  for (int u = 0; u < kWidePer; u++) {
    const int b = b0 + u * kWideThreads + tid;
    const bool valid = b < nb;
    uint32_t k = 0;
    if (valid) {
      B.num[b] = c0[u];
      B.cnt[b] = 0;
      B.num[nb + b] = c1[u];
      B.cnt[nb + b] = 0;
      tot += c0[u] + c1[u];
      k = order_class(c0[u] + 16u * c1[u]);
    }
    const uint64_t em = __ballot(valid && k == 0);
    if (em && lane == __ffsll((unsigned long long)em) - 1) atomicAdd(&hist[0], (uint32_t)__popcll(em));
    if (valid && k != 0) atomicAdd(&hist[k], 1u);
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) tot += __shfl_xor(tot, o);
  __syncthreads();
  if (tid < kOrderClasses && hist[tid]) atomicAdd(&sc->hist[tid], hist[tid]);
  if (lane == 0 && tot) atomicAdd(&sc->total, tot);
}

__global__ __launch_bounds__(kWideThreads) void k_order_scatter(AxrBins B, OrderScratch* sc,
                                                                uint32_t* __restrict__ ctr,
                                                                int ctr_total) {
  __shared__ uint32_t offs[kOrderClasses];  // heavy-first offset of each class
  __shared__ uint32_t cnt[kOrderClasses];   // this block's bins per class
  const int tid = threadIdx.x, lane = tid & 63;
  const int nb = B.nb;
  if (tid < kOrderClasses) {
    const uint32_t h = __hip_atomic_load(&sc->hist[tid], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    uint32_t sfx = h;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
      const uint32_t t = __shfl_down(sfx, o);
      if (lane + o < 64) sfx += t;
    }
    offs[tid] = sfx - h;  // classes above this one come first
    cnt[tid] = 0;
  }
  if (tid == 0 && blockIdx.x == 0)
    ctr[ctr_total] = __hip_atomic_load(&sc->total, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  __syncthreads();
  // the histogram is read: the last block to get here clears it
  if (tid == 0) {
    __threadfence();
    if (atomicAdd(&sc->ticket, 1u) == gridDim.x - 1) {
      for (int k = 0; k < kOrderClasses; k++) sc->hist[k] = 0;
      sc->total = 0;
      sc->ticket = 0;
    }
  }
  const int b0 = blockIdx.x * kWideBins;
  uint32_t kk[kWidePer], rank[kWidePer];
#pragma unroll
  for (int u = 0; u < kWidePer; u++) {
    const int b = b0 + u * kWideThreads + tid;
    const bool valid = b < nb;
    const uint32_t k = valid ? order_class(B.num[b] + 16u * B.num[nb + b]) : 0u;
    kk[u] = k;
    // empty bins take their ranks a wave at a time
    const uint64_t em = __ballot(valid && k == 0);
    uint32_t base = 0;
    const int leader = em ? __ffsll((unsigned long long)em) - 1 : 0;
    if (em && lane == leader) base = atomicAdd(&cnt[0], (uint32_t)__popcll(em));
    base = __shfl(base, leader);
    rank[u] = !valid ? 0u : k == 0 ? base + (uint32_t)__popcll(em & lanemask_lt())
                                   : atomicAdd(&cnt[k], 1u);
  }
  __syncthreads();
  if (tid < kOrderClasses && cnt[tid]) offs[tid] += atomicAdd(&sc->cur[tid], cnt[tid]);
  __syncthreads();
#pragma unroll
  for (int u = 0; u < kWidePer; u++) {
    const int b = b0 + u * kWideThreads + tid;
    if (b < nb) B.order[offs[kk[u]] + rank[u]] = (uint32_t)b;
  }
}

// The lanes of a wave in each of four classes: of(k) is the mask of the
// valid lanes whose class is k (k < 4).
struct ClassMasks {
  uint64_t b0, b1, vm;
  __device__ __forceinline__ ClassMasks(bool valid, uint32_t c)
      : b0(__ballot(valid && (c & 1u))), b1(__ballot(valid && (c & 2u))), vm(__ballot(valid)) {}
  __device__ __forceinline__ uint64_t of(uint32_t k) const {
    return ((k & 1u) ? b0 : vm & ~b0) & ((k & 2u) ? b1 : vm & ~b1);
  }
};

// Entries a shadow bin sorts in LDS; a bin with more (or with spilled
// entries) walks them in bucket order.
#ifndef AXR_SHADOW_SORT
#define AXR_SHADOW_SORT 8192
#endif
#ifndef AXR_SHADOW_SORT_MIN
#define AXR_SHADOW_SORT_MIN 64
#endif

// K3: one block per 64x64 texel bin of every pending map, heavy bins first;
// minimum depth keys in LDS, written once.  A bin starts from its map
// (from_map) or from the clear depth.
__global__ __launch_bounds__(AXR_SHADOW_THREADS) void k_shadow_bins(AxrBins B, AxrPassSet ps) {
  __shared__ uint32_t lk[AXR_SHADOW_BIN * AXR_SHADOW_BIN];
  __shared__ uint32_t srt[AXR_SHADOW_SORT];  // the bin's entries by class
  __shared__ uint32_t ccur[4];               // class counts, then cursors
  if (threadIdx.x < 4) ccur[threadIdx.x] = 0;
  const uint32_t slot = B.order[blockIdx.x];
  // the pass owning the slot (slot0 ascends); selected field by field so the
  // by-value pass table is never indexed dynamically
  const AxrShRec* __restrict__ rec = ps.p[0].rec;
  uint32_t* __restrict__ map = ps.p[0].map;
  int dim = ps.p[0].dim, nbx = ps.p[0].nbx, from_map = ps.p[0].from_map;
  uint32_t slot0 = 0;
#pragma unroll
  for (int p = 1; p < AXR_MAX_SHADOWS; p++) {
    if (p < ps.n && slot >= ps.p[p].slot0) {
      rec = ps.p[p].rec;
      map = ps.p[p].map;
      dim = ps.p[p].dim;
      nbx = ps.p[p].nbx;
      from_map = ps.p[p].from_map;
      slot0 = ps.p[p].slot0;
    }
  }
  const int b = (int)(slot - slot0);
  const uint32_t n = B.num[slot];
  if (from_map && n == 0) return;
  const int x0 = (b % nbx) << AXR_SHADOW_SHIFT, y0 = (b / nbx) << AXR_SHADOW_SHIFT;
  const int xe = imin(x0 + AXR_SHADOW_BIN, dim), ye = imin(y0 + AXR_SHADOW_BIN, dim);
  const int tid = (int)threadIdx.x, nthr = (int)blockDim.x, lane = tid & 63;
  // A bin of at most AXR_SHADOW_SORT entries, none spilled, is walked with
  // its casters ordered by work class, largest first (a counting sort into
  // LDS), so that a wave walks casters of like row counts instead of
  // waiting on its longest one.  Its entries are loaded once, up front.
  const bool sorted = n > AXR_SHADOW_SORT_MIN && n <= B.cap0 && n <= AXR_SHADOW_SORT;
  constexpr int kPer = AXR_SHADOW_SORT / AXR_SHADOW_THREADS;
  static_assert(kPer * AXR_SHADOW_THREADS == AXR_SHADOW_SORT, "whole rounds of the block");
  uint32_t v[kPer];
  if (sorted) {
    const uint32_t* __restrict__ ent = B.ent + bin_base(B, slot);
#pragma unroll
    for (int u = 0; u < kPer; u++) {
      const uint32_t e = (uint32_t)(u * AXR_SHADOW_THREADS + tid);
      v[u] = e < n ? ent[e] : 0u;
    }
  }
  const uint32_t clear = ordf(AXR_DEPTH_CLEAR);
  for (int i = tid; i < AXR_SHADOW_BIN * AXR_SHADOW_BIN; i += nthr) {
    const int x = x0 + (i & (AXR_SHADOW_BIN - 1)), y = y0 + (i >> AXR_SHADOW_SHIFT);
    lk[i] = (from_map && x < xe && y < ye) ? map[(size_t)y * dim + x] : clear;
  }
  __syncthreads();  // lk and the class counts start cleared
  auto walk = [&](uint32_t entry) {
    const AxrShRec r = rec[entry & AXR_SH_ENTRY_MASK];
    // texels lie in [x0, xe) x [y0, ye): shadow_walk clips every row to
    // the bin, so the tile index stays inside the LDS tile
    shadow_walk(
        r, dim, x0, y0, xe, ye,
        [&](int ax, int cy, float tz) {
          if (tz != tz) return;
          atomicMin(&lk[((cy - y0) << AXR_SHADOW_SHIFT) + (ax - x0)], ordf(tz == 0.0f ? 0.0f : tz));
        },
        [&](int ax, int cy, uint32_t key) {
          atomicMin(&lk[((cy - y0) << AXR_SHADOW_SHIFT) + (ax - x0)], key);
        });
  };
  if (sorted) {
    // lanes 0..3 of a wave keep, per round, the wave's first position among
    // the bin's casters of class `lane`
    uint32_t wb[kPer];
#pragma unroll
    for (int u = 0; u < kPer; u++) {
      wb[u] = 0;
      if ((uint32_t)(u * AXR_SHADOW_THREADS) >= n) break;
      const uint32_t e = (uint32_t)(u * AXR_SHADOW_THREADS + tid);
      const ClassMasks m(e < n, v[u] >> 30);
      const uint64_t ml = m.of((uint32_t)lane);
      if (lane < 4 && ml) wb[u] = atomicAdd(&ccur[lane], (uint32_t)__popcll(ml));
    }
    __syncthreads();
    const uint32_t n3 = ccur[3], n2 = ccur[2], n1 = ccur[1];
#pragma unroll
    for (int u = 0; u < kPer; u++) {
      if ((uint32_t)(u * AXR_SHADOW_THREADS) >= n) break;
      const uint32_t e = (uint32_t)(u * AXR_SHADOW_THREADS + tid), c = v[u] >> 30;
      const ClassMasks m(e < n, c);
      // classes 3, 2, 1, 0 in turn
      const uint32_t off = c == 3 ? 0u : c == 2 ? n3 : c == 1 ? n3 + n2 : n3 + n2 + n1;
      const uint32_t base = (uint32_t)__shfl((int)wb[u], (int)c);
      if (e < n) srt[off + base + (uint32_t)__popcll(m.of(c) & lanemask_lt())] = v[u];
    }
    __syncthreads();
    for (uint32_t e = (uint32_t)tid; e < n; e += (uint32_t)nthr) walk(srt[e]);
  } else {
    for_bin_entries(B, slot, n, tid, nthr, walk);
  }
  __syncthreads();
  for (int i = threadIdx.x; i < AXR_SHADOW_BIN * AXR_SHADOW_BIN; i += blockDim.x) {
    const int x = x0 + (i & (AXR_SHADOW_BIN - 1)), y = y0 + (i >> AXR_SHADOW_SHIFT);
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
    c += ok != nullptr && ok[i] != 0;
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

hipError_t launch_setup(const float4* vpos, const float4* vattr, const AxrGroupDev* groups,
                        const AxrDrawConst& dc, int do_main, AxrTriRec* rec, uint8_t* cls,
                        const AxrBins& MB, int NBX, const AxrPassSet& ps, const AxrBins& SB,
                        const float4* chunks, uint32_t t0, uint32_t t1, uint32_t* large,
                        uint32_t* large_ctl, hipStream_t s) {
  t0 &= ~63u;  // whole 64-triangle chunks per wave
  if (t1 > dc.T) t1 = dc.T;
  if (t1 <= t0 || (!do_main && ps.n == 0)) return hipSuccess;
  hipLaunchKernelGGL(k_setup, dim3((t1 - t0 + kBlock - 1) / kBlock), dim3(kBlock), 0, s, vpos,
                     vattr, groups, dc, do_main, rec, cls, MB, NBX, ps, SB, chunks, t0, t1,
                     large, large_ctl);
  if (do_main)
    hipLaunchKernelGGL(k_coarse, dim3(kCoarseBlocks), dim3(kCoarseThreads), 0, s, rec, large,
                       large_ctl, MB, dc.W, dc.WC, dc.HC, NBX);
  return hipGetLastError();
}

hipError_t launch_chunk_bounds(const float4* vpos, uint32_t T, float4* chunks, hipStream_t s) {
  if (T == 0) return hipSuccess;
  const uint32_t nch = (T + 63) / 64;
  hipLaunchKernelGGL(k_chunk_bounds, dim3((nch + 3) / 4), dim3(256), 0, s, vpos, T, chunks);
  return hipGetLastError();
}

hipError_t launch_bin_order(const AxrBins* main_bins, const AxrBins* shadow_bins,
                            uint32_t* ctr, uint32_t* clear, uint32_t* heavy, hipStream_t s) {
  OrderJob j[2];
  int n = 0;
  if (main_bins) j[n++] = OrderJob{*main_bins, 2, AXR_CTR_BIN_TOTAL, AXR_CTR_BIN_SPILL, heavy};
  if (shadow_bins) j[n++] = OrderJob{*shadow_bins, 1, AXR_CTR_SH_TOTAL, AXR_CTR_SH_SPILL, nullptr};
  if (n == 0) return hipSuccess;
  if (n == 1) j[1] = j[0];
  hipLaunchKernelGGL(k_bin_order, dim3(n), dim3(1024), 0, s, j[0], j[1], ctr,
                     main_bins ? clear : nullptr);
  return hipGetLastError();
}

hipError_t launch_bin_order_wide(const AxrBins& B, OrderScratch* sc, uint32_t* ctr,
                                 uint32_t* clear, hipStream_t s) {
  if (B.nb <= 0) return launch_bin_order(&B, nullptr, ctr, clear, nullptr, s);
  const unsigned grid = (unsigned)((B.nb + kWideBins - 1) / kWideBins);
  hipLaunchKernelGGL(k_order_count, dim3(grid), dim3(kWideThreads), 0, s, B, sc, ctr,
                     (int)AXR_CTR_BIN_SPILL, clear);
  hipLaunchKernelGGL(k_order_scatter, dim3(grid), dim3(kWideThreads), 0, s, B, sc, ctr,
                     (int)AXR_CTR_BIN_TOTAL);
  return hipGetLastError();
}

hipError_t launch_shadow_bins(const AxrBins& B, const AxrPassSet& ps, hipStream_t s) {
  if (ps.n == 0 || B.nb == 0) return hipSuccess;
  hipLaunchKernelGGL(k_shadow_bins, dim3(B.nb), dim3(AXR_SHADOW_THREADS), 0, s, B, ps);
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
