// axr_bin.h -- bin appends and the per-bin depth walk + shading, shared by
// the single-frame kernels and the batched (many-job) kernels.
#pragma once
#include "axr_pixel.h"

namespace axr {

__device__ __forceinline__ uint64_t lanemask_lt() {
  const unsigned lane = __lane_id();
  return lane ? (~0ull >> (64 - lane)) : 0ull;
}

// wave_bin_add in two steps, so that several appends' atomics can be in
// flight at once: wave_bin_add_issue groups the lanes and sends one atomic
// per group; finish() waits for it and gives each active lane its position.
struct WaveAdd {
  uint32_t base;   // the atomic's result, on the first lane of each group
  int leader_of;   // first lane of this lane's group
  uint32_t rank;   // rank in the group
  bool active;
  __device__ __forceinline__ uint32_t finish() const {
    const uint32_t b = __shfl(base, leader_of);
    return active ? b + rank : 0u;
  }
};
__device__ __forceinline__ WaveAdd wave_bin_add_issue(uint32_t* __restrict__ ctr, bool active,
                                                      uint32_t bin) {
  uint64_t pending = __ballot(active);
  WaveAdd w{0u, 0, 0u, active};
  uint32_t cnt = 0;  // the group's size on its first lane
  while (pending) {
    const int leader = __ffsll((unsigned long long)pending) - 1;
    const uint32_t b = __shfl(bin, leader);
    const uint64_t grp = __ballot(active && bin == b) & pending;
    if ((grp >> __lane_id()) & 1ull) {
      w.leader_of = leader;
      w.rank = (uint32_t)__popcll(grp & lanemask_lt());
    }
    if ((int)__lane_id() == leader) cnt = (uint32_t)__popcll(grp);
    pending &= ~grp;
  }
  if (cnt) w.base = atomicAdd(&ctr[bin], cnt);
  return w;
}
// Adds one per active lane to ctr[bin] with one atomic per distinct bin in
// the wave; returns the counter's previous value plus the lane's rank among
// the lanes that share its bin.  All lanes of the wave must call it.  The
// ballot loop only groups the lanes (no memory access); the groups' first
// lanes then issue their atomics in one instruction, so the wave waits for
// one round trip however many bins it touches.
__device__ __forceinline__ uint32_t wave_bin_add(uint32_t* __restrict__ ctr, bool active,
                                                 uint32_t bin) {
  return wave_bin_add_issue(ctr, active, bin).finish();
}

// Bin rectangle (bins of 1 << shift pixels) of a pixel rectangle clipped to
// [0,w) x [0,h).
struct BinRect {
  int bx0, by0, nbx, n;
};
__device__ __forceinline__ BinRect bin_rect(int x0, int y0, int x1, int y1, int w, int h,
                                            int shift) {
  BinRect r = {0, 0, 1, 0};
  x0 = imax(x0, 0);
  y0 = imax(y0, 0);
  x1 = imin(x1, w - 1);
  y1 = imin(y1, h - 1);
  if (x0 <= x1 && y0 <= y1) {
    r.bx0 = x0 >> shift;
    r.by0 = y0 >> shift;
    r.nbx = (x1 >> shift) - r.bx0 + 1;
    r.n = r.nbx * ((y1 >> shift) - r.by0 + 1);
  }
  return r;
}
// The bins of a BinRect row by row, column kx of row ky, stepped without
// dividing: bin k of the rectangle is (k / nbx, k % nbx).
struct BinCursor {
  int kx = 0, ky = 0;
  __device__ __forceinline__ uint32_t bin(const BinRect& r, int nbx_total) const {
    return (uint32_t)((r.by0 + ky) * nbx_total + r.bx0 + kx);
  }
  __device__ __forceinline__ void next(const BinRect& r) {
    if (++kx == r.nbx) {
      kx = 0;
      ky++;
    }
  }
};

// Appends a large triangle's reference 16-px tile t to the bins of
// 1 << SHIFT pixels that hold it: a 32-px bin gets (gi << 2) | the tile's
// quadrant; a 16-px bin is the tile; each 8-px bin of the tile gets gi << 2.
template <int SHIFT>
__device__ __forceinline__ void append_tile(const AxrBins& B, uint32_t slot_base, int tx, int ty,
                                            int NBX, uint32_t gi) {
  if (SHIFT == 5) {
    const uint32_t slot = slot_base + (uint32_t)((ty >> 1) * NBX + (tx >> 1));
    const uint32_t sub = (uint32_t)(((ty & 1) << 1) | (tx & 1));
    bin_store(B, slot, atomicAdd(&B.cnt[slot], 1u), (gi << 2) | sub);
  } else if (SHIFT == 4) {
    const uint32_t slot = slot_base + (uint32_t)(ty * NBX + tx);
    bin_store(B, slot, atomicAdd(&B.cnt[slot], 1u), gi << 2);
  } else {
    for (int q = 0; q < 4; q++) {
      const uint32_t slot =
          slot_base + (uint32_t)(((ty << 1) + (q >> 1)) * NBX + (tx << 1) + (q & 1));
      bin_store(B, slot, atomicAdd(&B.cnt[slot], 1u), gi << 2);
    }
  }
}

// The walked 16-px tile (walked_tile) of a large-triangle entry v of bin
// (bx, by), as append_tile stored it.
template <int SHIFT>
__device__ __forceinline__ TileBox entry_box(uint32_t v, int bx, int by) {
  if (SHIFT == 5) {
    const uint32_t sub = v & 3u;
    return tile_box_xy((bx << 1) + (int)(sub & 1u), (by << 1) + (int)(sub >> 1));
  }
  if (SHIFT == 4) return tile_box_xy(bx, by);
  return tile_box_xy(bx >> 1, by >> 1);
}

// ---------------------------------------------- wave-cooperative span walk
// A row span of the small walk: the row's edge and depth positions (as the
// serial walk holds them at that row), the fragment sequence number and the
// pixel range [xa, xa + len) already clipped to the bin.
struct SpanRec {
  float cx1, cx2, cz1, cz2;
  uint32_t seq;
  int cy, xa;
  float y;  // RN(1/(cx1 - cx2)) for a small_fast triangle's span, else 0
  int fast;  // the span's pixels take the fast exact path (small_fast)
};
// Triangles whose box inside the bin has more pixels than this walk
// spread over the wave; smaller ones walk on their own lane (48, 64 and 96
// measured: 64 and 96 take 16 us off the six-face cube pass, 1080p unchanged).
#ifndef AXR_COOP_AREA
#define AXR_COOP_AREA 64
#endif
struct WaveSpans {
  SpanRec r[64];
  uint32_t end[64];  // inclusive prefix sums of the spans' lengths
};

// One step of a cooperative walk: every lane offers one span (len may be
// 0); the wave walks all their pixels 64 at a time -- prefix sum of the
// lengths, binary search for each pixel's span -- calling frag(ax, cy, tz,
// seq) with the span's own arithmetic (span_t, span_tz), or, for a span of
// a small_fast triangle, frag_key(ax, cy, ordf(tz), seq) with the walk's
// fast exact arithmetic (small_walk's row: div_rn/rcp_rn, a zero-width row
// clamped to the exact 0 or 1).  All lanes of the wave must call it.
template <class F, class K>
__device__ __forceinline__ void coop_spans(WaveSpans& ws, int len, const SpanRec& sr, F&& frag,
                                           K&& frag_key) {
  const int lane = (int)__lane_id();
  uint32_t incl = (uint32_t)len;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const uint32_t v = __shfl_up(incl, (unsigned)d);
    if (lane >= d) incl += v;
  }
  const uint32_t total = __shfl(incl, 63);
  if (total == 0) return;
  ws.end[lane] = incl;
  if (len) ws.r[lane] = sr;
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  for (uint32_t base = 0; base < total; base += 64) {
    const uint32_t i = base + (uint32_t)lane;
    if (i < total) {
      int lo = 0, hi = 63;
      while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (ws.end[mid] > i)
          hi = mid;
        else
          lo = mid + 1;
      }
      const uint32_t start = lo ? ws.end[lo - 1] : 0u;
      const SpanRec r = ws.r[lo];
      const int ax = r.xa + (int)(i - start);
      if (r.fast) {
        const float d = r.cx1 - r.cx2, u = (float)ax - r.cx2;
        const float sv = u < 0.0f ? 0.0f : 1.0f;
        const bool flat = d == 0.0f;
        const float t = __builtin_amdgcn_fmed3f(div_rn(u, d, r.y), flat ? sv : 0.0f,
                                                flat ? sv : 1.0f);
        const uint32_t flip = r.cz1 > 0.0f ? 0x80000000u : 0xffffffffu;
        frag_key(ax, r.cy, __float_as_uint(rcp_rn((1 - t) * r.cz2 + t * r.cz1)) ^ flip, r.seq);
      } else {
        const float t = span_t(ax, r.cx1, r.cx2);
        frag(ax, r.cy, span_tz(t, r.cz1, r.cz2), r.seq);
      }
    }
  }
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// The small walk (a7) of one triangle per lane, spread over the wave: each
// step every lane still walking produces its next row span inside the
// rectangle [xlo, xhi) x [ylo, yhi) -- accumulating its edges row by row
// exactly as small_walk does, skipped rows included -- and the wave then
// takes the pixels of all the step's spans 64 at a time (prefix sum of the
// lengths, binary search for each pixel's span).  frag(ax, cy, tz, seq)
// sees the same fragments as small_walk; only their order differs.  All
// lanes of the wave must call it; `valid` marks lanes with a triangle.
template <class F, class K>
__device__ __forceinline__ void coop_small_walk(WaveSpans& ws, bool valid, const Geom<int>& g,
                                                uint32_t order, int xlo, int ylo, int xhi,
                                                int yhi, F&& frag, K&& frag_key) {
  int half = valid ? 0 : 2;
  const int fast = valid && small_fast(g) ? 1 : 0;
  Edges e;
  float cx1 = 0.f, cx2 = 0.f, cz1 = 0.f, cz2 = 0.f;
  int cy = 0;
  if (valid) {
    e = small_edges(g, 0);
    cx1 = cx2 = (float)g.x1;
    cz1 = cz2 = g.z1;
    cy = g.y1;
  }
  while (__ballot(half < 2)) {
    int len = 0;
    SpanRec sr;
    if (half == 0) {
      while (cy <= g.y2 && cy < yhi && cy < ylo) {  // rows above the rectangle
        cx1 += e.s1;
        cx2 += e.s2;
        cz1 += e.sz1;
        cz2 += e.sz2;
        cy++;
      }
      if (cy <= g.y2 && cy < yhi) {
        const int xa = imax(f2i(cx2), xlo);
        if (cx1 >= (float)xa) len = imin(xhi - 1, f2i(floorf(cx1))) - xa + 1;
        len = imax(len, 0);
        sr = SpanRec{cx1, cx2, cz1, cz2, frag_seq(order, 0), cy, xa,
                     fast ? rcp_rn(cx1 - cx2) : 0.0f, fast};
        cx1 += e.s1;
        cx2 += e.s2;
        cz1 += e.sz1;
        cz2 += e.sz2;
        cy++;
      } else {
        half = 1;
        e = small_edges(g, 1);
        cx1 = cx2 = (float)g.x3;
        cz1 = cz2 = g.z3;
        cy = g.y3;
      }
    } else if (half == 1) {
      while (cy >= g.y2 && cy >= ylo && cy >= yhi) {  // rows below the rectangle
        cx1 -= e.s1;
        cx2 -= e.s2;
        cz1 -= e.sz1;
        cz2 -= e.sz2;
        cy--;
      }
      if (cy >= g.y2 && cy >= ylo) {
        const int xa = imax(f2i(cx1), xlo);
        if (cx2 >= (float)xa) len = imin(xhi - 1, f2i(floorf(cx2))) - xa + 1;
        len = imax(len, 0);
        sr = SpanRec{cx1, cx2, cz1, cz2, frag_seq(order, 1), cy, xa,
                     fast ? rcp_rn(cx1 - cx2) : 0.0f, fast};
        cx1 -= e.s1;
        cx2 -= e.s2;
        cz1 -= e.sz1;
        cz2 -= e.sz2;
        cy--;
      } else {
        half = 2;
      }
    }
    coop_spans(ws, len, sr, frag, frag_key);
  }
}

// The tiled walk (a9) of one tile entry per lane, spread over the wave the
// same way: each lane emits its tile's row spans (both halves, each row's
// edges by the tile walk's own arithmetic), the wave walks their pixels.
// frag(ax, cy, tz, seq) sees tile_walk's fragments.  All lanes must call it.
template <class F>
__device__ __forceinline__ void coop_tile_walk(WaveSpans& ws, bool valid, const Geom<float>& g,
                                               const TileBox& b, uint32_t order, F&& frag) {
  int half = 2;
  Edges e;
  float cx1 = 0.f, cx2 = 0.f, cz1 = 0.f, cz2 = 0.f, yEnd = 0.f;
  int cy = 0;
  auto start = [&](int h) {
    e = tile_edges(g, h);
    const float yv = h == 0 ? g.y1 : g.y3;
    cy = f2i(oclamp(yv, b.yMin, b.yMax));
    const float dy = (float)cy - yv;
    const float xv = h == 0 ? g.x1 : g.x3, zv = h == 0 ? g.z1 : g.z3;
    cx1 = xv + dy * e.s1;
    cx2 = xv + dy * e.s2;
    cz1 = zv + dy * e.sz1;
    cz2 = zv + dy * e.sz2;
    yEnd = oclamp(g.y2, b.yMin, b.yMax);
  };
  if (valid) {
    if (g.y2 > b.yMin) {
      start(0);
      half = 0;
    } else if (g.y2 < b.yMax) {
      start(1);
      half = 1;
    }
  }
  while (__ballot(half < 2)) {
    int len = 0;
    SpanRec sr;
    if (half == 0) {
      if ((float)cy <= yEnd) {
        if (cx2 < b.xMax && cx1 > b.xMin) {
          const float xe = oclamp(cx1, b.xMin, b.xMax);
          const int xa = f2i(oclamp(cx2, b.xMin, b.xMax));
          if ((float)xa <= xe) len = f2i(floorf(xe)) - xa + 1;
          sr = SpanRec{cx1, cx2, cz1, cz2, frag_seq(order, 0), cy, xa, 0.0f, 0};
        }
        cx1 += e.s1;
        cx2 += e.s2;
        cz1 += e.sz1;
        cz2 += e.sz2;
        cy++;
      } else if (g.y2 < b.yMax) {
        start(1);
        half = 1;
      } else {
        half = 2;
      }
    } else if (half == 1) {
      if ((float)cy >= yEnd) {
        if (cx1 < b.xMax && cx2 > b.xMin) {
          const float xe = oclamp(cx2, b.xMin, b.xMax);
          const int xa = f2i(oclamp(cx1, b.xMin, b.xMax));
          if ((float)xa <= xe) len = f2i(floorf(xe)) - xa + 1;
          sr = SpanRec{cx1, cx2, cz1, cz2, frag_seq(order, 1), cy, xa, 0.0f, 0};
        }
        cx1 -= e.s1;
        cx2 -= e.s2;
        cz1 -= e.sz1;
        cz2 -= e.sz2;
        cy--;
      } else {
        half = 2;
      }
    }
    coop_spans(ws, len, sr, frag, frag);
  }
}

// Strict-test passes (FIRST) only: true when no fragment of the triangle
// can be nearer than zfar, the farthest depth already in the bin.  A
// fragment's depth interpolates the vertices' (within a few ulps of their
// range), so a nearest vertex depth beyond zfar by 0.1 % rules it out.
__device__ __forceinline__ bool behind_bin(const float z[3], float zfar) {
  const float zmax = omax(z[0], omax(z[1], z[2]));  // z = 1/depth
  return zmax > 0.0f && (1.0f / zmax) * 0.999f > zfar;
}

// Every fragment of the bin's entries takes an LDS min of its depth key:
// small triangles of slot s_small walk drawSmall's scanlines clipped to the
// bin (a7), large-triangle tile entries of slot s_large walk the tiled pass
// on their reference tile (a9).  FIRST: the key's low word is seq, so the
// first fragment in draw order wins a depth tie (a strict test); otherwise
// ~seq, the last one wins (the reference's F >= tz).
// quad >= 0 (16-px bins only): the slots are those of the 32-px bin
// holding this one as its quadrant quad, and only the large-triangle tile
// entries of that quadrant are walked; small triangles clip to the bin.
template <int SHIFT, bool FIRST, bool DEAL = false>
__device__ __forceinline__ void walk_bin(unsigned long long* lk, const AxrTriRec* __restrict__ rec,
                                         const AxrBins& B, uint32_t s_small, uint32_t s_large,
                                         int bx, int by, const AxrDrawConst& dc, int tid,
                                         int nthr, float zfar = 0.0f, int quad = -1) {
  constexpr int BIN = 1 << SHIFT, NPX = BIN * BIN;
  const int x0 = bx << SHIFT, y0 = by << SHIFT;
  const int xe = imin(x0 + BIN, dc.W), ye = imin(y0 + BIN, dc.H);
  auto key = [](float tz, uint32_t seq) -> unsigned long long {
    return FIRST ? make_key_first(tz, seq) : make_key(tz, seq);
  };
  // one span table per wave of the block: blocks of the bin kernels have
  // min(bin pixels, 512) threads
  constexpr int kWaves = (NPX < 512 ? NPX : 512) / 64;
  static_assert(AXR_RASTER_THREADS <= 512, "span tables are sized for 512-thread blocks");
  __shared__ WaveSpans spans[kWaves];
  const int wv = (int)(threadIdx.x >> 6);  // the block's wave (threads split into bins keep theirs)
  WaveSpans& ws = spans[wv < kWaves ? wv : kWaves - 1];
  auto put = [&](int ax, int cy, float tz, uint32_t seq) {
    if (!(tz <= AXR_DEPTH_CLEAR)) return;
    const unsigned lc = (unsigned)(((cy - y0) << SHIFT) + (ax - x0));
    if (lc < (unsigned)NPX) atomicMin(&lk[lc], key(tz, seq));
  };
  // A fragment given as its ordered depth ord = ordf(tz) (tz finite and
  // nonzero): no test against the clear depth is needed, since a key above
  // the bin's initial one (the clear depth's, or for the strict test the
  // opaque depth's with a zero tie word) leaves the minimum unchanged.  The
  // small walk keeps its pixels inside the bin; the index clamp only keeps a
  // stray one inside the LDS tile.
  auto put_key = [&](int ax, int cy, uint32_t ord, uint32_t seq) {
    const unsigned lc = (unsigned)(((cy - y0) << SHIFT) + (ax - x0));
    atomicMin(&lk[umin(lc, (unsigned)NPX - 1u)],
              ((unsigned long long)ord << 32) | (FIRST ? seq : ~seq));
  };
  // A triangle whose box inside the bin is small walks on its lane; larger
  // ones walk spread over the wave.
  for_bin_entries_waves(B, s_small, B.num[s_small], tid, nthr, [&](uint32_t gi, bool valid) {
    Geom<int> g{};
    bool big = false;
    AxrTriRec r;
    if (valid) {
      r = rec[gi];
      if (FIRST && behind_bin(r.z, zfar)) valid = false;  // cannot pass the strict test
    }
    if (valid) {
      g = small_geom(r.px, r.py, r.z);
      const int bw = imin(imax(g.x1, imax(g.x2, g.x3)), xe - 1) - imax(imin(g.x1, imin(g.x2, g.x3)), x0);
      const int bh = imin(g.y3, ye - 1) - imax(g.y1, y0);
      big = (bw + 1) * (bh + 1) > AXR_COOP_AREA;
      if (!big)
        small_walk(
            g, x0, y0, xe, ye,
            [&](int ax, int cy, float tz, int half) { put(ax, cy, tz, frag_seq(gi, half)); },
            [&](int ax, int cy, uint32_t ord, int half) {
              put_key(ax, cy, ord, frag_seq(gi, half));
            });
    }
    coop_small_walk(ws, valid && big, g, gi, x0, y0, xe, ye, put, put_key);
  });
  auto put_tile = [&](int ax, int cy, float tz, uint32_t seq) {
    if (!(tz <= AXR_DEPTH_CLEAR)) return;
    // a tile is larger than an 8-px bin: keep the bin's part of it
    if (SHIFT < 4 && ((unsigned)(ax - x0) >= (unsigned)BIN || (unsigned)(cy - y0) >= (unsigned)BIN))
      return;
    const unsigned lc = (unsigned)(((cy - y0) << SHIFT) + (ax - x0));
    if (lc < (unsigned)NPX) atomicMin(&lk[lc], key(tz, seq));
  };
  // a bin holds few large-triangle tiles, each walked by the whole wave
  // (coop_tile_walk): DEAL spreads them over the waves so they walk in
  // parallel -- shorter bins when the device has idle waves (a single
  // frame), more instructions when it has none (the persistent batch)
  for_bin_entries_waves<DEAL>(B, s_large, B.num[s_large], tid, nthr, [&](uint32_t v, bool valid) {
    Geom<float> g{};
    TileBox box{};
    uint32_t order = 0;
    bool big = false;
    AxrTriRec r;
    if (quad >= 0 && (int)(v & 3u) != quad) valid = false;  // another quadrant's tile
    if (valid) {
      r = rec[v >> 2];
      if (FIRST && behind_bin(r.z, zfar)) valid = false;  // cannot pass the strict test
    }
    if (valid) {
      const uint32_t gi = v >> 2;
      box = entry_box<SHIFT>(v, bx, by);
      g = tile_geom(r.px, r.py, r.z);
      order = dc.T + gi;
      // the triangle's box inside the tile
      const float bw = omin(omax(g.x1, omax(g.x2, g.x3)), box.xMax) -
                       omax(omin(g.x1, omin(g.x2, g.x3)), box.xMin);
      const float bh = omin(g.y3, box.yMax) - omax(g.y1, box.yMin);
      big = (bw + 1.0f) * (bh + 1.0f) > (float)AXR_COOP_AREA;
      if (!big)
        tile_walk(g, box, [&](int ax, int cy, float tz, int half) {
          put_tile(ax, cy, tz, frag_seq(order, half));
        });
    }
    coop_tile_walk(ws, valid && big, g, box, order, put_tile);
  });
}

// Depth walk and shading of one bin of (1 << SHIFT)^2 pixels by one block.
// Small triangles of slot s_small walk drawSmall's scanlines clipped to the
// bin (a7); large-triangle tile entries of slot s_large walk the tiled pass
// on their reference tile (a9); each fragment takes an LDS min of
// (orderable tz << 32) | ~seq.  Covered pixels are then compacted and
// shaded once from their winning key; empty ones get the clear colour.
// `empty` draws only the background.  lk, list, ncov are the bin's LDS;
// quad >= 0: a 16-px quadrant of a 32-px bin's slots (walk_bin);
// threads tid of nthr work on it (a block may hold two bins, each with half
// of its threads: every group passes the same barriers, an inactive group
// with no work).
template <int SHIFT, int MODE, bool DEAL = false>
__device__ __forceinline__ void raster_bin(
    unsigned long long* lk, uint16_t* list, uint32_t* ncov, const AxrTriRec* __restrict__ rec,
    const AxrBins& B, uint32_t s_small, uint32_t s_large, int empty, int bx, int by,
    const float4* __restrict__ vpos, const float4* __restrict__ vattr,
    const AxrGroupDev* __restrict__ groups, const AxrDrawConst& dc, float exposure,
    float* __restrict__ depth, uint16_t* __restrict__ ro, uint16_t* __restrict__ go,
    uint16_t* __restrict__ bo, uint32_t* __restrict__ rgba, uint32_t* __restrict__ ids,
    int tid, int nthr, bool active = true, int quad = -1) {
  constexpr int BIN = 1 << SHIFT, NPX = BIN * BIN;
  if (!active) {  // a thread group without a bin: the same barriers, no work
    __syncthreads();
    __syncthreads();
    if (MODE != SHADE_IDS) __syncthreads();
    return;
  }
  const int x0 = bx << SHIFT, y0 = by << SHIFT;
  const int xe = imin(x0 + BIN, dc.W), ye = imin(y0 + BIN, dc.H);
  constexpr bool ACC = MODE == SHADE_ACCUM;
  for (int i = tid; i < NPX; i += nthr) {
    uint64_t k0 = key_clear();
    if (ACC) {  // the frame's depth with the lowest order: any later fragment wins a tie
      const int x = x0 + (i & (BIN - 1)), y = y0 + (i >> SHIFT);
      if (x < xe && y < ye) k0 = ((uint64_t)ordf(depth[(size_t)y * dc.W + x]) << 32) | 0xFFFFFFFFull;
    }
    lk[i] = k0;
  }
  if (tid == 0) *ncov = 0;
  __syncthreads();
  if (!empty)
    walk_bin<SHIFT, false, DEAL>(lk, rec, B, s_small, s_large, bx, by, dc, tid, nthr, 0.0f, quad);
  __syncthreads();
  // nthr is a multiple of the wave size, so every lane of a wave runs the
  // same rounds and the covered pixels are appended with one LDS atomic a
  // wave (rank by the ballot below the lane)
  for (int i0 = 0; i0 < NPX; i0 += nthr) {
    const int i = i0 + tid;
    const int x = x0 + (i & (BIN - 1)), y = y0 + (i >> SHIFT);
    const bool here = i < NPX && x < xe && y < ye;
    const size_t p = (size_t)y * dc.W + x;
    const uint64_t key = here ? lk[i] : 0ull;
    const uint32_t low = (uint32_t)key;
    bool covered = false;
    if (!here) {
    } else if (MODE == SHADE_IDS) {
      uint32_t id = AXR_NONE;
      if (low != 0xFFFFFFFFu) {
        const uint32_t order = ((~low) - 1u) >> 1;
        id = order >= dc.T ? order - dc.T : order;
      }
      ids[p] = id;
      if (depth) depth[p] = unordf((uint32_t)(key >> 32));  // null: ids alone
    } else if (low == 0xFFFFFFFFu) {
      if (!ACC) {  // composing: the pixel keeps what the frame holds
        uint16_t rgb[3] = {dc.clear_rgb[0], dc.clear_rgb[1], dc.clear_rgb[2]};
        put_px<MODE>(p, AXR_DEPTH_CLEAR, rgb, exposure, depth, ro, go, bo, rgba);
      }
    } else {
      covered = true;
    }
    if (MODE != SHADE_IDS) {
      const uint64_t m = __ballot(covered);
      if (m != 0ull) {  // uniform over the wave
        const int first = __ffsll((unsigned long long)m) - 1;
        uint32_t base = 0;
        if ((int)__lane_id() == first) base = atomicAdd(ncov, (uint32_t)__popcll(m));
        base = __shfl(base, first);
        if (covered) list[base + (uint32_t)__popcll(m & lanemask_lt())] = (uint16_t)i;
      }
    }
  }
  if (MODE == SHADE_IDS) return;
  __syncthreads();
  const uint32_t n = *ncov;
  for (uint32_t j = tid; j < n; j += nthr) {
    const int i = list[j];
    const int x = x0 + (i & (BIN - 1)), y = y0 + (i >> SHIFT);
    const uint64_t key = lk[i];
    uint16_t rgb[3];
    resolve_pixel(key, x, y, rec, vpos, vattr, groups, dc, rgb);
    const size_t p = (size_t)y * dc.W + x;
    put_px<MODE>(p, unordf((uint32_t)(key >> 32)), rgb, exposure, depth, ro, go, bo, rgba);
    if (ACC) {
      const uint32_t order = ((~(uint32_t)key) - 1u) >> 1;
      ids[p] = order >= dc.T ? order - dc.T : order;
    }
  }
}

}  // namespace axr
