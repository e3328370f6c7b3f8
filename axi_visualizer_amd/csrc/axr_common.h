// axr_common.h -- shared types and device arithmetic of the gfx950 raster path.
//
// Every float operation here is one IEEE binary32 operation (the library is
// compiled with -ffp-contract=off and HIP's default correctly rounded
// division and sqrt), so results are bit-identical to the serial restatement
// of the reference's OpenCL kernels.  Decrees where OpenCL C is loose:
// dot(a,b) = (a.x*b.x + a.y*b.y) + a.z*b.z, min(x,y) = y<x?y:x,
// max(x,y) = x<y?y:x, float->int truncates and saturates (NaN -> 0),
// float->ushort saturates, normalize(v) = v * (1/sqrt(v.v)),
// pown by binary exponentiation.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define AXR_SH_MIP 1u
#define AXR_SH_SHADOW 2u
#define AXR_SH_PHONG 4u
#define AXR_SH_CULL 8u
#define AXR_GRP_DRAW 0x100u  // group takes part in the current drawAll

#define AXR_MAX_GROUPS 128

#define AXR_HD __host__ __device__ __forceinline__

#define AXR_MAX_SHADOWS 4
#define AXR_SHADOW_BIND 2  // cascades sampled by one drawAll
#define AXR_NONE 0xFFFFFFFFu

// Camera: rows of vm are (forward, horizontal, vertical), as VIEWMAT
// (Compute.py:313-314); scale/cax/cay as setScaleCull (Ops.py:231-233).
struct AxrCam {
  float pos[3];
  float vm[9];
  float scale, cax, cay;
};

// Constants of the chunk culling tests (chunk_outside_view/ortho), derived
// from a camera once per draw on the host: the view cone's side planes
// widened by 1e-4 -- R - kx*F, -R - kx*F, U - ky*F, -U - ky*F with kx =
// 1.0001*caX, ky = 1.0001*caY (rows F, R, U of vm) -- and their lengths;
// |F|; |R| and |U| times |scale| (the ortho test's pixel radii); the L1 norm
// of the eye.  ok = 0, as a zero-initialised block has, culls nothing.
struct AxrCull {
  float n[4][3];
  float nn[4];
  float nf, nrs, nus, l1eye;
  int ok;
};
inline AxrCull make_cull(const AxrCam& c) {
  auto row = [&](int r, float (&v)[3]) {
    for (int k = 0; k < 3; k++) v[k] = c.vm[3 * r + k];
  };
  auto len = [](const float (&v)[3]) { return sqrtf((v[0] * v[0] + v[1] * v[1]) + v[2] * v[2]); };
  float F[3], R[3], U[3];
  row(0, F);
  row(1, R);
  row(2, U);
  const float kx = c.cax * 1.0001f, ky = c.cay * 1.0001f;
  AxrCull k{};
  for (int a = 0; a < 3; a++) {
    k.n[0][a] = R[a] - F[a] * kx;
    k.n[1][a] = R[a] * -1.0f - F[a] * kx;
    k.n[2][a] = U[a] - F[a] * ky;
    k.n[3][a] = U[a] * -1.0f - F[a] * ky;
  }
  for (int i = 0; i < 4; i++) k.nn[i] = len(k.n[i]);
  const float s = fabsf(c.scale);
  k.nf = len(F);
  k.nrs = len(R) * s;
  k.nus = len(U) * s;
  k.l1eye = fabsf(c.pos[0]) + fabsf(c.pos[1]) + fabsf(c.pos[2]);
  k.ok = 1;
  return k;
}

// Per-group device table; re-uploaded only when a group or its flags change.
struct AxrGroupDev {
  uint32_t base;   // first global triangle index
  uint32_t ntri;
  uint32_t flags;  // AXR_SH_* | AXR_GRP_DRAW of the current drawAll
  int tex_size;    // N, or log2(N) for a mip chain (texSize, Ops.py:371-374)
  float mip_bias;
  int spec_pow;
  const ushort4* tex;  // texels (R,G,B,0) u16; mip chain smallest first
  const ushort4* refl;  // SSR groups: reflection cube strip (R,G,B,0), 2*refl_dim x 12*refl_dim
  int refl_dim;
};

struct AxrShadowDev {
  const uint32_t* map;  // orderable depth keys, dim*dim
  int dim;
  int wS;  // dim/2
  AxrCam cam;
};

// Per-draw constants, passed by value to the raster and shade kernels.
struct AxrDrawConst {
  AxrCam view;
  AxrCull cull;  // make_cull(view)
  float lint[3];
  float ldir[3];
  float amb;
  int W, H, WC, HC;
  uint32_t T;        // total triangles in the context (order space)
  uint32_t ngroups;
  int nshadow;
  int pcf_taps;
  AxrShadowDev sh[AXR_SHADOW_BIND];
  uint16_t clear_rgb[4];
  uint32_t flag_mask;  // shader flags honoured (a batch job may drop shadows)
};

// Setup output of one visible triangle (indexed by its global index).
struct AxrTriRec {
  float px[3], py[3], z[3];
};

// Ortho setup output of one shadow caster (int coords).
struct AxrShRec {
  int x[3], y[3];
  float z[3];
};

// Groups casting into one shadow pass (Ops.py:1094-1095 whichCast).
struct AxrCastMask {
  uint32_t bits[AXR_MAX_GROUPS / 32];
};

// One shadow pass (shadowMap(i), Ops.py:1089-1141) as the kernels run it.
// Passes are deferred and run together: one setup launch (sharing the
// position reads with the drawAll setup when there is one), one order
// launch and one depth launch over the bins of every pending map.
struct AxrShadowPass {
  AxrCam cam;
  AxrCull cull;  // make_cull(cam)
  int dim, nbx;
  float bias;
  uint32_t slot0;  // first bin slot of this map in the shadow bin store
  int from_map;    // 0: bins start from the clear depth (map cleared before)
  AxrCastMask cast;
  AxrShRec* rec;
  uint8_t* ok;
  uint32_t* map;
};
struct AxrPassSet {
  int n;
  AxrShadowPass p[AXR_MAX_SHADOWS];
};
// Bit g of the mask without dynamic indexing of a by-value argument (which
// would spill the argument to scratch).
AXR_HD bool cast_bit(const AxrCastMask& m, int g) {
  const uint32_t w = g < 32 ? m.bits[0] : g < 64 ? m.bits[1] : g < 96 ? m.bits[2] : m.bits[3];
  return (w >> (g & 31)) & 1u;
}

// Screens are binned into 32 x 32 pixel bins (two by two of the
// reference's 16-px tiles), shadow maps into 64 x 64 texel bins (their
// casters are cheaper to walk than to append to many bins).  A bin's
// triangles are walked by one block with the bin's depth keys in LDS.
#define AXR_BIN 32
#define AXR_BIN_SHIFT 5
#ifndef AXR_SHADOW_SHIFT
#define AXR_SHADOW_SHIFT 6
#endif
#define AXR_SHADOW_BIN (1 << AXR_SHADOW_SHIFT)

// Threads of one bin's block in the per-bin depth walks (screen and shadow).
#ifndef AXR_RASTER_THREADS
#define AXR_RASTER_THREADS 512
#endif
#ifndef AXR_SHADOW_THREADS
#define AXR_SHADOW_THREADS 1024
#endif

// Device counter block.
#define AXR_CTR_BIN_TOTAL 0  // bin entries of the last drawAll
#define AXR_CTR_BIN_SPILL 1  // most screen-bin entries spilled by one pass since the last check
#define AXR_CTR_SH_TOTAL 2   // bin entries of the last shadow pass
#define AXR_CTR_SH_SPILL 3   // the same as BIN_SPILL for shadow bins
#define AXR_CTR_STATS 4      // 3 words filled on request: small, large, casters
#define AXR_CTR_LOST 7       // entries dropped since the last check (spill list full)
#define AXR_CTR_WORDS 8

// Bin storage of one pass kind (screen or shadow).  A bin slot s is
// seg * nb + bin; segment 0 holds small triangles (or casters), segment 1
// the large triangles' 16-px tiles.  Setup appends straight into fixed
// capacity buckets (cap[seg] entries per slot); entries past a bucket's
// capacity go to the spill list as (slot, entry) pairs, which the raster
// reads only for slots whose count exceeds the capacity.
struct AxrBins {
  uint32_t* cnt;        // [nseg*nb] append cursors (zeroed by the order pass)
  uint32_t* num;        // [nseg*nb] entries of the current pass
  uint32_t* ent;        // bucket storage, segment 0 then segment 1
  uint32_t* order;      // [nb] heavy-first launch order
  uint2* spill;         // [spill_cap] (slot, entry)
  uint32_t* spill_ctr;  // [2]: append cursor, entries of the current pass
  uint32_t* head;       // work-queue head of a persistent raster (or null)
  uint32_t cap0, cap1, spill_cap;
  int nb;
};

// One job of a batch -- a resident scene seen by one camera at one
// resolution -- as the batched kernels read it from the device job table.
struct AxrJobDev {
  AxrDrawConst dc;  // camera, lights, size, bound shadow maps, flag mask
  const float4* vpos;
  const float4* vattr;
  const AxrGroupDev* groups;
  const float4* chunks;  // bounding sphere per 64-triangle chunk of the scene
  AxrTriRec* rec;  // [dc.T] setup records of this job
  float* depth;    // [W*H] outputs
  uint32_t* rgba;
  uint32_t* ids;
  uint16_t *r, *g, *b;  // HDR planes (batches rendered for HDR)
  float exposure;
  uint32_t slot0;  // first screen-bin slot of this job
  uint32_t blk0;   // first setup block of this job's scene group
  int nbx;         // bins per row
};

AXR_HD size_t bin_base(const AxrBins& B, uint32_t s) {
  return s < (uint32_t)B.nb ? (size_t)s * B.cap0
                            : (size_t)B.nb * B.cap0 + (size_t)(s - (uint32_t)B.nb) * B.cap1;
}

// Entry v at append position pos of slot s.
__device__ __forceinline__ void bin_store(const AxrBins& B, uint32_t s, uint32_t pos, uint32_t v) {
  const uint32_t cap = s < (uint32_t)B.nb ? B.cap0 : B.cap1;
  if (pos < cap) {
    B.ent[bin_base(B, s) + pos] = v;
  } else {
    const uint32_t q = atomicAdd(&B.spill_ctr[0], 1u);
    if (q < B.spill_cap) B.spill[q] = make_uint2(s, v);
  }
}

// The position of thread tid within a round of nthr entries (nthr a
// multiple of the wave size) when the entries are dealt to the waves in
// turn: a bin with fewer entries than threads then spreads them over every
// wave (each walking a few) instead of filling the first waves while the
// others wait at the next barrier.
__device__ __forceinline__ uint32_t dealt_slot(int tid, int nthr) {
  const int nw = nthr >> 6;
  return (uint32_t)((tid & 63) * nw + (tid >> 6));
}

// f(entry) for the n entries of slot s, strided over the block's threads:
// the bucket, then -- only when it overflowed -- the slot's spilled entries.
template <class F>
__device__ __forceinline__ void for_bin_entries(const AxrBins& B, uint32_t s, uint32_t n, int tid,
                                                int nthr, F&& f) {
  const uint32_t cap = s < (uint32_t)B.nb ? B.cap0 : B.cap1;
  const uint32_t* __restrict__ ent = B.ent + bin_base(B, s);
  const uint32_t nb = n < cap ? n : cap;
  const uint32_t k = (uint32_t)tid;
  for (uint32_t e = k; e < nb; e += (uint32_t)nthr) f(ent[e]);
  if (n > cap) {
    const uint32_t ns = B.spill_ctr[1];
    for (uint32_t q = k; q < ns; q += (uint32_t)nthr) {
      const uint2 p = B.spill[q];
      if (p.x == s) f(p.y);
    }
  }
}

// f(entry, valid) for the n entries of slot s in rounds of the block's
// threads, every lane of every wave calling f in every round (valid marks
// lanes holding an entry), so f may use wave-wide operations.  DEAL: the
// entries of a round go to the waves in turn (dealt_slot), else each wave
// takes 64 consecutive ones.
template <bool DEAL = false, class F>
__device__ __forceinline__ void for_bin_entries_waves(const AxrBins& B, uint32_t s, uint32_t n,
                                                      int tid, int nthr, F&& f) {
  const uint32_t cap = s < (uint32_t)B.nb ? B.cap0 : B.cap1;
  const uint32_t* __restrict__ ent = B.ent + bin_base(B, s);
  const uint32_t nb = n < cap ? n : cap;
  const uint32_t k = DEAL ? dealt_slot(tid, nthr) : (uint32_t)tid;
  for (uint32_t e0 = 0; e0 < nb; e0 += (uint32_t)nthr) {
    const uint32_t e = e0 + k;
    f(e < nb ? ent[e] : 0u, e < nb);
  }
  if (n > cap) {
    const uint32_t ns = B.spill_ctr[1];
    for (uint32_t q0 = 0; q0 < ns; q0 += (uint32_t)nthr) {
      const uint32_t q = q0 + k;
      bool ok = false;
      uint32_t v = 0;
      if (q < ns) {
        const uint2 p = B.spill[q];
        ok = p.x == s;
        v = p.y;
      }
      f(v, ok);
    }
  }
}

// -------------------------------------------------------------- helpers

AXR_HD float omin(float x, float y) { return y < x ? y : x; }
AXR_HD float omax(float x, float y) { return x < y ? y : x; }
AXR_HD int imin(int x, int y) { return y < x ? y : x; }
AXR_HD unsigned umin(unsigned a, unsigned b) { return a < b ? a : b; }
AXR_HD int imax(int x, int y) { return x < y ? y : x; }
AXR_HD float oclamp(float x, float lo, float hi) { return omin(omax(x, lo), hi); }
// ---------------------------------------------- correctly rounded division
// rcp_rn: RN(1/b) from the hardware's approximate reciprocal and one Newton
// step, exact for every b with |b| in [2^-60, 2^60] (axr_check_arith
// compares it with IEEE division for each such b).  div_rn: from y = RN(1/b)
// and q = RN(a*y), which lies within one ulp of a/b, Markstein's step
// q + (a - b*q)*y rounds to RN(a/b) when no intermediate under- or
// overflows (P. Markstein, IBM J. Res. Dev. 34(1):111-119, 1990).  The
// remainder a - b*q is exact for |a|, |b| in [2^-30, 2^30] and for b an
// integer in [1, 2^32] with a an integer-valued float; a = +0 gives the
// IEEE signed zero, a = -0 needs div_rn0.
__device__ __forceinline__ float rcp_rn(float b) {
  const float y0 = __builtin_amdgcn_rcpf(b);
  return __builtin_fmaf(__builtin_fmaf(-b, y0, 1.0f), y0, y0);
}
__device__ __forceinline__ float div_rn(float a, float b, float y) {
  const float q = a * y;
  return __builtin_fmaf(__builtin_fmaf(-b, q, a), y, q);
}
// IEEE 1/b for any b: rcp_rn where it is exact, IEEE division (kept inside
// the branch by the empty asm) for |b| outside [2^-60, 2^60], inf and NaN.
__device__ __forceinline__ float rcp_any(float b) {
  const float a = fabsf(b);
  if ((a >= 0x1p-60f) & (a <= 0x1p60f)) return rcp_rn(b);
  asm volatile("" : "+v"(b));
  return 1.0f / b;
}
// div_rn for a numerator that may be -0 (the correction step would turn
// -0/b into +0; a*y carries the quotient's signed zero).
__device__ __forceinline__ float div_rn0(float a, float b, float y) {
  return a == 0.0f ? a * y : div_rn(a, b, y);
}
// Numerators that divide exactly by an integer b in [1, 2^32] through
// div_rn0: zero, or finite with |a| >= 2^-94, so the quotient stays normal
// and the remainder a - b*q, a multiple of ulp(q), is representable.
__device__ __forceinline__ bool rn_num(float a) {
  const float m = fabsf(a);
  return (a == 0.0f) | ((m >= 0x1p-94f) & (m <= 0x1.fffffep127f));
}

// Saturating conversions: NaN -> 0, out-of-range -> the nearest limit,
// else truncation toward zero.  On the device these are the hardware's
// v_cvt_i32_f32 / v_cvt_u32_f32, which saturate exactly so (axr_check_arith
// compares them with the formulas below for every float); one instruction
// and one register where the formulas take branches or selects.
AXR_HD int f2i_ref(float x) {
  if (x != x) return 0;
  if (x >= 2147483648.0f) return 2147483647;
  if (x < -2147483648.0f) return (-2147483647 - 1);
  return (int)x;
}
AXR_HD uint16_t f2u16_ref(float x) {
  if (!(x > 0.0f)) return 0;
  if (x >= 65535.0f) return 65535;
  return (uint16_t)x;
}
AXR_HD int f2i(float x) {
#ifdef __HIP_DEVICE_COMPILE__
  int r;
  asm("v_cvt_i32_f32 %0, %1" : "=v"(r) : "v"(x));
  return r;
#else
  return f2i_ref(x);
#endif
}
AXR_HD uint16_t f2u16(float x) {
#ifdef __HIP_DEVICE_COMPILE__
  uint32_t r;
  asm("v_cvt_u32_f32 %0, %1" : "=v"(r) : "v"(x));
  return (uint16_t)(r < 65535u ? r : 65535u);
#else
  return f2u16_ref(x);
#endif
}
AXR_HD int abs_mask(int v, int m) {
  unsigned u = v < 0 ? 0u - (unsigned)v : (unsigned)v;
  return (int)(u & (unsigned)m);
}

// Orderable 32-bit key of a float: unsigned order == float order (non-NaN).
AXR_HD uint32_t ordf(float f) {
  uint32_t u;
  __builtin_memcpy(&u, &f, 4);
  return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
AXR_HD float unordf(uint32_t k) {
  uint32_t u = (k & 0x80000000u) ? (k & 0x7FFFFFFFu) : ~k;
  float f;
  __builtin_memcpy(&f, &u, 4);
  return f;
}

// 64-bit depth-test key: (orderable depth << 32) | ~seq.  The minimum key
// is the minimum depth and, among equal depths, the largest seq, i.e. the
// last fragment in pass order -- exactly the serial `F >= tz` semantics.
// seq = 2*order + half + 1, order = (large ? T : 0) + global triangle index.
#define AXR_DEPTH_CLEAR 65535.0f
AXR_HD uint64_t key_clear() { return ((uint64_t)ordf(AXR_DEPTH_CLEAR) << 32) | 0xFFFFFFFFull; }
AXR_HD uint64_t make_key(float tz, uint32_t seq) {
  float d = (tz == 0.0f) ? 0.0f : tz;
  return ((uint64_t)ordf(d) << 32) | (uint64_t)(~seq);
}
// Key of a strict depth test (the first fragment wins a tie).
AXR_HD uint64_t make_key_first(float tz, uint32_t seq) {
  float d = (tz == 0.0f) ? 0.0f : tz;
  return ((uint64_t)ordf(d) << 32) | seq;
}
AXR_HD uint32_t frag_seq(uint32_t order, int half) { return 2u * order + (uint32_t)half + 1u; }

// exp(x): Cody-Waite reduction x = k ln2 + r (ln 2 split into a 16-bit head
// and a tail, so k * head is exact) and the degree-7 Taylor polynomial of
// e^r in Horner form, scaled by 2^k.  0 for x <= -80 (and NaN), +inf above
// 88.  The same operation sequence on host and device gives the same bits.
AXR_HD float exp_cw(float x) {
  if (!(x > -80.0f)) return 0.0f;
  if (x > 88.0f) return __builtin_inff();
  const float k = rintf(x * 1.44269504f);
  const float r = (x - k * 0.693145751953125f) - k * 1.428606765330187e-6f;
  float p = 0.0001984127f;
  p = p * r + 0.0013888889f;
  p = p * r + 0.008333334f;
  p = p * r + 0.041666668f;
  p = p * r + 0.16666667f;
  p = p * r + 0.5f;
  p = p * r + 1.0f;
  p = p * r + 1.0f;
  return ldexpf(p, (int)k);
}

// log2(x): x = m 2^e with m in [sqrt(1/2), sqrt(2)); log2 x = e + ln(m) / ln 2
// with ln m = 2 atanh(s), s = (m - 1) / (m + 1), by its odd series through
// s^9 in Horner form.  Exact at powers of two; -inf at 0, NaN below 0 or at
// NaN, +inf at +inf.  The same operation sequence on host and device gives
// the same bits (the vendor log2 of host and device libraries differ).
AXR_HD float log2_cw(float x) {
  if (x != x || x < 0.0f) return __builtin_nanf("");
  if (x == 0.0f) return -__builtin_inff();
  if (x == __builtin_inff()) return x;
  int e = 0;
  float m = frexpf(x, &e);
  if (m < 0.70710678f) {
    m = m * 2.0f;
    e -= 1;
  }
  const float s = (m - 1.0f) / (m + 1.0f);
  const float s2 = s * s;
  float q = 0.11111111f;
  q = q * s2 + 0.14285715f;
  q = q * s2 + 0.2f;
  q = q * s2 + 0.33333334f;
  q = q * s2 + 1.0f;
  return (float)e + ((2.0f * s) * q) * 1.44269504f;
}
