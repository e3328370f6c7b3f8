// ssr.hip -- the screen-space-reflection water pass (Shaders_src/av_SSR.c,
// expanded as Shaders/drawtexcolsmshplerp_SSR_water.c) over a finished frame.
//
// Water groups are set up and binned like any draw, after the opaque frame
// is shaded.  One block per 32x32 bin: the bin's keys start at the opaque
// depth with the lowest sequence number, so only fragments strictly nearer
// than the frame survive (the shader's F > tz), and the first fragment in
// draw order wins a tie among water fragments.  Each surviving pixel replays
// its winner's walk and runs the shader: Fresnel term, transmittance toward
// the opaque colour, a screen-space ray march of the reflected ray against
// the depth buffer, and the reflection cube map where the march misses or
// fades.  Every read of the frame (own pixel, march, hit colour) is of the
// opaque frame as it was before the pass -- a copy taken first -- so the
// result does not depend on the order in which bins run.
#include "axr_bin.h"

using namespace axr;

namespace {

// av_SSR.c:6-13 (REFL_DEPTH, REFL_VSIZE, REFL_DBIAS, REFL_FADE, maxattn);
// REFL_LENGTH, REFL_STEP and ABSORB come in AxrSsrConst.
constexpr float kReflDepth = 2.0f;
constexpr float kReflVSize = 2.0f;
constexpr float kReflDBias = 0.5f;
constexpr float kReflFade = 80.0f;
constexpr float kMaxAttn = 0.1f;

// Iterations of the ray march whose positions are computed, and depths
// loaded, together (the loads of one group are in flight at once).
#ifndef AXR_SSR_MARCH
#define AXR_SSR_MARCH 4
#endif

// av_SSR.c:37-148 for one fragment; returns the colour, the caller stores
// it and the fragment depth.
__device__ __forceinline__ void ssr_px(const AxrDrawConst& dc, const AxrGroupDev& grp,
                                       const AxrSsrConst& S, int ax, int cy, float t, float tz,
                                       const Attr& c1, const Attr& c2, uint16_t rgb[3]) {
  const int W = dc.W, H = dc.H;
  const size_t p = (size_t)cy * W + ax;
  const float u = 1 - t;
  const V3 pos = v3((u * c2.px + t * c1.px) * tz, (u * c2.py + t * c1.py) * tz,
                    (u * c2.pz + t * c1.pz) * tz);
  const V3 norm = normalize(v3((u * c2.nx + t * c1.nx) * tz, (u * c2.ny + t * c1.ny) * tz,
                               (u * c2.nz + t * c1.nz) * tz));
  const int lenT = grp.tex_size;
  const int tex1 = abs_mask(f2i(((u * c2.u + t * c1.u) * tz) * (float)lenT), lenT - 1);
  const int tex2 = abs_mask(f2i(((float)lenT * (u * c2.v + t * c1.v)) * tz), lenT - 1);
  const ushort4 tx = grp.tex[tex1 + lenT * tex2];
  const V3 vp = v3(dc.view.pos[0], dc.view.pos[1], dc.view.pos[2]);
  float fr = 1.0f;
  if (S.fresnel) {
    fr = 1 - omax(0.0f, dot(normalize(sub(pos, vp)), norm)) * kMaxAttn;
    fr *= fr; fr *= fr; fr *= fr; fr *= fr; fr *= fr; fr *= fr;
  }
  const float scatter = exp_cw(S.absorb * (tz - S.F0[p]));
  const float ts[3] = {scatter * (float)S.R0[p] + (1 - scatter) * (float)tx.x,
                       scatter * (float)S.G0[p] + (1 - scatter) * (float)tx.y,
                       scatter * (float)S.B0[p] + (1 - scatter) * (float)tx.z};

  const V3 a = sub(pos, vp);
  const float nd = dot(a, norm);
  const V3 refl = scl(normalize(sub(a, scl(norm, 2 * nd))), kReflVSize);
  const bool down = refl.y < 0;
  const V3 ar = add(a, refl);
  const float rpz = dot(ar, cam_row(dc.view, 0));
  const float rpx = dot(ar, cam_row(dc.view, 1)) * -dc.view.scale / rpz + (float)(W / 2);
  const float rpy = dot(ar, cam_row(dc.view, 2)) * dc.view.scale / rpz + (float)(H / 2);
  const float rx = (float)ax, ry = (float)cy;
  const float dt = omax(fabsf(rpx - rx), fabsf(rpy - ry));
  const float vx = (float)S.step;
  const float slopex = (rpx - rx) / dt * vx;
  const float slopey = (rpy - ry) / dt * vx;
  const float slopez = (1.f / rpz - 1.f / tz) / dt * vx;
  // The march, kMarch iterations at a time: their positions first (the
  // shader's exact update sequence), then their depth loads together, then
  // the hit tests in order.  Iteration i runs while i * step < steps, the
  // position is on screen and sz > 0; a hit at iteration i ends the march
  // with rn = (i + 1) * step, else rn = (first iteration not run) * step.
  constexpr int kMarch = AXR_SSR_MARCH;
  float sx = rx, sy = ry, sz = 1.f / tz, sd = kReflDepth;
  float iz = sz > 0 ? 1.f / sz : 0.f;  // 1/sz of the current iteration
  int rn = 0, hit = 0, hx = 0, hy = 0;
  for (int it = 0, live = 1; live && !hit; it += kMarch) {
    int px_[kMarch];
    float iz_[kMarch], sd_[kMarch];
    bool on[kMarch];
#pragma unroll
    for (int j = 0; j < kMarch; j++) {
      on[j] = live && (it + j) * S.step < S.steps && sx >= 0 && sx < (float)W && sy >= 0 &&
              sy < (float)H && sz > 0;
      if (!on[j] && live) {
        live = 0;
        rn = (it + j) * S.step;
      }
      px_[j] = on[j] ? W * f2i(sy) + f2i(sx) : 0;
      iz_[j] = iz;
      sd_[j] = sd;
      if (on[j]) {
        sx += slopex;
        sy += slopey;
        const float sz0 = sz, iz0 = iz;
        sz += slopez;
        iz = rcp_any(sz);
        // 1/(sz - slopez): usually sz - slopez gives back sz0 exactly, whose
        // reciprocal is at hand (the same division, the same bits)
        const float back = sz - slopez;
        sd = fabsf(iz - (back == sz0 ? iz0 : rcp_any(back))) + kReflDBias;
      }
    }
    float f_[kMarch];
#pragma unroll
    for (int j = 0; j < kMarch; j++) f_[j] = on[j] ? S.F0[px_[j]] : 0.f;
#pragma unroll
    for (int j = 0; j < kMarch; j++) {
      if (!hit && on[j] && f_[j] < iz_[j] && f_[j] > iz_[j] - sd_[j]) {
        hit = 1;
        hx = px_[j] % W;
        hy = px_[j] / W;
        rn = (it + j + 1) * S.step;
      }
    }
  }

  // cube map face and texel (nearest), as the shader picks them
  const float m = omax(fabsf(refl.x), omax(fabsf(refl.y), fabsf(refl.z)));
  int face = 0;
  float qx = rx, qy = ry;
  if (m == fabsf(refl.x)) {
    face = refl.x > 0 ? 0 : 3;
    qx = refl.z / refl.x;
    qy = refl.y / refl.x;
  }
  if (m == fabsf(refl.y)) {
    face = refl.y > 0 ? 1 : 4;
    qx = refl.x / refl.y;
    qy = refl.z / refl.y;
  }
  if (m == fabsf(refl.z)) {
    face = refl.z > 0 ? 2 : 5;
    qx = refl.y / refl.z;
    qy = refl.x / refl.z;
  }
  const int dimR = grp.refl_dim;
  qx = omin((qx + 1) * (float)dimR, (float)(dimR * 2 - 1));
  qy = omin((qy + 1) * (float)dimR, (float)(dimR * 2 - 1));
  const int lenR = 2 * dimR;
  const ushort4 cube = grp.refl[(f2i(qx) + face * lenR) + 6 * lenR * f2i(qy)];
  const float cubec[3] = {(float)cube.x, (float)cube.y, (float)cube.z};
  const float fade = omin(1.f, (float)(S.steps - rn) / kReflFade);
  if (down) fr = 0.f;
  if (hit) {
    const size_t h = (size_t)W * hy + hx;
    const float so[3] = {(float)S.R0[h], (float)S.G0[h], (float)S.B0[h]};
#pragma unroll
    for (int c = 0; c < 3; c++)
      rgb[c] = f2u16((1 - fr) * ts[c] + fr * (fade * so[c] + (1 - fade) * cubec[c]));
  } else {
#pragma unroll
    for (int c = 0; c < 3; c++) rgb[c] = f2u16((1 - fr) * ts[c] + fr * cubec[c]);
  }
}

// Threads of an SSR block (one bin each): fewer than the opaque raster's,
// so more bins' marches are in flight.
#ifndef AXR_SSR_THREADS
#define AXR_SSR_THREADS 256
#endif
static_assert(AXR_SSR_THREADS % 64 == 0 && AXR_SSR_THREADS <= 512, "whole waves, span tables");

__global__ __launch_bounds__(AXR_SSR_THREADS) void k_ssr_raster(
    const AxrTriRec* __restrict__ rec, AxrBins B, const float4* __restrict__ vpos,
    const float4* __restrict__ vattr, const AxrGroupDev* __restrict__ groups, AxrDrawConst dc,
    AxrSsrConst S, int NBX, float* __restrict__ depth, uint16_t* __restrict__ ro,
    uint16_t* __restrict__ go, uint16_t* __restrict__ bo) {
  constexpr int NPX = AXR_BIN * AXR_BIN;
  __shared__ unsigned long long lk[NPX];
  __shared__ uint16_t list[NPX];
  __shared__ uint32_t ncov, bin_far;
  const int b = (int)B.order[blockIdx.x];
  const int bx = b % NBX, by = b / NBX;
  const int x0 = bx << AXR_BIN_SHIFT, y0 = by << AXR_BIN_SHIFT;
  const int xe = imin(x0 + AXR_BIN, dc.W), ye = imin(y0 + AXR_BIN, dc.H);
  const int tid = threadIdx.x, nthr = blockDim.x;
  if (B.num[b] == 0 && B.num[B.nb + b] == 0) return;  // no water here
  float zfar = 0.0f;  // farthest opaque depth of the bin (depths are positive)
  for (int i = tid; i < NPX; i += nthr) {
    const int x = x0 + (i & (AXR_BIN - 1)), y = y0 + (i >> AXR_BIN_SHIFT);
    const bool in = x < xe && y < ye;
    const float f = in ? S.F0[(size_t)y * dc.W + x] : 0.0f;
    lk[i] = in ? make_key_first(f, 0u) : 0ull;
    zfar = omax(zfar, f);
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) zfar = omax(zfar, __shfl_xor(zfar, o));
  if (tid == 0) {
    ncov = 0;
    bin_far = 0u;
  }
  __syncthreads();
  if ((tid & 63) == 0) atomicMax(&bin_far, __float_as_uint(zfar));  // non-negative: bits order
  __syncthreads();
  walk_bin<AXR_BIN_SHIFT, true, true>(lk, rec, B, (uint32_t)b, (uint32_t)(B.nb + b), bx, by, dc, tid,
                                nthr, __uint_as_float(bin_far));
  __syncthreads();
  for (int i = tid; i < NPX; i += nthr) {
    const int x = x0 + (i & (AXR_BIN - 1)), y = y0 + (i >> AXR_BIN_SHIFT);
    if (x < xe && y < ye && (uint32_t)lk[i] != 0u) list[atomicAdd(&ncov, 1u)] = (uint16_t)i;
  }
  __syncthreads();
  const uint32_t n = ncov;
  for (uint32_t j = tid; j < n; j += nthr) {
    const int i = list[j];
    const int x = x0 + (i & (AXR_BIN - 1)), y = y0 + (i >> AXR_BIN_SHIFT);
    FragState f;
    replay_frag((uint32_t)lk[i], x, y, rec, vpos, vattr, groups, dc, f);
    uint16_t rgb[3];
    ssr_px(dc, f.grp, S, x, y, f.t, f.tz, f.c1, f.c2, rgb);
    const size_t p = (size_t)y * dc.W + x;
    ro[p] = rgb[0];
    go[p] = rgb[1];
    bo[p] = rgb[2];
    depth[p] = f.tz;
  }
}

// The opaque frame's snapshot for the pass in one launch: thread i < n8
// copies pixels 8i..8i+7, depth as two 16-byte vectors and each colour plane
// as one; threads n8..n8+tail-1 copy the last np % 8 pixels one each.  The
// snapshot's planes start 16-byte aligned (a stride rounded up to 8 pixels).
__global__ void k_snap_frame(const float* __restrict__ depth, const uint16_t* __restrict__ r,
                             const uint16_t* __restrict__ g, const uint16_t* __restrict__ b,
                             size_t n8, size_t tail, float* __restrict__ F0,
                             uint16_t* __restrict__ R0, uint16_t* __restrict__ G0,
                             uint16_t* __restrict__ B0) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n8) {
    const float4* d4 = (const float4*)depth;
    const float4 d0 = d4[2 * i], d1 = d4[2 * i + 1];
    const uint4 rr = ((const uint4*)r)[i], gg = ((const uint4*)g)[i], bb = ((const uint4*)b)[i];
    ((float4*)F0)[2 * i] = d0;
    ((float4*)F0)[2 * i + 1] = d1;
    ((uint4*)R0)[i] = rr;
    ((uint4*)G0)[i] = gg;
    ((uint4*)B0)[i] = bb;
  } else if (i < n8 + tail) {
    const size_t p = 8 * n8 + (i - n8);
    F0[p] = depth[p];
    R0[p] = r[p];
    G0[p] = g[p];
    B0[p] = b[p];
  }
}

// Six faces' HDR planes [6][S][S] into the strip the shader samples:
// texel (x, y) of face f at column f * S + x of row y, as (R,G,B,0).
__global__ void k_pack_cube(const uint16_t* __restrict__ r, const uint16_t* __restrict__ g,
                            const uint16_t* __restrict__ b, int S, ushort4* __restrict__ strip) {
  const size_t n = (size_t)6 * S * S;
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int x = (int)(i % (size_t)(6 * S)), y = (int)(i / (size_t)(6 * S));
  const int f = x / S;
  const size_t s = ((size_t)f * S + y) * S + (x - f * S);
  strip[i] = make_ushort4(r[s], g[s], b[s], 0);
}

}  // namespace

namespace axr {

hipError_t launch_ssr(const AxrTriRec* rec, const AxrBins& B, const float4* vpos,
                      const float4* vattr, const AxrGroupDev* groups, const AxrDrawConst& dc,
                      const AxrSsrConst& S, int NBX, float* depth, uint16_t* r, uint16_t* g,
                      uint16_t* b, hipStream_t s) {
  if (B.nb == 0) return hipSuccess;
  hipLaunchKernelGGL(k_ssr_raster, dim3((unsigned)B.nb), dim3(AXR_SSR_THREADS), 0, s, rec, B,
                     vpos, vattr, groups, dc, S, NBX, depth, r, g, b);
  return hipGetLastError();
}

hipError_t launch_snap_frame(const float* depth, const uint16_t* r, const uint16_t* g,
                             const uint16_t* b, size_t np, float* F0, uint16_t* R0,
                             uint16_t* G0, uint16_t* B0, hipStream_t s) {
  const size_t n8 = np / 8, tail = np - 8 * n8;
  if (np == 0) return hipSuccess;
  hipLaunchKernelGGL(k_snap_frame, dim3((unsigned)((n8 + tail + 255) / 256)), dim3(256), 0, s,
                     depth, r, g, b, n8, tail, F0, R0, G0, B0);
  return hipGetLastError();
}

hipError_t launch_pack_cube(const uint16_t* r, const uint16_t* g, const uint16_t* b, int S,
                            ushort4* strip, hipStream_t s) {
  const size_t n = (size_t)6 * S * S;
  hipLaunchKernelGGL(k_pack_cube, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, r, g, b, S,
                     strip);
  return hipGetLastError();
}

}  // namespace axr
