// post.hip -- post passes on the HDR frame: the reference's blur
// (CLDraw.blur, Ops.py:1009-1020) and the multi-level fp16 bloom of the
// full-frame pipeline.
//
// blur: four launches, one per phase of Shaders/Post/blur.c and blur2.c run
// serially (the kernels' global barriers do not order work-groups):
// brightness downsample, 19-tap horizontal, 19-tap vertical, bilinear
// upsample added to the frame.  Decrees where the reference text is
// undefined: literals are float, stores saturate toward zero, rows the
// kernels never write are zero, vertical taps past the half-res rows are
// skipped.
//
// bloom: level 1 = max(avg2x2(frame / white) - threshold, 0); level l+1 =
// avg2x2(level l); going up, up_l = level_l + bilinear(up_{l+1}); the frame
// gains white * strength * bilinear(up_1).  Levels are fp16 (round to
// nearest even), arithmetic fp32.
#include <hip/hip_fp16.h>

#include "axr_launch.h"
#include "axr_tri.h"

using namespace axr;

namespace {

__device__ __forceinline__ float tap_w(int ao) {
  return ao == 0 ? 0.10125f : ao == 1 ? 0.09815f : ao == 2 ? 0.08941f : ao == 3 ? 0.07654f
       : ao == 4 ? 0.06157f : ao == 5 ? 0.04654f : ao == 6 ? 0.03306f : ao == 7 ? 0.02207f
       : ao == 8 ? 0.01384f : 0.00816f;
}

// blur.c:23-36: brightness of the (as indexed) 2-px sample of row 2cy.
__global__ void k_blur_down(const uint16_t* __restrict__ R, const uint16_t* __restrict__ G,
                            const uint16_t* __restrict__ B, int wF, int hF,
                            uint16_t* __restrict__ D) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= wF * hF) return;
  const int cx = i % wF, cy = i / wF;
  const size_t n = (size_t)wF * hF;
  if (cy >= hF - 2) {
    D[i] = 0;
    D[n + i] = 0;
    D[2 * n + i] = 0;
    return;
  }
  const int y1 = imin(wF * (hF - 1) * 4, wF * cy * 4 + 1);
  const int x1 = imin(wF * 2 - 1, cx * 2 + 1);
  const int a = wF * cy * 4 + cx * 2, b = y1 + cx * 2, c = wF * cy * 4 + x1, d = y1 + x1;
  const float cr = (float)(R[a] / 4 + R[b] / 4 + R[c] / 4 + R[d] / 4);
  const float cg = (float)(G[a] / 4 + G[b] / 4 + G[c] / 4 + G[d] / 4);
  const float cb = (float)(B[a] / 4 + B[b] / 4 + B[c] / 4 + B[d] / 4);
  const float lum = (0.3626f * cr + 0.5152f * cg) + 0.1222f * cb;
  D[i] = f2u16(cr * lum / 256 / 256);
  D[n + i] = f2u16(cg * lum / 256 / 256);
  D[2 * n + i] = f2u16(cb * lum / 256 / 256);
}

// blur.c:40-64 (horizontal) and blur2.c:21-60 (vertical): 19 taps, /0.9.
template <bool VERT>
__global__ void k_blur_taps(const uint16_t* __restrict__ S, int wF, int hF,
                            uint16_t* __restrict__ O) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= wF * hF) return;
  const int cx = i % wF, cy = i / wF;
  const size_t n = (size_t)wF * hF;
  if (cy >= hF - 1) {
    O[i] = 0;
    O[n + i] = 0;
    O[2 * n + i] = 0;
    return;
  }
#pragma unroll
  for (int k = 0; k < 3; k++) {
    const uint16_t* P = S + k * n;
    float a = 0.0f;
#pragma unroll
    for (int o = -9; o <= 9; o++) {
      const int ao = o < 0 ? -o : o;
      if (!VERT) {
        if (o < 0 && cx <= -o - 1) continue;
        if (o > 0 && cx >= wF - o) continue;
        a = a + tap_w(ao) * (float)P[(size_t)wF * cy + cx + o];
      } else {
        if (o < 0 && cy <= -o - 1) continue;
        if (o > 0 && (cy >= wF - o || cy + o >= hF)) continue;
        a = a + tap_w(ao) * (float)P[(size_t)wF * (cy + o) + cx];
      }
    }
    a /= 0.9f;
    O[k * n + i] = f2u16(a);
  }
}

// blur2.c:62-88: bilinear upsample of U added to the frame.
__global__ void k_blur_up(uint16_t* __restrict__ R, uint16_t* __restrict__ G,
                          uint16_t* __restrict__ B, int W, int H, const uint16_t* __restrict__ U) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= W * H) return;
  const int cx = i % W, cy = i / W;
  const int lenT = W / 2;
  const size_t n = (size_t)(W / 2) * (H / 2);
  float texr1 = (float)cx / 2.f;
  const int tex1 = (int)texr1;
  texr1 -= (float)tex1;
  float texr2 = (float)cy / 2.f;
  const int tex2 = (int)texr2;
  texr2 -= (float)tex2;
  const int t00 = tex1 + lenT * tex2;
  const int t10 = imin(tex1 + 1, lenT - 1) + lenT * tex2;
  const int t01 = tex1 + lenT * imin(tex2 + 1, H / 2 - 2);
  const int t11 = imin(tex1 + 1, lenT - 1) + lenT * imin(tex2 + 1, H / 2 - 2);
  const float texi1 = 1 - texr1, texi2 = 1 - texr2;
  uint16_t* P[3] = {R, G, B};
#pragma unroll
  for (int k = 0; k < 3; k++) {
    const uint16_t* S = U + k * n;
    const float up = ((texi1 * texi2 * (float)S[t00] + texr1 * texi2 * (float)S[t10]) +
                      texi1 * texr2 * (float)S[t01]) +
                     texr1 * texr2 * (float)S[t11];
    P[k][i] = f2u16((float)P[k][i] + up);
  }
}

// ----------------------------------------------------------- bloom
__device__ __forceinline__ float hf(const __half* p, size_t i) { return __half2float(p[i]); }

// bilinear sample of a (sw x sh) level at the position of texel (x, y) of
// the level twice its size: ((x + 0.5) / 2 - 0.5, (y + 0.5) / 2 - 0.5).
__device__ __forceinline__ float bilerp_up(const __half* __restrict__ s, int sw, int sh, int x,
                                           int y) {
  float fx = ((float)x + 0.5f) / 2.0f - 0.5f, fy = ((float)y + 0.5f) / 2.0f - 0.5f;
  fx = omax(fx, 0.0f);
  fy = omax(fy, 0.0f);
  const int x0 = imin((int)fx, sw - 1), y0 = imin((int)fy, sh - 1);
  const float ax = fx - (float)x0, ay = fy - (float)y0;
  const int x1 = imin(x0 + 1, sw - 1), y1 = imin(y0 + 1, sh - 1);
  const float bx = 1.0f - ax, by = 1.0f - ay;
  return ((by * bx * hf(s, (size_t)y0 * sw + x0) + by * ax * hf(s, (size_t)y0 * sw + x1)) +
          ay * bx * hf(s, (size_t)y1 * sw + x0)) +
         ay * ax * hf(s, (size_t)y1 * sw + x1);
}

__global__ void k_bloom_prefilter(const uint16_t* __restrict__ R, const uint16_t* __restrict__ G,
                                  const uint16_t* __restrict__ B, int W, int H, float white,
                                  float thr, __half* __restrict__ L1) {
  const int w = W / 2, h = H / 2;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= w * h) return;
  const int x = i % w, y = i / w;
  const size_t n = (size_t)w * h;
  const uint16_t* P[3] = {R, G, B};
#pragma unroll
  for (int k = 0; k < 3; k++) {
    const uint16_t* S = P[k];
    const float a00 = (float)S[(size_t)(2 * y) * W + 2 * x] / white;
    const float a01 = (float)S[(size_t)(2 * y) * W + 2 * x + 1] / white;
    const float a10 = (float)S[(size_t)(2 * y + 1) * W + 2 * x] / white;
    const float a11 = (float)S[(size_t)(2 * y + 1) * W + 2 * x + 1] / white;
    const float avg = (((a00 + a01) + a10) + a11) * 0.25f;
    L1[k * n + i] = __float2half_rn(omax(avg - thr, 0.0f));
  }
}

__global__ void k_bloom_down(const __half* __restrict__ S, int sw, int sh,
                             __half* __restrict__ D) {
  const int w = sw / 2, h = sh / 2;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= w * h) return;
  const int x = i % w, y = i / w;
  const size_t sn = (size_t)sw * sh, n = (size_t)w * h;
#pragma unroll
  for (int k = 0; k < 3; k++) {
    const __half* s = S + k * sn;
    const float avg = (((hf(s, (size_t)(2 * y) * sw + 2 * x) + hf(s, (size_t)(2 * y) * sw + 2 * x + 1)) +
                        hf(s, (size_t)(2 * y + 1) * sw + 2 * x)) +
                       hf(s, (size_t)(2 * y + 1) * sw + 2 * x + 1)) *
                      0.25f;
    D[k * n + i] = __float2half_rn(avg);
  }
}

// up = level + bilinear(coarser up), at the level's size w x h.
__global__ void k_bloom_up(const __half* __restrict__ lv, const __half* __restrict__ up2, int w,
                           int h, __half* __restrict__ up) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= w * h) return;
  const int x = i % w, y = i / w;
  const size_t n = (size_t)w * h, n2 = (size_t)(w / 2) * (h / 2);
#pragma unroll
  for (int k = 0; k < 3; k++)
    up[k * n + i] = __float2half_rn(hf(lv + k * n, i) + bilerp_up(up2 + k * n2, w / 2, h / 2, x, y));
}

__global__ void k_bloom_composite(uint16_t* __restrict__ R, uint16_t* __restrict__ G,
                                  uint16_t* __restrict__ B, int W, int H,
                                  const __half* __restrict__ up1, float white, float strength) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= W * H) return;
  const int x = i % W, y = i / W;
  const size_t n1 = (size_t)(W / 2) * (H / 2);
  const float ws = white * strength;
  uint16_t* P[3] = {R, G, B};
#pragma unroll
  for (int k = 0; k < 3; k++) {
    const float add = bilerp_up(up1 + k * n1, W / 2, H / 2, x, y) * ws;
    P[k][i] = f2u16((float)P[k][i] + add);
  }
}

inline unsigned blocks_for(size_t n) { return (unsigned)((n + 255) / 256); }

}  // namespace

namespace axr {

hipError_t launch_blur(uint16_t* R, uint16_t* G, uint16_t* B, int W, int H, uint16_t* tmp,
                       hipStream_t s) {
  const int wF = W / 2, hF = H / 2;
  const size_t n = (size_t)wF * hF;
  uint16_t *D = tmp, *T = tmp + 3 * n;
  hipLaunchKernelGGL(k_blur_down, dim3(blocks_for(n)), dim3(256), 0, s, R, G, B, wF, hF, D);
  hipLaunchKernelGGL(k_blur_taps<false>, dim3(blocks_for(n)), dim3(256), 0, s, D, wF, hF, T);
  hipLaunchKernelGGL(k_blur_taps<true>, dim3(blocks_for(n)), dim3(256), 0, s, T, wF, hF, D);
  hipLaunchKernelGGL(k_blur_up, dim3(blocks_for((size_t)W * H)), dim3(256), 0, s, R, G, B, W, H, D);
  return hipGetLastError();
}

size_t bloom_halves(int W, int H, int levels) {
  size_t n = 0;
  for (int l = 1; l <= levels; l++) n += 3 * (size_t)(W >> l) * (H >> l);
  return 2 * n;  // levels and the upward chain
}

hipError_t launch_bloom(uint16_t* R, uint16_t* G, uint16_t* B, int W, int H, int levels,
                        float threshold, float white, float strength, void* tmp, hipStream_t s) {
  __half* base = (__half*)tmp;
  __half* lv[16];
  __half* up[16];
  size_t off = 0;
  for (int l = 1; l <= levels; l++) {
    lv[l] = base + off;
    off += 3 * (size_t)(W >> l) * (H >> l);
  }
  for (int l = 1; l <= levels; l++) {
    up[l] = base + off;
    off += 3 * (size_t)(W >> l) * (H >> l);
  }
  hipLaunchKernelGGL(k_bloom_prefilter, dim3(blocks_for((size_t)(W / 2) * (H / 2))), dim3(256), 0,
                     s, R, G, B, W, H, white, threshold, lv[1]);
  for (int l = 1; l < levels; l++)
    hipLaunchKernelGGL(k_bloom_down, dim3(blocks_for((size_t)(W >> (l + 1)) * (H >> (l + 1)))),
                       dim3(256), 0, s, lv[l], W >> l, H >> l, lv[l + 1]);
  __half* cur = lv[levels];
  for (int l = levels - 1; l >= 1; l--) {
    hipLaunchKernelGGL(k_bloom_up, dim3(blocks_for((size_t)(W >> l) * (H >> l))), dim3(256), 0, s,
                       lv[l], cur, W >> l, H >> l, up[l]);
    cur = up[l];
  }
  hipLaunchKernelGGL(k_bloom_composite, dim3(blocks_for((size_t)W * H)), dim3(256), 0, s, R, G, B,
                     W, H, cur, white, strength);
  return hipGetLastError();
}

}  // namespace axr

// ------------------------------------------------ bloom without allocation
// axr_bloom_planes_noalloc: the bloom passes on caller-owned device planes,
// on the caller's stream, with the level pyramid in a file-scope device
// array instead of memory the library allocates.  Each GPU holds its own
// copy of the array (until the process ends); its address on a device is
// looked up once, with that device current, and kept per device: host
// code passes that address to the kernels, never the array's name.  Every
// level is written before it is read, so nothing needs clearing between
// calls; calls that overlap on one GPU share the array and race.
#include <mutex>

#include "axr_ctx.h"

namespace {

// 32 MiB: the fp16 levels and upward chain of a 3840x2160 frame with four
// levels (33,048,000 bytes) fit.
constexpr size_t kNoallocScratchBytes = (size_t)32 << 20;
__device__ __align__(256) unsigned char scratch[kNoallocScratchBytes];

constexpr int kNoallocMaxDevices = 64;
void* g_scratch_ptr[kNoallocMaxDevices] = {};
std::mutex g_scratch_mu;

}  // namespace

extern "C" int axr_bloom_planes_noalloc(uint16_t* r, uint16_t* g, uint16_t* b, int w, int h,
                                        int levels, float threshold, float white,
                                        float strength, void* stream) {
  using namespace axri;
  if (!r || !g || !b) return fail(AXR_EINVAL, "bloom_noalloc: null plane");
  if (w <= 0 || h <= 0) return fail(AXR_EINVAL, "bloom_noalloc: frame %dx%d", w, h);
  if (levels < 1 || levels > 12)
    return fail(AXR_EINVAL, "bloom_noalloc: levels must be 1..12, got %d", levels);
  if ((w % (1 << levels)) || (h % (1 << levels)))
    return fail(AXR_EINVAL, "bloom_noalloc: %dx%d is not a multiple of 2^%d", w, h, levels);
  if (!(white > 0.0f) || !(strength >= 0.0f) || threshold != threshold)
    return fail(AXR_EINVAL, "bloom_noalloc: white must be > 0 and strength >= 0");
  const size_t need = 2 * bloom_halves(w, h, levels);
  if (need > kNoallocScratchBytes)
    return fail(AXR_EINVAL,
                "bloom_noalloc: %dx%d with %d levels needs %zu bytes of scratch; the "
                "no-allocation entry point holds %zu (up to 3840x2160 with 4 levels)",
                w, h, levels, need, kNoallocScratchBytes);
  int dev = 0;
  HIPCHK(hipGetDevice(&dev));
  if (dev < 0 || dev >= kNoallocMaxDevices)
    return fail(AXR_EINVAL, "bloom_noalloc: device %d beyond %d", dev, kNoallocMaxDevices);
  void* p = nullptr;
  {
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
      void* q = nullptr;
      const hipError_t e = hipGetSymbolAddress(&q, HIP_SYMBOL(scratch));
      if (e != hipSuccess || !q)
        return fail(AXR_EHIP, "bloom_noalloc: scratch address on device %d: %s", dev,
                    hipGetErrorString(e));
      g_scratch_ptr[dev] = q;
    }
    p = g_scratch_ptr[dev];
  }
  HIPCHK(launch_bloom(r, g, b, w, h, levels, threshold, white, strength, p,
                      (hipStream_t)stream));
  return AXR_OK;
}
