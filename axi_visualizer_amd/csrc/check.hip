// Self-check of the device arithmetic the raster kernels rely on beyond
// plain IEEE operations (axr_common.h): the fast correctly rounded division
// (rcp_rn, div_rn, div_rn0) against the device's IEEE division on the
// operand ranges the kernels guard it with, and the hardware saturating
// conversions (f2i, f2u16) against their formulas for every float.
#include "axr_common.h"
#include "axr_ctx.h"

namespace {

__device__ __forceinline__ uint32_t mix32(uint64_t x) {
  x ^= x >> 33;
  x *= 0xff51afd7ed558ccdull;
  x ^= x >> 33;
  x *= 0xc4ceb9fe1a85ec53ull;
  x ^= x >> 33;
  return (uint32_t)x;
}

// A float of random sign and mantissa with exponent in [e0, e0 + ne).
__device__ __forceinline__ float rand_float(uint32_t h, uint32_t m, int e0, int ne) {
  const int e = e0 + (int)(h % (uint32_t)ne);
  return __uint_as_float(((h >> 31) << 31) | ((uint32_t)(e + 127) << 23) | (m & 0x7fffffu));
}

// Every b with |b| in [2^-60, 2^60], both signs.
__global__ void k_check_rcp(unsigned long long* __restrict__ bad) {
  constexpr uint32_t lo = 0x21800000u, hi = 0x5d800000u;  // 2^-60, 2^60
  unsigned long long n = 0;
  for (uint32_t i = lo + blockIdx.x * blockDim.x + threadIdx.x; i <= hi;
       i += gridDim.x * blockDim.x) {
#pragma unroll
    for (uint32_t s = 0; s < 2; s++) {
      const float b = __uint_as_float(i | (s << 31));
      n += __float_as_uint(rcp_rn(b)) != __float_as_uint(1.0f / b);
    }
  }
  if (n) atomicAdd(bad, n);
}

// Random quotients from the two regimes of div_rn: |a|, |b| in
// [2^-30, 2^30] (half of them span numerators k - c with k an integer),
// and b an integer in [1, 2^32] with a an integer, a finite float with
// |a| >= 2^-94, or a signed zero (div_rn0).
__global__ void k_check_div(uint64_t seed, uint32_t per, unsigned long long* __restrict__ tested,
                            unsigned long long* __restrict__ bad) {
  const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  unsigned long long nt = 0, nb = 0;
  for (uint32_t k = 0; k < per; k++) {
    const uint64_t s = (seed + t * per + k) * 4;
    const uint32_t h0 = mix32(s), h1 = mix32(s + 1), h2 = mix32(s + 2), h3 = mix32(s + 3);
    float a, b, q;
    if (k & 1) {
      b = rand_float(h0, h1, -30, 60);
      a = rand_float(h2, h3, -30, 60);
      if (h3 & 1) {
        const float c = a;
        a = (float)(int)(h2 & 0xfffu) - c;
        const float m = fabsf(a);
        if (!(a == 0.0f || (m >= 0x1p-30f && m <= 0x1p30f))) continue;
      }
      q = div_rn(a, b, rcp_rn(b));
    } else {
      const uint32_t bits = 1u + h0 % 32u;
      b = (float)(uint32_t)(1u + (h1 & (bits == 32u ? 0xffffffffu : (1u << bits) - 1u)));
      if (!(b >= 1.0f && b <= 0x1p32f)) continue;
      const uint32_t kind = h2 % 3u;
      if (kind == 0) {
        a = (float)(int32_t)h3;
        q = div_rn(a, b, rcp_rn(b));
      } else {
        a = kind == 1 ? rand_float(h3, h2 >> 2, -94, 222) : __uint_as_float((h3 >> 31) << 31);
        q = div_rn0(a, b, rcp_rn(b));
      }
    }
    nt++;
    nb += __float_as_uint(q) != __float_as_uint(a / b);
  }
  atomicAdd(tested, nt);
  if (nb) atomicAdd(bad, nb);
}

// Every float bit pattern: f2i and f2u16 against f2i_ref and f2u16_ref.
__global__ void k_check_cvt(unsigned long long* __restrict__ bad) {
  unsigned long long n = 0;
  const uint32_t stride = gridDim.x * blockDim.x;
  for (uint64_t i = blockIdx.x * blockDim.x + threadIdx.x; i <= 0xffffffffull; i += stride) {
    const float x = __uint_as_float((uint32_t)i);
    n += f2i(x) != f2i_ref(x);
    n += f2u16(x) != f2u16_ref(x);
  }
  if (n) atomicAdd(bad, n);
}

}  // namespace

extern "C" int axr_check_arith(uint64_t seed, uint32_t rounds, uint64_t out[4]) {
  if (!out) return fail(AXR_EINVAL, "check_arith: out is NULL");
  if (rounds > 1024) return fail(AXR_EINVAL, "check_arith: rounds %u > 1024", rounds);
  unsigned long long* d = nullptr;
  HIPCHK(hipMalloc(&d, 4 * sizeof(unsigned long long)));
  hipError_t e = hipMemset(d, 0, 4 * sizeof(unsigned long long));
  constexpr unsigned kBlocks = 4096, kThreads = 256, kPer = 64;
  if (e == hipSuccess) {
    hipLaunchKernelGGL(k_check_rcp, dim3(kBlocks), dim3(kThreads), 0, 0, d);
    for (uint32_t r = 0; r < rounds; r++)
      hipLaunchKernelGGL(k_check_div, dim3(kBlocks), dim3(kThreads), 0, 0,
                         seed + (uint64_t)r * kBlocks * kThreads * kPer, kPer, d + 1, d + 2);
    hipLaunchKernelGGL(k_check_cvt, dim3(kBlocks), dim3(kThreads), 0, 0, d + 3);
    e = hipGetLastError();
  }
  unsigned long long h[4] = {0, 0, 0, 0};
  if (e == hipSuccess) e = hipMemcpy(h, d, sizeof(h), hipMemcpyDeviceToHost);
  const hipError_t ef = hipFree(d);
  if (e != hipSuccess) return fail(AXR_EHIP, "check_arith: %s", hipGetErrorString(e));
  if (ef != hipSuccess) return fail(AXR_EHIP, "check_arith: %s", hipGetErrorString(ef));
  for (int k = 0; k < 4; k++) out[k] = h[k];
  return AXR_OK;
}
