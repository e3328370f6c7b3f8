// Host build of the per-chunk culling tests in axr_tri.h
// (chunk_outside_view, chunk_outside_ortho) and of their camera constants
// (make_cull, axr_common.h), compiled by tests/test_cull_host.py with
// hipcc's host pass: the code the setup kernels run, checked on the CPU.
#include "axr_tri.h"

using namespace axr;

namespace {
AxrCam cam_of(const float* pos, const float* vm, float scale, float cax, float cay) {
  AxrCam c{};
  for (int k = 0; k < 3; k++) c.pos[k] = pos[k];
  for (int k = 0; k < 9; k++) c.vm[k] = vm[k];
  c.scale = scale;
  c.cax = cax;
  c.cay = cay;
  return c;
}
}  // namespace

// out[i] = 1 when sphere i (x, y, z, r) is outside the perspective view.
extern "C" void host_outside_view(const float* spheres, int n, const float* pos, const float* vm,
                                  float scale, float cax, float cay, int* out) {
  const AxrCam c = cam_of(pos, vm, scale, cax, cay);
  const AxrCull k = make_cull(c);
  for (int i = 0; i < n; i++) {
    const float* s = spheres + 4 * i;
    out[i] = chunk_outside_view(make_float4(s[0], s[1], s[2], s[3]), c, k) ? 1 : 0;
  }
}

// out[i] = 1 when sphere i is outside an ortho shadow camera's accept test.
extern "C" void host_outside_ortho(const float* spheres, int n, const float* pos, const float* vm,
                                   float scale, int dim, float bias, int* out) {
  const AxrCam c = cam_of(pos, vm, scale, 0.0f, 0.0f);
  const AxrCull k = make_cull(c);
  for (int i = 0; i < n; i++) {
    const float* s = spheres + 4 * i;
    out[i] = chunk_outside_ortho(make_float4(s[0], s[1], s[2], s[3]), c, k, dim, bias) ? 1 : 0;
  }
}
