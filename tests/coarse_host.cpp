// Host build of the coarse binning enumerations in axr_tri.h (compiled by
// tests/test_coarse_host.py with hipcc's host pass): the same template code
// the k_coarse kernel and the batched setup run, here on the CPU so that
// its tile indices can be checked against the grid and against the oracle
// without a GPU.
#include "axr_tri.h"

using namespace axr;

// coarse_block_tiles over every thread of a block of nthreads; tiles as
// cy * WC + column in thread order.  Returns the count (entries past cap
// are counted, not stored).
extern "C" int host_coarse_block(const float* px, const float* py, int WC, int HC, int nthreads,
                                 int* tiles, int cap) {
  int n = 0;
  for (int tid = 0; tid < nthreads; tid++)
    coarse_block_tiles(px, py, WC, HC, tid, nthreads, [&](int cy, int ax, int G, float xe) {
      for (int j = 0; j < 8; j++) {
        const int tx = ax + j * G;
        if (!((float)tx <= xe)) continue;
        if (n < cap) tiles[n] = cy * WC + tx;
        n++;
      }
    });
  return n;
}

// The serial walk (batched setup): tiles in emission order.
extern "C" int host_coarse_walk(const float* px, const float* py, int WC, int HC, int* tiles,
                                int cap) {
  int n = 0;
  coarse_walk(px, py, WC, HC, [&](int t) {
    if (n < cap) tiles[n] = t;
    n++;
  });
  return n;
}

// The rows coarse_rows reports to one thread walking every row: (cy,
// first, xe) per row.
extern "C" int host_coarse_rows(const float* px, const float* py, int WC, int HC, int* cy,
                                int* first, float* xe, int cap) {
  int n = 0;
  coarse_rows(px, py, WC, HC, 0, 1, [&](int y, int f, float e) {
    if (n < cap) {
      cy[n] = y;
      first[n] = f;
      xe[n] = e;
    }
    n++;
  });
  return n;
}
