// Host-side helpers shared by the C-ABI translation units of libedge_hip.so.
#pragma once
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <string>

#include "edge_device.h"

namespace edge {

// Thread-local last error message (edge_last_error()).
void set_error(const char* fmt, ...);
void clear_error();

// Status codes returned by every extern "C" entry point (0 = ok).
enum Status : int {
  EDGE_OK = 0,
  EDGE_EINVAL = -1,     // bad argument (null pointer, bad dims, unsupported size)
  EDGE_EHIP = -2,       // a HIP runtime call failed
  EDGE_ENOMEM = -3,     // allocation failed
  EDGE_ENOCONV = -4,    // the solver did not converge (results still written)
};

inline Grid make_grid(int nx, int ny, int nz, float dx) {
  Grid g;
  g.nx = nx;
  g.ny = ny;
  g.nz = nz;
  g.dx = dx;
  g.idx = 1.0f / dx;
  g.cz = make_fdiv((uint32_t)nz);
  g.cy = make_fdiv((uint32_t)ny);
  for (int a = 0; a < 3; ++a) {
    g.fz[a] = make_fdiv((uint32_t)(nz + (a == 2)));
    g.fy[a] = make_fdiv((uint32_t)(ny + (a == 1)));
    g.nf[a] = (uint32_t)((int64_t)(nx + (a == 0)) * (ny + (a == 1)) * (nz + (a == 2)));
  }
  int n[3] = {nx, ny, nz};
  for (int a = 0; a < 3; ++a) {
    g.hi_c[a] = (float)((double)n[a] - 2.0 - 1e-6);
    g.hi_f[a] = (float)((double)(n[a] + 1) - 2.0 - 1e-6);
    g.hi1_c[a] = (float)((double)n[a] - 1.0 - 1e-6);
  }
  return g;
}

inline bool dims_ok(int nx, int ny, int nz) {
  // the 3-tap B-spline and the Hermite corner pair need at least 4 cells per axis
  // index decodes use 32-bit arithmetic: every face grid together below 2^31
  // buffer-resource gathers take a 32-bit byte range: one component's Hermite
  // table (32 B per cell) below 4 GiB, i.e. fewer than 2^27 cells
  const int64_t nfaces = (int64_t)(nx + 1) * ny * nz + (int64_t)nx * (ny + 1) * nz + (int64_t)nx * ny * (nz + 1);
  const int64_t ncells = (int64_t)nx * ny * nz;
  return nx >= 4 && ny >= 4 && nz >= 4 && nfaces < (int64_t)1 << 31 && ncells < (int64_t)1 << 27;
}

inline int blocks_for(int64_t n, int bs) { return (int)((n + bs - 1) / bs); }

}  // namespace edge

#define EDGE_CHECK_HIP(expr)                                                           \
  do {                                                                                 \
    hipError_t _e = (expr);                                                            \
    if (_e != hipSuccess) {                                                            \
      edge::set_error("%s failed: %s (%s:%d)", #expr, hipGetErrorString(_e), __FILE__, \
                      __LINE__);                                                       \
      return edge::EDGE_EHIP;                                                          \
    }                                                                                  \
  } while (0)

#define EDGE_REQUIRE(cond, ...)        \
  do {                                 \
    if (!(cond)) {                     \
      edge::set_error(__VA_ARGS__);    \
      return edge::EDGE_EINVAL;        \
    }                                  \
  } while (0)

// Check the launch that was just enqueued.
#define EDGE_CHECK_LAUNCH() EDGE_CHECK_HIP(hipGetLastError())

// XCD-aware block order: the hardware deals consecutive workgroups to the 8
// XCDs round-robin; remapping makes spatially consecutive tiles share one
// XCD's L2 (gathers of neighbouring cells overlap).
__device__ __forceinline__ int xcd_block(int b, int nb) {
  const int X = 8;
  int per = nb / X;
  if (per == 0 || b >= per * X) return b;
  return (b % X) * per + b / X;
}

#ifndef EDGE_BAND
#define EDGE_BAND 0
#endif
// Cell kernels whose threads gather around their own cell (the marches): each
// XCD still takes a contiguous slab of whole planes (xcd_block), but walks it
// in bands of W blocks of every plane in turn instead of plane by plane, so the
// blocks in flight on one XCD cover a narrow band through many planes and the
// +-1-plane halo of their gathers stays in that XCD's L2.  bpp: blocks per
// plane (0: off).
__device__ __forceinline__ int band_block(int b, int nb, int bpp, int W) {
  const int X = 8;
  const int per = nb / X;
  if (W <= 0 || bpp <= 0 || per == 0 || b >= per * X || per % bpp || bpp % W) return xcd_block(b, nb);
  const int planes = per / bpp;
  const int L = b / X;
  const int band = L / (W * planes), rem = L % (W * planes);
  return (b % X) * per + (rem / W) * bpp + band * W + rem % W;
}
