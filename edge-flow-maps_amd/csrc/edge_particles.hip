// Laden particles (sim/lp.py): gather of the MAC velocity at the particles,
// implicit drag, wall rules, the drag's scatter back onto the faces (P2G) and
// the viscous force.  One thread per particle (per face for the viscous
// force).  Particle vectors are SoA: 3 planes of n (the capacity); only
// particles q < num are live, as in the reference.
#include <algorithm>
#include <cstring>
#include <map>
#include <mutex>

#include <rocprim/device/device_radix_sort.hpp>

#include "edge_common.h"

using namespace edge;

namespace {

struct P3 {
  float* __restrict__ v[3];
};
struct CP3 {
  const float* __restrict__ v[3];
};

__device__ __forceinline__ void particle_u(const Grid& g, const Mac& U, const CP3& pos, int64_t q, float u[3]) {
  float p[3] = {pos.v[0][q], pos.v[1][q], pos.v[2][q]};
  interp_u_mac<false>(g, U, p, u, nullptr);
}

// sim/lp.py:6-29 update_particles_acc: acc = (g e_y - k (v - u(p))) / (1 + k dt), k = 1/St;
// acc = 0 for q >= num (the reference's acc.fill(0)).
// order (nullable): a permutation of the first num particles to visit them in
// (cell order from the last scatter): neighbouring threads then gather nearby
// velocities.  Each particle's result does not depend on it.
__global__ void __launch_bounds__(256) k_lp_acc(Grid g, Mac U, CP3 vel, CP3 pos, P3 acc, int64_t n, int64_t num,
                                               float dt, float St, float gravity, const uint32_t* __restrict__ order) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t q = (order && t < num) ? (int64_t)order[t] : t;
  if (q >= n) return;
  if (q >= num) {
#pragma unroll
    for (int a = 0; a < 3; ++a) acc.v[a][q] = 0.f;
    return;
  }
  const float k = 1.0f / St;
  float u[3];
  particle_u(g, U, pos, q, u);
  float a[3] = {0.f, gravity, 0.f};
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    a[c] -= k * (vel.v[c][q] - u[c]);
    a[c] /= (1.0f + k * dt);
    acc.v[c][q] = a[c];
  }
}

// sim/lp.py:32-70 update_particles_velocity: v += a dt, then the wall rules
// (the lower y wall uses dx, the others 3 dx).
__global__ void __launch_bounds__(256) k_lp_vel(CP3 acc, P3 vel, CP3 pos, int64_t num, float dt, float dx, float Lx,
                                               float Ly, float Lz) {
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= num) return;
  float v[3], p[3];
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    v[c] = vel.v[c][q] + acc.v[c][q] * dt;
    p[c] = pos.v[c][q];
  }
  if (p[1] < dx && v[1] < 0.f) v[1] = 0.f;
  if (p[1] > Ly - 3.f * dx && v[1] > 0.f) v[1] = 0.f;
  if (p[0] < 3.f * dx && v[0] < 0.f) v[0] = 0.f;
  if (p[0] > Lx - 3.f * dx && v[0] > 0.f) v[0] = 0.f;
  if (p[2] < 3.f * dx && v[2] < 0.f) v[2] = 0.f;
  if (p[2] > Lz - 3.f * dx && v[2] > 0.f) v[2] = 0.f;
#pragma unroll
  for (int c = 0; c < 3; ++c) vel.v[c][q] = v[c];
}

// sim/lp.py:73-97 calculate_velocity_diff: k (v - u(p)), k = density_ratio / (St r^3).
__global__ void __launch_bounds__(256) k_lp_vel_diff(Grid g, Mac U, CP3 vel, P3 diff, CP3 pos, int64_t num, float St,
                                                    float density_ratio, float radius_ratio,
                                                    const uint32_t* __restrict__ order) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= num) return;
  const int64_t q = order ? (int64_t)order[t] : t;
  const float er = 1.0f / (radius_ratio * radius_ratio * radius_ratio);
  const float k = 1.0f / St * density_ratio * er;
  float u[3];
  particle_u(g, U, pos, q, u);
#pragma unroll
  for (int c = 0; c < 3; ++c) diff.v[c][q] = k * (vel.v[c][q] - u[c]);
}

// ---------------------------------------------------------------------------
// sim/lp.py:100-156 laden_particles_back_with_computing_particle: the drag
// diff (r_p / r)^3 spread onto face grid a with the quadratic B-spline
// weights; particles within 3 dx of a wall are skipped, so a particle's three
// non-zero taps per axis are inside the grid.
// Particles are grouped by 8x8x8-cell tile (a radix sort of tile keys); one
// workgroup per tile accumulates its particles' taps in LDS with LDS atomics,
// then adds each touched face of the tile's box to the grid with one global
// atomic.  Summation order differs from a sequential loop by rounding only.
// ---------------------------------------------------------------------------
constexpr int TILE = 8;
constexpr int BOX = TILE + 4;   // face indices c0-2 .. c0+9 of a tile starting at cell c0

// the reference's skip test (NaN positions skipped)
__device__ __forceinline__ bool p2g_live(const float P[3], float dx, float Lx, float Ly, float Lz) {
  return P[0] >= 3.f * dx && P[1] >= 3.f * dx && P[2] >= 3.f * dx && P[0] <= Lx - 3.f * dx &&
         P[1] <= Ly - 3.f * dx && P[2] <= Lz - 3.f * dx;
}

struct Tiles {
  int tx, ty, tz;   // tiles per axis
};

__global__ void __launch_bounds__(256) k_p2g_keys(Grid g, CP3 pos, int64_t num, float Lx, float Ly, float Lz, Tiles T,
                                                 uint32_t* __restrict__ keys, uint32_t* __restrict__ vals) {
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= num) return;
  const float P[3] = {pos.v[0][q], pos.v[1][q], pos.v[2][q]};
  uint32_t key = (uint32_t)T.tx * T.ty * T.tz * (TILE * TILE * TILE);   // sentinel: not scattered
  if (p2g_live(P, g.dx, Lx, Ly, Lz)) {
    const int n[3] = {g.nx, g.ny, g.nz};
    int c[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) c[d] = min(max((int)floorf(P[d] / g.dx), 0), n[d] - 1);
    const uint32_t tile = ((uint32_t)(c[0] / TILE) * T.ty + (uint32_t)(c[1] / TILE)) * T.tz + (uint32_t)(c[2] / TILE);
    key = tile * (TILE * TILE * TILE) + (uint32_t)(((c[0] % TILE) * TILE + c[1] % TILE) * TILE + c[2] % TILE);
  }
  keys[q] = key;
  vals[q] = (uint32_t)q;
}

// start/end of each cell's run in the sorted keys (both zero for empty cells)
__global__ void __launch_bounds__(256) k_p2g_ranges(const uint32_t* __restrict__ keys, int64_t num, uint32_t ntiles,
                                                   uint32_t* __restrict__ start, uint32_t* __restrict__ end) {
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= num) return;
  const uint32_t k = keys[q];
  if (k >= ntiles) return;
  if (q == 0 || keys[q - 1] != k) start[k] = (uint32_t)q;
  if (q == num - 1 || keys[q + 1] != k) end[k] = (uint32_t)(q + 1);
}

// One workgroup per tile, one thread per cell: the thread sums its cell's
// particles into a register window of the faces they reach (per face grid a:
// 4 faces along a, starting at the cell's index - 1, and 3 along the other
// axes), adds the window to the tile's LDS box, and the box goes to the grid.
constexpr int TCELLS = TILE * TILE * TILE;
__global__ void __launch_bounds__(TCELLS) k_p2g_cells(Grid g, P3 drag, CP3 diff, CP3 pos, const float* __restrict__ rr,
                                                     float radius, Tiles T, const uint32_t* __restrict__ order,
                                                     const uint32_t* __restrict__ start,
                                                     const uint32_t* __restrict__ end) {
  __shared__ float box[3][BOX][BOX][BOX];
  const uint32_t tile = blockIdx.x;
  const int lc = threadIdx.x;
  const uint32_t key = tile * TCELLS + (uint32_t)lc;
  const uint32_t b = start[key], e = end[key];
  if (__syncthreads_count(b < e) == 0) return;
  const int tz = (int)(tile % (uint32_t)T.tz), ty = (int)((tile / (uint32_t)T.tz) % (uint32_t)T.ty),
            tx = (int)(tile / ((uint32_t)T.tz * T.ty));
  const int o[3] = {tx * TILE - 2, ty * TILE - 2, tz * TILE - 2};   // face index of box entry 0
  const int c[3] = {tx * TILE + (lc >> 6), ty * TILE + ((lc >> 3) & 7), tz * TILE + (lc & 7)};
  float* flat = &box[0][0][0][0];
  for (int t = threadIdx.x; t < 3 * BOX * BOX * BOX; t += blockDim.x) flat[t] = 0.f;
  __syncthreads();
  float acc[3][36];
#pragma unroll
  for (int a = 0; a < 3; ++a)
#pragma unroll
    for (int m = 0; m < 36; ++m) acc[a][m] = 0.f;
  const float dx = g.dx;
  for (uint32_t idx = b; idx < e; ++idx) {
    const uint32_t q = order[idx];
    const float P[3] = {pos.v[0][q], pos.v[1][q], pos.v[2][q]};
    const float r = rr[q];
    const float scale = r * r * r / (radius * radius * radius);
    const float sp[3] = {P[0] / dx, P[1] / dx, P[2] / dx};
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      const Axis3 X[3] = {bspline_axis(sp[0] - (a == 0 ? 0.f : 0.5f), 1e30f),
                          bspline_axis(sp[1] - (a == 1 ? 0.f : 0.5f), 1e30f),
                          bspline_axis(sp[2] - (a == 2 ? 0.f : 0.5f), 1e30f)};
      const float d = diff.v[a][q] * scale;
      // window offsets: 0 or 1 along a, 0 along the others
      int off[3];
      bool in_window = true;
#pragma unroll
      for (int ax = 0; ax < 3; ++ax) {
        off[ax] = X[ax].i0 - (c[ax] - 1);
        in_window = in_window && (ax == a ? (off[ax] == 0 || off[ax] == 1) : off[ax] == 0);
      }
      if (in_window) {
        // weight of window entry m along axis ax
        auto wt = [&](int ax, int m) {
          if (ax != a) return X[ax].w[m];
          return off[ax] == 0 ? (m < 3 ? X[ax].w[m < 3 ? m : 2] : 0.f) : (m > 0 ? X[ax].w[m > 0 ? m - 1 : 0] : 0.f);
        };
        constexpr int W = 3;
#pragma unroll
        for (int m0 = 0; m0 < (a == 0 ? 4 : 3); ++m0)
#pragma unroll
          for (int m1 = 0; m1 < (a == 1 ? 4 : 3); ++m1) {
            const float w01 = wt(0, m0) * wt(1, m1);
#pragma unroll
            for (int m2 = 0; m2 < (a == 2 ? 4 : 3); ++m2) {
              const int n1 = (a == 1 ? 4 : W), n2 = (a == 2 ? 4 : W);
              acc[a][(m0 * n1 + m1) * n2 + m2] += d * (w01 * wt(2, m2));
            }
          }
      } else {
        // a tap start outside the expected window (rounding at a cell edge): straight to LDS
        const int li = X[0].i0 - o[0], lj = X[1].i0 - o[1], lk = X[2].i0 - o[2];
        if (li < 0 || lj < 0 || lk < 0 || li + 2 >= BOX || lj + 2 >= BOX || lk + 2 >= BOX) continue;
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
          for (int j = 0; j < 3; ++j)
#pragma unroll
            for (int k = 0; k < 3; ++k)
              atomicAdd(&box[a][li + i][lj + j][lk + k], d * (X[0].w[i] * X[1].w[j] * X[2].w[k]));
      }
    }
  }
  if (b < e) {
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      const int n0 = (a == 0 ? 4 : 3), n1 = (a == 1 ? 4 : 3), n2 = (a == 2 ? 4 : 3);
#pragma unroll
      for (int m0 = 0; m0 < 4; ++m0)
#pragma unroll
        for (int m1 = 0; m1 < 4; ++m1)
#pragma unroll
          for (int m2 = 0; m2 < 4; ++m2) {
            if (m0 >= n0 || m1 >= n1 || m2 >= n2) continue;
            const float v = acc[a][(m0 * n1 + m1) * n2 + m2];
            if (v != 0.f) atomicAdd(&box[a][c[0] - 1 + m0 - o[0]][c[1] - 1 + m1 - o[1]][c[2] - 1 + m2 - o[2]], v);
          }
    }
  }
  __syncthreads();
  for (int t = threadIdx.x; t < 3 * BOX * BOX * BOX; t += blockDim.x) {
    const float v = flat[t];
    if (v == 0.f) continue;
    const int a = t / (BOX * BOX * BOX);
    const int rem = t - a * BOX * BOX * BOX;
    const int fi = o[0] + rem / (BOX * BOX), fj = o[1] + (rem / BOX) % BOX, fk = o[2] + rem % BOX;
    const int d0 = g.nx + (a == 0), d1 = g.ny + (a == 1), d2 = g.nz + (a == 2);
    if (fi < 0 || fj < 0 || fk < 0 || fi >= d0 || fj >= d1 || fk >= d2) continue;
    atomicAdd(drag.v[a] + ((int64_t)fi * d1 + fj) * d2 + fk, v);
  }
}

// Sort workspace per device, grown on demand and kept for the process.
struct P2GWork {
  size_t cap = 0, tiles_cap = 0, temp_bytes = 0;
  uint32_t *k0 = nullptr, *k1 = nullptr, *v0 = nullptr, *v1 = nullptr, *start = nullptr, *end = nullptr;
  void* temp = nullptr;
  int64_t perm_n = 0;   // v1 holds a permutation of 0..perm_n-1 (the last scatter's cell order)
};
std::map<int, P2GWork> g_p2g;

// The last scatter's cell order when it covers exactly m particles (else null):
// the gathers visit particles in it.  Stream-ordered: a gather sees the order of
// the last scatter launched before it on the same stream.
const uint32_t* cell_order(int64_t m) {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return nullptr;
  auto it = g_p2g.find(dev);
  return (it != g_p2g.end() && it->second.perm_n == m && m > 0) ? it->second.v1 : nullptr;
}

int p2g_workspace(P2GWork& w, int64_t num, uint32_t ntiles, int end_bit, hipStream_t st) {
  if ((size_t)num > w.cap) {
    w.perm_n = 0;
    for (uint32_t* p : {w.k0, w.k1, w.v0, w.v1})
      if (p) (void)hipFree(p);
    w.k0 = w.k1 = w.v0 = w.v1 = nullptr;
    w.cap = 0;
    const size_t bytes = (size_t)num * sizeof(uint32_t);
    EDGE_CHECK_HIP(hipMalloc(&w.k0, bytes));
    EDGE_CHECK_HIP(hipMalloc(&w.k1, bytes));
    EDGE_CHECK_HIP(hipMalloc(&w.v0, bytes));
    EDGE_CHECK_HIP(hipMalloc(&w.v1, bytes));
    w.cap = (size_t)num;
  }
  if (ntiles > w.tiles_cap) {
    if (w.start) (void)hipFree(w.start);
    if (w.end) (void)hipFree(w.end);
    w.start = w.end = nullptr;
    w.tiles_cap = 0;
    EDGE_CHECK_HIP(hipMalloc(&w.start, (size_t)ntiles * sizeof(uint32_t)));
    EDGE_CHECK_HIP(hipMalloc(&w.end, (size_t)ntiles * sizeof(uint32_t)));
    w.tiles_cap = ntiles;
  }
  size_t need = 0;
  EDGE_CHECK_HIP(rocprim::radix_sort_pairs(nullptr, need, w.k0, w.k1, w.v0, w.v1, (size_t)num, 0u, (unsigned)end_bit, st));
  if (need > w.temp_bytes) {
    if (w.temp) (void)hipFree(w.temp);
    w.temp = nullptr;
    w.temp_bytes = 0;
    EDGE_CHECK_HIP(hipMalloc(&w.temp, need));
    w.temp_bytes = need;
  }
  return EDGE_OK;
}

__global__ void __launch_bounds__(256) k_lp_zero(float* __restrict__ f, int64_t n) {
  int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) f[t] = 0.f;
}

int64_t blocks64(int64_t n) { return (n + 255) / 256; }

// One scatter's validated arguments and derived sizes.
struct P2GCall {
  Grid g;
  CP3 P, D;
  P3 drag;
  const float* rr;
  float radius, Lx, Ly, Lz;
  Tiles T;
  uint32_t ntiles, nkeys;
  int end_bit;
  int64_t m;
  hipStream_t st;
};

// The sort keys and values (double-buffered), the cell ranges and the sort's
// temporary storage of one scatter.
struct P2GBufs {
  uint32_t *k0, *k1, *v0, *v1, *start, *end;
  void* temp;
  size_t temp_bytes;
};

// Argument checks (host only, before any launch) and the sizes the sort and
// the tiles need.
int p2g_prepare(const char* who, float* dragx, float* dragy, float* dragz, const float* diff, const float* pos,
                const float* ref_radius, long n, long num, float dx, float radius, float Lx, float Ly, float Lz, int nx,
                int ny, int nz, void* stream, P2GCall& c) {
  EDGE_REQUIRE(dims_ok(nx, ny, nz), "%s: bad dims", who);
  EDGE_REQUIRE(dragx && dragy && dragz && diff && pos && ref_radius && n >= 0 && num >= 0 && n < (1L << 31),
               "%s: bad arguments", who);
  EDGE_REQUIRE(radius != 0.0f && dx > 0.0f, "%s: radius and dx must be non-zero", who);
  // the 3 dx wall margin is what keeps every tap in the grid: the domain must match the grid
  EDGE_REQUIRE(Lx <= (float)nx * dx * 1.0001f && Ly <= (float)ny * dx * 1.0001f && Lz <= (float)nz * dx * 1.0001f,
               "%s: domain (%g, %g, %g) larger than the grid (%d, %d, %d) x %g", who, Lx, Ly, Lz, nx, ny, nz, dx);
  const Tiles T{(nx + TILE - 1) / TILE, (ny + TILE - 1) / TILE, (nz + TILE - 1) / TILE};
  EDGE_REQUIRE((int64_t)T.tx * T.ty * T.tz * TCELLS < (1LL << 31), "%s: grid too large for 32-bit cell keys", who);
  c.st = (hipStream_t)stream;
  c.g = make_grid(nx, ny, nz, dx);
  float* dr[3] = {dragx, dragy, dragz};
  for (int a = 0; a < 3; ++a) {
    c.P.v[a] = pos + a * n;
    c.D.v[a] = diff + a * n;
    c.drag.v[a] = dr[a];
  }
  c.rr = ref_radius;
  c.radius = radius;
  c.Lx = Lx;
  c.Ly = Ly;
  c.Lz = Lz;
  c.T = T;
  c.ntiles = (uint32_t)T.tx * T.ty * T.tz;
  c.nkeys = c.ntiles * TCELLS;   // tile-major cell ids; nkeys = not scattered
  c.end_bit = 1;
  while ((1u << c.end_bit) <= c.nkeys) ++c.end_bit;
  c.m = std::min<int64_t>(num, n);
  return EDGE_OK;
}

// The drag grids zeroed (every call, also without live particles).
int p2g_zero(const P2GCall& c) {
  for (int a = 0; a < 3; ++a)
    hipLaunchKernelGGL(k_lp_zero, dim3(blocks64(c.g.nf[a])), dim3(256), 0, c.st, c.drag.v[a], (int64_t)c.g.nf[a]);
  EDGE_CHECK_LAUNCH();
  return EDGE_OK;
}

// Keys, the stable sort by cell, the cells' runs and the tile scatter.
int p2g_launch(const P2GCall& c, const P2GBufs& b) {
  hipLaunchKernelGGL(k_p2g_keys, dim3(blocks64(c.m)), dim3(256), 0, c.st, c.g, c.P, c.m, c.Lx, c.Ly, c.Lz, c.T, b.k0,
                     b.v0);
  EDGE_CHECK_LAUNCH();
  size_t temp = b.temp_bytes;
  EDGE_CHECK_HIP(rocprim::radix_sort_pairs(b.temp, temp, b.k0, b.k1, b.v0, b.v1, (size_t)c.m, 0u,
                                           (unsigned)c.end_bit, c.st));
  hipLaunchKernelGGL(k_lp_zero, dim3(blocks64(c.nkeys)), dim3(256), 0, c.st, (float*)b.start, (int64_t)c.nkeys);
  hipLaunchKernelGGL(k_lp_zero, dim3(blocks64(c.nkeys)), dim3(256), 0, c.st, (float*)b.end, (int64_t)c.nkeys);
  hipLaunchKernelGGL(k_p2g_ranges, dim3(blocks64(c.m)), dim3(256), 0, c.st, b.k1, c.m, c.nkeys, b.start, b.end);
  hipLaunchKernelGGL(k_p2g_cells, dim3(c.ntiles), dim3(TCELLS), 0, c.st, c.g, c.drag, c.D, c.P, c.rr, c.radius, c.T,
                     b.v1, b.start, b.end);
  EDGE_CHECK_LAUNCH();
  return EDGE_OK;
}

// ---------------------------------------------------------------------------
// The allocation-free scatter's storage: one file-scope device array per GPU
// (each GPU holds its own copy from the library's load until the process
// ends).  Layout: sort keys and values, double-buffered (4 x
// NOALLOC_MAX_PARTICLES words), the cells' run starts and ends (2 x
// NOALLOC_MAX_CELLS words), then the sort's temporary storage.  Every call on
// a GPU shares it: calls that overlap on one GPU (two streams) race.  The
// kernels clear what must start at zero (the run ends) on every call.
// ---------------------------------------------------------------------------
constexpr int64_t NOALLOC_MAX_PARTICLES = (int64_t)1 << 20;   // 1,048,576
constexpr int64_t NOALLOC_MAX_CELLS = (int64_t)1 << 22;       // 4,194,304 tile-padded cells (256x128x128)
constexpr size_t NOALLOC_SORT_OFFSET = (size_t)(4 * NOALLOC_MAX_PARTICLES + 2 * NOALLOC_MAX_CELLS) * 4;
constexpr size_t NOALLOC_SORT_BYTES = (size_t)8 << 20;
constexpr size_t NOALLOC_BYTES = NOALLOC_SORT_OFFSET + NOALLOC_SORT_BYTES;   // 56 MiB
static_assert(NOALLOC_SORT_OFFSET % 256 == 0, "sort storage must stay 256-byte aligned");

static __device__ __align__(256) unsigned char scratch[NOALLOC_BYTES];

std::mutex g_scratch_mu;
std::map<int, unsigned char*> g_scratch;   // device -> the array's address on that device

// The array's device address on the current device, looked up once per device
// (with that device current) and kept: the array's name in host code is only a
// host-side stand-in, never a device address.
int noalloc_scratch(unsigned char** out) {
  int dev = 0;
  EDGE_CHECK_HIP(hipGetDevice(&dev));
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  auto it = g_scratch.find(dev);
  if (it == g_scratch.end()) {
    void* p = nullptr;
    EDGE_CHECK_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    EDGE_REQUIRE(p != nullptr && ((uintptr_t)p & 255) == 0, "edge_lp_p2g_noalloc: scratch address %p unusable", p);
    it = g_scratch.emplace(dev, (unsigned char*)p).first;
  }
  *out = it->second;
  return EDGE_OK;
}

// sim/lp.py:158-181 update_particles_position: x += v dt, clamped to the domain.
__global__ void __launch_bounds__(256) k_lp_pos(CP3 vel, P3 pos, int64_t num, float dt, float Lx, float Ly, float Lz) {
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= num) return;
  const float L[3] = {Lx, Ly, Lz};
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    float x = pos.v[c][q] + vel.v[c][q] * dt;
    if (x < 0.f) x = 0.f;
    if (x > L[c]) x = L[c];
    pos.v[c][q] = x;
  }
}

// sim/lp.py:183-218 calculate_viscous_force on the three face grids: (1/Re)
// 7-point Laplacian with sample()'s clamped indices (a neighbour outside the
// grid reads the edge value).
__global__ void __launch_bounds__(256) k_lp_viscous(Grid g, Mac U, P3 vis, float Re) {
  int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  int a, i, j, k;
  int64_t l, nfa;
  if (!face_split(g, t, a, l, nfa, i, j, k)) return;
  const float* __restrict__ u = a == 0 ? U.ux : (a == 1 ? U.uy : U.uz);
  const int d0 = g.nx + (a == 0), d1 = g.ny + (a == 1), d2 = g.nz + (a == 2);
  auto at = [&](int ii, int jj, int kk) {
    ii = max(0, min(ii, d0 - 1));
    jj = max(0, min(jj, d1 - 1));
    kk = max(0, min(kk, d2 - 1));
    return u[((int64_t)ii * d1 + jj) * d2 + kk];
  };
  const float visc = 1.0f / Re;
  float lap = at(i + 1, j, k) + at(i - 1, j, k) + at(i, j + 1, k) + at(i, j - 1, k) + at(i, j, k + 1) +
              at(i, j, k - 1) + -6.0f * at(i, j, k);
  vis.v[a][l] = visc * lap / g.dx / g.dx;
}

// run_edge.py:258-262 calc_max_particle_speed: max(1e-3, max |v|) over all n
// particles (the float bits of a positive maximum order like its int bits).
__global__ void __launch_bounds__(256) k_lp_max_speed(CP3 vel, int64_t n, unsigned int* __restrict__ out) {
  float m = 1.e-3f;
  for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += (int64_t)gridDim.x * blockDim.x) {
    float a = vel.v[0][q], b = vel.v[1][q], c = vel.v[2][q];
    m = fmaxf(m, sqrtf(a * a + b * b + c * c));
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) m = fmaxf(m, __shfl_xor(m, off, 64));
  __shared__ float wm[4];
  if ((threadIdx.x & 63) == 0) wm[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) atomicMax(out, __float_as_uint(fmaxf(fmaxf(wm[0], wm[1]), fmaxf(wm[2], wm[3]))));
}

__global__ void k_lp_set_u32(unsigned int* p, unsigned int v) { *p = v; }


}  // namespace

extern "C" {

// vel/pos/acc/diff: 3 planes of n floats each (plane c at base + c*n).
int edge_lp_acc(const float* ux, const float* uy, const float* uz, int nx, int ny, int nz, const float* vel,
                const float* pos, float* acc, long n, long num, float dx, float dt, float St, float gravity,
                void* stream) {
  clear_error();
  EDGE_REQUIRE(dims_ok(nx, ny, nz), "edge_lp_acc: bad dims");
  EDGE_REQUIRE(ux && uy && uz && vel && pos && acc, "edge_lp_acc: null pointer");
  EDGE_REQUIRE(n >= 0 && num >= 0 && n < (1L << 31), "edge_lp_acc: bad particle counts");
  if (n == 0) return EDGE_OK;
  Grid g = make_grid(nx, ny, nz, dx);
  hipLaunchKernelGGL(k_lp_acc, dim3(blocks64(n)), dim3(256), 0, (hipStream_t)stream, g, Mac{ux, uy, uz},
                     CP3{{vel, vel + n, vel + 2 * n}}, CP3{{pos, pos + n, pos + 2 * n}}, P3{{acc, acc + n, acc + 2 * n}},
                     (int64_t)n, (int64_t)std::min(num, n), dt, St, gravity, cell_order(std::min((int64_t)num, (int64_t)n)));
  EDGE_CHECK_LAUNCH();
  return EDGE_OK;
}

int edge_lp_vel(const float* acc, float* vel, const float* pos, long n, long num, float dt, float dx, float Lx,
                float Ly, float Lz, void* stream) {
  clear_error();
  EDGE_REQUIRE(acc && vel && pos && n >= 0 && num >= 0 && n < (1L << 31), "edge_lp_vel: bad arguments");
  int64_t m = std::min(num, n);
  if (m == 0) return EDGE_OK;
  hipLaunchKernelGGL(k_lp_vel, dim3(blocks64(m)), dim3(256), 0, (hipStream_t)stream, CP3{{acc, acc + n, acc + 2 * n}},
                     P3{{vel, vel + n, vel + 2 * n}}, CP3{{pos, pos + n, pos + 2 * n}}, m, dt, dx, Lx, Ly, Lz);
  EDGE_CHECK_LAUNCH();
  return EDGE_OK;
}

int edge_lp_vel_diff(const float* ux, const float* uy, const float* uz, int nx, int ny, int nz, const float* vel,
                     float* diff, const float* pos, long n, long num, float dx, float St, float density_ratio,
                     float radius_ratio, void* stream) {
  clear_error();
  EDGE_REQUIRE(dims_ok(nx, ny, nz), "edge_lp_vel_diff: bad dims");
  EDGE_REQUIRE(ux && uy && uz && vel && diff && pos && n >= 0 && num >= 0 && n < (1L << 31),
               "edge_lp_vel_diff: bad arguments");
  int64_t m = std::min(num, n);
  if (m == 0) return EDGE_OK;
  Grid g = make_grid(nx, ny, nz, dx);
  hipLaunchKernelGGL(k_lp_vel_diff, dim3(blocks64(m)), dim3(256), 0, (hipStream_t)stream, g, Mac{ux, uy, uz},
                     CP3{{vel, vel + n, vel + 2 * n}}, P3{{diff, diff + n, diff + 2 * n}},
                     CP3{{pos, pos + n, pos + 2 * n}}, m, St, density_ratio, radius_ratio, cell_order(m));
  EDGE_CHECK_LAUNCH();
  return EDGE_OK;
}

// drag faces are zeroed here (the reference's fill(0.0) at the top of the kernel).
int edge_lp_p2g(float* dragx, float* dragy, float* dragz, const float* diff, const float* pos, const float* ref_radius,
                long n, long num, float dx, float radius, float Lx, float Ly, float Lz, int nx, int ny, int nz,
                void* stream) {
  clear_error();
  P2GCall c;
  int rc = p2g_prepare("edge_lp_p2g", dragx, dragy, dragz, diff, pos, ref_radius, n, num, dx, radius, Lx, Ly, Lz, nx,
                       ny, nz, stream, c);
  if (!rc) rc = p2g_zero(c);
  if (rc || c.m == 0) return rc;
  int dev = 0;
  EDGE_CHECK_HIP(hipGetDevice(&dev));
  P2GWork& w = g_p2g[dev];
  rc = p2g_workspace(w, c.m, c.nkeys, c.end_bit, c.st);
  if (rc) return rc;
  rc = p2g_launch(c, P2GBufs{w.k0, w.k1, w.v0, w.v1, w.start, w.end, w.temp, w.temp_bytes});
  if (rc) return rc;
  w.perm_n = c.m;
  return EDGE_OK;
}

int edge_lp_p2g_noalloc(float* dragx, float* dragy, float* dragz, const float* diff, const float* pos,
                        const float* ref_radius, long n, long num, float dx, float radius, float Lx, float Ly,
                        float Lz, int nx, int ny, int nz, void* stream) {
  clear_error();
  P2GCall c;
  int rc = p2g_prepare("edge_lp_p2g_noalloc", dragx, dragy, dragz, diff, pos, ref_radius, n, num, dx, radius, Lx, Ly,
                       Lz, nx, ny, nz, stream, c);
  if (rc) return rc;
  EDGE_REQUIRE(c.m <= NOALLOC_MAX_PARTICLES,
               "edge_lp_p2g_noalloc: %lld particles exceed its maximum of %lld (use edge_lp_p2g)", (long long)c.m,
               (long long)NOALLOC_MAX_PARTICLES);
  EDGE_REQUIRE((int64_t)c.nkeys <= NOALLOC_MAX_CELLS,
               "edge_lp_p2g_noalloc: the grid's %u tile-padded cells exceed its maximum of %lld (use edge_lp_p2g)",
               c.nkeys, (long long)NOALLOC_MAX_CELLS);
  size_t need = 0;
  EDGE_CHECK_HIP(rocprim::radix_sort_pairs(nullptr, need, (uint32_t*)nullptr, (uint32_t*)nullptr, (uint32_t*)nullptr,
                                           (uint32_t*)nullptr, (size_t)c.m, 0u, (unsigned)c.end_bit, c.st));
  EDGE_REQUIRE(need <= NOALLOC_SORT_BYTES, "edge_lp_p2g_noalloc: the sort needs %zu bytes of its %zu", need,
               (size_t)NOALLOC_SORT_BYTES);
  unsigned char* base = nullptr;
  rc = noalloc_scratch(&base);
  if (!rc) rc = p2g_zero(c);
  if (rc || c.m == 0) return rc;
  uint32_t* u = (uint32_t*)base;
  const P2GBufs b{u, u + NOALLOC_MAX_PARTICLES, u + 2 * NOALLOC_MAX_PARTICLES, u + 3 * NOALLOC_MAX_PARTICLES,
                  u + 4 * NOALLOC_MAX_PARTICLES, u + 4 * NOALLOC_MAX_PARTICLES + NOALLOC_MAX_CELLS,
                  (void*)(base + NOALLOC_SORT_OFFSET), (size_t)NOALLOC_SORT_BYTES};
  return p2g_launch(c, b);
}

int edge_lp_pos(const float* vel, float* pos, long n, long num, float dt, float Lx, float Ly, float Lz, void* stream) {
  clear_error();
  EDGE_REQUIRE(vel && pos && n >= 0 && num >= 0 && n < (1L << 31), "edge_lp_pos: bad arguments");
  int64_t m = std::min(num, n);
  if (m == 0) return EDGE_OK;
  hipLaunchKernelGGL(k_lp_pos, dim3(blocks64(m)), dim3(256), 0, (hipStream_t)stream, CP3{{vel, vel + n, vel + 2 * n}},
                     P3{{pos, pos + n, pos + 2 * n}}, m, dt, Lx, Ly, Lz);
  EDGE_CHECK_LAUNCH();
  return EDGE_OK;
}

int edge_lp_viscous(const float* ux, const float* uy, const float* uz, float* visx, float* visy, float* visz, int nx,
                    int ny, int nz, float dx, float Re, void* stream) {
  clear_error();
  EDGE_REQUIRE(dims_ok(nx, ny, nz), "edge_lp_viscous: bad dims");
  EDGE_REQUIRE(ux && uy && uz && visx && visy && visz && Re != 0.0f, "edge_lp_viscous: bad arguments");
  Grid g = make_grid(nx, ny, nz, dx);
  int64_t nf = (int64_t)g.nf[0] + g.nf[1] + g.nf[2];
  hipLaunchKernelGGL(k_lp_viscous, dim3(blocks64(nf)), dim3(256), 0, (hipStream_t)stream, g, Mac{ux, uy, uz},
                     P3{{visx, visy, visz}}, Re);
  EDGE_CHECK_LAUNCH();
  return EDGE_OK;
}

// out_dev: one device float (max(1e-3, max |v|) over the n particles).
int edge_lp_max_speed(const float* vel, long n, float* out_dev, void* stream) {
  clear_error();
  EDGE_REQUIRE(vel && out_dev && n >= 0 && n < (1L << 31), "edge_lp_max_speed: bad arguments");
  hipStream_t st = (hipStream_t)stream;
  const float floor_speed = 1.e-3f;
  unsigned int bits;
  std::memcpy(&bits, &floor_speed, sizeof bits);
  hipLaunchKernelGGL(k_lp_set_u32, dim3(1), dim3(1), 0, st, (unsigned int*)out_dev, bits);
  EDGE_CHECK_LAUNCH();
  if (n == 0) return EDGE_OK;
  int nb = (int)std::min<int64_t>(blocks64(n), 1024);
  hipLaunchKernelGGL(k_lp_max_speed, dim3(nb), dim3(256), 0, st, CP3{{vel, vel + n, vel + 2 * n}}, (int64_t)n,
                     (unsigned int*)out_dev);
  EDGE_CHECK_LAUNCH();
  return EDGE_OK;
}

}  // extern "C"
