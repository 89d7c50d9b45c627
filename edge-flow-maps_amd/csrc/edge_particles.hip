// Laden particles (sim/lp.py): gather of the MAC velocity at the particles,
// implicit drag, wall rules, the drag's scatter back onto the faces (P2G) and
// the viscous force.  One thread per particle (per face for the viscous
// force).  Particle vectors are SoA: 3 planes of n (the capacity); only
// particles q < num are live, as in the reference.
#include <algorithm>
#include <cstring>

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
__global__ void __launch_bounds__(256) k_lp_acc(Grid g, Mac U, CP3 vel, CP3 pos, P3 acc, int64_t n, int64_t num,
                                               float dt, float St, float gravity) {
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
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
                                                    float density_ratio, float radius_ratio) {
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= num) return;
  const float er = 1.0f / (radius_ratio * radius_ratio * radius_ratio);
  const float k = 1.0f / St * density_ratio * er;
  float u[3];
  particle_u(g, U, pos, q, u);
#pragma unroll
  for (int c = 0; c < 3; ++c) diff.v[c][q] = k * (vel.v[c][q] - u[c]);
}

// sim/lp.py:100-156 laden_particles_back_with_computing_particle: the drag
// diff (r_p / r)^3 spread onto face grid a with the quadratic B-spline
// weights; particles within 3 dx of a wall are skipped.  Such a particle's
// three non-zero taps per axis are inside the grid, so no bounds test binds.
// Accumulated with FP32 atomics (summation order differs from a sequential
// loop by rounding only).  The faces must be zeroed first.
__global__ void __launch_bounds__(256) k_lp_p2g(Grid g, P3 drag, CP3 diff, CP3 pos, const float* __restrict__ rr,
                                               int64_t num, float radius, float Lx, float Ly, float Lz) {
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= num) return;
  const float dx = g.dx;
  float P[3] = {pos.v[0][q], pos.v[1][q], pos.v[2][q]};
  // the reference's skip test, written so that a NaN position is skipped too
  if (!(P[0] >= 3.f * dx && P[1] >= 3.f * dx && P[2] >= 3.f * dx && P[0] <= Lx - 3.f * dx &&
        P[1] <= Ly - 3.f * dx && P[2] <= Lz - 3.f * dx))
    return;
  const float r = rr[q];
  const float scale = r * r * r / (radius * radius * radius);
  float s[3] = {P[0] / dx, P[1] / dx, P[2] / dx};
  const uint32_t ny = g.ny, nz = g.nz;
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    // face grid a: offset 0 on axis a, 0.5 on the others
    Axis3 A = bspline_axis(s[0] - (a == 0 ? 0.f : 0.5f), 1e30f);
    Axis3 B = bspline_axis(s[1] - (a == 1 ? 0.f : 0.5f), 1e30f);
    Axis3 C = bspline_axis(s[2] - (a == 2 ? 0.f : 0.5f), 1e30f);
    const uint32_t d1 = ny + (a == 1), d2 = nz + (a == 2);
    const float d = diff.v[a][q] * scale;
    float* __restrict__ f = drag.v[a];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int j = 0; j < 3; ++j) {
        const float wij = A.w[i] * B.w[j];
        float* row = f + ((uint32_t)(A.i0 + i) * d1 + (uint32_t)(B.i0 + j)) * d2 + (uint32_t)C.i0;
#pragma unroll
        for (int k = 0; k < 3; ++k) atomicAdd(row + k, d * (wij * C.w[k]));
      }
  }
}

__global__ void __launch_bounds__(256) k_lp_zero(float* __restrict__ f, int64_t n) {
  int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) f[t] = 0.f;
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

int64_t blocks64(int64_t n) { return (n + 255) / 256; }

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
                     (int64_t)n, (int64_t)std::min(num, n), dt, St, gravity);
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
                     CP3{{pos, pos + n, pos + 2 * n}}, m, St, density_ratio, radius_ratio);
  EDGE_CHECK_LAUNCH();
  return EDGE_OK;
}

// drag faces are zeroed here (the reference's fill(0.0) at the top of the kernel).
int edge_lp_p2g(float* dragx, float* dragy, float* dragz, const float* diff, const float* pos, const float* ref_radius,
                long n, long num, float dx, float radius, float Lx, float Ly, float Lz, int nx, int ny, int nz,
                void* stream) {
  clear_error();
  EDGE_REQUIRE(dims_ok(nx, ny, nz), "edge_lp_p2g: bad dims");
  EDGE_REQUIRE(dragx && dragy && dragz && diff && pos && ref_radius && n >= 0 && num >= 0 && n < (1L << 31),
               "edge_lp_p2g: bad arguments");
  EDGE_REQUIRE(radius != 0.0f && dx > 0.0f, "edge_lp_p2g: radius and dx must be non-zero");
  // the 3 dx wall margin is what keeps every tap in the grid: the domain must match the grid
  EDGE_REQUIRE(Lx <= (float)nx * dx * 1.0001f && Ly <= (float)ny * dx * 1.0001f && Lz <= (float)nz * dx * 1.0001f,
               "edge_lp_p2g: domain (%g, %g, %g) larger than the grid (%d, %d, %d) x %g", Lx, Ly, Lz, nx, ny, nz, dx);
  hipStream_t st = (hipStream_t)stream;
  Grid g = make_grid(nx, ny, nz, dx);
  float* dr[3] = {dragx, dragy, dragz};
  for (int a = 0; a < 3; ++a)
    hipLaunchKernelGGL(k_lp_zero, dim3(blocks64(g.nf[a])), dim3(256), 0, st, dr[a], (int64_t)g.nf[a]);
  EDGE_CHECK_LAUNCH();
  int64_t m = std::min(num, n);
  if (m == 0) return EDGE_OK;
  hipLaunchKernelGGL(k_lp_p2g, dim3(blocks64(m)), dim3(256), 0, st, g, P3{{dragx, dragy, dragz}},
                     CP3{{diff, diff + n, diff + 2 * n}}, CP3{{pos, pos + n, pos + 2 * n}}, ref_radius, m, radius, Lx,
                     Ly, Lz);
  EDGE_CHECK_LAUNCH();
  return EDGE_OK;
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
