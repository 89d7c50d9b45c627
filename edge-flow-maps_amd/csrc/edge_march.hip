// Backward-map marches (sim/hermite.py) and the forward face march
// (run_edge.py:330-357), one thread per cell or face, everything between the
// reads of the old map and the writes of the new one kept in registers.
#include <algorithm>

#include "edge_common.h"

using namespace edge;

namespace {

struct MapOut {
  float* psi;
  float* T[3];
  float* dT[3];
  float* d2T[3];
};

__device__ __forceinline__ void write_map_row(const MapOut& o, int64_t N, int64_t q, const float psi[3],
                                              const float T[3][3]) {
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    o.psi[c * N + q] = psi[c];
#pragma unroll
    for (int b = 0; b < 3; ++b) o.T[c][b * N + q] = T[c][b];
  }
}

// sim/hermite.py:1426-1492 advect_psi_garm_ed4_fusion_unroll_kernel.
// Point 0 is the cell centre; points 1..4 the tetrahedral eps-points
// (-,-,-), (+,+,-), (-,+,+), (+,-,+) of sim/hermite.py:1446-1449.  Each point is
// RK4-marched with its Jacobian (-dt), composed with the old map, and the four
// eps-points give dT by calc_ed4_garm_func (sim/hermite.py:1049-1075).
// With tabn non-null the low planes {psi_c, T_c} of the new map's garm table
// are written too (the words k_pack_table would write from the stored map);
// the high planes {dT_c, d2T_c} need the neighbours' dT and are k_pack_hi's
// (edge_tables.hip).
__global__ void __launch_bounds__(256) EDGE_MARCH_ATTR k_march_edge(Grid g, Mac U, const float4* __restrict__ tab,
                                                    const float* __restrict__ X, MapOut o, float dt, float eps,
                                                    float4* __restrict__ tabn) {
  const int64_t N = ncell(g);
  const int plane = g.ny * g.nz;
  const int64_t q =
      (int64_t)band_block(blockIdx.x, gridDim.x, plane % 256 ? 0 : plane / 256, EDGE_BAND) * blockDim.x + threadIdx.x;
  if (q >= N) return;
  const float x0[3] = {X[q], X[N + q], X[2 * N + q]};
  // dT accumulators: dT[c][0] = sum s_z T[c][1] + s_y T[c][2], etc.
  float acc[3][3] = {{0.f, 0.f, 0.f}, {0.f, 0.f, 0.f}, {0.f, 0.f, 0.f}};
#pragma unroll 1
  for (int pt = 0; pt < 5; ++pt) {
    // sign pattern of point pt (pt = 0: centre)
    float sx = 0.f, sy = 0.f, sz = 0.f;
    if (pt == 1) { sx = -1.f; sy = -1.f; sz = -1.f; }
    if (pt == 2) { sx = 1.f; sy = 1.f; sz = -1.f; }
    if (pt == 3) { sx = -1.f; sy = 1.f; sz = 1.f; }
    if (pt == 4) { sx = 1.f; sy = -1.f; sz = 1.f; }
    float p[3] = {x0[0] + sx * eps, x0[1] + sy * eps, x0[2] + sz * eps};
    if (pt == 0) { p[0] = x0[0]; p[1] = x0[1]; p[2] = x0[2]; }
    float T[3][3] = {{1.f, 0.f, 0.f}, {0.f, 1.f, 0.f}, {0.f, 0.f, 1.f}};
    rk4_psi_T(g, U, p, T, -dt);
    compose(g, tab, p, T);
    if (pt == 0) {
      write_map_row(o, N, q, p, T);
      if (tabn) {
#pragma unroll
        for (int c = 0; c < 3; ++c) tabn[2 * c * N + q] = make_float4(p[c], T[c][0], T[c][1], T[c][2]);
      }
    } else {
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        acc[c][0] += sz * T[c][1] + sy * T[c][2];
        acc[c][1] += sz * T[c][0] + sx * T[c][2];
        acc[c][2] += sy * T[c][0] + sx * T[c][1];
      }
    }
  }
#pragma unroll
  for (int c = 0; c < 3; ++c)
#pragma unroll
    for (int b = 0; b < 3; ++b) o.dT[c][b * N + q] = acc[c][b] / 8.0f / eps;
}

// sim/hermite.py:1290-1338 advect_ps_hermite_ed8 fused: 8 points X +- eps
// (bit 0 = x sign, bit 1 = y, bit 2 = z; sim/hermite.py:704-725), each
// advance_map_0 (-dt) then interp_psi_hermite through the old map (full
// Hermite), then calc_ed8_3 (sim/hermite.py:1098-1151).
__global__ void __launch_bounds__(256) EDGE_ED8_ATTR k_march_ed8(Grid g, Mac U, const float4* __restrict__ tab,
                                                   const float* __restrict__ X, MapOut o, float dt, float eps) {
  const int64_t N = ncell(g);
  const int64_t q = (int64_t)xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
  if (q >= N) return;
  const float x0[3] = {X[q], X[N + q], X[2 * N + q]};
  // per component: sum, x, y, z, yz, zx, xy, xyz moments
  float m[3][8];
#pragma unroll
  for (int c = 0; c < 3; ++c)
#pragma unroll
    for (int t = 0; t < 8; ++t) m[c][t] = 0.f;
#pragma unroll 1
  for (int e = 0; e < 8; ++e) {
    float sx = (e & 1) ? 1.f : -1.f, sy = (e & 2) ? 1.f : -1.f, sz = (e & 4) ? 1.f : -1.f;
    float p[3] = {x0[0] + sx * eps, x0[1] + sy * eps, x0[2] + sz * eps};
    rk4_psi(g, U, p, -dt);
    float v[3];
    map_eval<false>(g, tab, p, v, nullptr);
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      m[c][0] += v[c];
      m[c][1] += sx * v[c];
      m[c][2] += sy * v[c];
      m[c][3] += sz * v[c];
      m[c][4] += (sy * sz) * v[c];
      m[c][5] += (sz * sx) * v[c];
      m[c][6] += (sx * sy) * v[c];
      m[c][7] += (sx * sy * sz) * v[c];
    }
  }
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    o.psi[c * N + q] = m[c][0] / 8.0f;
    o.T[c][q] = m[c][1] / 8.0f / eps;
    o.T[c][N + q] = m[c][2] / 8.0f / eps;
    o.T[c][2 * N + q] = m[c][3] / 8.0f / eps;
    o.dT[c][q] = m[c][4] / 8.0f / eps / eps;
    o.dT[c][N + q] = m[c][5] / 8.0f / eps / eps;
    o.dT[c][2 * N + q] = m[c][6] / 8.0f / eps / eps;
    o.d2T[c][q] = m[c][7] / 8.0f / eps / eps / eps;
  }
}

// sim/hermite.py:1398-1421 advect_ps_garm up to the copies: reset_to_identity_1,
// advance_map_1_transpose (-dt), interp_psi_hermite_garm (full Hermite).
// dT / d2T follow by calc_fd_3_from_grad (edge_calc_fd3).
__global__ void __launch_bounds__(256) EDGE_MARCH_ATTR k_march_ge(Grid g, Mac U, const float4* __restrict__ tab,
                                                  const float* __restrict__ X, MapOut o, float dt) {
  const int64_t N = ncell(g);
  const int64_t q = (int64_t)xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
  if (q >= N) return;
  float p[3] = {X[q], X[N + q], X[2 * N + q]};
  float T[3][3] = {{1.f, 0.f, 0.f}, {0.f, 1.f, 0.f}, {0.f, 0.f, 1.f}};
  rk4_psi_T(g, U, p, T, -dt);
  compose(g, tab, p, T);
  write_map_row(o, N, q, p, T);
}

// sim/hermite.py:1375-1395 advect_ps_hermite_ed4_fd_unroll fused: 4 points,
// advance_map_0 (-dt), interp_psi_hermite_fd_unroll (table built from derived
// dT/d2T), calc_ed4_1 (sim/hermite.py:1154-1177).
__global__ void __launch_bounds__(256) k_march_ed4(Grid g, Mac U, const float4* __restrict__ tab,
                                                   const float* __restrict__ X, MapOut o, float dt, float eps) {
  const int64_t N = ncell(g);
  const int64_t q = (int64_t)xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
  if (q >= N) return;
  const float x0[3] = {X[q], X[N + q], X[2 * N + q]};
  float m[3][4];
#pragma unroll
  for (int c = 0; c < 3; ++c)
#pragma unroll
    for (int t = 0; t < 4; ++t) m[c][t] = 0.f;
#pragma unroll 1
  for (int e = 0; e < 4; ++e) {
    // 000 011 110 101 -> (-,-,-), (+,+,-), (-,+,+), (+,-,+)
    float sx = (e == 1 || e == 3) ? 1.f : -1.f;
    float sy = (e == 1 || e == 2) ? 1.f : -1.f;
    float sz = (e >= 2) ? 1.f : -1.f;
    float p[3] = {x0[0] + sx * eps, x0[1] + sy * eps, x0[2] + sz * eps};
    rk4_psi(g, U, p, -dt);
    float v[3];
    map_eval<false>(g, tab, p, v, nullptr);
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      m[c][0] += v[c];
      m[c][1] += sx * v[c];
      m[c][2] += sy * v[c];
      m[c][3] += sz * v[c];
    }
  }
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    o.psi[c * N + q] = m[c][0] / 4.0f;
    o.T[c][q] = m[c][1] / 4.0f / eps;
    o.T[c][N + q] = m[c][2] / 4.0f / eps;
    o.T[c][2 * N + q] = m[c][3] / 4.0f / eps;
  }
}

// run_edge.py:330-357 march_phi_grid: RK4 (+dt) of face points phi_a and column F_a.
struct FaceMap {
  float* phi[3];
  float* F[3];
};

__global__ void __launch_bounds__(256) EDGE_FACES_ATTR k_advance_faces(Grid g, Mac U, FaceMap fm, float dt) {
  const int64_t nf0 = (int64_t)(g.nx + 1) * g.ny * g.nz;
  const int64_t nf1 = (int64_t)g.nx * (g.ny + 1) * g.nz;
  const int64_t nf2 = (int64_t)g.nx * g.ny * (g.nz + 1);
  int64_t t = (int64_t)xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
  int a;
  int64_t nf;
  if (t < nf0) { a = 0; nf = nf0; }
  else if (t < nf0 + nf1) { a = 1; t -= nf0; nf = nf1; }
  else if (t < nf0 + nf1 + nf2) { a = 2; t -= nf0 + nf1; nf = nf2; }
  else return;
  float* phi = fm.phi[a];
  float* F = fm.F[a];
  float p[3] = {phi[t], phi[nf + t], phi[2 * nf + t]};
  float f[3] = {F[t], F[nf + t], F[2 * nf + t]};
  rk4_phi_F(g, U, p, f, dt);
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    phi[c * nf + t] = p[c];
    F[c * nf + t] = f[c];
  }
}

// A sequence of velocity fields for the EFM backtrack, passed by value (one
// launch takes up to EFM_SEQ of them; longer sequences chain launches).
constexpr int EFM_SEQ = 32;
struct VelSeq {
  const float* ux[EFM_SEQ];
  const float* uy[EFM_SEQ];
  const float* uz[EFM_SEQ];
  float dt[EFM_SEQ];
  int n;
};

// run_efm.py:170-185 backtrack_psi_grid fused: every face point and its column
// (from the identity when `from_identity`, else from psi/T in memory) marched
// by advance_map through velocity 0, 1, .., n-1 with dt[s], kept in registers
// between the marches and written once.  The same per-face operations, in the
// same order, as one k_reset_faces and n k_advance_faces launches.
__global__ void __launch_bounds__(256) EDGE_FACES_ATTR k_backtrack_faces(Grid g, VelSeq vs, FaceMap fm,
                                                                        int from_identity) {
  const int64_t nf0 = (int64_t)(g.nx + 1) * g.ny * g.nz;
  const int64_t nf1 = (int64_t)g.nx * (g.ny + 1) * g.nz;
  const int64_t nf2 = (int64_t)g.nx * g.ny * (g.nz + 1);
  int64_t t = (int64_t)xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
  int a;
  int64_t nf;
  if (t < nf0) { a = 0; nf = nf0; }
  else if (t < nf0 + nf1) { a = 1; t -= nf0; nf = nf1; }
  else if (t < nf0 + nf1 + nf2) { a = 2; t -= nf0 + nf1; nf = nf2; }
  else return;
  float* phi = fm.phi[a];
  float* F = fm.F[a];
  float p[3], f[3];
  if (from_identity) {
    int i, j, k;
    face_ijk(g, a, (uint32_t)t, i, j, k);
    face_coord(a, i, j, k, g.dx, p);
    f[0] = a == 0 ? 1.f : 0.f;
    f[1] = a == 1 ? 1.f : 0.f;
    f[2] = a == 2 ? 1.f : 0.f;
  } else {
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      p[c] = phi[c * nf + t];
      f[c] = F[c * nf + t];
    }
  }
#pragma unroll 1
  for (int s = 0; s < vs.n; ++s) rk4_phi_F(g, Mac{vs.ux[s], vs.uy[s], vs.uz[s]}, p, f, vs.dt[s]);
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    phi[c * nf + t] = p[c];
    F[c * nf + t] = f[c];
  }
}

// Generic point marches on n points (3 planes of n): advance_map_0 (mode 0) or
// advance_map_1_transpose (mode 1; T_c = 3 planes of n holding row c).
struct Rows {
  float* T[3];
};

__device__ __forceinline__ void load_rows(const Rows& R, int64_t n, int64_t t, float TT[3][3]) {
#pragma unroll
  for (int a = 0; a < 3; ++a)
#pragma unroll
    for (int b = 0; b < 3; ++b) TT[a][b] = R.T[a][b * n + t];
}
__device__ __forceinline__ void store_rows(const Rows& R, int64_t n, int64_t t, const float TT[3][3]) {
#pragma unroll
  for (int a = 0; a < 3; ++a)
#pragma unroll
    for (int b = 0; b < 3; ++b) R.T[a][b * n + t] = TT[a][b];
}

__global__ void __launch_bounds__(256) k_advance_points(Grid g, Mac U, float* __restrict__ pts, Rows R, int64_t n,
                                                        float dt, int mode) {
  int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  float p[3] = {pts[t], pts[n + t], pts[2 * n + t]};
  if (mode == 0) {
    rk4_psi(g, U, p, dt);
  } else {
    float TT[3][3];
    load_rows(R, n, t, TT);
    rk4_psi_T(g, U, p, TT, dt);
    store_rows(R, n, t, TT);
  }
  pts[t] = p[0];
  pts[n + t] = p[1];
  pts[2 * n + t] = p[2];
}

// Passive dye particles (no reference counterpart: the 512x256x256 torus
// workload's 16M dye particles, which only follow the flow).  Each live
// particle q < num moves like a point of the forward face map, the RK4 of
// run_edge.py:336-357 with +dt and the B-spline MAC velocity at every stage
// (G2P), and is then clamped into the domain [0, L] as sim/lp.py:158-181
// clamps laden particles.  pos: 3 planes of n (the capacity).
__global__ void __launch_bounds__(256) k_dye_advance(Grid g, Mac U, float* __restrict__ pos, int64_t n, int64_t num,
                                                     float dt, float Lx, float Ly, float Lz) {
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= num) return;
  float p[3] = {pos[q], pos[n + q], pos[2 * n + q]};
  rk4_psi(g, U, p, dt);
  const float L[3] = {Lx, Ly, Lz};
#pragma unroll
  for (int a = 0; a < 3; ++a) pos[a * n + q] = fminf(fmaxf(p[a], 0.f), L[a]);
}

// Compose n points (and their rows T, if given) with a map table
// (sim/hermite.py:921-1046).
__global__ void __launch_bounds__(256) k_compose_points(Grid g, const float4* __restrict__ tab,
                                                        float* __restrict__ pts, Rows R, int64_t n, int with_T) {
  int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  float p[3] = {pts[t], pts[n + t], pts[2 * n + t]};
  if (with_T) {
    float TT[3][3];
    load_rows(R, n, t, TT);
    compose(g, tab, p, TT);
    store_rows(R, n, t, TT);
  } else {
    float v[3];
    map_eval<false>(g, tab, p, v, nullptr);
    p[0] = v[0];
    p[1] = v[1];
    p[2] = v[2];
  }
  pts[t] = p[0];
  pts[n + t] = p[1];
  pts[2 * n + t] = p[2];
}

// sim/hermite.py:1049-1075 calc_ed4_garm on stored point Jacobians:
// P[p].T[c] = rows of point p (p = 0..3: (-,-,-), (+,+,-), (-,+,+), (+,-,+)).
struct Rows4 {
  const float* T[4][3];
};
__global__ void __launch_bounds__(256) k_calc_ed4_garm(Rows4 P, Rows O, int64_t n, float eps) {
  int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const float sx[4] = {-1.f, 1.f, -1.f, 1.f}, sy[4] = {-1.f, 1.f, 1.f, -1.f}, sz[4] = {-1.f, -1.f, 1.f, 1.f};
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    float a0 = 0.f, a1 = 0.f, a2 = 0.f;
#pragma unroll
    for (int p = 0; p < 4; ++p) {
      float t0 = P.T[p][c][t], t1 = P.T[p][c][n + t], t2 = P.T[p][c][2 * n + t];
      a0 += sz[p] * t1 + sy[p] * t2;
      a1 += sz[p] * t0 + sx[p] * t2;
      a2 += sy[p] * t0 + sx[p] * t1;
    }
    O.T[c][t] = a0 / 8.0f / eps;
    O.T[c][n + t] = a1 / 8.0f / eps;
    O.T[c][2 * n + t] = a2 / 8.0f / eps;
  }
}

// sim/hermite.py:1098-1151 calc_ed8_3 / 1154-1177 calc_ed4_1 on stored point
// fields (np = 8 or 4).  ED8 point e has x sign bit 0, y bit 1, z bit 2; ED4
// points are (-,-,-), (+,+,-), (-,+,+), (+,-,+).
struct Pts8 {
  const float* P[8];
};
__global__ void __launch_bounds__(256) k_calc_ed(Pts8 S, int np, MapOut o, int64_t n, float eps) {
  int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  float inv = (float)np;
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    float m[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    for (int e = 0; e < np; ++e) {
      float sx, sy, sz;
      if (np == 8) {
        sx = (e & 1) ? 1.f : -1.f; sy = (e & 2) ? 1.f : -1.f; sz = (e & 4) ? 1.f : -1.f;
      } else {
        sx = (e == 1 || e == 3) ? 1.f : -1.f; sy = (e == 1 || e == 2) ? 1.f : -1.f; sz = (e >= 2) ? 1.f : -1.f;
      }
      float v = S.P[e][c * n + t];
      m[0] += v; m[1] += sx * v; m[2] += sy * v; m[3] += sz * v;
      m[4] += (sy * sz) * v; m[5] += (sz * sx) * v; m[6] += (sx * sy) * v; m[7] += (sx * sy * sz) * v;
    }
    o.psi[c * n + t] = m[0] / inv;
    o.T[c][t] = m[1] / inv / eps;
    o.T[c][n + t] = m[2] / inv / eps;
    o.T[c][2 * n + t] = m[3] / inv / eps;
    if (np == 8) {
      o.dT[c][t] = m[4] / inv / eps / eps;
      o.dT[c][n + t] = m[5] / inv / eps / eps;
      o.dT[c][2 * n + t] = m[6] / inv / eps / eps;
      o.d2T[c][t] = m[7] / inv / eps / eps / eps;
    }
  }
}

// dst = src + (ox, oy, oz) on a 3-plane field (the eps-point resets of
// sim/hermite.py:704-783).
__global__ void __launch_bounds__(256) k_offset_points(const float* __restrict__ src, float* __restrict__ dst, int64_t n,
                                                       float ox, float oy, float oz) {
  int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  dst[t] = src[t] + ox;
  dst[n + t] = src[n + t] + oy;
  dst[2 * n + t] = src[2 * n + t] + oz;
}

// Point-evaluation hooks (tests): pts (n,3) interleaved as the oracle's.
// what 0: B-spline value(+grad) of one field with BL; 1: interp_u_MAC(_grad);
// 2: map value(+grad) from a table.
__global__ void __launch_bounds__(256) k_eval_points(Grid g, int what, const float* __restrict__ f, int d0, int d1,
                                                     int d2, float bl0, float bl1, float bl2, Mac U,
                                                     const float4* __restrict__ tab, const float* __restrict__ pts,
                                                     int64_t n, float* __restrict__ val, float* __restrict__ grad) {
  int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  float p[3] = {pts[3 * t], pts[3 * t + 1], pts[3 * t + 2]};
  if (what == 0) {
    Axis3 A = bspline_axis(p[0] * g.idx - bl0, (float)((double)d0 - 2.0 - 1e-6));
    Axis3 B = bspline_axis(p[1] * g.idx - bl1, (float)((double)d1 - 2.0 - 1e-6));
    Axis3 C = bspline_axis(p[2] * g.idx - bl2, (float)((double)d2 - 2.0 - 1e-6));
    float gr[3];
    if (grad) {
      val[t] = bspline_eval<true>(make_rsrc(f, 4u * d0 * d1 * d2), (uint32_t)d1 * d2, d2, A, B, C, gr);
      float inv = g.idx;
      grad[3 * t] = gr[0] * inv;
      grad[3 * t + 1] = gr[1] * inv;
      grad[3 * t + 2] = gr[2] * inv;
    } else {
      val[t] = bspline_eval<false>(make_rsrc(f, 4u * d0 * d1 * d2), (uint32_t)d1 * d2, d2, A, B, C, gr);
    }
  } else if (what == 1) {
    float u[3], J[3][3];
    if (grad) {
      interp_u_mac<true>(g, U, p, u, J);
#pragma unroll
      for (int a = 0; a < 3; ++a)
#pragma unroll
        for (int b = 0; b < 3; ++b) grad[9 * t + 3 * a + b] = J[a][b];
    } else {
      interp_u_mac<false>(g, U, p, u, J);
    }
#pragma unroll
    for (int a = 0; a < 3; ++a) val[3 * t + a] = u[a];
  } else {
    float v[3], G[3][3];
    if (grad) {
      map_eval<true>(g, tab, p, v, G);
#pragma unroll
      for (int a = 0; a < 3; ++a)
#pragma unroll
        for (int b = 0; b < 3; ++b) grad[9 * t + 3 * a + b] = G[a][b];
    } else {
      map_eval<false>(g, tab, p, v, G);
    }
#pragma unroll
    for (int a = 0; a < 3; ++a) val[3 * t + a] = v[a];
  }
}

// reset_to_identity_1/2/3 (sim/hermite.py:642-701): psi = X, T = I, dT = 0, d2T = 0
// (only the given outputs are written).
__global__ void __launch_bounds__(256) k_reset_identity(Grid g, const float* __restrict__ X, MapOut o) {
  const int64_t N = ncell(g);
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= N) return;
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    if (o.psi) o.psi[c * N + q] = X[c * N + q];
#pragma unroll
    for (int b = 0; b < 3; ++b) {
      if (o.T[c]) o.T[c][b * N + q] = (b == c) ? 1.f : 0.f;
      if (o.dT[c]) o.dT[c][b * N + q] = 0.f;
    }
    if (o.d2T[c]) o.d2T[c][q] = 0.f;
  }
}

// run_edge.py:306-320 reset_to_identity on the faces: phi_a = X_a, F_a = e_a.
__global__ void __launch_bounds__(256) k_reset_faces(Grid g, FaceMap fm) {
  const int64_t nf0 = (int64_t)(g.nx + 1) * g.ny * g.nz;
  const int64_t nf1 = (int64_t)g.nx * (g.ny + 1) * g.nz;
  const int64_t nf2 = (int64_t)g.nx * g.ny * (g.nz + 1);
  int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  int a;
  int64_t nf;
  if (t < nf0) { a = 0; nf = nf0; }
  else if (t < nf0 + nf1) { a = 1; t -= nf0; nf = nf1; }
  else if (t < nf0 + nf1 + nf2) { a = 2; t -= nf0 + nf1; nf = nf2; }
  else return;
  int d[3];
  face_dims(g, a, d);
  int i, j, k;
  face_ijk(g, a, (uint32_t)t, i, j, k);
  float Xf[3];
  face_coord(a, i, j, k, g.dx, Xf);
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    if (fm.phi[a]) fm.phi[a][c * nf + t] = Xf[c];
    if (fm.F[a]) fm.F[a][c * nf + t] = (c == a) ? 1.f : 0.f;
  }
}

// util/taichi_utils.py:23-44 coordinate generators (mode 0 cells, 1..3 faces x/y/z).
__global__ void __launch_bounds__(256) k_coords(Grid g, int mode, float* __restrict__ out) {
  int d[3];
  if (mode == 0) { d[0] = g.nx; d[1] = g.ny; d[2] = g.nz; }
  else face_dims(g, mode - 1, d);
  const int64_t n = (int64_t)d[0] * d[1] * d[2];
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  int k = (int)(t % d[2]);
  int j = (int)((t / d[2]) % d[1]);
  int i = (int)(t / ((int64_t)d[1] * d[2]));
  float Xf[3];
  if (mode == 0) {
    Xf[0] = ((float)i + 0.5f) * g.dx;
    Xf[1] = ((float)j + 0.5f) * g.dx;
    Xf[2] = ((float)k + 0.5f) * g.dx;
  } else {
    face_coord(mode - 1, i, j, k, g.dx, Xf);
  }
  out[t] = Xf[0];
  out[n + t] = Xf[1];
  out[2 * n + t] = Xf[2];
}

}  // namespace

// ---------------------------------------------------------------------------
// C-ABI
// ---------------------------------------------------------------------------
extern "C" {

#define EDGE_MAC_OK(ux, uy, uz) EDGE_REQUIRE((ux) && (uy) && (uz), "null velocity pointer")

int edge_march_edge(const float* ux, const float* uy, const float* uz, const float* tab_old, const float* X,
                    float* psi, float* Tx, float* Ty, float* Tz, float* dTx, float* dTy, float* dTz, int nx, int ny,
                    int nz, float dx, float dt, float eps, void* stream) {
  clear_error();
  EDGE_REQUIRE(dims_ok(nx, ny, nz), "edge_march_edge: bad dims");
  EDGE_MAC_OK(ux, uy, uz);
  EDGE_REQUIRE(tab_old && X && psi && Tx && Ty && Tz && dTx && dTy && dTz, "edge_march_edge: null pointer");
  EDGE_REQUIRE(eps != 0.0f, "edge_march_edge: eps must be non-zero");
  Grid g = make_grid(nx, ny, nz, dx);
  MapOut o{psi, {Tx, Ty, Tz}, {dTx, dTy, dTz}, {nullptr, nullptr, nullptr}};
  hipLaunchKernelGGL(k_march_edge, dim3(blocks_for(ncell(g), 256)), dim3(256), 0, (hipStream_t)stream, g,
                     Mac{ux, uy, uz}, (const float4*)tab_old, X, o, dt, eps, (float4*)nullptr);
  EDGE_CHECK_LAUNCH();
  return EDGE_OK;
}

int edge_march_edge_table(const float* ux, const float* uy, const float* uz, const float* tab_old, const float* X,
                          float* psi, float* Tx, float* Ty, float* Tz, float* dTx, float* dTy, float* dTz,
                          float* tab_new, int nx, int ny, int nz, float dx, float dt, float eps, void* stream) {
  clear_error();
  EDGE_REQUIRE(dims_ok(nx, ny, nz), "edge_march_edge_table: bad dims");
  EDGE_MAC_OK(ux, uy, uz);
  EDGE_REQUIRE(tab_old && X && psi && Tx && Ty && Tz && dTx && dTy && dTz && tab_new,
               "edge_march_edge_table: null pointer");
  EDGE_REQUIRE(tab_new != tab_old, "edge_march_edge_table: the new table must not be the old one");
  EDGE_REQUIRE(eps != 0.0f, "edge_march_edge_table: eps must be non-zero");
  Grid g = make_grid(nx, ny, nz, dx);
  MapOut o{psi, {Tx, Ty, Tz}, {dTx, dTy, dTz}, {nullptr, nullptr, nullptr}};
  hipLaunchKernelGGL(k_march_edge, dim3(blocks_for(ncell(g), 256)), dim3(256), 0, (hipStream_t)stream, g,
                     Mac{ux, uy, uz}, (const float4*)tab_old, X, o, dt, eps, (float4*)tab_new);
  EDGE_CHECK_LAUNCH();
  return EDGE_OK;
}

int edge_march_ed8(const float* ux, const float* uy, const float* uz, const float* tab_old, const float* X,
                   float* psi, float* Tx, float* Ty, float* Tz, float* dTx, float* dTy, float* dTz, float* d2Tx,
                   float* d2Ty, float* d2Tz, int nx, int ny, int nz, float dx, float dt, float eps, void* stream) {
  clear_error();
  EDGE_REQUIRE(dims_ok(nx, ny, nz), "edge_march_ed8: bad dims");
  EDGE_MAC_OK(ux, uy, uz);
  EDGE_REQUIRE(tab_old && X && psi && Tx && Ty && Tz && dTx && dTy && dTz && d2Tx && d2Ty && d2Tz,
               "edge_march_ed8: null pointer");
  EDGE_REQUIRE(eps != 0.0f, "edge_march_ed8: eps must be non-zero");
  Grid g = make_grid(nx, ny, nz, dx);
  MapOut o{psi, {Tx, Ty, Tz}, {dTx, dTy, dTz}, {d2Tx, d2Ty, d2Tz}};
  hipLaunchKernelGGL(k_march_ed8, dim3(blocks_for(ncell(g), 256)), dim3(256), 0, (hipStream_t)stream, g,
                     Mac{ux, uy, uz}, (const float4*)tab_old, X, o, dt, eps);
  EDGE_CHECK_LAUNCH();
  return EDGE_OK;
}

int edge_march_ge(const float* ux, const float* uy, const float* uz, const float* tab_old, const float* X,
                  float* psi, float* Tx, float* Ty, float* Tz, int nx, int ny, int nz, float dx, float dt,
                  void* stream) {
  clear_error();
  EDGE_REQUIRE(dims_ok(nx, ny, nz), "edge_march_ge: bad dims");
  EDGE_MAC_OK(ux, uy, uz);
  EDGE_REQUIRE(tab_old && X && psi && Tx && Ty && Tz, "edge_march_ge: null pointer");
  Grid g = make_grid(nx, ny, nz, dx);
  MapOut o{psi, {Tx, Ty, Tz}, {nullptr, nullptr, nullptr}, {nullptr, nullptr, nullptr}};
  hipLaunchKernelGGL(k_march_ge, dim3(blocks_for(ncell(g), 256)), dim3(256), 0, (hipStream_t)stream, g,
                     Mac{ux, uy, uz}, (const float4*)tab_old, X, o, dt);
  EDGE_CHECK_LAUNCH();
  return EDGE_OK;
}

int edge_march_ed4(const float* ux, const float* uy, const float* uz, const float* tab_old, const float* X,
                   float* psi, float* Tx, float* Ty, float* Tz, int nx, int ny, int nz, float dx, float dt,
                   float eps, void* stream) {
  clear_error();
  EDGE_REQUIRE(dims_ok(nx, ny, nz), "edge_march_ed4: bad dims");
  EDGE_MAC_OK(ux, uy, uz);
  EDGE_REQUIRE(tab_old && X && psi && Tx && Ty && Tz, "edge_march_ed4: null pointer");
  EDGE_REQUIRE(eps != 0.0f, "edge_march_ed4: eps must be non-zero");
  Grid g = make_grid(nx, ny, nz, dx);
  MapOut o{psi, {Tx, Ty, Tz}, {nullptr, nullptr, nullptr}, {nullptr, nullptr, nullptr}};
  hipLaunchKernelGGL(k_march_ed4, dim3(blocks_for(ncell(g), 256)), dim3(256), 0, (hipStream_t)stream, g,
                     Mac{ux, uy, uz}, (const float4*)tab_old, X, o, dt, eps);
  EDGE_CHECK_LAUNCH();
  return EDGE_OK;
}

int edge_advance_map_faces(const float* ux, const float* uy, const float* uz, float* phix, float* phiy, float* phiz,
                           float* Fx, float* Fy, float* Fz, int nx, int ny, int nz, float dx, float dt,
                           void* stream) {
  clear_error();
  EDGE_REQUIRE(dims_ok(nx, ny, nz), "edge_advance_map_faces: bad dims");
  EDGE_MAC_OK(ux, uy, uz);
  EDGE_REQUIRE(phix && phiy && phiz && Fx && Fy && Fz, "edge_advance_map_faces: null pointer");
  Grid g = make_grid(nx, ny, nz, dx);
  int64_t nf = (int64_t)(nx + 1) * ny * nz + (int64_t)nx * (ny + 1) * nz + (int64_t)nx * ny * (nz + 1);
  FaceMap fm{{phix, phiy, phiz}, {Fx, Fy, Fz}};
  hipLaunchKernelGGL(k_advance_faces, dim3(blocks_for(nf, 256)), dim3(256), 0, (hipStream_t)stream, g,
                     Mac{ux, uy, uz}, fm, dt);
  EDGE_CHECK_LAUNCH();
  return EDGE_OK;
}

int edge_backtrack_faces(const float* const* ux, const float* const* uy, const float* const* uz, const float* dts,
                         int n, float* phix, float* phiy, float* phiz, float* Fx, float* Fy, float* Fz, int nx, int ny,
                         int nz, float dx, void* stream) {
  clear_error();
  EDGE_REQUIRE(dims_ok(nx, ny, nz), "edge_backtrack_faces: bad dims");
  EDGE_REQUIRE(n >= 0 && (n == 0 || (ux && uy && uz && dts)), "edge_backtrack_faces: bad velocity list");
  EDGE_REQUIRE(phix && phiy && phiz && Fx && Fy && Fz, "edge_backtrack_faces: null pointer");
  for (int s = 0; s < n; ++s) EDGE_MAC_OK(ux[s], uy[s], uz[s]);
  Grid g = make_grid(nx, ny, nz, dx);
  int64_t nf = (int64_t)(nx + 1) * ny * nz + (int64_t)nx * (ny + 1) * nz + (int64_t)nx * ny * (nz + 1);
  FaceMap fm{{phix, phiy, phiz}, {Fx, Fy, Fz}};
  int s0 = 0;
  do {  // n == 0: one launch that only writes the identity
    VelSeq vs;
    vs.n = std::min(EFM_SEQ, n - s0);
    for (int s = 0; s < vs.n; ++s) {
      vs.ux[s] = ux[s0 + s];
      vs.uy[s] = uy[s0 + s];
      vs.uz[s] = uz[s0 + s];
      vs.dt[s] = dts[s0 + s];
    }
    for (int s = vs.n; s < EFM_SEQ; ++s) {
      vs.ux[s] = vs.uy[s] = vs.uz[s] = nullptr;
      vs.dt[s] = 0.f;
    }
    hipLaunchKernelGGL(k_backtrack_faces, dim3(blocks_for(nf, 256)), dim3(256), 0, (hipStream_t)stream, g, vs, fm,
                       s0 == 0 ? 1 : 0);
    EDGE_CHECK_LAUNCH();
    s0 += vs.n;
  } while (s0 < n);
  return EDGE_OK;
}

int edge_advance_points(int mode, const float* ux, const float* uy, const float* uz, float* pts, float* Tx,
                        float* Ty, float* Tz, long n, int nx, int ny, int nz, float dx, float dt, void* stream) {
  clear_error();
  EDGE_REQUIRE(dims_ok(nx, ny, nz), "edge_advance_points: bad dims");
  EDGE_MAC_OK(ux, uy, uz);
  EDGE_REQUIRE(pts && (mode == 0 || (mode == 1 && Tx && Ty && Tz)), "edge_advance_points: bad mode/pointers");
  if (n <= 0) return EDGE_OK;
  Grid g = make_grid(nx, ny, nz, dx);
  hipLaunchKernelGGL(k_advance_points, dim3(blocks_for(n, 256)), dim3(256), 0, (hipStream_t)stream, g,
                     Mac{ux, uy, uz}, pts, Rows{{Tx, Ty, Tz}}, (int64_t)n, dt, mode);
  EDGE_CHECK_LAUNCH();
  return EDGE_OK;
}

int edge_dye_advance(const float* ux, const float* uy, const float* uz, float* pos, long n, long num, int nx, int ny,
                     int nz, float dx, float dt, void* stream) {
  clear_error();
  EDGE_REQUIRE(dims_ok(nx, ny, nz), "edge_dye_advance: bad dims");
  EDGE_MAC_OK(ux, uy, uz);
  EDGE_REQUIRE(pos && n >= 0 && num >= 0 && n < (1L << 31) && dx > 0.0f, "edge_dye_advance: bad arguments");
  const int64_t m = std::min<int64_t>(num, n);
  if (m == 0) return EDGE_OK;
  Grid g = make_grid(nx, ny, nz, dx);
  hipLaunchKernelGGL(k_dye_advance, dim3(blocks_for(m, 256)), dim3(256), 0, (hipStream_t)stream, g, Mac{ux, uy, uz},
                     pos, (int64_t)n, m, dt, (float)nx * dx, (float)ny * dx, (float)nz * dx);
  EDGE_CHECK_LAUNCH();
  return EDGE_OK;
}

int edge_compose_points(const float* tab, float* pts, float* Tx, float* Ty, float* Tz, long n, int nx, int ny, int nz,
                        float dx, void* stream) {
  clear_error();
  EDGE_REQUIRE(dims_ok(nx, ny, nz), "edge_compose_points: bad dims");
  EDGE_REQUIRE(tab && pts, "edge_compose_points: null pointer");
  EDGE_REQUIRE((Tx && Ty && Tz) || (!Tx && !Ty && !Tz), "edge_compose_points: give all or none of T");
  if (n <= 0) return EDGE_OK;
  Grid g = make_grid(nx, ny, nz, dx);
  hipLaunchKernelGGL(k_compose_points, dim3(blocks_for(n, 256)), dim3(256), 0, (hipStream_t)stream, g,
                     (const float4*)tab, pts, Rows{{Tx, Ty, Tz}}, (int64_t)n, Tx ? 1 : 0);
  EDGE_CHECK_LAUNCH();
  return EDGE_OK;
}

// T_p_c: point p (0..3), row c; out dT_c.
int edge_calc_ed4_garm(const float* T0x, const float* T0y, const float* T0z, const float* T1x, const float* T1y,
                       const float* T1z, const float* T2x, const float* T2y, const float* T2z, const float* T3x,
                       const float* T3y, const float* T3z, float* dTx, float* dTy, float* dTz, long n, float eps,
                       void* stream) {
  clear_error();
  EDGE_REQUIRE(T0x && T0y && T0z && T1x && T1y && T1z && T2x && T2y && T2z && T3x && T3y && T3z && dTx && dTy && dTz,
               "edge_calc_ed4_garm: null pointer");
  EDGE_REQUIRE(eps != 0.0f, "edge_calc_ed4_garm: eps must be non-zero");
  if (n <= 0) return EDGE_OK;
  Rows4 P{{{T0x, T0y, T0z}, {T1x, T1y, T1z}, {T2x, T2y, T2z}, {T3x, T3y, T3z}}};
  hipLaunchKernelGGL(k_calc_ed4_garm, dim3(blocks_for(n, 256)), dim3(256), 0, (hipStream_t)stream, P,
                     Rows{{dTx, dTy, dTz}}, (int64_t)n, eps);
  EDGE_CHECK_LAUNCH();
  return EDGE_OK;
}

// np = 8: calc_ed8_3 (all outputs); np = 4: calc_ed4_1 (psi, T only).
int edge_calc_ed(int np, const float* const* pts, float* psi, float* Tx, float* Ty, float* Tz, float* dTx, float* dTy,
                 float* dTz, float* d2Tx, float* d2Ty, float* d2Tz, long n, float eps, void* stream) {
  clear_error();
  EDGE_REQUIRE(np == 8 || np == 4, "edge_calc_ed: np must be 4 or 8");
  EDGE_REQUIRE(pts && psi && Tx && Ty && Tz, "edge_calc_ed: null pointer");
  EDGE_REQUIRE(np == 4 || (dTx && dTy && dTz && d2Tx && d2Ty && d2Tz), "edge_calc_ed: ED8 needs dT and d2T");
  EDGE_REQUIRE(eps != 0.0f, "edge_calc_ed: eps must be non-zero");
  Pts8 S;
  for (int e = 0; e < 8; ++e) S.P[e] = e < np ? pts[e] : nullptr;
  for (int e = 0; e < np; ++e) EDGE_REQUIRE(S.P[e], "edge_calc_ed: null point field %d", e);
  if (n <= 0) return EDGE_OK;
  MapOut o{psi, {Tx, Ty, Tz}, {dTx, dTy, dTz}, {d2Tx, d2Ty, d2Tz}};
  hipLaunchKernelGGL(k_calc_ed, dim3(blocks_for(n, 256)), dim3(256), 0, (hipStream_t)stream, S, np, o, (int64_t)n,
                     eps);
  EDGE_CHECK_LAUNCH();
  return EDGE_OK;
}

int edge_offset_points(const float* src, float* dst, long n, float ox, float oy, float oz, void* stream) {
  clear_error();
  EDGE_REQUIRE(src && dst, "edge_offset_points: null pointer");
  if (n <= 0) return EDGE_OK;
  hipLaunchKernelGGL(k_offset_points, dim3(blocks_for(n, 256)), dim3(256), 0, (hipStream_t)stream, src, dst,
                     (int64_t)n, ox, oy, oz);
  EDGE_CHECK_LAUNCH();
  return EDGE_OK;
}

int edge_eval_points(int what, const float* f, int d0, int d1, int d2, const float* BL, const float* ux,
                     const float* uy, const float* uz, const float* tab, int nx, int ny, int nz, const float* pts,
                     long n, float dx, float* val, float* grad, void* stream) {
  clear_error();
  EDGE_REQUIRE(what >= 0 && what <= 2, "edge_eval_points: bad what");
  EDGE_REQUIRE(pts && val, "edge_eval_points: null pointer");
  if (what == 0) EDGE_REQUIRE(f && BL && d0 >= 4 && d1 >= 4 && d2 >= 4, "edge_eval_points: bad field");
  if (what == 1) EDGE_MAC_OK(ux, uy, uz);
  if (what == 2) EDGE_REQUIRE(tab, "edge_eval_points: null table");
  if (what != 0) EDGE_REQUIRE(dims_ok(nx, ny, nz), "edge_eval_points: bad dims");
  if (n <= 0) return EDGE_OK;
  Grid g = make_grid(what == 0 ? 4 : nx, what == 0 ? 4 : ny, what == 0 ? 4 : nz, dx);
  float bl[3] = {0.f, 0.f, 0.f};
  if (what == 0) { bl[0] = BL[0]; bl[1] = BL[1]; bl[2] = BL[2]; }
  hipLaunchKernelGGL(k_eval_points, dim3(blocks_for(n, 256)), dim3(256), 0, (hipStream_t)stream, g, what, f, d0, d1,
                     d2, bl[0], bl[1], bl[2], Mac{ux, uy, uz}, (const float4*)tab, pts, (int64_t)n, val, grad);
  EDGE_CHECK_LAUNCH();
  return EDGE_OK;
}

int edge_reset_identity(const float* X, float* psi, float* Tx, float* Ty, float* Tz, float* dTx, float* dTy,
                        float* dTz, float* d2Tx, float* d2Ty, float* d2Tz, int nx, int ny, int nz, void* stream) {
  clear_error();
  EDGE_REQUIRE(dims_ok(nx, ny, nz), "edge_reset_identity: bad dims");
  EDGE_REQUIRE(!psi || X, "edge_reset_identity: psi needs X");
  Grid g = make_grid(nx, ny, nz, 1.0f);
  MapOut o{psi, {Tx, Ty, Tz}, {dTx, dTy, dTz}, {d2Tx, d2Ty, d2Tz}};
  hipLaunchKernelGGL(k_reset_identity, dim3(blocks_for(ncell(g), 256)), dim3(256), 0, (hipStream_t)stream, g, X,
                     o);
  EDGE_CHECK_LAUNCH();
  return EDGE_OK;
}

int edge_reset_faces(float* phix, float* phiy, float* phiz, float* Fx, float* Fy, float* Fz, int nx, int ny, int nz,
                     float dx, void* stream) {
  clear_error();
  EDGE_REQUIRE(dims_ok(nx, ny, nz), "edge_reset_faces: bad dims");
  Grid g = make_grid(nx, ny, nz, dx);
  int64_t nf = (int64_t)(nx + 1) * ny * nz + (int64_t)nx * (ny + 1) * nz + (int64_t)nx * ny * (nz + 1);
  FaceMap fm{{phix, phiy, phiz}, {Fx, Fy, Fz}};
  hipLaunchKernelGGL(k_reset_faces, dim3(blocks_for(nf, 256)), dim3(256), 0, (hipStream_t)stream, g, fm);
  EDGE_CHECK_LAUNCH();
  return EDGE_OK;
}

int edge_coords(int mode, float* out, int nx, int ny, int nz, float dx, void* stream) {
  clear_error();
  EDGE_REQUIRE(dims_ok(nx, ny, nz) && out && mode >= 0 && mode <= 3, "edge_coords: bad arguments");
  Grid g = make_grid(nx, ny, nz, dx);
  int64_t n = (int64_t)(nx + (mode == 1)) * (ny + (mode == 2)) * (nz + (mode == 3));
  hipLaunchKernelGGL(k_coords, dim3(blocks_for(n, 256)), dim3(256), 0, (hipStream_t)stream, g, mode, out);
  EDGE_CHECK_LAUNCH();
  return EDGE_OK;
}

}  // extern "C"
