// Step-loop driver kernels of run_edge.py / run_ed8.py, the semi-Lagrangian
// velocity advection of sim/hermite.py, the field arithmetic of
// util/taichi_utils.py and the vortex-ring initial condition of
// sim/init_conditions.py.  Face kernels cover the three face grids in one
// launch (run_edge.py's three `for I in ti.grouped(u_*1)` loops).
#include "edge_common.h"

using namespace edge;

namespace {

struct Faces {
  float* f[3];
};
struct CFaces {
  const float* f[3];
};

// Decode a flat index over the concatenated face grids x, y, z.
__device__ __forceinline__ bool face_of(const Grid& g, int64_t t, int& a, int64_t& local, int64_t& nf, int& i, int& j,
                                        int& k) {
  return face_split(g, t, a, local, nf, i, j, k);
}

__device__ __forceinline__ int64_t nfaces(const Grid& g) {
  return (int64_t)(g.nx + 1) * g.ny * g.nz + (int64_t)g.nx * (g.ny + 1) * g.nz + (int64_t)g.nx * g.ny * (g.nz + 1);
}

// min and max of the six neighbours of face (i, j, k) on face grid a of u1
// (indices clamped to the grid, as clamp_u's sample() does, run_edge.py:508-519).
__device__ __forceinline__ void face_nbr_range(const Grid& g, int a, int i, int j, int k, const float* __restrict__ u1,
                                               float& mn, float& mx) {
  int d[3];
  face_dims(g, a, d);
  auto at = [&](int ii, int jj, int kk) {
    ii = max(0, min(ii, d[0] - 1));
    jj = max(0, min(jj, d[1] - 1));
    kk = max(0, min(kk, d[2] - 1));
    return u1[((int64_t)ii * d[1] + jj) * d[2] + kk];
  };
  float l_ = at(i - 1, j, k), r_ = at(i + 1, j, k), b_ = at(i, j - 1, k), t_ = at(i, j + 1, k),
        a_ = at(i, j, k - 1), c_ = at(i, j, k + 1);
  mx = fmaxf(fmaxf(fmaxf(l_, r_), fmaxf(b_, t_)), fmaxf(a_, c_));
  mn = fminf(fminf(fminf(l_, r_), fminf(b_, t_)), fminf(a_, c_));
}

// run_edge.py:360-373 map_u: u1_a = F_a . interp_u_MAC(u0, phi_a).  With I
// given (I.f[0] non-null) the stored value is sc * (u1_a + m * I_a), rounded as
// add_fields(u1, I, u1, m) then scale_field(u1, sc, u1) round it (the BFECC
// error (map - init) / 2 without the extra pass over u1).  With C given
// (C.f[0] non-null) the stored value is clamp(C_a - u1_a) to the min and max of
// C_a's six neighbours: the BFECC correction u - map(err) and clamp_u in the same
// pass (run_edge.py:683-693 with BFECC_clamp; the map's value is the tmp_u that
// the separate passes would store and reload).
__global__ void __launch_bounds__(256) k_map_u(Grid g, Mac U0, Faces U1, CFaces F, CFaces PHI, CFaces I, float m,
                                               float sc, CFaces C) {
  int64_t t = (int64_t)xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
  int a, i, j, k;
  int64_t l, nf;
  if (!face_of(g, t, a, l, nf, i, j, k)) return;
  const float* ph = PHI.f[a];
  const float* Fa = F.f[a];
  float p[3] = {ph[l], ph[nf + l], ph[2 * nf + l]};
  float v[3];
  interp_u_mac<false>(g, U0, p, v, nullptr);
  float r = Fa[l] * v[0] + Fa[nf + l] * v[1] + Fa[2 * nf + l] * v[2];
  if (I.f[0]) {
    r = r + m * I.f[a][l];
    r = sc * r;
  }
  if (C.f[0]) {
    float mn, mx;
    face_nbr_range(g, a, i, j, k, C.f[a], mn, mx);
    r = fminf(fmaxf(C.f[a][l] - r, mn), mx);
  }
  U1.f[a][l] = r;
}

// run_edge.py:376-417 map_u_psi: u1_a(X_a) = sum_c d psi_c/d x_a (X_a) u0_c(psi(X_a)).
__global__ void __launch_bounds__(256) k_map_u_psi(Grid g, Mac U0, Faces U1, const float4* __restrict__ tab) {
  int64_t t = (int64_t)xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
  int a, i, j, k;
  int64_t l, nf;
  if (!face_of(g, t, a, l, nf, i, j, k)) return;
  float X[3], psi[3], Gc[3], v[3];
  face_coord(a, i, j, k, g.dx, X);
  // only column a of grad psi enters
  if (a == 0)
    map_eval_dir<0>(g, tab, X, psi, Gc);
  else if (a == 1)
    map_eval_dir<1>(g, tab, X, psi, Gc);
  else
    map_eval_dir<2>(g, tab, X, psi, Gc);
  interp_u_mac<false>(g, U0, psi, v, nullptr);
  U1.f[a][l] = Gc[0] * v[0] + Gc[1] * v[1] + Gc[2] * v[2];
}

// Face (i, j, k) of face grid D: phi = psi(X_D), F = column D of grad psi.
template <int D>
__device__ __forceinline__ void split_face(const Grid& g, const float4* __restrict__ tab, int i, int j, int k,
                                           const Faces& PHI, const Faces& F) {
  float X[3], psi[3], Gc[3];
  face_coord(D, i, j, k, g.dx, X);
  map_eval_dir<D, true>(g, tab, X, psi, Gc);
  const int64_t nf = g.nf[D];
  const int64_t d1 = g.ny + (D == 1), d2 = g.nz + (D == 2);
  const int64_t l = ((int64_t)i * d1 + j) * d2 + k;
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    PHI.f[D][c * nf + l] = psi[c];
    F.f[D][c * nf + l] = Gc[c];
  }
}

// run_ed8.py:190-223 split_psi_hermite: phi_a = psi(X_a), F_a = column a of grad psi.
// Thread t < N takes cell t's three low faces (x at i, y at j, z at k), so one
// pass reads each table row once from memory and its neighbours from cache;
// the threads after N take the three top planes (x at nx, y at ny, z at nz).
__global__ void __launch_bounds__(256) k_split_psi(Grid g, const float4* __restrict__ tab, Faces PHI, Faces F) {
  const int64_t t = (int64_t)xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
  const int64_t N = ncell(g);
  if (t < N) {
    int i, j, k;
    cell_ijk(g, t, i, j, k);
    split_face<0>(g, tab, i, j, k, PHI, F);
    split_face<1>(g, tab, i, j, k, PHI, F);
    split_face<2>(g, tab, i, j, k, PHI, F);
    return;
  }
  int64_t r = t - N;
  const int64_t px = (int64_t)g.ny * g.nz, py = (int64_t)g.nx * g.nz, pz = (int64_t)g.nx * g.ny;
  if (r < px) {
    split_face<0>(g, tab, g.nx, (int)(r / g.nz), (int)(r % g.nz), PHI, F);
    return;
  }
  r -= px;
  if (r < py) {
    split_face<1>(g, tab, (int)(r / g.nz), g.ny, (int)(r % g.nz), PHI, F);
    return;
  }
  r -= py;
  if (r < pz) split_face<2>(g, tab, (int)(r / g.ny), (int)(r % g.ny), g.nz, PHI, F);
}

// util/taichi_utils.py:171-227: trilinear value or 8-tap min/max of a
// cell-centred field with nc <= 4 channels, BL 0.5.
// A smoke field keeps its NC channels together per cell (channel c of cell q
// at q * NC + c): a 4-channel tap is one 16-byte load.
template <int NC>
__device__ __forceinline__ void ld_cell(const float* __restrict__ s, int64_t q, float v[NC]);
template <>
__device__ __forceinline__ void ld_cell<1>(const float* __restrict__ s, int64_t q, float v[1]) {
  v[0] = s[q];
}
template <>
__device__ __forceinline__ void ld_cell<4>(const float* __restrict__ s, int64_t q, float v[4]) {
  const float4 t = ((const float4*)s)[q];
  v[0] = t.x;
  v[1] = t.y;
  v[2] = t.z;
  v[3] = t.w;
}
template <int NC>
__device__ __forceinline__ void st_cell(float* __restrict__ s, int64_t q, const float v[NC]);
template <>
__device__ __forceinline__ void st_cell<1>(float* __restrict__ s, int64_t q, const float v[1]) {
  s[q] = v[0];
}
template <>
__device__ __forceinline__ void st_cell<4>(float* __restrict__ s, int64_t q, const float v[4]) {
  ((float4*)s)[q] = make_float4(v[0], v[1], v[2], v[3]);
}

template <int NC>
__device__ __forceinline__ void linear_taps(const Grid& g, const float* __restrict__ s0, const float p[3], int mode,
                                            float out[NC], float mn[NC], float mx[NC]) {
  float ss[3];
  int i0[3], i1[3];
  float fr[3];
  const int dims[3] = {g.nx, g.ny, g.nz};
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    float u = p[d] * g.idx - 0.5f;
    ss[d] = __builtin_amdgcn_fmed3f(u, 1.0f, g.hi1_c[d]);   // max(1, min(u, hi)), hi >= 1
    float fl = floorf(ss[d]);
    fr[d] = ss[d] - fl;
    i0[d] = min((int)fl, dims[d] - 1);
    i1[d] = min((int)fl + 1, dims[d] - 1);
  }
#pragma unroll
  for (int c = 0; c < NC; ++c) out[c] = 0.f;
  bool first = true;
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b)
#pragma unroll
      for (int cz = 0; cz < 2; ++cz) {
        int ii = a ? i1[0] : i0[0], jj = b ? i1[1] : i0[1], kk = cz ? i1[2] : i0[2];
        // N_1(x) = 1 - |x| with x = s - (iu + a)
        float wx = a ? fr[0] : 1.0f - fr[0];
        float wy = b ? fr[1] : 1.0f - fr[1];
        float wz = cz ? fr[2] : 1.0f - fr[2];
        float tv[NC];
        ld_cell<NC>(s0, cell_id(g, ii, jj, kk), tv);
#pragma unroll
        for (int c = 0; c < NC; ++c) {
          float v = tv[c];
          if (mode == 0) {
            out[c] += v * wx * wy * wz;
          } else {
            mn[c] = first ? v : fminf(mn[c], v);
            mx[c] = first ? v : fmaxf(mx[c], v);
          }
        }
        first = false;
      }
}

// run_edge.py:420-427 map_smoke (mode 0) and 469-477 clamp_smoke (mode 1):
// psi_c = mean of the six face points of phi.
// With s2 given, mode 0 stores sc * (map + m * s2) and mode 1 clamps s1 + m * s2,
// rounded as add_fields then scale_field (mode 0) or add_fields (mode 1) would
// round them in a pass of their own.  Mode 2 clamps s1 + m * map(s2): s2 is
// first pulled back through the same map (mode 0's value, kept in a register
// instead of a field), so one pass does map_smoke(s2, tmp) and the clamp of
// s1 - tmp (run_edge.py:684-689).
template <int NC>
__global__ void __launch_bounds__(256) k_map_smoke(Grid g, int mode, const float* __restrict__ s0, float* __restrict__ s1,
                                                   CFaces PHI, const float* __restrict__ s2, float m, float sc) {
  const int64_t N = ncell(g);
  const int64_t q = (int64_t)xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
  if (q >= N) return;
  int i, j, k;
  cell_ijk(g, q, i, j, k);
  const int64_t n0 = (int64_t)(g.nx + 1) * g.ny * g.nz, n1 = (int64_t)g.nx * (g.ny + 1) * g.nz,
                n2 = (int64_t)g.nx * g.ny * (g.nz + 1);
  int64_t fx0 = ((int64_t)i * g.ny + j) * g.nz + k, fx1 = fx0 + (int64_t)g.ny * g.nz;
  int64_t fy0 = ((int64_t)i * (g.ny + 1) + j) * g.nz + k, fy1 = fy0 + g.nz;
  int64_t fz0 = ((int64_t)i * g.ny + j) * (g.nz + 1) + k, fz1 = fz0 + 1;
  float p[3];
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    float s = PHI.f[0][c * n0 + fx0] + PHI.f[0][c * n0 + fx1] + PHI.f[1][c * n1 + fy0] + PHI.f[1][c * n1 + fy1] +
              PHI.f[2][c * n2 + fz0] + PHI.f[2][c * n2 + fz1];
    p[c] = (1.0f / 6.0f) * s;
  }
  float out[NC], mn[NC], mx[NC];
  float v2[NC], res[NC];
  if (mode == 2) {
    linear_taps<NC>(g, s2, p, 0, v2, mn, mx);     // v2 = map(s2)
    linear_taps<NC>(g, s0, p, 1, out, mn, mx);    // the clamp range from s0
  } else {
    linear_taps<NC>(g, s0, p, mode, out, mn, mx);
    if (s2) ld_cell<NC>(s2, q, v2);
  }
  if (mode == 0) {
#pragma unroll
    for (int c = 0; c < NC; ++c) {
      float r = out[c];
      if (s2) {
        r = r + m * v2[c];
        r = sc * r;
      }
      res[c] = r;
    }
  } else {
    float v1[NC];
    ld_cell<NC>(s1, q, v1);
#pragma unroll
    for (int c = 0; c < NC; ++c) {
      float v = v1[c];
      if (s2) v = v + m * v2[c];
      res[c] = fminf(fmaxf(v, mn[c]), mx[c]);
    }
  }
  st_cell<NC>(s1, q, res);
}

// run_edge.py:443-466 map_smoke_psi (mode 0) and 480-505 clamp_smoke_psi
// (mode 1): psi_c = mean of the map's values at the six face centres.
template <int NC>
__global__ void __launch_bounds__(256) k_map_smoke_psi(Grid g, int mode, const float* __restrict__ s0,
                                                       float* __restrict__ s1, const float4* __restrict__ tab) {
  const int64_t N = ncell(g);
  const int64_t q = (int64_t)xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
  if (q >= N) return;
  int i, j, k;
  cell_ijk(g, q, i, j, k);
  float acc[3] = {0.f, 0.f, 0.f};
#pragma unroll 1
  for (int f = 0; f < 6; ++f) {
    int a = f >> 1, up = f & 1;
    float X[3], v[3];
    face_coord(a, i + (a == 0 ? up : 0), j + (a == 1 ? up : 0), k + (a == 2 ? up : 0), g.dx, X);
    map_eval<false>(g, tab, X, v, nullptr);
    acc[0] += v[0];
    acc[1] += v[1];
    acc[2] += v[2];
  }
  float p[3] = {(1.0f / 6.0f) * acc[0], (1.0f / 6.0f) * acc[1], (1.0f / 6.0f) * acc[2]};
  float out[NC], mn[NC], mx[NC];
  linear_taps<NC>(g, s0, p, mode, out, mn, mx);
  if (mode == 1) {
    float v1[NC];
    ld_cell<NC>(s1, q, v1);
#pragma unroll
    for (int c = 0; c < NC; ++c) out[c] = fminf(fmaxf(v1[c], mn[c]), mx[c]);
  }
  st_cell<NC>(s1, q, out);
}


// run_edge.py:508-519 clamp_u over all three face grids: u2 = clamp(u, min, max
// of the six neighbours in u1).
// SUB: the clamped value is U1 - U2 on the same face (the BFECC correction
// u - tmp_u, computed exactly as add_fields(u, tmp, err, -1) rounds it) and the
// result overwrites U2 in place; each thread reads and writes only its own face
// of U2, so the in-place update has no race.
template <bool SUB>
__global__ void __launch_bounds__(256) k_clamp_u(Grid g, CFaces U, CFaces U1, Faces U2) {
  int64_t t = (int64_t)xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
  int a, i, j, k;
  int64_t l, nf;
  if (!face_of(g, t, a, l, nf, i, j, k)) return;
  const float* u1 = U1.f[a];
  float mn, mx;
  face_nbr_range(g, a, i, j, k, u1, mn, mx);
  float v = SUB ? u1[l] - U2.f[a][l] : U.f[a][l];
  U2.f[a][l] = fminf(fmaxf(v, mn), mx);
}

// sim/hermite.py:1578-1621 advect_u_semi_lagrangian: RK4 backtrace from the face
// centre, then component a of interp_u_MAC at the foot.  The face's own
// component is read from 27 taps held in registers (OwnTaps) wherever the
// stage point or foot falls on them, which is everywhere the displacement is
// under half a cell: 9 loads instead of 45 per face for that component, the
// same values.
#ifndef EDGE_SL_OWN
#define EDGE_SL_OWN 1
#endif
#ifndef EDGE_SL_HOLD
#define EDGE_SL_HOLD 1
#endif
template <int A>
__device__ __forceinline__ void advect_face(const Grid& g, const Mac& U0, float* __restrict__ out, float dt, int i,
                                            int j, int k) {
  float p1[3], q[3], v1[3], v2[3], v3[3], v4[3];
  face_coord(A, i, j, k, g.dx, p1);
#if EDGE_SL_OWN
  const MacComp<A> M(g, U0);
  const OwnTaps<A> W(g, M, i, j, k);
  interp_u_mac_own<A>(g, U0, p1, M, W, v1);
#pragma unroll
  for (int d = 0; d < 3; ++d) q[d] = p1[d] - v1[d] * dt * 0.5f;
#if EDGE_SL_HOLD
  // the other two components: stage 2's taps held for stages 3 and 4 (the
  // stage points move the same way from the face, so they usually share them)
  Taps27 T1, T2;
  interp_u_mac_own<A, 1>(g, U0, q, M, W, v2, &T1, &T2);
#pragma unroll
  for (int d = 0; d < 3; ++d) q[d] = p1[d] - v2[d] * dt * 0.5f;
  interp_u_mac_own<A, 2>(g, U0, q, M, W, v3, &T1, &T2);
#pragma unroll
  for (int d = 0; d < 3; ++d) q[d] = p1[d] - v3[d] * dt;
  interp_u_mac_own<A, 2>(g, U0, q, M, W, v4, &T1, &T2);
#else
  interp_u_mac_own<A>(g, U0, q, M, W, v2);
#pragma unroll
  for (int d = 0; d < 3; ++d) q[d] = p1[d] - v2[d] * dt * 0.5f;
  interp_u_mac_own<A>(g, U0, q, M, W, v3);
#pragma unroll
  for (int d = 0; d < 3; ++d) q[d] = p1[d] - v3[d] * dt;
  interp_u_mac_own<A>(g, U0, q, M, W, v4);
#endif
#pragma unroll
  for (int d = 0; d < 3; ++d) q[d] = p1[d] - (v1[d] + 2.0f * v2[d] + 2.0f * v3[d] + v4[d]) / 6.0f * dt;
  *out = interp_u_comp_own<A>(g, q, M, W);   // the face's own component at the departure point
#else
  interp_u_mac<false>(g, U0, p1, v1, nullptr);
#pragma unroll
  for (int d = 0; d < 3; ++d) q[d] = p1[d] - v1[d] * dt * 0.5f;
  interp_u_mac<false>(g, U0, q, v2, nullptr);
#pragma unroll
  for (int d = 0; d < 3; ++d) q[d] = p1[d] - v2[d] * dt * 0.5f;
  interp_u_mac<false>(g, U0, q, v3, nullptr);
#pragma unroll
  for (int d = 0; d < 3; ++d) q[d] = p1[d] - v3[d] * dt;
  interp_u_mac<false>(g, U0, q, v4, nullptr);
#pragma unroll
  for (int d = 0; d < 3; ++d) q[d] = p1[d] - (v1[d] + 2.0f * v2[d] + 2.0f * v3[d] + v4[d]) / 6.0f * dt;
  *out = interp_u_comp(g, U0, q, A);
#endif
}

#ifndef EDGE_WPE_SL
#define EDGE_WPE_SL 3   // 167 VGPRs, no spills; without the bound 171 (2 waves per SIMD), 0.454 ms instead of 0.356
#endif
#if EDGE_WPE_SL
#define EDGE_SL_ATTR __attribute__((amdgpu_waves_per_eu(EDGE_WPE_SL, 8)))
#else
#define EDGE_SL_ATTR
#endif
__global__ void __launch_bounds__(256) EDGE_SL_ATTR k_advect_sl(Grid g, Mac U0, Faces U1, float dt) {
  int64_t t = (int64_t)xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
  int a, i, j, k;
  int64_t l, nf;
  if (!face_of(g, t, a, l, nf, i, j, k)) return;
  // a is uniform except in the two waves that straddle grid boundaries
  if (a == 0)
    advect_face<0>(g, U0, U1.f[0] + l, dt, i, j, k);
  else if (a == 1)
    advect_face<1>(g, U0, U1.f[1] + l, dt, i, j, k);
  else
    advect_face<2>(g, U0, U1.f[2] + l, dt, i, j, k);
}

// run_edge.py:295-303 calc_max_speed: max over cells of the centred speed, floor 1e-3.
// out holds the float bits; positive floats order like their int bits.
__global__ void __launch_bounds__(256) k_max_speed(Grid g, Mac U, unsigned int* __restrict__ out) {
  const int64_t N = ncell(g);
  float m = 1.e-3f;
  for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < N; q += (int64_t)gridDim.x * blockDim.x) {
    int i, j, k;
    cell_ijk(g, q, i, j, k);
    float u = 0.5f * (U.ux[q] + U.ux[q + (int64_t)g.ny * g.nz]);
    int64_t qy = ((int64_t)i * (g.ny + 1) + j) * g.nz + k;
    float v = 0.5f * (U.uy[qy] + U.uy[qy + g.nz]);
    int64_t qz = ((int64_t)i * g.ny + j) * (g.nz + 1) + k;
    float w = 0.5f * (U.uz[qz] + U.uz[qz + 1]);
    m = fmaxf(m, sqrtf(u * u + v * v + w * w));
  }
  // wave64 reduction, block maximum in LDS, one atomic per block
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) m = fmaxf(m, __shfl_xor(m, off, 64));
  __shared__ float wm[4];
  if ((threadIdx.x & 63) == 0) wm[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    float b = fmaxf(fmaxf(wm[0], wm[1]), fmaxf(wm[2], wm[3]));
    atomicMax(out, __float_as_uint(b));
  }
}

__global__ void k_set_u32(unsigned int* p, unsigned int v) { *p = v; }

// util/taichi_utils.py:12-15 add_fields (dest = f1 + m f2) and 6-9 scale_field.
// The step calls them in place (dest == f1, r == a): no __restrict__ here.
__global__ void __launch_bounds__(256) k_axpy(const float* f1, const float* f2, float* dest, float m, int64_t n) {
  int64_t t = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) * 4;
  if (t + 3 < n) {
    float4 a = *(const float4*)(f1 + t), b = *(const float4*)(f2 + t);
    float4 r = {a.x + m * b.x, a.y + m * b.y, a.z + m * b.z, a.w + m * b.w};
    *(float4*)(dest + t) = r;
  } else {
    for (; t < n; ++t) dest[t] = f1[t] + m * f2[t];
  }
}

__global__ void __launch_bounds__(256) k_scale(const float* a, float alpha, float* r, int64_t n) {
  int64_t t = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) * 4;
  if (t + 3 < n) {
    float4 v = *(const float4*)(a + t);
    float4 o = {alpha * v.x, alpha * v.y, alpha * v.z, alpha * v.w};
    *(float4*)(r + t) = o;
  } else {
    for (; t < n; ++t) r[t] = alpha * a[t];
  }
}

// add_fields then scale_field in one pass: dest = s (f1 + m f2), rounded as the
// two calls round (f1 + m f2 to FP32, then the product).
__global__ void __launch_bounds__(256) k_axpy_scale(const float* f1, const float* f2, float* dest, float m, float sc,
                                                    int64_t n) {
  int64_t t = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) * 4;
  if (t + 3 < n) {
    float4 a = *(const float4*)(f1 + t), b = *(const float4*)(f2 + t);
    float4 r;
    r.x = a.x + m * b.x;
    r.y = a.y + m * b.y;
    r.z = a.z + m * b.z;
    r.w = a.w + m * b.w;
    r = {sc * r.x, sc * r.y, sc * r.z, sc * r.w};
    *(float4*)(dest + t) = r;
  } else {
    for (; t < n; ++t) {
      const float v = f1[t] + m * f2[t];
      dest[t] = sc * v;
    }
  }
}

__global__ void __launch_bounds__(256) k_fill(float* __restrict__ p, float v, int64_t n) {
  int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) p[t] = v;
}

// sim/init_conditions.py:13-36 add_vortex_ring_and_smoke (pf = cell centres).
// The sample points and their vorticity vectors do not depend on the cell: a
// block computes them once per chunk of RING_CHUNK samples into LDS (one
// sample per thread, the same expressions), then every thread sums the chunk
// for its cell in sample order.
constexpr int RING_CHUNK = 256;
__global__ void __launch_bounds__(RING_CHUNK) k_vortex_ring(Grid g, float w, float rad, float delta, float3 c,
                                                            float3 ex, float3 ey, const float* __restrict__ pf,
                                                            float* __restrict__ vf, float* __restrict__ smokef,
                                                            float3 color, int num_samples) {
  __shared__ float sp[6][RING_CHUNK];   // sample position (3) and vorticity vector (3)
  const int64_t N = ncell(g);
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const bool live = q < N;
  const float pi = 3.14159265358979323846f;
  const float curve_length = (2.0f * pi * rad) / (float)num_samples;
  float P[3] = {0.f, 0.f, 0.f};
  if (live) {
    P[0] = pf[q];
    P[1] = pf[N + q];
    P[2] = pf[2 * N + q];
  }
  float acc[3] = {0.f, 0.f, 0.f}, sm = 0.f;
  for (int l0 = 0; l0 < num_samples; l0 += RING_CHUNK) {
    const int l = l0 + (int)threadIdx.x;
    if (l < num_samples) {
      const float theta = (float)l / (float)num_samples * 2.0f * pi;
      const float ct = cosf(theta), st = sinf(theta);
      sp[0][threadIdx.x] = rad * (ct * ex.x + st * ey.x) + c.x;
      sp[1][threadIdx.x] = rad * (ct * ex.y + st * ey.y) + c.y;
      sp[2][threadIdx.x] = rad * (ct * ex.z + st * ey.z) + c.z;
      sp[3][threadIdx.x] = w * (-st * ex.x + ct * ey.x);
      sp[4][threadIdx.x] = w * (-st * ex.y + ct * ey.y);
      sp[5][threadIdx.x] = w * (-st * ex.z + ct * ey.z);
    }
    __syncthreads();
    const int m = min(RING_CHUNK, num_samples - l0);
    if (live) {
      for (int s = 0; s < m; ++s) {
        const float pd[3] = {P[0] - sp[0][s], P[1] - sp[1][s], P[2] - sp[2][s]};
        const float wv[3] = {sp[3][s], sp[4][s], sp[5][s]};
        const float r = sqrtf(pd[0] * pd[0] + pd[1] * pd[1] + pd[2] * pd[2]);
        const float rd = r / delta;
        const float e = expf(-(rd * rd * rd));
        const float fac = curve_length * (-1.0f / (4.0f * pi * (r * r * r)) * (1.0f - e));
        acc[0] += fac * (pd[1] * wv[2] - pd[2] * wv[1]);
        acc[1] += fac * (pd[2] * wv[0] - pd[0] * wv[2]);
        acc[2] += fac * (pd[0] * wv[1] - pd[1] * wv[0]);
        sm += curve_length * e;
      }
    }
    __syncthreads();
  }
  if (!live) return;
  vf[q] += acc[0];
  vf[N + q] += acc[1];
  vf[2 * N + q] += acc[2];
  float4* sq = (float4*)smokef + q;   // the 4 channels of cell q, together
  float s3 = sq->w + sm;
  if (s3 > 0.002f)
    *sq = make_float4(color.x, color.y, color.z, 1.0f);
  else
    *sq = make_float4(0.f, 0.f, 0.f, 0.f);
}

// util/taichi_utils.py:103-116 split_central_vector: cell vector -> MAC faces.
__global__ void __launch_bounds__(256) k_split_central(Grid g, const float* __restrict__ vc, Faces U) {
  int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  int a, i, j, k;
  int64_t l, nf;
  if (!face_of(g, t, a, l, nf, i, j, k)) return;
  const int64_t N = ncell(g);
  auto at = [&](int ii, int jj, int kk) {
    ii = max(0, min(ii, g.nx - 1));
    jj = max(0, min(jj, g.ny - 1));
    kk = max(0, min(kk, g.nz - 1));
    return vc[(int64_t)a * N + cell_id(g, ii, jj, kk)];
  };
  float r = at(i, j, k);
  float lft = at(i - (a == 0), j - (a == 1), k - (a == 2));
  U.f[a][l] = 0.5f * (r + lft);
}

// util/taichi_utils.py:95-100 get_central_vector: vc_a = mean of the cell's two
// a-faces.  vc: 3 planes of N cells.
__global__ void __launch_bounds__(256) k_central_vector(Grid g, Mac U, float* __restrict__ vc) {
  const int64_t N = ncell(g);
  const int64_t q = (int64_t)xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
  if (q >= N) return;
  int i, j, k;
  cell_ijk(g, q, i, j, k);
  const int64_t qx = q;  // x-face (i, j, k); (i + 1, j, k) is one (ny nz) plane further
  const int64_t qy = ((int64_t)i * (g.ny + 1) + j) * g.nz + k;
  const int64_t qz = ((int64_t)i * g.ny + j) * (g.nz + 1) + k;
  vc[q] = 0.5f * (U.ux[qx + (int64_t)g.ny * g.nz] + U.ux[qx]);
  vc[N + q] = 0.5f * (U.uy[qy + g.nz] + U.uy[qy]);
  vc[2 * N + q] = 0.5f * (U.uz[qz + 1] + U.uz[qz]);
}

// util/taichi_utils.py:70-92 curl: central differences of a cell-centred vector
// field with sample()'s clamp of the neighbour index to the grid.
__global__ void __launch_bounds__(256) k_curl(Grid g, const float* __restrict__ vf, float* __restrict__ cf,
                                              float inv_dist) {
  const int64_t N = ncell(g);
  const int64_t q = (int64_t)xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
  if (q >= N) return;
  int i, j, k;
  cell_ijk(g, q, i, j, k);
  const int64_t sx = (int64_t)g.ny * g.nz, sy = g.nz;
  const int64_t r = i + 1 < g.nx ? q + sx : q, l = i > 0 ? q - sx : q;   // vr, vl
  const int64_t t = j + 1 < g.ny ? q + sy : q, b = j > 0 ? q - sy : q;   // vt, vb
  const int64_t c = k + 1 < g.nz ? q + 1 : q, a = k > 0 ? q - 1 : q;     // vc, va
  const float* vx = vf;
  const float* vy = vf + N;
  const float* vz = vf + 2 * N;
  const float d_vx_dz = inv_dist * (vx[c] - vx[a]);
  const float d_vx_dy = inv_dist * (vx[t] - vx[b]);
  const float d_vy_dx = inv_dist * (vy[r] - vy[l]);
  const float d_vy_dz = inv_dist * (vy[c] - vy[a]);
  const float d_vz_dx = inv_dist * (vz[r] - vz[l]);
  const float d_vz_dy = inv_dist * (vz[t] - vz[b]);
  cf[q] = d_vz_dy - d_vy_dz;
  cf[N + q] = d_vx_dz - d_vz_dx;
  cf[2 * N + q] = d_vy_dx - d_vx_dy;
}

// The device part of a frame output (run_edge.py:720-725 with
// util/io_utils.py:53-75): |curl u| of the cell-centred velocity and the smoke,
// both in VTK point order (x fastest), the smoke's nc channels interleaved per
// point.  Thread t is VTK point t.  The central vector and the curl are
// k_central_vector's and k_curl's expressions; the norm is numpy's float32
// norm over the last axis, sqrt((x*x + y*y) + z*z), without contraction.
__global__ void __launch_bounds__(256) k_vtk_frame(Grid g, Mac U, const float* __restrict__ smoke, int nc,
                                                   float* __restrict__ wn, float* __restrict__ sm, float inv_dist) {
  const int64_t N = ncell(g);
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= N) return;
  const uint32_t t32 = (uint32_t)t;
  const uint32_t rest = t32 / (uint32_t)g.nx;
  const int i = (int)(t32 - rest * (uint32_t)g.nx);
  const int k = (int)(rest / (uint32_t)g.ny);
  const int j = (int)(rest - (uint32_t)k * (uint32_t)g.ny);
  const int64_t sx = (int64_t)g.ny * g.nz;
  // component a of the central vector at cell (ii, jj, kk), clamped like sample()
  auto vcomp = [&](int a, int ii, int jj, int kk) {
    ii = max(0, min(ii, g.nx - 1));
    jj = max(0, min(jj, g.ny - 1));
    kk = max(0, min(kk, g.nz - 1));
    if (a == 0) {
      const int64_t q = ((int64_t)ii * g.ny + jj) * g.nz + kk;
      return 0.5f * (U.ux[q + sx] + U.ux[q]);
    }
    if (a == 1) {
      const int64_t q = ((int64_t)ii * (g.ny + 1) + jj) * g.nz + kk;
      return 0.5f * (U.uy[q + g.nz] + U.uy[q]);
    }
    const int64_t q = ((int64_t)ii * g.ny + jj) * (g.nz + 1) + kk;
    return 0.5f * (U.uz[q + 1] + U.uz[q]);
  };
  const float d_vx_dz = inv_dist * (vcomp(0, i, j, k + 1) - vcomp(0, i, j, k - 1));
  const float d_vx_dy = inv_dist * (vcomp(0, i, j + 1, k) - vcomp(0, i, j - 1, k));
  const float d_vy_dx = inv_dist * (vcomp(1, i + 1, j, k) - vcomp(1, i - 1, j, k));
  const float d_vy_dz = inv_dist * (vcomp(1, i, j, k + 1) - vcomp(1, i, j, k - 1));
  const float d_vz_dx = inv_dist * (vcomp(2, i + 1, j, k) - vcomp(2, i - 1, j, k));
  const float d_vz_dy = inv_dist * (vcomp(2, i, j + 1, k) - vcomp(2, i, j - 1, k));
  const float wx = d_vz_dy - d_vy_dz, wy = d_vx_dz - d_vz_dx, wz = d_vy_dx - d_vx_dy;
  float w2;
  {
#pragma clang fp contract(off)
    w2 = (wx * wx + wy * wy) + wz * wz;
  }
  wn[t] = sqrtf(w2);   // correctly rounded (hipcc's default for f32 sqrt)
  const int64_t q = ((int64_t)i * g.ny + j) * g.nz + k;
  for (int c = 0; c < nc; ++c) sm[t * nc + c] = smoke[q * nc + c];
}

}  // namespace

extern "C" {

int edge_map_u(const float* u0x, const float* u0y, const float* u0z, float* u1x, float* u1y, float* u1z,
               const float* Fx, const float* Fy, const float* Fz, const float* phix, const float* phiy,
               const float* phiz, int nx, int ny, int nz, float dx, void* stream) {
  clear_error();
  EDGE_REQUIRE(dims_ok(nx, ny, nz), "edge_map_u: bad dims");
  EDGE_REQUIRE(u0x && u0y && u0z && u1x && u1y && u1z && Fx && Fy && Fz && phix && phiy && phiz,
               "edge_map_u: null pointer");
  Grid g = make_grid(nx, ny, nz, dx);
  hipLaunchKernelGGL(k_map_u, dim3(blocks_for(
                                  (int64_t)(nx + 1) * ny * nz + (int64_t)nx * (ny + 1) * nz + (int64_t)nx * ny * (nz + 1),
                                  256)),
                     dim3(256), 0, (hipStream_t)stream, g, Mac{u0x, u0y, u0z}, Faces{{u1x, u1y, u1z}},
                     CFaces{{Fx, Fy, Fz}}, CFaces{{phix, phiy, phiz}}, CFaces{{nullptr, nullptr, nullptr}}, 0.f,
                     0.f, CFaces{{nullptr, nullptr, nullptr}});
  EDGE_CHECK_LAUNCH();
  return EDGE_OK;
}

int edge_map_u_add_scale(const float* u0x, const float* u0y, const float* u0z, float* u1x, float* u1y, float* u1z,
                         const float* Fx, const float* Fy, const float* Fz, const float* phix, const float* phiy,
                         const float* phiz, const float* ix, const float* iy, const float* iz, float m, float s,
                         int nx, int ny, int nz, float dx, void* stream) {
  clear_error();
  EDGE_REQUIRE(dims_ok(nx, ny, nz), "edge_map_u_add_scale: bad dims");
  EDGE_REQUIRE(u0x && u0y && u0z && u1x && u1y && u1z && Fx && Fy && Fz && phix && phiy && phiz && ix && iy && iz,
               "edge_map_u_add_scale: null pointer");
  EDGE_REQUIRE(ix != u1x && iy != u1y && iz != u1z, "edge_map_u_add_scale: the added field must not be the output");
  Grid g = make_grid(nx, ny, nz, dx);
  hipLaunchKernelGGL(k_map_u, dim3(blocks_for(
                                  (int64_t)(nx + 1) * ny * nz + (int64_t)nx * (ny + 1) * nz + (int64_t)nx * ny * (nz + 1),
                                  256)),
                     dim3(256), 0, (hipStream_t)stream, g, Mac{u0x, u0y, u0z}, Faces{{u1x, u1y, u1z}},
                     CFaces{{Fx, Fy, Fz}}, CFaces{{phix, phiy, phiz}}, CFaces{{ix, iy, iz}}, m, s,
                     CFaces{{nullptr, nullptr, nullptr}});
  EDGE_CHECK_LAUNCH();
  return EDGE_OK;
}

int edge_map_u_sub_clamp(const float* u0x, const float* u0y, const float* u0z, float* u1x, float* u1y, float* u1z,
                         const float* Fx, const float* Fy, const float* Fz, const float* phix, const float* phiy,
                         const float* phiz, const float* cx, const float* cy, const float* cz, int nx, int ny, int nz,
                         float dx, void* stream) {
  clear_error();
  EDGE_REQUIRE(dims_ok(nx, ny, nz), "edge_map_u_sub_clamp: bad dims");
  EDGE_REQUIRE(u0x && u0y && u0z && u1x && u1y && u1z && Fx && Fy && Fz && phix && phiy && phiz && cx && cy && cz,
               "edge_map_u_sub_clamp: null pointer");
  EDGE_REQUIRE(cx != u1x && cy != u1y && cz != u1z && u0x != u1x && u0y != u1y && u0z != u1z,
               "edge_map_u_sub_clamp: the output must differ from the pulled-back and the clamped fields");
  Grid g = make_grid(nx, ny, nz, dx);
  hipLaunchKernelGGL(k_map_u, dim3(blocks_for(
                                  (int64_t)(nx + 1) * ny * nz + (int64_t)nx * (ny + 1) * nz + (int64_t)nx * ny * (nz + 1),
                                  256)),
                     dim3(256), 0, (hipStream_t)stream, g, Mac{u0x, u0y, u0z}, Faces{{u1x, u1y, u1z}},
                     CFaces{{Fx, Fy, Fz}}, CFaces{{phix, phiy, phiz}}, CFaces{{nullptr, nullptr, nullptr}}, 0.f, 0.f,
                     CFaces{{cx, cy, cz}});
  EDGE_CHECK_LAUNCH();
  return EDGE_OK;
}

int edge_map_u_psi(const float* u0x, const float* u0y, const float* u0z, float* u1x, float* u1y, float* u1z,
                   const float* tab, int nx, int ny, int nz, float dx, void* stream) {
  clear_error();
  EDGE_REQUIRE(dims_ok(nx, ny, nz), "edge_map_u_psi: bad dims");
  EDGE_REQUIRE(u0x && u0y && u0z && u1x && u1y && u1z && tab, "edge_map_u_psi: null pointer");
  Grid g = make_grid(nx, ny, nz, dx);
  int64_t nf = (int64_t)(nx + 1) * ny * nz + (int64_t)nx * (ny + 1) * nz + (int64_t)nx * ny * (nz + 1);
  hipLaunchKernelGGL(k_map_u_psi, dim3(blocks_for(nf, 256)), dim3(256), 0, (hipStream_t)stream, g,
                     Mac{u0x, u0y, u0z}, Faces{{u1x, u1y, u1z}}, (const float4*)tab);
  EDGE_CHECK_LAUNCH();
  return EDGE_OK;
}

int edge_split_psi_hermite(const float* tab, float* phix, float* phiy, float* phiz, float* Fx, float* Fy, float* Fz,
                           int nx, int ny, int nz, float dx, void* stream) {
  clear_error();
  EDGE_REQUIRE(dims_ok(nx, ny, nz), "edge_split_psi_hermite: bad dims");
  EDGE_REQUIRE(tab && phix && phiy && phiz && Fx && Fy && Fz, "edge_split_psi_hermite: null pointer");
  Grid g = make_grid(nx, ny, nz, dx);
  const int64_t nt = ncell(g) + (int64_t)ny * nz + (int64_t)nx * nz + (int64_t)nx * ny;
  hipLaunchKernelGGL(k_split_psi, dim3(blocks_for(nt, 256)), dim3(256), 0, (hipStream_t)stream, g,
                     (const float4*)tab, Faces{{phix, phiy, phiz}}, Faces{{Fx, Fy, Fz}});
  EDGE_CHECK_LAUNCH();
  return EDGE_OK;
}

// NULL counts as aligned (an absent optional field)
static bool aligned16(const void* p) { return ((uintptr_t)p & 15u) == 0; }

static int map_smoke_launch(const char* who, int mode, const float* s0, float* s1, const float* s2, float m,
                            float sc, int nc, const float* phix, const float* phiy, const float* phiz, int nx, int ny,
                            int nz, float dx, void* stream) {
  EDGE_REQUIRE(dims_ok(nx, ny, nz), "%s: bad dims", who);
  EDGE_REQUIRE(s0 && s1 && phix && phiy && phiz && (mode == 0 || mode == 1 || (mode == 2 && s2)),
               "%s: bad arguments", who);
  EDGE_REQUIRE(nc == 4 || nc == 1, "%s: nc must be 1 or 4 (got %d)", who, nc);
  EDGE_REQUIRE(s0 != s1 && s2 != s1 && s2 != s0, "%s: the smoke fields must be distinct", who);
  EDGE_REQUIRE(nc == 1 || (aligned16(s0) && aligned16(s1) && aligned16(s2)),
               "%s: 4-channel smoke fields must be 16-byte aligned", who);
  Grid g = make_grid(nx, ny, nz, dx);
  dim3 grid(blocks_for(ncell(g), 256));
  if (nc == 4)
    hipLaunchKernelGGL(k_map_smoke<4>, grid, dim3(256), 0, (hipStream_t)stream, g, mode, s0, s1,
                       CFaces{{phix, phiy, phiz}}, s2, m, sc);
  else
    hipLaunchKernelGGL(k_map_smoke<1>, grid, dim3(256), 0, (hipStream_t)stream, g, mode, s0, s1,
                       CFaces{{phix, phiy, phiz}}, s2, m, sc);
  EDGE_CHECK_LAUNCH();
  return EDGE_OK;
}

// mode 0: map_smoke, 1: clamp_smoke (smoke1 updated in place).  nc in 1..4.
int edge_map_smoke(int mode, const float* s0, float* s1, int nc, const float* phix, const float* phiy,
                   const float* phiz, int nx, int ny, int nz, float dx, void* stream) {
  clear_error();
  return map_smoke_launch("edge_map_smoke", mode, s0, s1, nullptr, 0.f, 0.f, nc, phix, phiy, phiz, nx, ny, nz, dx,
                          stream);
}

int edge_map_smoke_add_scale(int mode, const float* s0, float* s1, const float* s2, float m, float s, int nc,
                             const float* phix, const float* phiy, const float* phiz, int nx, int ny, int nz, float dx,
                             void* stream) {
  clear_error();
  EDGE_REQUIRE(s2, "edge_map_smoke_add_scale: null pointer");
  return map_smoke_launch("edge_map_smoke_add_scale", mode, s0, s1, s2, m, s, nc, phix, phiy, phiz, nx, ny, nz, dx,
                          stream);
}

// mode 0: map_smoke_psi, 1: clamp_smoke_psi.
int edge_map_smoke_psi(int mode, const float* s0, float* s1, int nc, const float* tab, int nx, int ny, int nz,
                       float dx, void* stream) {
  clear_error();
  EDGE_REQUIRE(dims_ok(nx, ny, nz), "edge_map_smoke_psi: bad dims");
  EDGE_REQUIRE(s0 && s1 && tab && (mode == 0 || mode == 1), "edge_map_smoke_psi: bad arguments");
  EDGE_REQUIRE(nc == 4 || nc == 1, "edge_map_smoke_psi: nc must be 1 or 4 (got %d)", nc);
  EDGE_REQUIRE(nc == 1 || (aligned16(s0) && aligned16(s1)),
               "edge_map_smoke_psi: 4-channel smoke fields must be 16-byte aligned");
  Grid g = make_grid(nx, ny, nz, dx);
  dim3 grid(blocks_for(ncell(g), 256));
  if (nc == 4)
    hipLaunchKernelGGL(k_map_smoke_psi<4>, grid, dim3(256), 0, (hipStream_t)stream, g, mode, s0, s1,
                       (const float4*)tab);
  else
    hipLaunchKernelGGL(k_map_smoke_psi<1>, grid, dim3(256), 0, (hipStream_t)stream, g, mode, s0, s1,
                       (const float4*)tab);
  EDGE_CHECK_LAUNCH();
  return EDGE_OK;
}

int edge_clamp_u(const float* ux, const float* uy, const float* uz, const float* u1x, const float* u1y,
                 const float* u1z, float* u2x, float* u2y, float* u2z, int nx, int ny, int nz, void* stream) {
  clear_error();
  EDGE_REQUIRE(dims_ok(nx, ny, nz), "edge_clamp_u: bad dims");
  EDGE_REQUIRE(ux && uy && uz && u1x && u1y && u1z && u2x && u2y && u2z, "edge_clamp_u: null pointer");
  Grid g = make_grid(nx, ny, nz, 1.0f);
  int64_t nf = (int64_t)(nx + 1) * ny * nz + (int64_t)nx * (ny + 1) * nz + (int64_t)nx * ny * (nz + 1);
  hipLaunchKernelGGL(k_clamp_u<false>, dim3(blocks_for(nf, 256)), dim3(256), 0, (hipStream_t)stream, g,
                     CFaces{{ux, uy, uz}}, CFaces{{u1x, u1y, u1z}}, Faces{{u2x, u2y, u2z}});
  EDGE_CHECK_LAUNCH();
  return EDGE_OK;
}

int edge_sub_clamp_u(const float* ux, const float* uy, const float* uz, float* tx, float* ty, float* tz, int nx,
                     int ny, int nz, void* stream) {
  clear_error();
  EDGE_REQUIRE(dims_ok(nx, ny, nz), "edge_sub_clamp_u: bad dims");
  EDGE_REQUIRE(ux && uy && uz && tx && ty && tz, "edge_sub_clamp_u: null pointer");
  Grid g = make_grid(nx, ny, nz, 1.0f);
  int64_t nf = (int64_t)(nx + 1) * ny * nz + (int64_t)nx * (ny + 1) * nz + (int64_t)nx * ny * (nz + 1);
  hipLaunchKernelGGL(k_clamp_u<true>, dim3(blocks_for(nf, 256)), dim3(256), 0, (hipStream_t)stream, g,
                     CFaces{{ux, uy, uz}}, CFaces{{ux, uy, uz}}, Faces{{tx, ty, tz}});
  EDGE_CHECK_LAUNCH();
  return EDGE_OK;
}

int edge_advect_u_sl(const float* u0x, const float* u0y, const float* u0z, float* u1x, float* u1y, float* u1z,
                     int nx, int ny, int nz, float dx, float dt, void* stream) {
  clear_error();
  EDGE_REQUIRE(dims_ok(nx, ny, nz), "edge_advect_u_sl: bad dims");
  EDGE_REQUIRE(u0x && u0y && u0z && u1x && u1y && u1z, "edge_advect_u_sl: null pointer");
  EDGE_REQUIRE(u0x != u1x && u0y != u1y && u0z != u1z, "edge_advect_u_sl: input and output must differ");
  Grid g = make_grid(nx, ny, nz, dx);
  int64_t nf = (int64_t)(nx + 1) * ny * nz + (int64_t)nx * (ny + 1) * nz + (int64_t)nx * ny * (nz + 1);
  hipLaunchKernelGGL(k_advect_sl, dim3(blocks_for(nf, 256)), dim3(256), 0, (hipStream_t)stream, g,
                     Mac{u0x, u0y, u0z}, Faces{{u1x, u1y, u1z}}, dt);
  EDGE_CHECK_LAUNCH();
  return EDGE_OK;
}

// max speed into a device float (out_dev); the caller reads it back.
int edge_max_speed(const float* ux, const float* uy, const float* uz, float* out_dev, int nx, int ny, int nz,
                   void* stream) {
  clear_error();
  EDGE_REQUIRE(dims_ok(nx, ny, nz), "edge_max_speed: bad dims");
  EDGE_REQUIRE(ux && uy && uz && out_dev, "edge_max_speed: null pointer");
  Grid g = make_grid(nx, ny, nz, 1.0f);
  hipLaunchKernelGGL(k_set_u32, dim3(1), dim3(1), 0, (hipStream_t)stream, (unsigned int*)out_dev, 0u);
  EDGE_CHECK_LAUNCH();
  int nb = std::min(blocks_for(ncell(g), 256), 1024);
  hipLaunchKernelGGL(k_max_speed, dim3(nb), dim3(256), 0, (hipStream_t)stream, g, Mac{ux, uy, uz},
                     (unsigned int*)out_dev);
  EDGE_CHECK_LAUNCH();
  return EDGE_OK;
}

int edge_axpy(const float* f1, const float* f2, float* dest, float m, long n, void* stream) {
  clear_error();
  EDGE_REQUIRE(f1 && f2 && dest && n >= 0, "edge_axpy: bad arguments");
  EDGE_REQUIRE((((uintptr_t)f1 | (uintptr_t)f2 | (uintptr_t)dest) & 15) == 0, "edge_axpy: pointers not 16-byte aligned");
  if (n == 0) return EDGE_OK;
  hipLaunchKernelGGL(k_axpy, dim3(blocks_for((n + 3) / 4, 256)), dim3(256), 0, (hipStream_t)stream, f1, f2, dest, m,
                     (int64_t)n);
  EDGE_CHECK_LAUNCH();
  return EDGE_OK;
}

int edge_axpy_scale(const float* f1, const float* f2, float* dest, float m, float s, long n, void* stream) {
  clear_error();
  EDGE_REQUIRE(f1 && f2 && dest && n >= 0, "edge_axpy_scale: bad arguments");
  EDGE_REQUIRE((((uintptr_t)f1 | (uintptr_t)f2 | (uintptr_t)dest) & 15) == 0,
               "edge_axpy_scale: pointers not 16-byte aligned");
  if (n == 0) return EDGE_OK;
  hipLaunchKernelGGL(k_axpy_scale, dim3(blocks_for((n + 3) / 4, 256)), dim3(256), 0, (hipStream_t)stream, f1, f2,
                     dest, m, s, (int64_t)n);
  EDGE_CHECK_LAUNCH();
  return EDGE_OK;
}

int edge_scale(const float* a, float alpha, float* r, long n, void* stream) {
  clear_error();
  EDGE_REQUIRE(a && r && n >= 0, "edge_scale: bad arguments");
  EDGE_REQUIRE((((uintptr_t)a | (uintptr_t)r) & 15) == 0, "edge_scale: pointers not 16-byte aligned");
  if (n == 0) return EDGE_OK;
  hipLaunchKernelGGL(k_scale, dim3(blocks_for((n + 3) / 4, 256)), dim3(256), 0, (hipStream_t)stream, a, alpha, r,
                     (int64_t)n);
  EDGE_CHECK_LAUNCH();
  return EDGE_OK;
}

int edge_copy(const float* src, float* dst, long n, void* stream) {
  clear_error();
  EDGE_REQUIRE(src && dst && n >= 0, "edge_copy: bad arguments");
  if (n == 0 || src == dst) return EDGE_OK;
  EDGE_CHECK_HIP(hipMemcpyAsync(dst, src, (size_t)n * sizeof(float), hipMemcpyDeviceToDevice, (hipStream_t)stream));
  return EDGE_OK;
}

int edge_fill(float* p, float v, long n, void* stream) {
  clear_error();
  EDGE_REQUIRE(p && n >= 0, "edge_fill: bad arguments");
  if (n == 0) return EDGE_OK;
  hipLaunchKernelGGL(k_fill, dim3(blocks_for(n, 256)), dim3(256), 0, (hipStream_t)stream, p, v, (int64_t)n);
  EDGE_CHECK_LAUNCH();
  return EDGE_OK;
}

int edge_vortex_ring(float w, float rad, float delta, const float* c, const float* ex, const float* ey,
                     const float* pf, float* vf, float* smokef, const float* color, int num_samples, int nx, int ny,
                     int nz, void* stream) {
  clear_error();
  EDGE_REQUIRE(dims_ok(nx, ny, nz), "edge_vortex_ring: bad dims");
  EDGE_REQUIRE(c && ex && ey && pf && vf && smokef && color && num_samples > 0, "edge_vortex_ring: bad arguments");
  EDGE_REQUIRE(aligned16(smokef), "edge_vortex_ring: the smoke field must be 16-byte aligned");
  Grid g = make_grid(nx, ny, nz, 1.0f);
  hipLaunchKernelGGL(k_vortex_ring, dim3(blocks_for(ncell(g), RING_CHUNK)), dim3(RING_CHUNK), 0, (hipStream_t)stream,
                     g, w, rad, delta, make_float3(c[0], c[1], c[2]), make_float3(ex[0], ex[1], ex[2]),
                     make_float3(ey[0], ey[1], ey[2]), pf, vf, smokef, make_float3(color[0], color[1], color[2]),
                     num_samples);
  EDGE_CHECK_LAUNCH();
  return EDGE_OK;
}

int edge_get_central_vector(const float* ux, const float* uy, const float* uz, float* vc, int nx, int ny, int nz,
                            void* stream) {
  clear_error();
  EDGE_REQUIRE(dims_ok(nx, ny, nz), "edge_get_central_vector: bad dims");
  EDGE_REQUIRE(ux && uy && uz && vc, "edge_get_central_vector: null pointer");
  Grid g = make_grid(nx, ny, nz, 1.0f);
  hipLaunchKernelGGL(k_central_vector, dim3(blocks_for(ncell(g), 256)), dim3(256), 0, (hipStream_t)stream, g,
                     Mac{ux, uy, uz}, vc);
  EDGE_CHECK_LAUNCH();
  return EDGE_OK;
}

int edge_curl(const float* vf, float* cf, int nx, int ny, int nz, float dx, void* stream) {
  clear_error();
  EDGE_REQUIRE(dims_ok(nx, ny, nz), "edge_curl: bad dims");
  EDGE_REQUIRE(vf && cf && vf != cf, "edge_curl: null or aliased pointers");
  EDGE_REQUIRE(dx > 0.f, "edge_curl: dx must be positive");
  Grid g = make_grid(nx, ny, nz, dx);
  // 1/(2 dx) in f32, as the kernel computes it from its f32 argument
  const float inv_dist = 1.0f / (2.0f * dx);
  hipLaunchKernelGGL(k_curl, dim3(blocks_for(ncell(g), 256)), dim3(256), 0, (hipStream_t)stream, g, vf, cf,
                     inv_dist);
  EDGE_CHECK_LAUNCH();
  return EDGE_OK;
}

int edge_vtk_frame(const float* ux, const float* uy, const float* uz, const float* smoke, int nc, float* wn,
                   float* sm, int nx, int ny, int nz, float dx, void* stream) {
  clear_error();
  EDGE_REQUIRE(dims_ok(nx, ny, nz), "edge_vtk_frame: bad dims");
  EDGE_REQUIRE(ux && uy && uz && wn && (nc == 0 || (smoke && sm)), "edge_vtk_frame: null pointer");
  EDGE_REQUIRE(nc >= 0 && nc <= 4, "edge_vtk_frame: nc must be in 0..4 (got %d)", nc);
  EDGE_REQUIRE(dx > 0.f, "edge_vtk_frame: dx must be positive");
  Grid g = make_grid(nx, ny, nz, dx);
  hipLaunchKernelGGL(k_vtk_frame, dim3(blocks_for(ncell(g), 256)), dim3(256), 0, (hipStream_t)stream, g,
                     Mac{ux, uy, uz}, smoke, nc, wn, sm, 1.0f / (2.0f * dx));
  EDGE_CHECK_LAUNCH();
  return EDGE_OK;
}

int edge_split_central_vector(const float* vc, float* ux, float* uy, float* uz, int nx, int ny, int nz,
                              void* stream) {
  clear_error();
  EDGE_REQUIRE(dims_ok(nx, ny, nz), "edge_split_central_vector: bad dims");
  EDGE_REQUIRE(vc && ux && uy && uz, "edge_split_central_vector: null pointer");
  Grid g = make_grid(nx, ny, nz, 1.0f);
  int64_t nf = (int64_t)(nx + 1) * ny * nz + (int64_t)nx * (ny + 1) * nz + (int64_t)nx * ny * (nz + 1);
  hipLaunchKernelGGL(k_split_central, dim3(blocks_for(nf, 256)), dim3(256), 0, (hipStream_t)stream, g, vc,
                     Faces{{ux, uy, uz}});
  EDGE_CHECK_LAUNCH();
  return EDGE_OK;
}

}  // extern "C"
