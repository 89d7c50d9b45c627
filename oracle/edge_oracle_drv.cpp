// CPU oracle, part 2: the step-loop driver kernels of run_edge.py / run_ed8.py,
// the semi-Lagrangian advection of sim/hermite.py, the vortex-ring initial
// condition of sim/init_conditions.py and the particle kernels of sim/lp.py.
// TEST INFRASTRUCTURE ONLY (see edge_oracle.cpp for the rules and the layout).
#include "edge_oracle.h"

using namespace orc;

namespace {

template <class R>
struct FaceSet {  // three face grids of a MAC quantity, nc planes each
  Fld<R> f[3];
};

template <class R>
FaceSet<R> mkfaces(const R* fx, const R* fy, const R* fz, int nx, int ny, int nz) {
  return FaceSet<R>{{mk(fx, nx + 1, ny, nz), mk(fy, nx, ny + 1, nz), mk(fz, nx, ny, nz + 1)}};
}

// run_edge.py:360-373  map_u: u1_a(f) = F_a(f) . u0(phi_a(f)).
template <class R>
void map_u(const Mac<R>& U0, const FaceSet<R>& U1, const FaceSet<R>& F, const FaceSet<R>& PHI, R dx) {
  for (int a = 0; a < 3; ++a) {
    const Fld<R>& o = U1.f[a];
#pragma omp parallel for collapse(2)
    for (int i = 0; i < o.n0; ++i)
      for (int j = 0; j < o.n1; ++j)
        for (int k = 0; k < o.n2; ++k) {
          R p[3] = {PHI.f[a].at(0, i, j, k), PHI.f[a].at(1, i, j, k), PHI.f[a].at(2, i, j, k)};
          R v[3];
          interp_u_MAC(U0, p, dx, v);
          o.at(0, i, j, k) = F.f[a].at(0, i, j, k) * v[0] + F.f[a].at(1, i, j, k) * v[1] +
                             F.f[a].at(2, i, j, k) * v[2];
        }
  }
}

// run_edge.py:376-417  map_u_psi: pull u0 back through the cell map at each face.
template <class R>
void map_u_psi(const Mac<R>& U0, const FaceSet<R>& U1, const Map<R>& M, R dx) {
  for (int a = 0; a < 3; ++a) {
    const Fld<R>& o = U1.f[a];
#pragma omp parallel for collapse(2)
    for (int i = 0; i < o.n0; ++i)
      for (int j = 0; j < o.n1; ++j)
        for (int k = 0; k < o.n2; ++k) {
          R X[3], psi[3], G[3][3], v[3];
          face_coord(a, i, j, k, dx, X);
          map_eval(HK_GARM_FD, M, X, dx, psi, G);
          interp_u_MAC(U0, psi, dx, v);
          // transpose_matrix, then column a: (dpsi_0/dx_a, dpsi_1/dx_a, dpsi_2/dx_a)
          o.at(0, i, j, k) = G[0][a] * v[0] + G[1][a] * v[1] + G[2][a] * v[2];
        }
  }
}

// run_edge.py:420-427 / 469-477: psi_c of a cell from the six face points of phi.
template <class R>
void cell_from_faces(const FaceSet<R>& PHI, int i, int j, int k, R out[3]) {
  for (int a = 0; a < 3; ++a) {
    R s = PHI.f[0].at(a, i, j, k) + PHI.f[0].at(a, i + 1, j, k) + PHI.f[1].at(a, i, j, k) +
          PHI.f[1].at(a, i, j + 1, k) + PHI.f[2].at(a, i, j, k) + PHI.f[2].at(a, i, j, k + 1);
    out[a] = R(1) / R(6) * s;
  }
}

// run_edge.py:430-440 split_to_mac + the averaging of map_smoke_psi (443-466).
template <class R>
void cell_from_map(const Map<R>& M, int i, int j, int k, R dx, R out[3]) {
  R acc[3] = {0, 0, 0};
  const int off[6][4] = {{0, 0, 0, 0}, {0, 1, 0, 0}, {1, 0, 0, 0}, {1, 0, 1, 0}, {2, 0, 0, 0}, {2, 0, 0, 1}};
  for (int f = 0; f < 6; ++f) {
    R X[3], v[3];
    face_coord(off[f][0], i + off[f][1], j + off[f][2], k + off[f][3], dx, X);
    map_eval(HK_GARM_FD, M, X, dx, v, (R(*)[3])nullptr);
    for (int a = 0; a < 3; ++a) acc[a] += v[a];
  }
  for (int a = 0; a < 3; ++a) out[a] = R(1) / R(6) * acc[a];
}

}  // namespace

#define DEF_DRV(R, S)                                                                              \
  extern "C" void orc_map_u_##S(const R* u0x, const R* u0y, const R* u0z, R* u1x, R* u1y, R* u1z,    \
                                const R* Fx, const R* Fy, const R* Fz, const R* px, const R* py,     \
                                const R* pz, int nx, int ny, int nz, R dx) {                        \
    map_u(mkmac(u0x, u0y, u0z, nx, ny, nz), mkfaces((const R*)u1x, (const R*)u1y, (const R*)u1z, nx, \
                                                   ny, nz),                                         \
          mkfaces(Fx, Fy, Fz, nx, ny, nz), mkfaces(px, py, pz, nx, ny, nz), dx);                    \
  }                                                                                                 \
  extern "C" void orc_map_u_psi_##S(const R* u0x, const R* u0y, const R* u0z, R* u1x, R* u1y,        \
                                    R* u1z, const R* psi, const R* T, const R* dT, int nx, int ny,   \
                                    int nz, R dx) {                                                 \
    map_u_psi(mkmac(u0x, u0y, u0z, nx, ny, nz),                                                     \
              mkfaces((const R*)u1x, (const R*)u1y, (const R*)u1z, nx, ny, nz),                     \
              mkmap(psi, T, dT, (const R*)nullptr, nx, ny, nz), dx);                                \
  }                                                                                                 \
  /* run_edge.py:420-427 map_smoke / 469-477 clamp_smoke (mode 0 map, 1 clamp) */                   \
  extern "C" void orc_map_smoke_##S(int mode, const R* s0, R* s1, int nc, const R* px, const R* py, \
                                    const R* pz, int nx, int ny, int nz, R dx) {                    \
    Fld<R> S0 = mk(s0, nx, ny, nz), S1 = mk((const R*)s1, nx, ny, nz);                              \
    FaceSet<R> P = mkfaces(px, py, pz, nx, ny, nz);                                                 \
    _Pragma("omp parallel for collapse(2)") for (int i = 0; i < nx; ++i) for (int j = 0; j < ny;    \
                                                                             ++j) for (int k = 0;   \
                                                                                       k < nz;      \
                                                                                       ++k) {       \
      R pc[3], v[8], mn[8], mx[8];                                                                  \
      cell_from_faces(P, i, j, k, pc);                                                              \
      if (mode == 0) {                                                                              \
        interp_1(S0, nc, pc, dx, v);                                                                \
        for (int c = 0; c < nc; ++c) S1.at(c, i, j, k) = v[c];                                      \
      } else {                                                                                      \
        sample_min_max_1(S0, nc, pc, dx, mn, mx);                                                   \
        for (int c = 0; c < nc; ++c)                                                                \
          S1.at(c, i, j, k) = std::min(std::max(S1.at(c, i, j, k), mn[c]), mx[c]);                  \
      }                                                                                             \
    }                                                                                               \
  }                                                                                                 \
  /* run_edge.py:443-466 map_smoke_psi / 480-505 clamp_smoke_psi */                                  \
  extern "C" void orc_map_smoke_psi_##S(int mode, const R* s0, R* s1, int nc, const R* psi,         \
                                        const R* T, const R* dT, int nx, int ny, int nz, R dx) {    \
    Fld<R> S0 = mk(s0, nx, ny, nz), S1 = mk((const R*)s1, nx, ny, nz);                              \
    Map<R> M = mkmap(psi, T, dT, (const R*)nullptr, nx, ny, nz);                                    \
    _Pragma("omp parallel for collapse(2)") for (int i = 0; i < nx; ++i) for (int j = 0; j < ny;    \
                                                                             ++j) for (int k = 0;   \
                                                                                       k < nz;      \
                                                                                       ++k) {       \
      R pc[3], v[8], mn[8], mx[8];                                                                  \
      cell_from_map(M, i, j, k, dx, pc);                                                            \
      if (mode == 0) {                                                                              \
        interp_1(S0, nc, pc, dx, v);                                                                \
        for (int c = 0; c < nc; ++c) S1.at(c, i, j, k) = v[c];                                      \
      } else {                                                                                      \
        sample_min_max_1(S0, nc, pc, dx, mn, mx);                                                   \
        for (int c = 0; c < nc; ++c)                                                                \
          S1.at(c, i, j, k) = std::min(std::max(S1.at(c, i, j, k), mn[c]), mx[c]);                  \
      }                                                                                             \
    }                                                                                               \
  }                                                                                                 \
  /* run_edge.py:508-519 clamp_u on one face grid of dims n0,n1,n2 */                                \
  extern "C" void orc_clamp_u_##S(const R* u, const R* u1, R* u2, int n0, int n1, int n2) {         \
    Fld<R> U = mk(u, n0, n1, n2), U1 = mk(u1, n0, n1, n2), U2 = mk((const R*)u2, n0, n1, n2);       \
    _Pragma("omp parallel for collapse(2)") for (int i = 0; i < n0; ++i) for (int j = 0; j < n1;    \
                                                                             ++j) for (int k = 0;   \
                                                                                       k < n2;      \
                                                                                       ++k) {       \
      R l = sample(U1, 0, i - 1, j, k), r = sample(U1, 0, i + 1, j, k);                             \
      R b = sample(U1, 0, i, j - 1, k), t = sample(U1, 0, i, j + 1, k);                             \
      R a = sample(U1, 0, i, j, k - 1), c = sample(U1, 0, i, j, k + 1);                             \
      R mx = std::max(std::max(std::max(l, r), std::max(b, t)), std::max(a, c));                    \
      R mn = std::min(std::min(std::min(l, r), std::min(b, t)), std::min(a, c));                    \
      U2.at(0, i, j, k) = std::min(std::max(U.at(0, i, j, k), mn), mx);                             \
    }                                                                                               \
  }                                                                                                 \
  /* sim/hermite.py:1578-1621 advect_u_semi_lagrangian (dt is the full argument) */                  \
  extern "C" void orc_advect_u_sl_##S(const R* u0x, const R* u0y, const R* u0z, R* u1x, R* u1y,     \
                                      R* u1z, int nx, int ny, int nz, R dx, R dt) {                 \
    Mac<R> U = mkmac(u0x, u0y, u0z, nx, ny, nz);                                                    \
    FaceSet<R> O = mkfaces((const R*)u1x, (const R*)u1y, (const R*)u1z, nx, ny, nz);                \
    for (int a = 0; a < 3; ++a) {                                                                   \
      const Fld<R>& o = O.f[a];                                                                     \
      _Pragma("omp parallel for collapse(2)") for (int i = 0; i < o.n0; ++i) for (                  \
          int j = 0; j < o.n1; ++j) for (int k = 0; k < o.n2; ++k) {                                \
        R p1[3], v1[3], v2[3], v3[3], v4[3], v5[3], q[3];                                           \
        face_coord(a, i, j, k, dx, p1);                                                             \
        interp_u_MAC(U, p1, dx, v1);                                                                \
        for (int d = 0; d < 3; ++d) q[d] = p1[d] - v1[d] * dt * R(0.5);                             \
        interp_u_MAC(U, q, dx, v2);                                                                 \
        for (int d = 0; d < 3; ++d) q[d] = p1[d] - v2[d] * dt * R(0.5);                             \
        interp_u_MAC(U, q, dx, v3);                                                                 \
        for (int d = 0; d < 3; ++d) q[d] = p1[d] - v3[d] * dt;                                      \
        interp_u_MAC(U, q, dx, v4);                                                                 \
        for (int d = 0; d < 3; ++d)                                                                 \
          q[d] = p1[d] - (v1[d] + R(2) * v2[d] + R(2) * v3[d] + v4[d]) / R(6.0) * dt;               \
        interp_u_MAC(U, q, dx, v5);                                                                 \
        o.at(0, i, j, k) = v5[a];                                                                   \
      }                                                                                             \
    }                                                                                               \
  }                                                                                                 \
  /* run_edge.py:295-303 calc_max_speed */                                                          \
  extern "C" R orc_max_speed_##S(const R* ux, const R* uy, const R* uz, int nx, int ny, int nz) {   \
    Mac<R> U = mkmac(ux, uy, uz, nx, ny, nz);                                                       \
    R m = R(1.e-3);                                                                                 \
    for (int i = 0; i < nx; ++i)                                                                    \
      for (int j = 0; j < ny; ++j)                                                                  \
        for (int k = 0; k < nz; ++k) {                                                              \
          R u = R(0.5) * (U.ux.at(0, i, j, k) + U.ux.at(0, i + 1, j, k));                           \
          R v = R(0.5) * (U.uy.at(0, i, j, k) + U.uy.at(0, i, j + 1, k));                           \
          R w = R(0.5) * (U.uz.at(0, i, j, k) + U.uz.at(0, i, j, k + 1));                           \
          m = std::max(m, (R)std::sqrt(u * u + v * v + w * w));                                     \
        }                                                                                           \
    return m;                                                                                       \
  }                                                                                                 \
  /* run_ed8.py:190-223 / run_ge.py:184-217 split_psi_hermite: face phi_a, F_a from the map */       \
  extern "C" void orc_split_psi_hermite_##S(int kind, const R* psi, const R* T, const R* dT,        \
                                            const R* d2T, R* phx, R* phy, R* phz, R* Fx, R* Fy,     \
                                            R* Fz, int nx, int ny, int nz, R dx) {                  \
    Map<R> M = mkmap(psi, T, dT, d2T, nx, ny, nz);                                                  \
    FaceSet<R> PH = mkfaces((const R*)phx, (const R*)phy, (const R*)phz, nx, ny, nz);               \
    FaceSet<R> FF = mkfaces((const R*)Fx, (const R*)Fy, (const R*)Fz, nx, ny, nz);                  \
    for (int a = 0; a < 3; ++a) {                                                                   \
      const Fld<R>& o = PH.f[a];                                                                    \
      _Pragma("omp parallel for collapse(2)") for (int i = 0; i < o.n0; ++i) for (                  \
          int j = 0; j < o.n1; ++j) for (int k = 0; k < o.n2; ++k) {                                \
        R X[3], v[3], G[3][3];                                                                      \
        face_coord(a, i, j, k, dx, X);                                                              \
        map_eval(kind, M, X, dx, v, G);                                                             \
        for (int c = 0; c < 3; ++c) {                                                               \
          PH.f[a].at(c, i, j, k) = v[c];                                                            \
          FF.f[a].at(c, i, j, k) = G[c][a];                                                         \
        }                                                                                           \
      }                                                                                             \
    }                                                                                               \
  }                                                                                                 \
  /* sim/init_conditions.py:13-36 add_vortex_ring_and_smoke (pf = X, cell field) */                  \
  extern "C" void orc_vortex_ring_##S(R w, R rad, R delta, const R* c, const R* ex, const R* ey,    \
                                      const R* pf, R* vf, R* smokef, const R* color,                \
                                      int num_samples, int nx, int ny, int nz) {                    \
    Fld<R> P = mk(pf, nx, ny, nz), V = mk((const R*)vf, nx, ny, nz),                                \
           SM = mk((const R*)smokef, nx, ny, nz);                                                   \
    const R pi = R(3.14159265358979323846);                                                        \
    R curve_length = (R(2) * pi * rad) / R(num_samples);                                            \
    _Pragma("omp parallel for collapse(2)") for (int i = 0; i < nx; ++i) for (int j = 0; j < ny;    \
                                                                             ++j) for (int k = 0;   \
                                                                                       k < nz;      \
                                                                                       ++k) {       \
      for (int l = 0; l < num_samples; ++l) {                                                       \
        R theta = R(l) / R(num_samples) * R(2) * pi;                                                \
        R ps[3], pd[3], wv[3];                                                                      \
        for (int d = 0; d < 3; ++d) {                                                               \
          ps[d] = rad * (std::cos(theta) * ex[d] + std::sin(theta) * ey[d]) + c[d];                 \
          pd[d] = P.at(d, i, j, k) - ps[d];                                                         \
          wv[d] = w * (-std::sin(theta) * ex[d] + std::cos(theta) * ey[d]);                         \
        }                                                                                           \
        R r = std::sqrt(pd[0] * pd[0] + pd[1] * pd[1] + pd[2] * pd[2]);                             \
        R rd = r / delta;                                                                           \
        R fac = curve_length * (R(-1) / (R(4) * pi * (r * r * r)) * (R(1) - std::exp(-(rd * rd * rd))));\
        R cr[3] = {pd[1] * wv[2] - pd[2] * wv[1], pd[2] * wv[0] - pd[0] * wv[2],                    \
                   pd[0] * wv[1] - pd[1] * wv[0]};                                                  \
        for (int d = 0; d < 3; ++d) V.at(d, i, j, k) += fac * cr[d];                                \
        SM.at(3, i, j, k) += curve_length * std::exp(-(rd * rd * rd));                              \
      }                                                                                             \
      if (SM.at(3, i, j, k) > R(0.002)) {                                                           \
        SM.at(3, i, j, k) = R(1);                                                                   \
        for (int d = 0; d < 3; ++d) SM.at(d, i, j, k) = color[d];                                   \
      } else {                                                                                      \
        for (int d = 0; d < 4; ++d) SM.at(d, i, j, k) = R(0);                                       \
      }                                                                                             \
    }                                                                                               \
  }
ORC_BOTH(DEF_DRV)

// --------------------------------------------------------------------------
// Particles (sim/lp.py).  Positions/velocities are AoS-free: 3 planes of n.
// --------------------------------------------------------------------------
namespace {
template <class R>
void particle_u(const Mac<R>& U, const R* pos, long n, long q, R dx, R out[3]) {
  R p[3] = {pos[q], pos[n + q], pos[2 * n + q]}, g[3];
  out[0] = interp_grad_2(U.ux, 0, p, dx, R(0), R(0.5), R(0.5), g);
  out[1] = interp_grad_2(U.uy, 0, p, dx, R(0.5), R(0), R(0.5), g);
  out[2] = interp_grad_2(U.uz, 0, p, dx, R(0.5), R(0.5), R(0), g);
}
}  // namespace

#define DEF_LP(R, S)                                                                               \
  /* sim/lp.py:6-29 */                                                                              \
  extern "C" void orc_lp_acc_##S(const R* ux, const R* uy, const R* uz, int nx, int ny, int nz,     \
                                 const R* vel, const R* pos, R* acc, long n, long num, R dx,        \
                                 R dt, R St, R gravity) {                                           \
    Mac<R> U = mkmac(ux, uy, uz, nx, ny, nz);                                                       \
    _Pragma("omp parallel for") for (long q = 0; q < n; ++q) {                                      \
      for (int a = 0; a < 3; ++a) acc[a * n + q] = 0;                                               \
      if (q < num) {                                                                                \
        R k = R(1) / St, uu[3];                                                                     \
        particle_u(U, pos, n, q, dx, uu);                                                           \
        acc[n + q] += gravity;                                                                      \
        for (int a = 0; a < 3; ++a) {                                                               \
          acc[a * n + q] -= k * (vel[a * n + q] - uu[a]);                                           \
          acc[a * n + q] /= (R(1) + k * dt);                                                        \
        }                                                                                           \
      }                                                                                             \
    }                                                                                               \
  }                                                                                                 \
  /* sim/lp.py:32-70 */                                                                             \
  extern "C" void orc_lp_vel_##S(const R* acc, R* vel, const R* pos, long n, long num, R dt, R dx,  \
                                 R Lx, R Ly, R Lz) {                                                \
    _Pragma("omp parallel for") for (long q = 0; q < std::min(num, n); ++q) {                           \
      for (int a = 0; a < 3; ++a) vel[a * n + q] += acc[a * n + q] * dt;                            \
      R* v = vel;                                                                                   \
      if (pos[n + q] < dx && v[n + q] < 0) v[n + q] = 0;                                            \
      if (pos[n + q] > Ly - R(3) * dx && v[n + q] > 0) v[n + q] = 0;                                \
      if (pos[q] < R(3) * dx && v[q] < 0) v[q] = 0;                                                 \
      if (pos[q] > Lx - R(3) * dx && v[q] > 0) v[q] = 0;                                            \
      if (pos[2 * n + q] < R(3) * dx && v[2 * n + q] < 0) v[2 * n + q] = 0;                         \
      if (pos[2 * n + q] > Lz - R(3) * dx && v[2 * n + q] > 0) v[2 * n + q] = 0;                    \
    }                                                                                               \
  }                                                                                                 \
  /* sim/lp.py:73-97 */                                                                             \
  extern "C" void orc_lp_vel_diff_##S(const R* ux, const R* uy, const R* uz, int nx, int ny,        \
                                      int nz, const R* vel, R* diff, const R* pos, long n,          \
                                      long num, R dx, R St, R density_ratio, R radius_ratio) {      \
    Mac<R> U = mkmac(ux, uy, uz, nx, ny, nz);                                                       \
    R er = R(1.0) / (radius_ratio * radius_ratio * radius_ratio);                                   \
    _Pragma("omp parallel for") for (long q = 0; q < std::min(num, n); ++q) {                           \
      R k = R(1.0) / St * density_ratio * er, uu[3];                                                \
      particle_u(U, pos, n, q, dx, uu);                                                             \
      for (int a = 0; a < 3; ++a) diff[a * n + q] = k * (vel[a * n + q] - uu[a]);                   \
    }                                                                                               \
  }                                                                                                 \
  /* sim/lp.py:100-156 P2G of the drag (sequential: the reference uses atomics) */                   \
  extern "C" void orc_lp_p2g_##S(R* dx_f, R* dy_f, R* dz_f, const R* diff, const R* pos,            \
                                 const R* ref_radius, long n, long num, R dx, R radius, R Lx,       \
                                 R Ly, R Lz, int nx, int ny, int nz) {                              \
    FaceSet<R> D = mkfaces((const R*)dx_f, (const R*)dy_f, (const R*)dz_f, nx, ny, nz);             \
    for (int a = 0; a < 3; ++a) std::fill(D.f[a].d, D.f[a].d + D.f[a].N(), R(0));                  \
    for (long q = 0; q < std::min(num, n); ++q) {                                                       \
      R P[3] = {pos[q], pos[n + q], pos[2 * n + q]};                                                \
      if (P[0] < R(3) * dx || P[1] < R(3) * dx || P[2] < R(3) * dx || P[0] > Lx - R(3) * dx ||      \
          P[1] > Ly - R(3) * dx || P[2] > Lz - R(3) * dx)                                           \
        continue;                                                                                   \
      R p[3] = {P[0] / dx, P[1] / dx, P[2] / dx};                                                   \
      R rr = ref_radius[q];                                                                         \
      R rr3 = rr * rr * rr, rad3 = radius * radius * radius;                                        \
      for (int a = 0; a < 3; ++a) {                                                                 \
        R sh[3] = {R(0.5), R(0.5), R(0.5)};                                                         \
        sh[a] = R(0);                                                                               \
        int base[3];                                                                                \
        for (int d = 0; d < 3; ++d) base[d] = (int)(p[d] - sh[d]);                                  \
        int lim[3] = {nx, ny, nz};                                                                  \
        lim[a] += 1;                                                                                \
        for (int oi = -1; oi < 3; ++oi)                                                             \
          for (int oj = -1; oj < 3; ++oj)                                                           \
            for (int ok = -1; ok < 3; ++ok) {                                                       \
              int f[3] = {base[0] + oi, base[1] + oj, base[2] + ok};                                \
              if (f[0] < 0 || f[0] >= lim[0] || f[1] < 0 || f[1] >= lim[1] || f[2] < 0 ||           \
                  f[2] >= lim[2])                                                                   \
                continue;                                                                           \
              R wgt = N_2(p[0] - R(f[0]) - sh[0]) * N_2(p[1] - R(f[1]) - sh[1]) *                   \
                      N_2(p[2] - R(f[2]) - sh[2]);                                                  \
              D.f[a].at(0, f[0], f[1], f[2]) += diff[a * n + q] * wgt * rr3 / rad3;                 \
            }                                                                                       \
      }                                                                                             \
    }                                                                                               \
  }                                                                                                 \
  /* sim/lp.py:158-181 */                                                                           \
  extern "C" void orc_lp_pos_##S(const R* vel, R* pos, long n, long num, R dt, R Lx, R Ly, R Lz) {  \
    R L[3] = {Lx, Ly, Lz};                                                                          \
    _Pragma("omp parallel for") for (long q = 0; q < std::min(num, n); ++q) {                           \
      for (int a = 0; a < 3; ++a) pos[a * n + q] += vel[a * n + q] * dt;                            \
      for (int a = 0; a < 3; ++a) {                                                                 \
        if (pos[a * n + q] < 0) pos[a * n + q] = 0;                                                 \
        if (pos[a * n + q] > L[a]) pos[a * n + q] = L[a];                                           \
      }                                                                                             \
    }                                                                                               \
  }                                                                                                 \
  /* sim/lp.py:183-218 on one face grid */                                                          \
  extern "C" void orc_lp_viscous_##S(const R* u, R* vis, int n0, int n1, int n2, R dx, R Re) {      \
    Fld<R> U = mk(u, n0, n1, n2), V = mk((const R*)vis, n0, n1, n2);                                \
    _Pragma("omp parallel for collapse(2)") for (int i = 0; i < n0; ++i) for (int j = 0; j < n1;    \
                                                                             ++j) for (int k = 0;   \
                                                                                       k < n2;      \
                                                                                       ++k) {       \
      R visc = R(1.0) / Re;                                                                         \
      V.at(0, i, j, k) = visc *                                                                     \
                         (sample(U, 0, i + 1, j, k) + sample(U, 0, i - 1, j, k) +                   \
                          sample(U, 0, i, j + 1, k) + sample(U, 0, i, j - 1, k) +                   \
                          sample(U, 0, i, j, k + 1) + sample(U, 0, i, j, k - 1) +                   \
                          R(-6) * sample(U, 0, i, j, k)) /                                          \
                         dx / dx;                                                                   \
    }                                                                                               \
  }
ORC_BOTH(DEF_LP)

// Frame output (run_edge.py:719-725): cell-centred velocity and its curl.
namespace {

// util/taichi_utils.py:95-100 get_central_vector: each component the mean of
// the cell's two faces on its axis.
template <class R>
void central_vector(const Mac<R>& U, const Fld<R>& V) {
#pragma omp parallel for collapse(2)
  for (int i = 0; i < V.n0; ++i)
    for (int j = 0; j < V.n1; ++j)
      for (int k = 0; k < V.n2; ++k) {
        V.at(0, i, j, k) = R(0.5) * (U.ux.at(0, i + 1, j, k) + U.ux.at(0, i, j, k));
        V.at(1, i, j, k) = R(0.5) * (U.uy.at(0, i, j + 1, k) + U.uy.at(0, i, j, k));
        V.at(2, i, j, k) = R(0.5) * (U.uz.at(0, i, j, k + 1) + U.uz.at(0, i, j, k));
      }
}

// util/taichi_utils.py:70-92 curl: inv_dist = 1/(2 dx) in R; the six
// neighbours through sample()'s index clamp.
template <class R>
void curl(const Fld<R>& vf, const Fld<R>& cf, R dx) {
  const R inv_dist = R(1) / (R(2) * dx);
#pragma omp parallel for collapse(2)
  for (int i = 0; i < cf.n0; ++i)
    for (int j = 0; j < cf.n1; ++j)
      for (int k = 0; k < cf.n2; ++k) {
        R vr[3], vl[3], vt[3], vb[3], vc[3], va[3];
        for (int c = 0; c < 3; ++c) {
          vr[c] = sample(vf, c, i + 1, j, k);
          vl[c] = sample(vf, c, i - 1, j, k);
          vt[c] = sample(vf, c, i, j + 1, k);
          vb[c] = sample(vf, c, i, j - 1, k);
          vc[c] = sample(vf, c, i, j, k + 1);
          va[c] = sample(vf, c, i, j, k - 1);
        }
        const R d_vx_dz = inv_dist * (vc[0] - va[0]);
        const R d_vx_dy = inv_dist * (vt[0] - vb[0]);
        const R d_vy_dx = inv_dist * (vr[1] - vl[1]);
        const R d_vy_dz = inv_dist * (vc[1] - va[1]);
        const R d_vz_dx = inv_dist * (vr[2] - vl[2]);
        const R d_vz_dy = inv_dist * (vt[2] - vb[2]);
        cf.at(0, i, j, k) = d_vz_dy - d_vy_dz;
        cf.at(1, i, j, k) = d_vx_dz - d_vz_dx;
        cf.at(2, i, j, k) = d_vy_dx - d_vx_dy;
      }
}

}  // namespace

#define DEF_OUT(R, S)                                                                              \
  extern "C" void orc_central_vector_##S(const R* ux, const R* uy, const R* uz, R* vc, int nx,      \
                                         int ny, int nz) {                                          \
    central_vector(mkmac(ux, uy, uz, nx, ny, nz), mk((const R*)vc, nx, ny, nz));                   \
  }                                                                                                 \
  extern "C" void orc_curl_##S(const R* vf, R* cf, int nx, int ny, int nz, R dx) {                 \
    curl(mk(vf, nx, ny, nz), mk((const R*)cf, nx, ny, nz), dx);                                     \
  }
ORC_BOTH(DEF_OUT)
