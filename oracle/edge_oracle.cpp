// CPU oracle for the EDGE flow-map time step.  TEST INFRASTRUCTURE ONLY.
//
// This file restates, literally and unfused, the reference algorithm of
// pfm-gatech/EDGE-Flow-Maps (Python + Taichi) so that the HIP kernels of the
// product can be checked against it.  Only tests/, __graft_entry__.smoke() and
// bench.py's cpu_baseline leg may load it, and only as the checker / the timed
// CPU baseline; the product path never calls it.
//
// Every function is templated on the real type R (float = the reference's
// default_fp=ti.f32, run_edge.py:38,161; double = the FP64 oracle run that each
// result is also compared against).  Citations are to the reference files.
//
// Layout (shared with the HIP side): a field of nc components on an n0*n1*n2
// grid is nc contiguous planes, plane c at offset c*n0*n1*n2, index
// (i*n1 + j)*n2 + k (k fastest; Taichi's ti.field index order, with the
// vector components split out into planes).
//
// Deviations that the reference leaves unspecified (documented, not hidden):
//  * Taichi's f32 atomic reductions (sim/mgpcg.py:127-130,153,169-175) have no
//    defined order; here every reduction accumulates sequentially in double.
//  * The host-side alpha/beta arithmetic (sim/mgpcg.py:239,272) is done in
//    double from the R-rounded dot products, then rounded to R, as the
//    reference does through its 0-d f32 fields.
#include "edge_oracle.h"

// ==========================================================================
// extern "C" entry points.  Suffix _f = float (FP32), _d = double (FP64).
// ==========================================================================
using namespace orc;


// ---- point evaluations (test hooks) -------------------------------------
#define DEF_POINTS(R, S)                                                                           \
  extern "C" void orc_eval_bspline_##S(const R* f, int n0, int n1, int n2, const R* BL,             \
                                       const R* pts, long n, R dx, R* val, R* grad) {               \
    Fld<R> F = mk(f, n0, n1, n2);                                                                   \
    _Pragma("omp parallel for") for (long q = 0; q < n; ++q) {                                      \
      const R* p = pts + 3 * q;                                                                     \
      if (grad)                                                                                     \
        val[q] = interp_grad_2(F, 0, p, dx, BL[0], BL[1], BL[2], grad + 3 * q);                     \
      else                                                                                          \
        val[q] = interp_2(F, 0, p, dx, BL[0], BL[1], BL[2]);                                        \
    }                                                                                               \
  }                                                                                                 \
  extern "C" void orc_eval_umac_##S(const R* ux, const R* uy, const R* uz, int nx, int ny, int nz,  \
                                    const R* pts, long n, R dx, R* val, R* J) {                     \
    Mac<R> U = mkmac(ux, uy, uz, nx, ny, nz);                                                       \
    _Pragma("omp parallel for") for (long q = 0; q < n; ++q) {                                      \
      if (J) {                                                                                      \
        R JJ[3][3];                                                                                 \
        interp_u_MAC_grad(U, pts + 3 * q, dx, val + 3 * q, JJ);                                     \
        for (int a = 0; a < 3; ++a)                                                                 \
          for (int b = 0; b < 3; ++b) J[9 * q + 3 * a + b] = JJ[a][b];                              \
      } else {                                                                                      \
        interp_u_MAC(U, pts + 3 * q, dx, val + 3 * q);                                              \
      }                                                                                             \
    }                                                                                               \
  }                                                                                                 \
  extern "C" void orc_eval_hermite_##S(int kind, const R* psi, const R* T, const R* dT,             \
                                       const R* d2T, int nx, int ny, int nz, const R* pts, long n,  \
                                       R dx, R* val, R* G) {                                        \
    Map<R> M = mkmap(psi, T, dT, d2T, nx, ny, nz);                                                  \
    _Pragma("omp parallel for") for (long q = 0; q < n; ++q) {                                      \
      R GG[3][3];                                                                                   \
      map_eval(kind, M, pts + 3 * q, dx, val + 3 * q, G ? GG : nullptr);                            \
      if (G)                                                                                        \
        for (int a = 0; a < 3; ++a)                                                                 \
          for (int b = 0; b < 3; ++b) G[9 * q + 3 * a + b] = GG[a][b];                              \
    }                                                                                               \
  }                                                                                                 \
  extern "C" void orc_eval_linear_##S(const R* f, int nc, int n0, int n1, int n2, const R* pts,     \
                                      long n, R dx, R* val, R* mini, R* maxi) {                     \
    Fld<R> F = mk(f, n0, n1, n2);                                                                   \
    _Pragma("omp parallel for") for (long q = 0; q < n; ++q) {                                      \
      if (val) interp_1(F, nc, pts + 3 * q, dx, val + nc * q);                                      \
      if (mini) sample_min_max_1(F, nc, pts + 3 * q, dx, mini + nc * q, maxi + nc * q);             \
    }                                                                                               \
  }
ORC_BOTH(DEF_POINTS)

// ---- map marches ---------------------------------------------------------
#define DEF_MARCH(R, S)                                                                            \
  /* sim/hermite.py:786-803; pts SoA (3 planes of n) */                                              \
  extern "C" void orc_advance_map_0_##S(R* pts, long n, const R* ux, const R* uy, const R* uz,      \
                                        int nx, int ny, int nz, R dx, R dt) {                       \
    Mac<R> U = mkmac(ux, uy, uz, nx, ny, nz);                                                       \
    _Pragma("omp parallel for") for (long q = 0; q < n; ++q) {                                      \
      R p[3] = {pts[q], pts[n + q], pts[2 * n + q]};                                                \
      advance_map_0_pt(p, U, dx, dt);                                                               \
      for (int a = 0; a < 3; ++a) pts[a * n + q] = p[a];                                            \
    }                                                                                               \
  }                                                                                                 \
  /* sim/hermite.py:863-918; psi 3 planes of n, T 9 planes (Tx,Ty,Tz) of n */                       \
  extern "C" void orc_advance_map_1_transpose_##S(R* psi, R* T, long n, const R* ux, const R* uy,   \
                                                  const R* uz, int nx, int ny, int nz, R dx, R dt) {\
    Mac<R> U = mkmac(ux, uy, uz, nx, ny, nz);                                                       \
    _Pragma("omp parallel for") for (long q = 0; q < n; ++q) {                                      \
      R p[3], TT[3][3];                                                                             \
      for (int a = 0; a < 3; ++a) {                                                                 \
        p[a] = psi[a * n + q];                                                                      \
        for (int b = 0; b < 3; ++b) TT[a][b] = T[(3 * a + b) * n + q];                              \
      }                                                                                             \
      advance_map_1_transpose_pt(p, TT, U, dx, dt);                                                 \
      for (int a = 0; a < 3; ++a) {                                                                 \
        psi[a * n + q] = p[a];                                                                      \
        for (int b = 0; b < 3; ++b) T[(3 * a + b) * n + q] = TT[a][b];                              \
      }                                                                                             \
    }                                                                                               \
  }                                                                                                 \
  /* run_edge.py:336-357 over one face grid of n points: phi 3 planes, F 3 planes */                \
  extern "C" void orc_advance_map_col_##S(R* phi, R* F, long n, const R* ux, const R* uy,           \
                                          const R* uz, int nx, int ny, int nz, R dx, R dt) {        \
    Mac<R> U = mkmac(ux, uy, uz, nx, ny, nz);                                                       \
    _Pragma("omp parallel for") for (long q = 0; q < n; ++q) {                                      \
      R p[3], f[3];                                                                                 \
      for (int a = 0; a < 3; ++a) { p[a] = phi[a * n + q]; f[a] = F[a * n + q]; }                   \
      advance_map_col_pt(p, f, U, dx, dt);                                                          \
      for (int a = 0; a < 3; ++a) { phi[a * n + q] = p[a]; F[a * n + q] = f[a]; }                   \
    }                                                                                               \
  }                                                                                                 \
  /* sim/hermite.py:1426-1492: fused EDGE march, old map in, new map out */                          \
  extern "C" void orc_march_edge_##S(const R* psi, const R* T, const R* dT, R* psi_o, R* T_o,       \
                                     R* dT_o, const R* ux, const R* uy, const R* uz, const R* X,    \
                                     int nx, int ny, int nz, R dx, R dt, R eps) {                   \
    Map<R> M = mkmap(psi, T, dT, (const R*)nullptr, nx, ny, nz);                                    \
    Map<R> O = mkmap((const R*)psi_o, (const R*)T_o, (const R*)dT_o, (const R*)nullptr, nx, ny, nz);\
    Mac<R> U = mkmac(ux, uy, uz, nx, ny, nz);                                                       \
    Fld<R> XX = mk(X, nx, ny, nz);                                                                  \
    _Pragma("omp parallel for collapse(2)") for (int i = 0; i < nx; ++i) for (int j = 0; j < ny;    \
                                                                             ++j) for (int k = 0;   \
                                                                                       k < nz; ++k) \
        edge_march_cell(M, O, U, XX, i, j, k, dx, dt, eps);                                         \
  }                                                                                                 \
  /* sim/hermite.py:921-1046 composition for n points with their rows T (9 planes) */                \
  extern "C" void orc_compose_##S(int kind, const R* psi, const R* T, const R* dT, const R* d2T,    \
                                  int nx, int ny, int nz, R* pts, R* Tp, long n, R dx) {            \
    Map<R> M = mkmap(psi, T, dT, d2T, nx, ny, nz);                                                  \
    _Pragma("omp parallel for") for (long q = 0; q < n; ++q) {                                      \
      R p[3], TT[3][3];                                                                             \
      for (int a = 0; a < 3; ++a) {                                                                 \
        p[a] = pts[a * n + q];                                                                      \
        if (Tp)                                                                                     \
          for (int b = 0; b < 3; ++b) TT[a][b] = Tp[(3 * a + b) * n + q];                           \
      }                                                                                             \
      if (Tp) {                                                                                     \
        compose_pt(kind, M, p, TT, dx);                                                             \
        for (int a = 0; a < 3; ++a)                                                                 \
          for (int b = 0; b < 3; ++b) Tp[(3 * a + b) * n + q] = TT[a][b];                           \
      } else {                                                                                      \
        R v[3];                                                                                     \
        map_eval(kind, M, p, dx, v, (R(*)[3])nullptr);                                              \
        for (int a = 0; a < 3; ++a) p[a] = v[a];                                                    \
      }                                                                                             \
      for (int a = 0; a < 3; ++a) pts[a * n + q] = p[a];                                            \
    }                                                                                               \
  }                                                                                                 \
  /* sim/hermite.py:1098-1151 calc_ed8_3 from 8 point fields (8 x 3 planes of N) */                  \
  extern "C" void orc_calc_ed8_##S(const R* P8, long N, R eps, R* psi, R* T, R* dT, R* d2T) {       \
    _Pragma("omp parallel for") for (long q = 0; q < N; ++q) {                                      \
      R P[8][3], ps[3], TT[3][3], DD[3][3], d2[3];                                                  \
      for (int e = 0; e < 8; ++e)                                                                   \
        for (int a = 0; a < 3; ++a) P[e][a] = P8[(size_t)(3 * e + a) * N + q];                     \
      calc_ed8_cell(P, eps, ps, TT, DD, d2);                                                        \
      for (int a = 0; a < 3; ++a) {                                                                 \
        psi[a * N + q] = ps[a];                                                                     \
        d2T[a * N + q] = d2[a];                                                                     \
        for (int b = 0; b < 3; ++b) {                                                               \
          T[(3 * a + b) * N + q] = TT[a][b];                                                        \
          dT[(3 * a + b) * N + q] = DD[a][b];                                                       \
        }                                                                                           \
      }                                                                                             \
    }                                                                                               \
  }                                                                                                 \
  /* sim/hermite.py:1154-1177 calc_ed4_1 from 4 point fields */                                      \
  extern "C" void orc_calc_ed4_1_##S(const R* P4, long N, R eps, R* psi, R* T) {                    \
    _Pragma("omp parallel for") for (long q = 0; q < N; ++q) {                                      \
      R P[4][3], ps[3], TT[3][3];                                                                   \
      for (int e = 0; e < 4; ++e)                                                                   \
        for (int a = 0; a < 3; ++a) P[e][a] = P4[(size_t)(3 * e + a) * N + q];                     \
      calc_ed4_1_cell(P, eps, ps, TT);                                                              \
      for (int a = 0; a < 3; ++a) {                                                                 \
        psi[a * N + q] = ps[a];                                                                     \
        for (int b = 0; b < 3; ++b) T[(3 * a + b) * N + q] = TT[a][b];                             \
      }                                                                                             \
    }                                                                                               \
  }
ORC_BOTH(DEF_MARCH)

// sim/hermite.py:1180-1287  calc_fd_3_from_grad (want_dT) / _from_grad_grad.
template <class R>
static void fd3(const R* T, R* dT, R* d2T, int nx, int ny, int nz, R dx, int want_dT) {
  size_t N = (size_t)nx * ny * nz;
  if (want_dT) {
    for (int c = 0; c < 3; ++c) {
      Fld<R> Tc = mk(T + 3 * N * c, nx, ny, nz);
      Fld<R> D = mk(dT + 3 * N * c, nx, ny, nz);
#pragma omp parallel for collapse(2)
      for (int i = 0; i < nx; ++i)
        for (int j = 0; j < ny; ++j)
          for (int k = 0; k < nz; ++k) {
            D.at(0, i, j, k) = ((sample(Tc, 1, i, j, k + 1) - sample(Tc, 1, i, j, k - 1)) +
                                (sample(Tc, 2, i, j + 1, k) - sample(Tc, 2, i, j - 1, k))) / R(4) / dx;
            D.at(1, i, j, k) = ((sample(Tc, 0, i, j, k + 1) - sample(Tc, 0, i, j, k - 1)) +
                                (sample(Tc, 2, i + 1, j, k) - sample(Tc, 2, i - 1, j, k))) / R(4) / dx;
            D.at(2, i, j, k) = ((sample(Tc, 0, i, j + 1, k) - sample(Tc, 0, i, j - 1, k)) +
                                (sample(Tc, 1, i + 1, j, k) - sample(Tc, 1, i - 1, j, k))) / R(4) / dx;
          }
    }
  }
  for (int c = 0; c < 3; ++c) {
    Fld<R> D = mk(dT + 3 * N * c, nx, ny, nz);
    Fld<R> E = mk(d2T + N * c, nx, ny, nz);
#pragma omp parallel for collapse(2)
    for (int i = 0; i < nx; ++i)
      for (int j = 0; j < ny; ++j)
        for (int k = 0; k < nz; ++k)
          E.at(0, i, j, k) = ((sample(D, 0, i + 1, j, k) - sample(D, 0, i - 1, j, k)) +
                              (sample(D, 1, i, j + 1, k) - sample(D, 1, i, j - 1, k)) +
                              (sample(D, 2, i, j, k + 1) - sample(D, 2, i, j, k - 1))) / R(6) / dx;
  }
}
extern "C" void orc_fd3_f(const float* T, float* dT, float* d2T, int nx, int ny, int nz, float dx, int want_dT) {
  fd3<float>(T, dT, d2T, nx, ny, nz, dx, want_dT);
}
extern "C" void orc_fd3_d(const double* T, double* dT, double* d2T, int nx, int ny, int nz, double dx, int want_dT) {
  fd3<double>(T, dT, d2T, nx, ny, nz, dx, want_dT);
}
