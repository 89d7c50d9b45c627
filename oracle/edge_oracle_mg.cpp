// CPU oracle, part 3: multigrid-preconditioned CG pressure projection,
// sim/mgpcg.py (MGPCG, MGPCG_3) and sim/mgpcg_solid.py (MGPCG_Solid,
// MGPCG_Solid_3).  TEST INFRASTRUCTURE ONLY (see edge_oracle.cpp).
//
// One entry point per precision runs a whole Poisson() call
// (apply_bc -> divergence [-> correct_divergence] -> init -> solve ->
// get_result -> subtract_grad_p -> apply_bc); two more expose the V-cycle and
// the operator so the GPU pieces can be checked one at a time.
#include <array>
#include <vector>

#include "edge_oracle.h"

using namespace orc;

namespace {

template <class R>
struct MGO {
  int L = 0;
  bool solid = false;
  int bt[3][2];                     // boundary_types (2 = wall)
  std::vector<std::array<int, 3>> dims;
  std::vector<std::vector<R>> r, z;
  std::vector<std::vector<int>> bm;  // solid only
  std::vector<R> x, p, Ap;

  size_t n(int l) const { return (size_t)dims[l][0] * dims[l][1] * dims[l][2]; }
  size_t id(int l, int i, int j, int k) const { return ((size_t)i * dims[l][1] + j) * dims[l][2] + k; }
  bool is_solid(int l, int i, int j, int k) const { return solid && bm[l][id(l, i, j, k)] > 0; }

  // sim/mgpcg.py:87-99 / sim/mgpcg_solid.py:99-111
  R neighbor_sum(int l, const std::vector<R>& v, int i, int j, int k) const {
    R ret = 0;
    int I[3] = {i, j, k};
    for (int a = 0; a < 3; ++a) {
      int Ip[3] = {i, j, k}, Im[3] = {i, j, k};
      Ip[a] += 1;
      Im[a] -= 1;
      if (I[a] < dims[l][a] - 1 && !is_solid(l, Ip[0], Ip[1], Ip[2])) ret += v[id(l, Ip[0], Ip[1], Ip[2])];
      if (I[a] > 0 && !is_solid(l, Im[0], Im[1], Im[2])) ret += v[id(l, Im[0], Im[1], Im[2])];
    }
    return ret;
  }
  // sim/mgpcg.py:101-110 / sim/mgpcg_solid.py:113-130
  R num_fluid_neighbors(int l, int i, int j, int k) const {
    R num = R(6);
    int I[3] = {i, j, k};
    for (int a = 0; a < 3; ++a) {
      int Ip[3] = {i, j, k}, Im[3] = {i, j, k};
      Ip[a] += 1;
      Im[a] -= 1;
      if (I[a] <= 0 && bt[a][0] == 2)
        num -= R(1);
      else if (solid && I[a] > 0 && is_solid(l, Im[0], Im[1], Im[2]))
        num -= R(1);
      if (I[a] >= dims[l][a] - 1 && bt[a][1] == 2)
        num -= R(1);
      else if (solid && I[a] < dims[l][a] - 1 && is_solid(l, Ip[0], Ip[1], Ip[2]))
        num -= R(1);
    }
    return num;
  }

  template <class Fn>
  void for_cells(int l, Fn fn) const {
#pragma omp parallel for collapse(2)
    for (int i = 0; i < dims[l][0]; ++i)
      for (int j = 0; j < dims[l][1]; ++j)
        for (int k = 0; k < dims[l][2]; ++k) fn(i, j, k);
  }

  void compute_Ap() {  // sim/mgpcg.py:112-117
    for_cells(0, [&](int i, int j, int k) {
      if (is_solid(0, i, j, k)) return;
      Ap[id(0, i, j, k)] = num_fluid_neighbors(0, i, j, k) * p[id(0, i, j, k)] - neighbor_sum(0, p, i, j, k);
    });
  }
  double reduce(const std::vector<R>& a, const std::vector<R>& b) const {  // sim/mgpcg.py:126-130
    double s = 0;
    for (int i = 0; i < dims[0][0]; ++i)
      for (int j = 0; j < dims[0][1]; ++j)
        for (int k = 0; k < dims[0][2]; ++k) {
          if (is_solid(0, i, j, k)) continue;
          size_t q = id(0, i, j, k);
          s += (double)(a[q] * b[q]);
        }
    return s;
  }
  void restrict_(int l) {  // sim/mgpcg.py:147-153
    for (int i = 0; i < dims[l][0]; ++i)
      for (int j = 0; j < dims[l][1]; ++j)
        for (int k = 0; k < dims[l][2]; ++k) {
          if (is_solid(l, i, j, k)) continue;
          size_t q = id(l, i, j, k);
          R res = r[l][q] - (num_fluid_neighbors(l, i, j, k) * z[l][q] - neighbor_sum(l, z[l], i, j, k));
          r[l + 1][id(l + 1, i / 2, j / 2, k / 2)] += res * R(1.0) / R(2.0);
        }
  }
  void prolongate(int l) {  // sim/mgpcg.py:155-158 (unmasked in both solvers)
    for_cells(l, [&](int i, int j, int k) { z[l][id(l, i, j, k)] += z[l + 1][id(l + 1, i / 2, j / 2, k / 2)]; });
  }
  void smooth(int l, int phase) {  // sim/mgpcg.py:160-167
    for_cells(l, [&](int i, int j, int k) {
      if (((i + j + k) & 1) != phase || is_solid(l, i, j, k)) return;
      size_t q = id(l, i, j, k);
      z[l][q] = (r[l][q] + neighbor_sum(l, z[l], i, j, k)) / num_fluid_neighbors(l, i, j, k);
    });
  }
  void recenter(std::vector<R>& v) {  // sim/mgpcg.py:169-175 (mean over all cells)
    double m = 0;
    double cnt = (double)n(0);
    for (size_t q = 0; q < n(0); ++q) m += (double)(v[q] / (R)cnt);
    R mean = (R)m;
    for (size_t q = 0; q < n(0); ++q) v[q] -= mean;
  }
  void precondition() {  // sim/mgpcg.py:177-195
    std::fill(z[0].begin(), z[0].end(), R(0));
    for (int l = 0; l < L - 1; ++l) {
      for (int s = 0; s < 2; ++s) {
        smooth(l, 0);
        smooth(l, 1);
      }
      std::fill(z[l + 1].begin(), z[l + 1].end(), R(0));
      std::fill(r[l + 1].begin(), r[l + 1].end(), R(0));
      restrict_(l);
    }
    for (int s = 0; s < 20; ++s) {
      smooth(L - 1, 0);
      smooth(L - 1, 1);
    }
    for (int l = L - 2; l >= 0; --l) {
      prolongate(l);
      for (int s = 0; s < 2; ++s) {
        smooth(l, 1);
        smooth(l, 0);
      }
    }
  }
  // sim/mgpcg_solid.py:87-97
  void downsample_bm() {
    for (int l = 1; l < L; ++l)
      for (int i = 0; i < dims[l][0]; ++i)
        for (int j = 0; j < dims[l][1]; ++j)
          for (int k = 0; k < dims[l][2]; ++k) {
            int all_solid = 1;
            for (int a = 0; a < 2; ++a)
              for (int b = 0; b < 2; ++b)
                for (int c = 0; c < 2; ++c)
                  if (bm[l - 1][id(l - 1, 2 * i + a, 2 * j + b, 2 * k + c)] <= 0) all_solid = 0;
            bm[l][id(l, i, j, k)] = all_solid;
          }
  }
  bool all_neumann() const {
    int s = 0;
    for (int a = 0; a < 3; ++a) s += bt[a][0] + bt[a][1];
    return s == 12;
  }

  // sim/mgpcg.py:197-283 / sim/mgpcg_solid.py:224-306.
  // stop_mode 0: the reference criterion (plain: rTr < max(abs_tol, rel_tol*z0'r0);
  // solid: rTr < abs_tol); stop_mode 1: sqrt(rTr) <= rel_tol * ||b|| with b the
  // (recentred) initial residual.  Returns the number of CG iterations run.
  int solve(int max_iters, int stop_mode, double abs_tol, double rel_tol, double* out_rTr, double* out_b2) {
    const double eps = 1e-12;
    if (solid) downsample_bm();
    bool neu = all_neumann();
    if (neu) recenter(r[0]);
    double b2 = reduce(r[0], r[0]);
    if (out_b2) *out_b2 = b2;
    precondition();
    R beta = 0;
    for_cells(0, [&](int i, int j, int k) {  // update_p, sim/mgpcg.py:142-145
      if (is_solid(0, i, j, k)) return;
      size_t q = id(0, i, j, k);
      p[q] = z[0][q] + beta * p[q];
    });
    double old_zTr = (double)(R)reduce(z[0], r[0]);
    double tol;
    if (stop_mode == 1)
      tol = rel_tol * rel_tol * b2;
    else
      tol = solid ? abs_tol : std::max(abs_tol, old_zTr * rel_tol);
    int it = 0;
    double rTr = b2;
    while (max_iters < 0 || it < max_iters) {
      compute_Ap();
      double pAp = (double)(R)reduce(p, Ap);
      R alpha = (R)(old_zTr / (pAp + eps));
      for_cells(0, [&](int i, int j, int k) {
        if (is_solid(0, i, j, k)) return;
        size_t q = id(0, i, j, k);
        x[q] += alpha * p[q];
      });
      for_cells(0, [&](int i, int j, int k) {
        if (is_solid(0, i, j, k)) return;
        size_t q = id(0, i, j, k);
        r[0][q] -= alpha * Ap[q];
      });
      rTr = (double)(R)reduce(r[0], r[0]);
      ++it;
      bool done = (stop_mode == 1) ? (rTr <= tol) : (rTr < tol);
      if (done) break;
      if (neu) recenter(r[0]);
      precondition();
      double new_zTr = (double)(R)reduce(z[0], r[0]);
      beta = (R)(new_zTr / (old_zTr + eps));
      for_cells(0, [&](int i, int j, int k) {
        if (is_solid(0, i, j, k)) return;
        size_t q = id(0, i, j, k);
        p[q] = z[0][q] + beta * p[q];
      });
      old_zTr = new_zTr;
    }
    if (out_rTr) *out_rTr = rTr;
    return it;
  }
};

template <class R>
void mg_setup(MGO<R>& M, int nx, int ny, int nz, int levels, const int* bt, const int* bm0) {
  M.L = levels;
  M.solid = (bm0 != nullptr);
  for (int a = 0; a < 3; ++a)
    for (int s = 0; s < 2; ++s) M.bt[a][s] = bt ? bt[2 * a + s] : 2;
  for (int l = 0; l < levels; ++l) M.dims.push_back({nx >> l, ny >> l, nz >> l});
  for (int l = 0; l < levels; ++l) {
    M.r.emplace_back(M.n(l), R(0));
    M.z.emplace_back(M.n(l), R(0));
    if (M.solid) M.bm.emplace_back(M.n(l), 0);
  }
  if (M.solid) std::copy(bm0, bm0 + M.n(0), M.bm[0].begin());
  M.x.assign(M.n(0), R(0));
  M.p.assign(M.n(0), R(0));
  M.Ap.assign(M.n(0), R(0));
}

// sim/mgpcg.py:373-392 apply_bc (+ sim/mgpcg_solid.py:347-356 for solids)
template <class R>
void apply_bc(const MGO<R>& M, R* ux, R* uy, R* uz, int nx, int ny, int nz, const int* bm, const R* bvel) {
  Mac<R> U = mkmac(ux, uy, uz, nx, ny, nz);
  for (int i = 0; i <= nx; ++i)
    for (int j = 0; j < ny; ++j)
      for (int k = 0; k < nz; ++k) {
        if (i == 0 && M.bt[0][0] == 2) U.ux.at(0, i, j, k) = 0;
        if (i == nx && M.bt[0][1] == 2) U.ux.at(0, i, j, k) = 0;
      }
  for (int i = 0; i < nx; ++i)
    for (int j = 0; j <= ny; ++j)
      for (int k = 0; k < nz; ++k) {
        if (j == 0 && M.bt[1][0] == 2) U.uy.at(0, i, j, k) = 0;
        if (j == ny && M.bt[1][1] == 2) U.uy.at(0, i, j, k) = 0;
      }
  for (int i = 0; i < nx; ++i)
    for (int j = 0; j < ny; ++j)
      for (int k = 0; k <= nz; ++k) {
        if (k == 0 && M.bt[2][0] == 2) U.uz.at(0, i, j, k) = 0;
        if (k == nz && M.bt[2][1] == 2) U.uz.at(0, i, j, k) = 0;
      }
  if (bm) {
    size_t N = (size_t)nx * ny * nz;
    for (int i = 0; i < nx; ++i)
      for (int j = 0; j < ny; ++j)
        for (int k = 0; k < nz; ++k) {
          size_t q = ((size_t)i * ny + j) * nz + k;
          if (bm[q] <= 0) continue;
          R vx = bvel[q], vy = bvel[N + q], vz = bvel[2 * N + q];
          U.ux.at(0, i, j, k) = vx;
          U.ux.at(0, i + 1, j, k) = vx;
          U.uy.at(0, i, j, k) = vy;
          U.uy.at(0, i, j + 1, k) = vy;
          U.uz.at(0, i, j, k) = vz;
          U.uz.at(0, i, j, k + 1) = vz;
        }
  }
}

// sim/mgpcg.py:406-432 subtract_grad_p (identical in sim/mgpcg_solid.py:386-411);
// when f is given, also f = -(grad p) (sim/mgpcg.py:448-479).
template <class R>
void subtract_grad_p(const std::vector<R>& P, R* ux, R* uy, R* uz, R* fx, R* fy, R* fz, int nx, int ny, int nz) {
  Fld<R> Pf = mk(P.data(), nx, ny, nz);
  Mac<R> U = mkmac(ux, uy, uz, nx, ny, nz);
  for (int i = 0; i <= nx; ++i)
    for (int j = 0; j < ny; ++j)
      for (int k = 0; k < nz; ++k) {
        R pr = sample(Pf, 0, i, j, k), pl = sample(Pf, 0, i - 1, j, k);
        if (i - 1 < 0) pl = 0;
        if (i >= nx) pr = 0;
        U.ux.at(0, i, j, k) -= (pr - pl);
        if (fx) fx[((size_t)i * ny + j) * nz + k] = -(pr - pl);
      }
  for (int i = 0; i < nx; ++i)
    for (int j = 0; j <= ny; ++j)
      for (int k = 0; k < nz; ++k) {
        R pt = sample(Pf, 0, i, j, k), pb = sample(Pf, 0, i, j - 1, k);
        if (j - 1 < 0) pb = 0;
        if (j >= ny) pt = 0;
        U.uy.at(0, i, j, k) -= pt - pb;
        if (fy) fy[((size_t)i * (ny + 1) + j) * nz + k] = -(pt - pb);
      }
  for (int i = 0; i < nx; ++i)
    for (int j = 0; j < ny; ++j)
      for (int k = 0; k <= nz; ++k) {
        R pc = sample(Pf, 0, i, j, k), pa = sample(Pf, 0, i, j, k - 1);
        if (k - 1 < 0) pa = 0;
        if (k >= nz) pc = 0;
        U.uz.at(0, i, j, k) -= pc - pa;
        if (fz) fz[((size_t)i * ny + j) * (nz + 1) + k] = -(pc - pa);
      }
}

// The whole Poisson() call.  stats[0] = iterations, [1] = final rTr,
// [2] = ||b||^2, [3] = fluid divergence sum (solid), [4] = 1 if corrected.
template <class R>
void poisson(R* ux, R* uy, R* uz, int nx, int ny, int nz, int levels, const int* bt, const int* bm,
             const R* bvel, int max_iters, int stop_mode, double abs_tol, double rel_tol, R* p_out,
             R* fx, R* fy, R* fz, double* stats) {
  MGO<R> M;
  mg_setup(M, nx, ny, nz, levels, bt, bm);
  apply_bc(M, ux, uy, uz, nx, ny, nz, bm, bvel);
  Mac<R> U = mkmac(ux, uy, uz, nx, ny, nz);
  // divergence: sim/mgpcg.py:394-404, sim/mgpcg_solid.py:358-375
  std::vector<R> div(M.n(0), R(0));
  double dsum = 0;
  long nfluid = 0;
  for (int i = 0; i < nx; ++i)
    for (int j = 0; j < ny; ++j)
      for (int k = 0; k < nz; ++k) {
        size_t q = ((size_t)i * ny + j) * nz + k;
        if (bm && bm[q] > 0) {
          div[q] = 0;
          continue;
        }
        R vl = sample(U.ux, 0, i, j, k), vr = sample(U.ux, 0, i + 1, j, k);
        R vb = sample(U.uy, 0, i, j, k), vt = sample(U.uy, 0, i, j + 1, k);
        R va = sample(U.uz, 0, i, j, k), vc = sample(U.uz, 0, i, j, k + 1);
        div[q] = vr - vl + vt - vb + vc - va;
        dsum += (double)div[q];
        ++nfluid;
      }
  stats[3] = dsum;
  stats[4] = 0;
  if (bm && dsum > 0.0) {  // sim/mgpcg_solid.py:377-383,422-424
    R avg = (R)(dsum / (double)nfluid);
    for (size_t q = 0; q < M.n(0); ++q)
      if (bm[q] <= 0) div[q] -= avg;
    stats[4] = 1;
  }
  // init(u_div, -1): sim/mgpcg.py:59-75
  for (size_t q = 0; q < M.n(0); ++q) M.r[0][q] = div[q] * R(-1);
  double rTr = 0, b2 = 0;
  int it = M.solve(max_iters, stop_mode, abs_tol, rel_tol, &rTr, &b2);
  stats[0] = it;
  stats[1] = rTr;
  stats[2] = b2;
  if (p_out) std::copy(M.x.begin(), M.x.end(), p_out);
  subtract_grad_p(M.x, ux, uy, uz, fx, fy, fz, nx, ny, nz);
  apply_bc(M, ux, uy, uz, nx, ny, nz, bm, bvel);
}

}  // namespace

#define DEF_MG(R, S)                                                                               \
  extern "C" void orc_poisson_##S(R* ux, R* uy, R* uz, int nx, int ny, int nz, int levels,          \
                                  const int* bt, const int* bm, const R* bvel, int max_iters,      \
                                  int stop_mode, double abs_tol, double rel_tol, R* p_out, R* fx,   \
                                  R* fy, R* fz, double* stats) {                                    \
    poisson<R>(ux, uy, uz, nx, ny, nz, levels, bt, bm, bvel, max_iters, stop_mode, abs_tol,         \
               rel_tol, p_out, fx, fy, fz, stats);                                                  \
  }                                                                                                 \
  /* sim/mgpcg.py:177-195: z = M^-1 r (one V-cycle); bm may be null */                             \
  extern "C" void orc_vcycle_##S(const R* r0, R* z0, int nx, int ny, int nz, int levels,           \
                                 const int* bt, const int* bm) {                                    \
    MGO<R> M;                                                                                       \
    mg_setup(M, nx, ny, nz, levels, bt, bm);                                                        \
    if (M.solid) M.downsample_bm();                                                                 \
    std::copy(r0, r0 + M.n(0), M.r[0].begin());                                                     \
    M.precondition();                                                                               \
    std::copy(M.z[0].begin(), M.z[0].end(), z0);                                                    \
  }                                                                                                 \
  /* sim/mgpcg.py:112-117: Ap for a given p; bm may be null */                                      \
  extern "C" void orc_apply_A_##S(const R* p, R* Ap, int nx, int ny, int nz, const int* bt,        \
                                  const int* bm) {                                                  \
    MGO<R> M;                                                                                       \
    mg_setup(M, nx, ny, nz, 1, bt, bm);                                                             \
    std::copy(p, p + M.n(0), M.p.begin());                                                          \
    M.compute_Ap();                                                                                 \
    std::copy(M.Ap.begin(), M.Ap.end(), Ap);                                                        \
  }                                                                                                 \
  /* CG on A x = b with b given (recentred first when all-Neumann) */                               \
  extern "C" int orc_mgpcg_solve_##S(const R* b, R* x, int nx, int ny, int nz, int levels,          \
                                     const int* bt, const int* bm, int max_iters, int stop_mode,    \
                                     double abs_tol, double rel_tol, double* stats) {               \
    MGO<R> M;                                                                                       \
    mg_setup(M, nx, ny, nz, levels, bt, bm);                                                        \
    std::copy(b, b + M.n(0), M.r[0].begin());                                                       \
    double rTr = 0, b2 = 0;                                                                         \
    int it = M.solve(max_iters, stop_mode, abs_tol, rel_tol, &rTr, &b2);                            \
    std::copy(M.x.begin(), M.x.end(), x);                                                           \
    if (stats) {                                                                                    \
      stats[0] = it;                                                                                \
      stats[1] = rTr;                                                                               \
      stats[2] = b2;                                                                                \
    }                                                                                               \
    return it;                                                                                      \
  }
ORC_BOTH(DEF_MG)
