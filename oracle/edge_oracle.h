// Shared templates of the CPU oracle (TEST INFRASTRUCTURE ONLY; see
// edge_oracle.cpp for the rules that govern its use and the layout).
#pragma once
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

namespace orc {

// --------------------------------------------------------------------------
// Field views
// --------------------------------------------------------------------------
template <class R>
struct Fld {
  R* d;
  int n0, n1, n2;
  size_t N() const { return (size_t)n0 * n1 * n2; }
  R& at(int c, int i, int j, int k) const {
    return d[(size_t)c * N() + ((size_t)i * n1 + j) * n2 + k];
  }
};

// Python int() truncates toward zero (util/taichi_utils.py:64-66).
template <class R>
static inline int trunc_idx(R u) { return (int)u; }
static inline int trunc_idx(int u) { return u; }

// util/taichi_utils.py:61-67  sample(): index-clamped fetch.
template <class R, class I>
static inline R sample(const Fld<R>& f, int c, I u, I v, I w) {
  int i = std::max(0, std::min(trunc_idx(u), f.n0 - 1));
  int j = std::max(0, std::min(trunc_idx(v), f.n1 - 1));
  int k = std::max(0, std::min(trunc_idx(w), f.n2 - 1));
  return f.at(c, i, j, k);
}

// dim - 2 - eps with eps = 1e-6 (util/taichi_utils.py:3), a Python double that
// Taichi folds into an R constant.
template <class R>
static inline R hi2(int dim) { return (R)((double)dim - 2.0 - 1e-6); }
template <class R>
static inline R hi1(int dim) { return (R)((double)dim - 1.0 - 1e-6); }

// --------------------------------------------------------------------------
// Quadratic B-spline (util/taichi_utils.py:230-251)
// --------------------------------------------------------------------------
template <class R>
static inline R N_2(R x) {
  R result = 0;
  R a = std::fabs(x);
  if (a < R(0.5))
    result = R(3.0 / 4.0) - a * a;
  else if (a < R(1.5))
    result = R(0.5) * ((R(1.5) - a) * (R(1.5) - a));
  return result;
}

template <class R>
static inline R dN_2(R x) {
  R result = 0;
  R a = std::fabs(x);
  if (a < R(0.5))
    result = R(-2) * a;
  else if (a < R(1.5))
    result = R(0.5) * (R(2) * a - R(3));
  if (x < R(0)) result = -result;
  return result;
}

// util/taichi_utils.py:254-280  interp_2 (64 taps, 27 with non-zero weight).
template <class R>
static R interp_2(const Fld<R>& vf, int c, const R p[3], R dx, R BLx, R BLy, R BLz) {
  R u = p[0] / dx - BLx, v = p[1] / dx - BLy, w = p[2] / dx - BLz;
  R s = std::max(R(1), std::min(u, hi2<R>(vf.n0)));
  R t = std::max(R(1), std::min(v, hi2<R>(vf.n1)));
  R l = std::max(R(1), std::min(w, hi2<R>(vf.n2)));
  R iu = std::floor(s), iv = std::floor(t), iw = std::floor(l);
  R interped = 0;
  for (int i = -1; i < 3; ++i)
    for (int j = -1; j < 3; ++j)
      for (int k = -1; k < 3; ++k) {
        R x = s - (iu + R(i)), y = t - (iv + R(j)), z = l - (iw + R(k));
        R value = sample(vf, c, iu + R(i), iv + R(j), iw + R(k));
        interped += value * N_2(x) * N_2(y) * N_2(z);
      }
  return interped;
}

// util/taichi_utils.py:283-319  interp_grad_2.
template <class R>
static R interp_grad_2(const Fld<R>& vf, int c, const R p[3], R dx, R BLx, R BLy, R BLz,
                       R grad[3]) {
  R u = p[0] / dx - BLx, v = p[1] / dx - BLy, w = p[2] / dx - BLz;
  R s = std::max(R(1), std::min(u, hi2<R>(vf.n0)));
  R t = std::max(R(1), std::min(v, hi2<R>(vf.n1)));
  R l = std::max(R(1), std::min(w, hi2<R>(vf.n2)));
  R iu = std::floor(s), iv = std::floor(t), iw = std::floor(l);
  R px = 0, py = 0, pz = 0, interped = 0;
  R inv = R(1) / dx;
  for (int i = -1; i < 3; ++i)
    for (int j = -1; j < 3; ++j)
      for (int k = -1; k < 3; ++k) {
        R x = s - (iu + R(i)), y = t - (iv + R(j)), z = l - (iw + R(k));
        R value = sample(vf, c, iu + R(i), iv + R(j), iw + R(k));
        px += inv * (value * dN_2(x) * N_2(y) * N_2(z));
        py += inv * (value * N_2(x) * dN_2(y) * N_2(z));
        pz += inv * (value * N_2(x) * N_2(y) * dN_2(z));
        interped += value * N_2(x) * N_2(y) * N_2(z);
      }
  grad[0] = px; grad[1] = py; grad[2] = pz;
  return interped;
}

// Face fields of the MAC grid: u_x (nx+1,ny,nz), u_y (nx,ny+1,nz), u_z (nx,ny,nz+1)
// (run_edge.py:193-195).
template <class R>
struct Mac {
  Fld<R> ux, uy, uz;
};

// util/taichi_utils.py:322-327  interp_u_MAC.
template <class R>
static inline void interp_u_MAC(const Mac<R>& u, const R p[3], R dx, R out[3]) {
  out[0] = interp_2(u.ux, 0, p, dx, R(0.0), R(0.5), R(0.5));
  out[1] = interp_2(u.uy, 0, p, dx, R(0.5), R(0.0), R(0.5));
  out[2] = interp_2(u.uz, 0, p, dx, R(0.5), R(0.5), R(0.0));
}

// util/taichi_utils.py:336-341  interp_u_MAC_grad; J rows are grad(u_a).
template <class R>
static inline void interp_u_MAC_grad(const Mac<R>& u, const R p[3], R dx, R out[3], R J[3][3]) {
  out[0] = interp_grad_2(u.ux, 0, p, dx, R(0.0), R(0.5), R(0.5), J[0]);
  out[1] = interp_grad_2(u.uy, 0, p, dx, R(0.5), R(0.0), R(0.5), J[1]);
  out[2] = interp_grad_2(u.uz, 0, p, dx, R(0.5), R(0.5), R(0.0), J[2]);
}

// --------------------------------------------------------------------------
// Linear interpolation (util/taichi_utils.py:119-227)
// --------------------------------------------------------------------------
template <class R>
static inline R N_1(R x) { return R(1) - std::fabs(x); }

// interp_1 for every component of an nc-component field (default BL 0.5).
template <class R>
static void interp_1(const Fld<R>& vf, int nc, const R p[3], R dx, R* out) {
  R u = p[0] / dx - R(0.5), v = p[1] / dx - R(0.5), w = p[2] / dx - R(0.5);
  R s = std::max(R(1), std::min(u, hi1<R>(vf.n0)));
  R t = std::max(R(1), std::min(v, hi1<R>(vf.n1)));
  R l = std::max(R(1), std::min(w, hi1<R>(vf.n2)));
  R iu = std::floor(s), iv = std::floor(t), iw = std::floor(l);
  for (int c = 0; c < nc; ++c) out[c] = R(0) * sample(vf, c, iu, iv, iw);
  for (int i = 0; i < 2; ++i)
    for (int j = 0; j < 2; ++j)
      for (int k = 0; k < 2; ++k) {
        R x = s - (iu + R(i)), y = t - (iv + R(j)), z = l - (iw + R(k));
        for (int c = 0; c < nc; ++c) {
          R value = sample(vf, c, iu + R(i), iv + R(j), iw + R(k));
          out[c] += value * N_1(x) * N_1(y) * N_1(z);
        }
      }
}

// sample_min_max_1 (util/taichi_utils.py:201-227).
template <class R>
static void sample_min_max_1(const Fld<R>& vf, int nc, const R p[3], R dx, R* mini, R* maxi) {
  R u = p[0] / dx - R(0.5), v = p[1] / dx - R(0.5), w = p[2] / dx - R(0.5);
  R s = std::max(R(1), std::min(u, hi1<R>(vf.n0)));
  R t = std::max(R(1), std::min(v, hi1<R>(vf.n1)));
  R l = std::max(R(1), std::min(w, hi1<R>(vf.n2)));
  R iu = std::floor(s), iv = std::floor(t), iw = std::floor(l);
  for (int c = 0; c < nc; ++c) mini[c] = maxi[c] = sample(vf, c, iu, iv, iw);
  for (int i = 0; i < 2; ++i)
    for (int j = 0; j < 2; ++j)
      for (int k = 0; k < 2; ++k)
        for (int c = 0; c < nc; ++c) {
          R value = sample(vf, c, iu + R(i), iv + R(j), iw + R(k));
          mini[c] = std::min(mini[c], value);
          maxi[c] = std::max(maxi[c], value);
        }
}

// --------------------------------------------------------------------------
// Cubic Hermite basis (sim/hermite.py:4-41)
// --------------------------------------------------------------------------
template <class R>
static inline R hermite_f(R x) {
  R r = 0, a = std::fabs(x);
  if (a < R(1)) r = R(2) * (a * a * a) - R(3) * (a * a) + R(1);
  return r;
}
template <class R>
static inline R hermite_df(R x) {
  R r = 0, a = std::fabs(x);
  if (a < R(1)) r = R(6) * (a * a) - R(6) * a;
  if (x < R(0)) r = -r;
  return r;
}
template <class R>
static inline R hermite_g(R x) {
  R r = 0, a = std::fabs(x);
  if (a < R(1)) r = (a * a * a) - R(2) * (a * a) + a;
  if (x < R(0)) r = -r;
  return r;
}
template <class R>
static inline R hermite_dg(R x) {
  R r = 0, a = std::fabs(x);
  if (a < R(1)) r = R(3) * (a * a) - R(4) * a + R(1);
  return r;
}

// The map data of one component c of psi: value plane, gradient T_c (3 planes),
// cross derivatives dT_c = [yz, xz, xy] (3 planes), and d2T_c (xyz, 1 plane).
// Which of them are used depends on the Hermite variant.
template <class R>
struct HermiteData {
  Fld<R> psi;  // 3 planes; component `index` is used
  Fld<R> T;    // 3 planes: grad psi_index (sim/hermite.py:610-620 "Hermite" layout)
  Fld<R> dT;   // 3 planes (garm, full)
  Fld<R> d2T;  // 1 plane (full only)
};

enum HermiteKind { HK_GARM_FD = 0, HK_FULL = 1, HK_FD = 2 };

// Generic Hermite value(+gradient) of psi[index] at p.  kind selects the
// variant:
//   HK_GARM_FD  interp(_grad)_hermite_garm_fd_unroll  sim/hermite.py:298-456
//   HK_FULL     interp(_grad)_hermite                 sim/hermite.py:459-607
//   HK_FD       interp(_grad)_hermite_fd_unroll       sim/hermite.py:64-295
// The loop bodies below are the reference's term lists; with want_grad false
// only `interped` is accumulated, which is the value-only reference function.
template <class R>
static R interp_hermite_any(int kind, const HermiteData<R>& H, int index, const R p[3], R dx,
                            R* grad /* may be null */) {
  const Fld<R>& vf = H.psi;
  const Fld<R>& g = H.T;
  const Fld<R>& gg = H.dT;
  R u = p[0] / dx - R(0.5), v = p[1] / dx - R(0.5), w = p[2] / dx - R(0.5);
  R s = std::max(R(1), std::min(u, hi2<R>(vf.n0)));
  R t = std::max(R(1), std::min(v, hi2<R>(vf.n1)));
  R l = std::max(R(1), std::min(w, hi2<R>(vf.n2)));
  R iu = std::floor(s), iv = std::floor(t), iw = std::floor(l);
  R px = 0, py = 0, pz = 0, interped = 0;
  R inv_dx = R(1) / dx;
  for (int i = 0; i < 2; ++i)
    for (int j = 0; j < 2; ++j)
      for (int k = 0; k < 2; ++k) {
        R x = s - (iu + R(i)), y = t - (iv + R(j)), z = l - (iw + R(k));
        R hxf = hermite_f(x), hxg = hermite_g(x), dhxf = hermite_df(x), dhxg = hermite_dg(x);
        R hyf = hermite_f(y), hyg = hermite_g(y), dhyf = hermite_df(y), dhyg = hermite_dg(y);
        R hzf = hermite_f(z), hzg = hermite_g(z), dhzf = hermite_df(z), dhzg = hermite_dg(z);
        R ii = iu + R(i), jj = iv + R(j), kk = iw + R(k);
        // one term: coefficient * basis(x) * basis(y) * basis(z)
        auto add = [&](R val, R bx, R by, R bz, R dbx, R dby, R dbz) {
          interped += val * bx * by * bz;
          px += val * dbx * by * bz;
          py += val * bx * dby * bz;
          pz += val * bx * by * dbz;
        };
        R value = sample(vf, index, ii, jj, kk);
        add(value, hxf, hyf, hzf, dhxf, dhyf, dhzf);
        R v1[3];
        for (int q = 0; q < 3; ++q) v1[q] = sample(g, q, ii, jj, kk) * dx;
        add(v1[0], hxg, hyf, hzf, dhxg, dhyf, dhzf);
        add(v1[1], hxf, hyg, hzf, dhxf, dhyg, dhzf);
        add(v1[2], hxf, hyf, hzg, dhxf, dhyf, dhzg);
        R v2[3], v3;
        if (kind == HK_FD) {
          // sim/hermite.py:111-158: cross terms from central differences of T.
          R one = R(1);
          v2[0] = ((sample(g, 1, ii, jj, kk + one) - sample(g, 1, ii, jj, kk - one)) +
                   (sample(g, 2, ii, jj + one, kk) - sample(g, 2, ii, jj - one, kk))) / R(4) * dx;
          v2[1] = ((sample(g, 0, ii, jj, kk + one) - sample(g, 0, ii, jj, kk - one)) +
                   (sample(g, 2, ii + one, jj, kk) - sample(g, 2, ii - one, jj, kk))) / R(4) * dx;
          v2[2] = ((sample(g, 0, ii, jj + one, kk) - sample(g, 0, ii, jj - one, kk)) +
                   (sample(g, 1, ii + one, jj, kk) - sample(g, 1, ii - one, jj, kk))) / R(4) * dx;
          R v30 = ((sample(g, 1, ii + one, jj, kk + one) - sample(g, 1, ii + one, jj, kk - one)) +
                   (sample(g, 2, ii + one, jj + one, kk) - sample(g, 2, ii + one, jj - one, kk))) / R(4);
          R v31 = ((sample(g, 1, ii - one, jj, kk + one) - sample(g, 1, ii - one, jj, kk - one)) +
                   (sample(g, 2, ii - one, jj + one, kk) - sample(g, 2, ii - one, jj - one, kk))) / R(4);
          R v32 = ((sample(g, 0, ii, jj + one, kk + one) - sample(g, 0, ii, jj + one, kk - one)) +
                   (sample(g, 2, ii + one, jj + one, kk) - sample(g, 2, ii - one, jj + one, kk))) / R(4);
          R v33 = ((sample(g, 0, ii, jj - one, kk + one) - sample(g, 0, ii, jj - one, kk - one)) +
                   (sample(g, 2, ii + one, jj - one, kk) - sample(g, 2, ii - one, jj - one, kk))) / R(4);
          R v34 = ((sample(g, 0, ii, jj + one, kk + one) - sample(g, 0, ii, jj - one, kk + one)) +
                   (sample(g, 1, ii + one, jj, kk + one) - sample(g, 1, ii - one, jj, kk + one))) / R(4);
          R v35 = ((sample(g, 0, ii, jj + one, kk - one) - sample(g, 0, ii, jj - one, kk - one)) +
                   (sample(g, 1, ii + one, jj, kk - one) - sample(g, 1, ii - one, jj, kk - one))) / R(4);
          v3 = (v30 - v31 + v32 - v33 + v34 - v35) / R(6) * dx;
        } else {
          for (int q = 0; q < 3; ++q) v2[q] = sample(gg, q, ii, jj, kk) * dx * dx;
          if (kind == HK_GARM_FD) {
            // sim/hermite.py:340-345 / 440-445: xyz term from central differences of dT.
            R one = R(1);
            v3 = ((sample(gg, 0, ii + one, jj, kk) - sample(gg, 0, ii - one, jj, kk)) +
                  (sample(gg, 1, ii, jj + one, kk) - sample(gg, 1, ii, jj - one, kk)) +
                  (sample(gg, 2, ii, jj, kk + one) - sample(gg, 2, ii, jj, kk - one))) / R(6) * dx * dx;
          } else {
            v3 = sample(H.d2T, 0, ii, jj, kk) * dx * dx * dx;  // sim/hermite.py:508,596
          }
        }
        add(v2[0], hxf, hyg, hzg, dhxf, dhyg, dhzg);
        add(v2[1], hxg, hyf, hzg, dhxg, dhyf, dhzg);
        add(v2[2], hxg, hyg, hzf, dhxg, dhyg, dhzf);
        add(v3, hxg, hyg, hzg, dhxg, dhyg, dhzg);
      }
  if (grad) {
    grad[0] = px * inv_dx;
    grad[1] = py * inv_dx;
    grad[2] = pz * inv_dx;
  }
  return interped;
}

// A full backward map: psi (3 planes) and, per component c, T_c, dT_c, d2T_c.
template <class R>
struct Map {
  Fld<R> psi;
  Fld<R> T[3];
  Fld<R> dT[3];
  Fld<R> d2T[3];
  HermiteData<R> comp(int c) const { return HermiteData<R>{psi, T[c], dT[c], d2T[c]}; }
};

// Hermite value and gradient of all three components of the map at p;
// G rows are grad psi_c.  Used by interp_psi_hermite(_garm)(_fd_unroll)
// (sim/hermite.py:921-1046), map_u_psi (run_edge.py:376-417) and
// split_psi_hermite (run_ed8.py:190-223).
template <class R>
static inline void map_eval(int kind, const Map<R>& M, const R p[3], R dx, R val[3], R G[3][3]) {
  for (int c = 0; c < 3; ++c) val[c] = interp_hermite_any(kind, M.comp(c), c, p, dx, G ? G[c] : nullptr);
}

// --------------------------------------------------------------------------
// RK4 map marches
// --------------------------------------------------------------------------
// sim/hermite.py:786-803  advance_map_0 body, one point.
template <class R>
static void advance_map_0_pt(R psi[3], const Mac<R>& U, R dx, R dt) {
  R u1[3], u2[3], u3[3], u4[3], q[3];
  interp_u_MAC(U, psi, dx, u1);
  for (int a = 0; a < 3; ++a) q[a] = psi[a] + R(0.5) * dt * u1[a];
  interp_u_MAC(U, q, dx, u2);
  for (int a = 0; a < 3; ++a) q[a] = psi[a] + R(0.5) * dt * u2[a];
  interp_u_MAC(U, q, dx, u3);
  for (int a = 0; a < 3; ++a) q[a] = psi[a] + dt * u3[a];
  interp_u_MAC(U, q, dx, u4);
  for (int a = 0; a < 3; ++a)
    psi[a] = psi[a] + dt * R(1) / R(6) * (u1[a] + R(2) * u2[a] + R(2) * u3[a] + u4[a]);
}

// sim/hermite.py:806-860  advance_map_1_transpose_func, one point.
// T[c] are rows (grad psi_c) on entry and exit; the march works on columns.
template <class R>
static void advance_map_1_transpose_pt(R psi[3], R T[3][3], const Mac<R>& U, R dx, R dt) {
  R C[3][3];  // C[c] = column c = d psi / d x_c  (transpose_matrix)
  for (int c = 0; c < 3; ++c)
    for (int a = 0; a < 3; ++a) C[c][a] = T[a][c];
  R uk[4][3], dC[4][3][3];
  R ps[3], Cs[3][3];
  for (int a = 0; a < 3; ++a) ps[a] = psi[a];
  for (int c = 0; c < 3; ++c)
    for (int a = 0; a < 3; ++a) Cs[c][a] = C[c][a];
  const R coef[3] = {R(0.5) * dt, R(0.5) * dt, dt};
  for (int st = 0; st < 4; ++st) {
    R J[3][3];
    interp_u_MAC_grad(U, ps, dx, uk[st], J);
    for (int c = 0; c < 3; ++c)
      for (int a = 0; a < 3; ++a)
        dC[st][c][a] = J[a][0] * Cs[c][0] + J[a][1] * Cs[c][1] + J[a][2] * Cs[c][2];
    if (st < 3) {
      for (int a = 0; a < 3; ++a) ps[a] = psi[a] + coef[st] * uk[st][a];
      for (int c = 0; c < 3; ++c)
        for (int a = 0; a < 3; ++a) Cs[c][a] = C[c][a] + coef[st] * dC[st][c][a];
    }
  }
  R h = dt * R(1) / R(6);
  for (int a = 0; a < 3; ++a)
    psi[a] = psi[a] + h * (uk[0][a] + R(2) * uk[1][a] + R(2) * uk[2][a] + uk[3][a]);
  for (int c = 0; c < 3; ++c)
    for (int a = 0; a < 3; ++a)
      C[c][a] = C[c][a] + h * (dC[0][c][a] + R(2) * dC[1][c][a] + R(2) * dC[2][c][a] + dC[3][c][a]);
  for (int c = 0; c < 3; ++c)
    for (int a = 0; a < 3; ++a) T[a][c] = C[c][a];
}

// run_edge.py:336-357  advance_map: RK4 of a face point phi and ONE column F
// (d F / dt = J F), no transpose.
template <class R>
static void advance_map_col_pt(R phi[3], R F[3], const Mac<R>& U, R dx, R dt) {
  R uk[4][3], dF[4][3], ps[3], Fs[3];
  for (int a = 0; a < 3; ++a) { ps[a] = phi[a]; Fs[a] = F[a]; }
  const R coef[3] = {R(0.5) * dt, R(0.5) * dt, R(1.0) * dt};
  for (int st = 0; st < 4; ++st) {
    R J[3][3];
    interp_u_MAC_grad(U, ps, dx, uk[st], J);
    for (int a = 0; a < 3; ++a) dF[st][a] = J[a][0] * Fs[0] + J[a][1] * Fs[1] + J[a][2] * Fs[2];
    if (st < 3)
      for (int a = 0; a < 3; ++a) {
        ps[a] = phi[a] + coef[st] * uk[st][a];
        Fs[a] = F[a] + coef[st] * dF[st][a];
      }
  }
  R h = dt * R(1) / R(6);
  for (int a = 0; a < 3; ++a) {
    phi[a] = phi[a] + h * (uk[0][a] + R(2) * uk[1][a] + R(2) * uk[2][a] + uk[3][a]);
    F[a] = F[a] + h * (dF[0][a] + R(2) * dF[1][a] + R(2) * dF[2][a] + dF[3][a]);
  }
}

// Compose a marched point (psi_step, rows T_step) with the old map:
// sim/hermite.py:959-980 (garm_fd_unroll) and 1019-1046 (full).
template <class R>
static void compose_pt(int kind, const Map<R>& M, R psi[3], R T[3][3], R dx) {
  R val[3], G[3][3];
  map_eval(kind, M, psi, dx, val, G);
  R Tn[3][3];
  for (int a = 0; a < 3; ++a)
    for (int b = 0; b < 3; ++b) Tn[a][b] = G[a][0] * T[0][b] + G[a][1] * T[1][b] + G[a][2] * T[2][b];
  for (int a = 0; a < 3; ++a) {
    psi[a] = val[a];
    for (int b = 0; b < 3; ++b) T[a][b] = Tn[a][b];
  }
}

// sim/hermite.py:1049-1075  calc_ed4_garm_func.  Tp[p][c] is row c of point p;
// points: 0 (-,-,-), 1 (+,+,-), 2 (-,+,+), 3 (+,-,+)  (sim/hermite.py:1446-1449).
template <class R>
static void calc_ed4_garm(const R Tp[4][3][3], R eps, R dT[3][3]) {
  for (int c = 0; c < 3; ++c) {
    const R* t0 = Tp[0][c]; const R* t1 = Tp[1][c]; const R* t2 = Tp[2][c]; const R* t3 = Tp[3][c];
    dT[c][0] = ((-t0[1] - t1[1] + t2[1] + t3[1]) + (-t0[2] + t1[2] + t2[2] - t3[2])) / R(8) / eps;
    dT[c][1] = ((-t0[0] - t1[0] + t2[0] + t3[0]) + (-t0[2] + t1[2] - t2[2] + t3[2])) / R(8) / eps;
    dT[c][2] = ((-t0[0] + t1[0] + t2[0] - t3[0]) + (-t0[1] + t1[1] - t2[1] + t3[1])) / R(8) / eps;
  }
}

// sim/hermite.py:1426-1492  advect_psi_garm_ed4_fusion_unroll_kernel, one cell.
// Reads the old map M (psi, T, dT), writes the new one into Out.
template <class R>
static void edge_march_cell(const Map<R>& M, const Map<R>& Out, const Mac<R>& U, const Fld<R>& X,
                            int i, int j, int k, R dx, R dt, R eps) {
  R x0[3] = {X.at(0, i, j, k), X.at(1, i, j, k), X.at(2, i, j, k)};
  static const int sg[4][3] = {{-1, -1, -1}, {1, 1, -1}, {-1, 1, 1}, {1, -1, 1}};
  R P[5][3], T[5][3][3];
  for (int q = 0; q < 5; ++q) {
    for (int a = 0; a < 3; ++a) {
      P[q][a] = x0[a];
      for (int b = 0; b < 3; ++b) T[q][a][b] = (a == b) ? R(1) : R(0);
    }
    if (q > 0)
      for (int a = 0; a < 3; ++a) P[q][a] = x0[a] + R(sg[q - 1][a]) * eps;
  }
  for (int q = 0; q < 5; ++q) advance_map_1_transpose_pt(P[q], T[q], U, dx, -dt);
  for (int q = 0; q < 5; ++q) compose_pt(HK_GARM_FD, M, P[q], T[q], dx);
  R dT[3][3];
  R Tp[4][3][3];
  for (int q = 0; q < 4; ++q)
    for (int a = 0; a < 3; ++a)
      for (int b = 0; b < 3; ++b) Tp[q][a][b] = T[q + 1][a][b];
  calc_ed4_garm(Tp, eps, dT);
  for (int a = 0; a < 3; ++a) {
    Out.psi.at(a, i, j, k) = P[0][a];
    for (int b = 0; b < 3; ++b) {
      Out.T[a].at(b, i, j, k) = T[0][a][b];
      Out.dT[a].at(b, i, j, k) = dT[a][b];
    }
  }
}

// sim/hermite.py:1098-1151  calc_ed8_3, from 8 points named by the reset
// pattern of sim/hermite.py:704-725 (bit 0 = x sign, bit 1 = y, bit 2 = z).
template <class R>
static void calc_ed8_cell(const R P[8][3], R eps, R psi[3], R T[3][3], R dT[3][3], R d2T[3]) {
  auto sgn = [](int q, int axis) { return ((q >> axis) & 1) ? 1 : -1; };
  for (int a = 0; a < 3; ++a) {
    R sum = 0, px = 0, py = 0, pz = 0, pyz = 0, pzx = 0, pxy = 0, pxyz = 0;
    // The reference spells each sum out term by term in point order 000..111.
    for (int q = 0; q < 8; ++q) {
      R v = P[q][a];
      int sx = sgn(q, 0), sy = sgn(q, 1), sz = sgn(q, 2);
      sum += v;
      px += R(sx) * v;
      py += R(sy) * v;
      pz += R(sz) * v;
      pyz += R(sy * sz) * v;
      pzx += R(sz * sx) * v;
      pxy += R(sx * sy) * v;
      pxyz += R(sx * sy * sz) * v;
    }
    psi[a] = sum / R(8);
    T[a][0] = px / R(8) / eps;
    T[a][1] = py / R(8) / eps;
    T[a][2] = pz / R(8) / eps;
    dT[a][0] = pyz / R(8) / eps / eps;
    dT[a][1] = pzx / R(8) / eps / eps;
    dT[a][2] = pxy / R(8) / eps / eps;
    d2T[a] = pxyz / R(8) / eps / eps / eps;
  }
}

// sim/hermite.py:1154-1177  calc_ed4_1: points 0 (-,-,-), 1 (+,+,-), 2 (-,+,+), 3 (+,-,+).
template <class R>
static void calc_ed4_1_cell(const R P[4][3], R eps, R psi[3], R T[3][3]) {
  for (int a = 0; a < 3; ++a) {
    psi[a] = (P[0][a] + P[1][a] + P[2][a] + P[3][a]) / R(4);
    T[a][0] = (-P[0][a] + P[1][a] - P[2][a] + P[3][a]) / R(4) / eps;
    T[a][1] = (-P[0][a] + P[1][a] + P[2][a] - P[3][a]) / R(4) / eps;
    T[a][2] = (-P[0][a] - P[1][a] + P[2][a] + P[3][a]) / R(4) / eps;
  }
}

// Cell-centre and face coordinates (util/taichi_utils.py:23-44).
template <class R>
static inline void face_coord(int axis, int i, int j, int k, R dx, R X[3]) {
  R h = R(0.5);
  X[0] = (axis == 0 ? R(i) : R(i) + h) * dx;
  X[1] = (axis == 1 ? R(j) : R(j) + h) * dx;
  X[2] = (axis == 2 ? R(k) : R(k) + h) * dx;
}

template <class R>
static Fld<R> mk(const R* d, int n0, int n1, int n2) { return Fld<R>{const_cast<R*>(d), n0, n1, n2}; }

template <class R>
static Mac<R> mkmac(const R* ux, const R* uy, const R* uz, int nx, int ny, int nz) {
  return Mac<R>{mk(ux, nx + 1, ny, nz), mk(uy, nx, ny + 1, nz), mk(uz, nx, ny, nz + 1)};
}

// map planes: psi[3], Tx[3], Ty[3], Tz[3], dTx[3], dTy[3], dTz[3], d2Tx, d2Ty, d2Tz.
template <class R>
static Map<R> mkmap(const R* psi, const R* T, const R* dT, const R* d2T, int nx, int ny, int nz) {
  size_t N = (size_t)nx * ny * nz;
  Map<R> M;
  M.psi = mk(psi, nx, ny, nz);
  for (int c = 0; c < 3; ++c) {
    M.T[c] = mk(T ? T + 3 * N * c : nullptr, nx, ny, nz);
    M.dT[c] = mk(dT ? dT + 3 * N * c : nullptr, nx, ny, nz);
    M.d2T[c] = mk(d2T ? d2T + N * c : nullptr, nx, ny, nz);
  }
  return M;
}

#define ORC_BOTH(X) X(float, f) X(double, d)

}  // namespace orc
