// Multigrid-preconditioned CG pressure projection (sim/mgpcg.py,
// sim/mgpcg_solid.py) on gfx950.
//
// Three modes share one algorithm skeleton:
//  * EDGE_MG_PLAIN (0): the reference V-cycle (2 pre/post red-black
//    Gauss-Seidel half-sweep pairs, 20 bottom pairs, restriction r[l+1] +=
//    res/2, injection prolongation; sim/mgpcg.py:147-195) in FP32.
//  * EDGE_MG_MIXED (1): the same smoother and transfers with the multigrid
//    levels stored in BF16 (FP32 arithmetic) and the 20 bottom pairs replaced
//    by a dense FP64 solve of the coarsest level.
//  * EDGE_MG_DENSE (2): FP32 levels with that dense coarsest solve (the
//    level-0 passes are latency-bound, so BF16's halved bytes buy ~7% per
//    iteration while its rounding costs about one CG iteration in twelve).
// All keep x, p, Ap and the CG residual in FP32, do every inner product in
// FP64 with a deterministic two-level reduction (per-block partials, the last
// block to finish sums them in a fixed order and updates the CG scalars on the
// device), and never read a scalar back except for the convergence flag.
//
// Fusions (none changes the arithmetic of a step, only where it happens):
//  * fill(0) + first red half-sweep of a level = one pass (z = r/diag on red
//    cells, 0 elsewhere), also fused into the restriction that produces r[l+1];
//  * prolongation + the first black post-smoothing half-sweep = one pass (the
//    black update reads its red neighbours as z + z_coarse on the fly; red
//    cells are overwritten by the next half-sweep before anything reads them);
//  * recenter of r (sim/mgpcg.py:169-175) is folded into the level-0 first pass;
//  * update_p (p = z + beta p) + compute_Ap + reduce(p, Ap) = one pass;
//  * update_x + update_r + reduce(r, r) + the sum for the next recenter = one pass;
//  * the last red half-sweep of level 0 + reduce(z, r) = one pass.
#include <hip/hip_bf16.h>

#include <cstdlib>
#include <cstring>
#include <functional>
#include <mutex>
#include <vector>

#include "edge_common.h"

using namespace edge;

namespace {

constexpr int MAXL = 12;
constexpr int MAXB = 8;        // largest CG batch between host convergence checks
constexpr int NB = 1024;       // blocks of every reduction kernel
constexpr int BS = 256;

// device scalar slots (double)
enum Slot {
  S_ZTR = 0,     // z.r of the current direction
  S_PAP,         // p.Ap
  S_RTR,         // r.r after the update
  S_B2,          // ||b||^2 (recentred)
  S_MEAN,        // mean to subtract from r in the next level-0 pass
  S_ALPHA,
  S_BETA,
  S_TOL2,        // stop when rTr < tol2 (reference) or <= tol2 (relative)
  S_DONE,        // 0 running, 1 converged, 2 hit max_iters
  S_ITERS,
  S_MAXIT,
  S_DSUM,        // solid: sum of fluid divergence
  S_NFLUID,
  S_CORR,        // solid: correction subtracted from fluid divergence
  S_R02,         // warm start: ||b - A x0||^2
  S_USEX0,
  S_STOPMODE,
  S_ABSTOL,
  S_RELTOL,
  S_SOLID,
  S_NCELL,
  S_ZTR0,
  S_NEU,         // all walls: recentre r (sim/mgpcg.py:211-217)
  S_EPS,         // the guard added to the alpha and beta denominators (see set_scalars)
  S_X0ACT,       // warm start, for the first level-0 pass: 0 nothing, 1 r = w (x0 kept), 2 x = 0 (x0 dropped)
  S_APMEAN,      // mean of Ap over all cells (all walls), for the recenter folded into the CG update
  S_COUNT
};

// finalize actions of the last block
enum Fin { F_NONE = 0, F_MEAN, F_B2, F_ZR_INIT, F_ZR, F_ALPHA, F_XR, F_DIVSUM, F_X0, F_SUM2, F_MEAN_B2 };

template <class S>
__device__ __forceinline__ float ld(const S* p, int64_t i);
template <>
__device__ __forceinline__ float ld<float>(const float* p, int64_t i) { return p[i]; }
template <>
__device__ __forceinline__ float ld<__hip_bfloat16>(const __hip_bfloat16* p, int64_t i) {
  return __bfloat162float(p[i]);
}
template <class S>
__device__ __forceinline__ void st(S* p, int64_t i, float v);
template <>
__device__ __forceinline__ void st<float>(float* p, int64_t i, float v) { p[i] = v; }
template <>
__device__ __forceinline__ void st<__hip_bfloat16>(__hip_bfloat16* p, int64_t i, float v) {
  p[i] = __float2bfloat16(v);
}

struct Lv {
  int n0, n1, n2;
  const int8_t* bm;   // 1 = solid (null when no solids)
  const uint8_t* nb;  // solid bits per cell for the row-chunk kernels (null when no solids)
  FDiv f1, f2, fh;    // divisions by n1, n2 and n2/2
  FDiv f8;            // division by n2/8 (row chunks; n2 a multiple of 8)
};

inline Lv make_lv(int n0, int n1, int n2) {
  Lv L;
  L.n0 = n0;
  L.n1 = n1;
  L.n2 = n2;
  L.bm = nullptr;
  L.nb = nullptr;
  L.f1 = make_fdiv((uint32_t)n1);
  L.f2 = make_fdiv((uint32_t)n2);
  L.fh = make_fdiv((uint32_t)(n2 > 1 ? n2 / 2 : 1));
  L.f8 = make_fdiv((uint32_t)(n2 >= 8 ? n2 / 8 : 1));
  return L;
}
// row-chunk kernels apply to a level when its rows split into 8-cell chunks
inline bool chunked(const Lv& L) { return L.n2 % 8 == 0; }

__device__ __forceinline__ int64_t lid(const Lv& L, int i, int j, int k) {
  return ((int64_t)i * L.n1 + j) * L.n2 + k;
}
__device__ __forceinline__ bool solid_at(const Lv& L, int i, int j, int k) {
  return L.bm && L.bm[lid(L, i, j, k)] > 0;
}

// sim/mgpcg.py:101-110 / sim/mgpcg_solid.py:113-130 (bt: 6 bits, bit 2a+s set
// when boundary_types[a][s] == 2).
__device__ __forceinline__ float num_fluid(const Lv& L, int bt, int i, int j, int k) {
  float num = 6.0f;
  int I[3] = {i, j, k};
  int n[3] = {L.n0, L.n1, L.n2};
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    int lo[3] = {i, j, k}, hi[3] = {i, j, k};
    lo[a] -= 1;
    hi[a] += 1;
    if (I[a] <= 0 && ((bt >> (2 * a)) & 1))
      num -= 1.0f;
    else if (L.bm && I[a] > 0 && solid_at(L, lo[0], lo[1], lo[2]))
      num -= 1.0f;
    if (I[a] >= n[a] - 1 && ((bt >> (2 * a + 1)) & 1))
      num -= 1.0f;
    else if (L.bm && I[a] < n[a] - 1 && solid_at(L, hi[0], hi[1], hi[2]))
      num -= 1.0f;
  }
  return num;
}

// sim/mgpcg.py:87-99 / sim/mgpcg_solid.py:99-111; the value of a neighbour is
// fetched by `get(i,j,k)` so callers can add the prolonged coarse value.
template <class G>
__device__ __forceinline__ float neighbor_sum(const Lv& L, int i, int j, int k, G get) {
  float ret = 0.f;
  if (i < L.n0 - 1 && !solid_at(L, i + 1, j, k)) ret += get(i + 1, j, k);
  if (i > 0 && !solid_at(L, i - 1, j, k)) ret += get(i - 1, j, k);
  if (j < L.n1 - 1 && !solid_at(L, i, j + 1, k)) ret += get(i, j + 1, k);
  if (j > 0 && !solid_at(L, i, j - 1, k)) ret += get(i, j - 1, k);
  if (k < L.n2 - 1 && !solid_at(L, i, j, k + 1)) ret += get(i, j, k + 1);
  if (k > 0 && !solid_at(L, i, j, k - 1)) ret += get(i, j, k - 1);
  return ret;
}

__device__ __forceinline__ void decode(const Lv& L, int64_t q64, int& i, int& j, int& k) {
  uint32_t q = (uint32_t)q64;
  uint32_t row = fdiv(q, L.f2);
  k = (int)(q - row * (uint32_t)L.n2);
  uint32_t ii = fdiv(row, L.f1);
  j = (int)(row - ii * (uint32_t)L.n1);
  i = (int)ii;
}

// cell of colour `phase` number t (n2 even)
__device__ __forceinline__ void colour_cell(const Lv& L, int64_t t64, int phase, int& i, int& j, int& k) {
  uint32_t t = (uint32_t)t64;
  uint32_t row = fdiv(t, L.fh);
  int kk = (int)(t - row * (uint32_t)(L.n2 >> 1));
  uint32_t ii = fdiv(row, L.f1);
  j = (int)(row - ii * (uint32_t)L.n1);
  i = (int)ii;
  k = 2 * kk + ((i + j + phase) & 1);
}

// Reduction kernels give each thread CPT cells (blockDim apart), so a block
// covers BS*CPT consecutive cells.
constexpr int CPT = 4;
inline int red_blocks(int64_t n) { return (int)((n + (int64_t)BS * CPT - 1) / ((int64_t)BS * CPT)); }
#define FOR_CELLS(q, N)                                                                          \
  for (int64_t _u = 0, q = (int64_t)blockIdx.x * (BS * CPT) + threadIdx.x; _u < CPT; ++_u, q += BS) \
    if (q < (N))

// ---------------------------------------------------------------------------
// Deterministic reductions: per-block FP64 partials, then one small kernel sums
// them in a fixed order and runs the finalize action.
// ---------------------------------------------------------------------------
struct Red {
  double* partials;      // 2 per block of the largest reduction launch
  double* sc;            // scalars
};

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

__device__ void finalize(double* sc, int fin, double s0, double s1) {
  switch (fin) {
    case F_MEAN:
      sc[S_MEAN] = s0 / sc[S_NCELL];
      sc[S_X0ACT] = 0.0;
      break;
    case F_MEAN_B2: {
      // the recenter folded into the first level-0 pass: the mean (all walls) stays
      // in S_MEAN for it, and ||r - mean||^2 = sum r^2 - n mean^2 (no solids here)
      const double m = sc[S_NEU] != 0.0 ? s0 / sc[S_NCELL] : 0.0;
      const double b2 = s1 - sc[S_NCELL] * m * m;
      sc[S_MEAN] = m;
      sc[S_B2] = b2;
      sc[S_TOL2] = sc[S_RELTOL] * sc[S_RELTOL] * b2;
      sc[S_X0ACT] = 0.0;
      break;
    }
    case F_SUM2:
      sc[S_B2] = s0;
      sc[S_MEAN] = 0.0;
      break;
    case F_B2:
      sc[S_B2] = s0;
      sc[S_MEAN] = 0.0;  // r is mean-free now
      sc[S_TOL2] = sc[S_RELTOL] * sc[S_RELTOL] * s0;  // relative rule; replaced after z.r in reference mode
      break;
    case F_ZR_INIT:
      sc[S_X0ACT] = 0.0;   // the first level-0 pass has applied it
      sc[S_MEAN] = 0.0;    // and recentred r (k_update_xr8's prediction starts from 0)
      sc[S_ZTR] = s0;
      sc[S_ZTR0] = s0;
      sc[S_BETA] = 0.0;
      if (sc[S_STOPMODE] == 0.0) {
        // sim/mgpcg.py:228 tol = max(abs_tol, z0'r0 * rel_tol); solid: abs_tol (sim/mgpcg_solid.py:281)
        sc[S_TOL2] = sc[S_SOLID] != 0.0 ? sc[S_ABSTOL] : fmax(sc[S_ABSTOL], s0 * sc[S_RELTOL]);
      }
      break;
    case F_ZR: {
      // sim/mgpcg.py:270-276 (S_EPS: the reference's 1e-12, or 0 with a zero guard)
      double nz = s0;
      double den = sc[S_ZTR] + sc[S_EPS];
      sc[S_BETA] = den != 0.0 ? nz / den : 0.0;
      sc[S_ZTR] = nz;
      break;
    }
    case F_ALPHA: {
      // sim/mgpcg.py:236-239; s1 = sum of Ap (the update's recenter, see k_update_xr8)
      sc[S_PAP] = s0;
      sc[S_APMEAN] = sc[S_NEU] != 0.0 ? s1 / sc[S_NCELL] : 0.0;
      double den = s0 + sc[S_EPS];
      sc[S_ALPHA] = den != 0.0 ? sc[S_ZTR] / den : 0.0;
      break;
    }
    case F_XR: {
      // sim/mgpcg.py:248-259: rTr and the stop test; s1 = sum of r for recenter
      sc[S_RTR] = s0;
      sc[S_MEAN] = sc[S_NEU] != 0.0 ? s1 / sc[S_NCELL] : 0.0;
      double it = sc[S_ITERS] + 1.0;
      sc[S_ITERS] = it;
      bool conv = sc[S_STOPMODE] == 0.0 ? (s0 < sc[S_TOL2]) : (s0 <= sc[S_TOL2]);
      if (conv)
        sc[S_DONE] = 1.0;
      else if (it >= sc[S_MAXIT])
        sc[S_DONE] = 2.0;
      break;
    }
    case F_DIVSUM:
      // sim/mgpcg_solid.py:372-383, 421-424: one-sided correction
      sc[S_DSUM] = s0;
      sc[S_NFLUID] = s1;
      sc[S_CORR] = (s0 > 0.0 && s1 > 0.0) ? s0 / s1 : 0.0;
      break;
    case F_X0:
      sc[S_R02] = s0;
      sc[S_USEX0] = (s0 <= sc[S_B2]) ? 1.0 : 0.0;
      sc[S_X0ACT] = (s0 <= sc[S_B2]) ? 1.0 : 2.0;
      if (s0 <= sc[S_B2]) sc[S_MEAN] = 0.0;   // r = b - A x0 is mean-free (1'A = 0 with walls all round)
      break;
    default:
      break;
  }
}

// Per-block partial sums (fixed order within the block).  The launch that
// follows, k_finish, adds the partials in a fixed order and runs the finalize
// action; the kernel boundary orders the two, so no fence or counter is needed.
__device__ void block_partial(const Red& R, double s0, double s1) {
  __shared__ double sh[2][BS / 64];
  s0 = wave_sum(s0);
  s1 = wave_sum(s1);
  int w = threadIdx.x >> 6, ln = threadIdx.x & 63;
  if (ln == 0) {
    sh[0][w] = s0;
    sh[1][w] = s1;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double a = 0, b = 0;
#pragma unroll
    for (int q = 0; q < BS / 64; ++q) {
      a += sh[0][q];
      b += sh[1][q];
    }
    R.partials[2 * blockIdx.x] = a;
    R.partials[2 * blockIdx.x + 1] = b;
  }
}

constexpr int FIN_THREADS = 1024;
// guard: skip when the solve has already stopped (the producing kernel skipped too).
__global__ void __launch_bounds__(FIN_THREADS) k_finish(Red R, int nb, int fin, int guard) {
  if (guard && R.sc[S_DONE] != 0.0) return;
  __shared__ double sh[2][FIN_THREADS / 64];
  double a = 0, b = 0;
  for (int q = threadIdx.x; q < nb; q += FIN_THREADS) {
    a += R.partials[2 * q];
    b += R.partials[2 * q + 1];
  }
  a = wave_sum(a);
  b = wave_sum(b);
  int w = threadIdx.x >> 6, ln = threadIdx.x & 63;
  if (ln == 0) {
    sh[0][w] = a;
    sh[1][w] = b;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double ta = 0, tb = 0;
    for (int q = 0; q < FIN_THREADS / 64; ++q) {
      ta += sh[0][q];
      tb += sh[1][q];
    }
    finalize(R.sc, fin, ta, tb);
  }
}

#define EXIT_IF_DONE(sc) \
  if ((sc)[S_DONE] != 0.0) return

// ---------------------------------------------------------------------------
// Boundary conditions, divergence, gradient (sim/mgpcg.py:373-432,
// sim/mgpcg_solid.py:326-411)
// ---------------------------------------------------------------------------
struct Faces3 {
  float* f[3];
};

// apply_bc: box walls zeroed, then the faces of solid cells set to boundary_vel;
// where two solid cells share a face the higher-index cell's value stands (the
// result of a sequential loop over cells).
// done_guard (when not null): skip unless the solve has stopped (S_DONE set),
// for the post-solve launches enqueued before the host has seen the flag.
__global__ void __launch_bounds__(BS) k_apply_bc(Grid g, int bt, Faces3 U, const int* __restrict__ bm,
                                                const float* __restrict__ bvel, const double* __restrict__ done_guard) {
  if (done_guard && done_guard[S_DONE] == 0.0) return;
  int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  int a, i, j, k;
  int64_t l, nfa;
  if (!face_split(g, t, a, l, nfa, i, j, k)) return;
  const int nx = g.nx, ny = g.ny, nz = g.nz;
  int I[3] = {i, j, k};
  int n[3] = {nx, ny, nz};
  float v = U.f[a][l];
  bool changed = false;
  if (I[a] == 0 && ((bt >> (2 * a)) & 1)) { v = 0.f; changed = true; }
  if (I[a] == n[a] && ((bt >> (2 * a + 1)) & 1)) { v = 0.f; changed = true; }
  if (bm) {
    const int64_t N = (int64_t)nx * ny * nz;
    // cell on the + side (index I[a]) and on the - side (I[a]-1)
    int hi[3] = {i, j, k}, lo[3] = {i, j, k};
    lo[a] -= 1;
    bool done = false;
    if (I[a] < n[a]) {
      int64_t q = ((int64_t)hi[0] * ny + hi[1]) * nz + hi[2];
      if (bm[q] > 0) { v = bvel[a * N + q]; done = true; changed = true; }
    }
    if (!done && I[a] > 0) {
      int64_t q = ((int64_t)lo[0] * ny + lo[1]) * nz + lo[2];
      if (bm[q] > 0) { v = bvel[a * N + q]; changed = true; }
    }
  }
  if (changed) U.f[a][l] = v;
}

// walls (bit 2a: low side of axis a, 2a+1: high side): when non-zero, a face on
// such a wall reads as 0, the value apply_bc would have written there first
// (the no-solid projection skips that pass: see edge_poisson).
__device__ __forceinline__ float divergence_at(int nx, int ny, int nz, const Faces3& U, int i, int j, int k,
                                               int walls = 0) {
  // sim/mgpcg.py:394-404 (the sample() clamps never bind for i+1 <= nx)
  int64_t qx = ((int64_t)i * ny + j) * nz + k;
  int64_t qy = ((int64_t)i * (ny + 1) + j) * nz + k;
  int64_t qz = ((int64_t)i * ny + j) * (nz + 1) + k;
  float vl = U.f[0][qx], vr = U.f[0][qx + (int64_t)ny * nz];
  float vb = U.f[1][qy], vt = U.f[1][qy + nz];
  float va = U.f[2][qz], vc = U.f[2][qz + 1];
  if (walls) {
    if (i == 0 && (walls & 1)) vl = 0.f;
    if (i == nx - 1 && (walls & 2)) vr = 0.f;
    if (j == 0 && (walls & 4)) vb = 0.f;
    if (j == ny - 1 && (walls & 8)) vt = 0.f;
    if (k == 0 && (walls & 16)) va = 0.f;
    if (k == nz - 1 && (walls & 32)) vc = 0.f;
  }
  return vr - vl + vt - vb + vc - va;
}

// solid: sum and count of the fluid divergence
__global__ void __launch_bounds__(BS) k_div_sum(Grid g, Faces3 U, const int8_t* __restrict__ bm0, Red R) {
  const int nx = g.nx, ny = g.ny, nz = g.nz;
  const int64_t N = (int64_t)nx * ny * nz;
  double s = 0, c = 0;
  FOR_CELLS(q, N) {
    if (bm0[q] > 0) continue;
    int i, j, k;
    cell_ijk(g, q, i, j, k);
    s += (double)divergence_at(nx, ny, nz, U, i, j, k);
    c += 1.0;
  }
  block_partial(R, s, c);
}

// init(u_div, -1) (sim/mgpcg.py:67-75): r = -(div - corr) on fluid cells (0 on
// solids), x = x0 or 0, p = 0; partial sums of r (the recenter's mean) and r^2.
// When b is given instead (solve API), r = b.
__global__ void __launch_bounds__(BS) k_init(Grid g, Faces3 U, const float* __restrict__ b,
                                            const int8_t* __restrict__ bm0, float* __restrict__ r, float* __restrict__ x,
                                            float* __restrict__ p, const float* __restrict__ x0, Red R, int walls) {
  const int nx = g.nx, ny = g.ny, nz = g.nz;
  const int64_t N = (int64_t)nx * ny * nz;
  const double* sc = R.sc;
  float corr = (float)sc[S_CORR];
  double s = 0, s2 = 0;
  FOR_CELLS(q, N) {
    float v;
    if (b) {
      v = b[q];
    } else if (bm0 && bm0[q] > 0) {
      v = 0.f;
    } else {
      int i, j, k;
      cell_ijk(g, q, i, j, k);
      float d = divergence_at(nx, ny, nz, U, i, j, k, walls);
      if (bm0) d -= corr;
      v = d * -1.0f;
    }
    r[q] = v;
    x[q] = x0 ? x0[q] : 0.f;
    p[q] = 0.f;
    s += (double)v;
    s2 += (double)v * (double)v;
  }
  block_partial(R, s, s2);
}

// r -= mean over all cells (sim/mgpcg.py:169-175) and ||r||^2 over fluid cells.
__global__ void __launch_bounds__(BS) k_recenter(Lv L, float* __restrict__ r, Red R) {
  const int64_t N = (int64_t)L.n0 * L.n1 * L.n2;
  float mean = (float)R.sc[S_MEAN];
  double s = 0;
  FOR_CELLS(q, N) {
    float v = r[q] - mean;
    r[q] = v;
    if (!(L.bm && L.bm[q] > 0)) s += (double)v * (double)v;
  }
  block_partial(R, s, 0.0);
}

// warm start: w = (r - mean) - A x0 and ||w||^2; decide whether to keep x0.
// mean: S_MEAN, 0 once r is recentred, else the recenter still to come.
__global__ void __launch_bounds__(BS) k_residual_x0(Lv L, int bt, const float* __restrict__ r, const float* __restrict__ x,
                                                   float* __restrict__ w, Red R) {
  const int64_t N = (int64_t)L.n0 * L.n1 * L.n2;
  const float mean = (float)R.sc[S_MEAN];
  double s = 0;
  FOR_CELLS(q, N) {
    int i, j, k;
    decode(L, q, i, j, k);
    if (L.bm && L.bm[q] > 0) {
      w[q] = r[q];
      continue;
    }
    float ax = num_fluid(L, bt, i, j, k) * x[q] -
               neighbor_sum(L, i, j, k, [&](int a, int b, int c) { return x[lid(L, a, b, c)]; });
    float v = (r[q] - mean) - ax;
    w[q] = v;
    s += (double)v * (double)v;
  }
  block_partial(R, s, 0.0);
}

__global__ void __launch_bounds__(BS) k_apply_x0(int64_t N, float* __restrict__ r, const float* __restrict__ w,
                                                float* __restrict__ x, const double* __restrict__ sc) {
  int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= N) return;
  if (sc[S_USEX0] != 0.0) r[q] = w[q];
  else x[q] = 0.f;
}

// subtract_grad_p (sim/mgpcg.py:406-432) with p = x, + apply_bc, and the
// optional force buffer -(grad p) (sim/mgpcg.py:447-479).
// walls: non-zero in the no-solid projection, whose closing apply_bc (box
// walls zeroed) is folded in here: a face on a wall gets 0, as apply_bc after
// the subtraction would leave it; the force buffer is -grad p everywhere.
__global__ void __launch_bounds__(BS) k_subtract_grad(Grid g, Faces3 U, Faces3 F, const float* __restrict__ P,
                                                     int walls, const double* __restrict__ done_guard) {
  if (done_guard && done_guard[S_DONE] == 0.0) return;
  int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  int a, i, j, k;
  int64_t l, nfa;
  if (!face_split(g, t, a, l, nfa, i, j, k)) return;
  const int nx = g.nx, ny = g.ny, nz = g.nz;
  int I[3] = {i, j, k};
  int n[3] = {nx, ny, nz};
  auto pat = [&](int ii, int jj, int kk) {
    ii = max(0, min(ii, nx - 1));
    jj = max(0, min(jj, ny - 1));
    kk = max(0, min(kk, nz - 1));
    return P[((int64_t)ii * ny + jj) * nz + kk];
  };
  float pr = pat(i, j, k);
  float pl = pat(i - (a == 0), j - (a == 1), k - (a == 2));
  if (I[a] - 1 < 0) pl = 0.f;
  if (I[a] >= n[a]) pr = 0.f;
  const bool wall = walls && ((I[a] == 0 && ((walls >> (2 * a)) & 1)) || (I[a] == n[a] && ((walls >> (2 * a + 1)) & 1)));
  if (wall)
    U.f[a][l] = 0.f;
  else
    U.f[a][l] -= (pr - pl);
  if (F.f[a]) F.f[a][l] = -(pr - pl);
}

// k_subtract_grad with four faces per thread for the x and y faces (the same
// values): aligned float4 runs along k (nz % 4 == 0, 16-byte aligned fields);
// z faces (rows of nz + 1, not 16-byte aligned) one per thread.
__global__ void __launch_bounds__(BS) k_subtract_grad4(Grid g, Faces3 U, Faces3 F, const float* __restrict__ P,
                                                      int walls, const double* __restrict__ done_guard) {
  if (done_guard && done_guard[S_DONE] == 0.0) return;
  const uint32_t nx = g.nx, ny = g.ny, nz = g.nz, q4 = nz >> 2;
  const uint32_t n0 = (nx + 1) * ny * q4, n1 = nx * (ny + 1) * q4;
  uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n0 + n1) {
    const int a = t < n0 ? 0 : 1;
    if (a == 1) t -= n0;
    const uint32_t row = t / q4, k0 = (t - row * q4) << 2;
    const uint32_t d1 = ny + (uint32_t)a;   // the face grid's j extent
    const uint32_t i = row / d1, j = row - i * d1;
    const uint32_t I = a == 0 ? i : j, n = a == 0 ? nx : ny;
    float4 pr = make_float4(0.f, 0.f, 0.f, 0.f), pl = pr;
    if (I < n) pr = *(const float4*)(P + ((size_t)i * ny + j) * nz + k0);
    if (I > 0) pl = *(const float4*)(P + ((size_t)(i - (a == 0)) * ny + (j - (a == 1))) * nz + k0);
    const size_t l = (size_t)row * nz + k0;
    const float4 d = make_float4(pr.x - pl.x, pr.y - pl.y, pr.z - pl.z, pr.w - pl.w);
    const bool wall = walls && ((I == 0 && ((walls >> (2 * a)) & 1)) || (I == n && ((walls >> (2 * a + 1)) & 1)));
    float4* u = (float4*)(U.f[a] + l);
    if (wall) {
      *u = make_float4(0.f, 0.f, 0.f, 0.f);
    } else {
      float4 v = *u;
      v.x -= d.x;
      v.y -= d.y;
      v.z -= d.z;
      v.w -= d.w;
      *u = v;
    }
    if (F.f[a]) *(float4*)(F.f[a] + l) = make_float4(-d.x, -d.y, -d.z, -d.w);
    return;
  }
  t -= n0 + n1;
  if (t >= nx * ny * (nz + 1)) return;
  const uint32_t row = t / (nz + 1), k = t - row * (nz + 1);
  const uint32_t i = row / ny, j = row - i * ny;
  const float* pc = P + ((size_t)i * ny + j) * nz;
  const float pr = k < nz ? pc[k] : 0.f;
  const float pl = k > 0 ? pc[k - 1] : 0.f;
  const bool wall = walls && ((k == 0 && ((walls >> 4) & 1)) || (k == nz && ((walls >> 5) & 1)));
  if (wall)
    U.f[2][t] = 0.f;
  else
    U.f[2][t] -= (pr - pl);
  if (F.f[2]) F.f[2][t] = -(pr - pl);
}

// dst = src (n floats), skipped unless the solve has stopped (see k_apply_bc)
__global__ void __launch_bounds__(BS) k_copy_done(int64_t n, const float* __restrict__ src, float* __restrict__ dst,
                                                 const double* __restrict__ done_guard) {
  if (done_guard && done_guard[S_DONE] == 0.0) return;
  for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += (int64_t)gridDim.x * blockDim.x)
    dst[q] = src[q];
}

// ---------------------------------------------------------------------------
// CG vector kernels
// ---------------------------------------------------------------------------
// p = z + beta p (fluid cells; sim/mgpcg.py:142-145), Ap = A p with neighbours'
// new p computed on the fly, and p.Ap (sim/mgpcg.py:112-130).
template <class S>
__global__ void __launch_bounds__(BS) k_p_Ap(Lv L, int bt, const S* __restrict__ z, float* __restrict__ p,
                                            float* __restrict__ Ap, Red R) {
  EXIT_IF_DONE(R.sc);
  const int64_t N = (int64_t)L.n0 * L.n1 * L.n2;
  const float beta = (float)R.sc[S_BETA];
  double s = 0;
  FOR_CELLS(q, N) {
    int i, j, k;
    decode(L, q, i, j, k);
    if (L.bm && L.bm[q] > 0) continue;  // solid p stays (0), Ap untouched
    auto pn = [&](int a, int b, int c) {
      int64_t m = lid(L, a, b, c);
      return ld(z, m) + beta * p[m];
    };
    float pq = pn(i, j, k);
    float ap = num_fluid(L, bt, i, j, k) * pq - neighbor_sum(L, i, j, k, pn);
    Ap[q] = ap;
    s += (double)pq * (double)ap;
  }
  // p itself is written by k_update_xr: neighbours above read the old p
  block_partial(R, s, 0.0);
}

// p = z + beta p, x += alpha p, r -= alpha Ap (fluid cells); r.r over fluid
// cells and the sum of r over all cells (sim/mgpcg.py:132-145, 248-249, 169-175).
template <class S>
__global__ void __launch_bounds__(BS) k_update_xr(Lv L, const S* __restrict__ z, float* __restrict__ x,
                                                 float* __restrict__ r, float* __restrict__ p,
                                                 const float* __restrict__ Ap, Red R) {
  EXIT_IF_DONE(R.sc);
  const int64_t N = (int64_t)L.n0 * L.n1 * L.n2;
  const float alpha = (float)R.sc[S_ALPHA];
  const float beta = (float)R.sc[S_BETA];
  double s2 = 0, s1 = 0;
  FOR_CELLS(q, N) {
    float rv = r[q];
    if (!(L.bm && L.bm[q] > 0)) {
      float pn = ld(z, q) + beta * p[q];
      p[q] = pn;
      x[q] += alpha * pn;
      rv -= alpha * Ap[q];
      r[q] = rv;
      s2 += (double)rv * (double)rv;
    }
    s1 += (double)rv;
  }
  block_partial(R, s2, s1);
}

// ---------------------------------------------------------------------------
// V-cycle kernels (sim/mgpcg.py:147-195)
// ---------------------------------------------------------------------------
// Level-0 entry: r <- r - mean (recenter) on every cell; the level's residual
// copy rl (== r in FP32 mode) is written; z = r/diag on red fluid cells, 0
// elsewhere (fill(0) + the first red half-sweep).
template <class S>
__global__ void __launch_bounds__(BS) k_first0(Lv L, int bt, float* __restrict__ r, S* __restrict__ rl,
                                              S* __restrict__ z, const double* __restrict__ sc, int ignore_done) {
  if (!ignore_done && sc[S_DONE] != 0.0) return;
  const int64_t N = (int64_t)L.n0 * L.n1 * L.n2;
  int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= N) return;
  float mean = (float)sc[S_MEAN];
  float v = r[q] - mean;
  r[q] = v;
  if ((void*)rl != (void*)r) st(rl, q, v);
  int i, j, k;
  decode(L, q, i, j, k);
  float zv = 0.f;
  if (((i + j + k) & 1) == 0 && !(L.bm && L.bm[q] > 0)) zv = v / num_fluid(L, bt, i, j, k);
  st(z, q, zv);
}

// One red-black Gauss-Seidel half-sweep (sim/mgpcg.py:160-167).
template <class S>
__global__ void __launch_bounds__(BS) k_smooth(Lv L, int bt, const S* __restrict__ r, S* __restrict__ z, int phase,
                                              const double* __restrict__ sc, int ignore_done) {
  if (!ignore_done && sc[S_DONE] != 0.0) return;
  const int64_t half = ((int64_t)L.n0 * L.n1 * L.n2) >> 1;
  int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= half) return;
  int i, j, k;
  colour_cell(L, t, phase, i, j, k);
  int64_t q = lid(L, i, j, k);
  if (L.bm && L.bm[q] > 0) return;
  float ns = neighbor_sum(L, i, j, k, [&](int a, int b, int c) { return ld(z, lid(L, a, b, c)); });
  st(z, q, (ld(r, q) + ns) / num_fluid(L, bt, i, j, k));
}

// Last red half-sweep of level 0 fused with reduce(z, r) over fluid cells:
// thread per (k even, k+1) pair, updates the red cell, reads the black one.
template <class S>
__global__ void __launch_bounds__(BS) k_smooth_last_dot(Lv L, int bt, const S* __restrict__ rl,
                                                       const float* __restrict__ r, S* __restrict__ z, Red R,
                                                       int ignore_done) {
  if (!ignore_done && R.sc[S_DONE] != 0.0) return;
  const int64_t pairs = ((int64_t)L.n0 * L.n1 * L.n2) >> 1;
  double s = 0;
  FOR_CELLS(t, pairs) {
    uint32_t row = fdiv((uint32_t)t, L.fh);
    int kk = (int)((uint32_t)t - row * (uint32_t)(L.n2 >> 1));
    uint32_t ii = fdiv(row, L.f1);
    int j = (int)(row - ii * (uint32_t)L.n1), i = (int)ii;
#pragma unroll
    for (int e = 0; e < 2; ++e) {
      int k = 2 * kk + e;
      int64_t q = lid(L, i, j, k);
      if (L.bm && L.bm[q] > 0) continue;
      float zv;
      if (((i + j + k) & 1) == 0) {
        float ns = neighbor_sum(L, i, j, k, [&](int a, int b, int c) { return ld(z, lid(L, a, b, c)); });
        zv = (ld(rl, q) + ns) / num_fluid(L, bt, i, j, k);
        st(z, q, zv);
        zv = ld(z, q);  // the stored value is what CG uses
      } else {
        zv = ld(z, q);
      }
      s += (double)zv * (double)r[q];
    }
  }
  block_partial(R, s, 0.0);
}

// Restriction (sim/mgpcg.py:147-153) in gather form: each coarse cell sums
// res/2 of its fluid children, then gets the fused fill(0) + first red
// half-sweep of its level.
template <class S>
__global__ void __launch_bounds__(BS) k_restrict(Lv F, Lv C, int bt, const S* __restrict__ rf, const S* __restrict__ zf,
                                                S* __restrict__ rc, S* __restrict__ zc, const double* __restrict__ sc,
                                                int ignore_done) {
  if (!ignore_done && sc[S_DONE] != 0.0) return;
  const int64_t NC = (int64_t)C.n0 * C.n1 * C.n2;
  int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= NC) return;
  int I, J, K;
  decode(C, q, I, J, K);
  float acc = 0.f;
#pragma unroll
  for (int c = 0; c < 8; ++c) {
    int i = 2 * I + (c >> 2), j = 2 * J + ((c >> 1) & 1), k = 2 * K + (c & 1);
    int64_t f = lid(F, i, j, k);
    if (F.bm && F.bm[f] > 0) continue;
    float ns = neighbor_sum(F, i, j, k, [&](int a, int b, int cc) { return ld(zf, lid(F, a, b, cc)); });
    float res = ld(rf, f) - (num_fluid(F, bt, i, j, k) * ld(zf, f) - ns);
    acc += res * 1.0f / 2.0f;
  }
  st(rc, q, acc);
  float zv = 0.f;
  if (((I + J + K) & 1) == 0 && !(C.bm && C.bm[q] > 0)) zv = ld(rc, q) / num_fluid(C, bt, I, J, K);
  st(zc, q, zv);
}

// Prolongation (sim/mgpcg.py:155-158) fused with the first black post-smoothing
// half-sweep: neighbours read as z + z_coarse.
template <class S>
__global__ void __launch_bounds__(BS) k_prolong_black(Lv F, Lv C, int bt, const S* __restrict__ rf, S* __restrict__ zf,
                                                     const S* __restrict__ zc, const double* __restrict__ sc,
                                                     int ignore_done) {
  if (!ignore_done && sc[S_DONE] != 0.0) return;
  const int64_t half = ((int64_t)F.n0 * F.n1 * F.n2) >> 1;
  int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= half) return;
  int i, j, k;
  colour_cell(F, t, 1, i, j, k);
  int64_t q = lid(F, i, j, k);
  // solid cells: the reference's unmasked prolongation (sim/mgpcg_solid.py:182-185)
  // changes z there, but no kernel ever reads z of a solid cell
  if (F.bm && F.bm[q] > 0) return;
  float ns = neighbor_sum(F, i, j, k, [&](int a, int b, int c) {
    return ld(zf, lid(F, a, b, c)) + ld(zc, lid(C, a >> 1, b >> 1, c >> 1));
  });
  st(zf, q, (ld(rf, q) + ns) / num_fluid(F, bt, i, j, k));
}

// ---------------------------------------------------------------------------
// Row-chunk kernels: one thread owns 8 consecutive cells along k of one row
// (n2 a multiple of 8): one index decode and 16/32-byte vector loads per 8
// cells instead of per cell.  Each cell's arithmetic is that of the per-cell
// kernels above (same neighbour order i+1, i-1, j+1, j-1, k+1, k-1, same
// boundary terms); solids enter through the per-cell bits of k_nbmask.
// ---------------------------------------------------------------------------
template <class S>
__device__ __forceinline__ void ld8(const S* p, float v[8]);
template <>
__device__ __forceinline__ void ld8<float>(const float* p, float v[8]) {
  const float4 a = ((const float4*)p)[0], b = ((const float4*)p)[1];
  v[0] = a.x; v[1] = a.y; v[2] = a.z; v[3] = a.w;
  v[4] = b.x; v[5] = b.y; v[6] = b.z; v[7] = b.w;
}
template <>
__device__ __forceinline__ void ld8<__hip_bfloat16>(const __hip_bfloat16* p, float v[8]) {
  const uint4 u = *(const uint4*)p;
  const uint32_t w[4] = {u.x, u.y, u.z, u.w};
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    v[2 * q] = __uint_as_float(w[q] << 16);
    v[2 * q + 1] = __uint_as_float(w[q] & 0xffff0000u);
  }
}
template <class S>
__device__ __forceinline__ void st8(S* p, const float v[8]);
template <>
__device__ __forceinline__ void st8<float>(float* p, const float v[8]) {
  ((float4*)p)[0] = make_float4(v[0], v[1], v[2], v[3]);
  ((float4*)p)[1] = make_float4(v[4], v[5], v[6], v[7]);
}
template <>
__device__ __forceinline__ void st8<__hip_bfloat16>(__hip_bfloat16* p, const float v[8]) {
  uint32_t w[4];
#pragma unroll
  for (int q = 0; q < 4; ++q)
    w[q] = (uint32_t)__bfloat16_as_ushort(__float2bfloat16(v[2 * q])) |
           ((uint32_t)__bfloat16_as_ushort(__float2bfloat16(v[2 * q + 1])) << 16);
  *(uint4*)p = make_uint4(w[0], w[1], w[2], w[3]);
}
// the value a store of v to S reads back as
template <class S>
__device__ __forceinline__ float through(float v);
template <>
__device__ __forceinline__ float through<float>(float v) { return v; }
template <>
__device__ __forceinline__ float through<__hip_bfloat16>(float v) { return __bfloat162float(__float2bfloat16(v)); }

__device__ __forceinline__ void chunk_ijk(const Lv& L, uint32_t t, int& i, int& j, int& k0) {
  const uint32_t row = fdiv(t, L.f8);
  k0 = 8 * (int)(t - row * (uint32_t)(L.n2 >> 3));
  const uint32_t ii = fdiv(row, L.f1);
  j = (int)(row - ii * (uint32_t)L.n1);
  i = (int)ii;
}

// Solid bits of a cell (k_nbmask): bit e set when neighbour e (in the order
// i+1, i-1, j+1, j-1, k+1, k-1) lies inside the level and is solid, bit 6 when
// the cell itself is solid.  num_fluid = diag_free - popcount(bits 0..5) and
// neighbor_sum skips the neighbours whose bit is set.
constexpr uint32_t NB_SELF = 64u;
// the bits of a chunk's 8 cells, byte c for cell k0 + c (0 without solids)
__device__ __forceinline__ uint64_t nb8(const Lv& L, int64_t q0) {
  return L.nb ? *(const uint64_t*)(L.nb + q0) : 0ull;
}
__device__ __forceinline__ uint32_t nbc(uint64_t m, int c) { return (uint32_t)(m >> (8 * c)) & 0xffu; }

// z at a chunk's 8 cells, its 4 neighbouring rows and the two k-neighbours
struct Nb8 {
  float c[8], ip[8], im[8], jp[8], jm[8];
  float kp, km;
};
template <class S>
__device__ __forceinline__ void load_nb(const Lv& L, const S* z, int i, int j, int k0, int64_t q0, Nb8& N) {
  const int64_t si = (int64_t)L.n1 * L.n2, sj = L.n2;
  ld8(z + q0, N.c);
#pragma unroll
  for (int c = 0; c < 8; ++c) N.ip[c] = N.im[c] = N.jp[c] = N.jm[c] = 0.f;
  if (i < L.n0 - 1) ld8(z + q0 + si, N.ip);
  if (i > 0) ld8(z + q0 - si, N.im);
  if (j < L.n1 - 1) ld8(z + q0 + sj, N.jp);
  if (j > 0) ld8(z + q0 - sj, N.jm);
  N.kp = k0 + 8 < L.n2 ? ld(z, q0 + 8) : 0.f;
  N.km = k0 > 0 ? ld(z, q0 - 1) : 0.f;
}
// neighbor_sum for chunk cell c with its solid bits mb
__device__ __forceinline__ float nsum8(const Lv& L, const Nb8& N, int i, int j, int k0, int c, uint32_t mb) {
  float ret = 0.f;
  if (i < L.n0 - 1 && !(mb & 1u)) ret += N.ip[c];
  if (i > 0 && !(mb & 2u)) ret += N.im[c];
  if (j < L.n1 - 1 && !(mb & 4u)) ret += N.jp[c];
  if (j > 0 && !(mb & 8u)) ret += N.jm[c];
  const int k = k0 + c;
  if (k < L.n2 - 1 && !(mb & 16u)) ret += (c < 7 ? N.c[c + 1] : N.kp);
  if (k > 0 && !(mb & 32u)) ret += (c > 0 ? N.c[c - 1] : N.km);
  return ret;
}
// num_fluid without solids
__device__ __forceinline__ float diag_free(const Lv& L, int bt, int i, int j, int k) {
  float num = 6.0f;
  if (i <= 0 && (bt & 1)) num -= 1.0f;
  if (i >= L.n0 - 1 && ((bt >> 1) & 1)) num -= 1.0f;
  if (j <= 0 && ((bt >> 2) & 1)) num -= 1.0f;
  if (j >= L.n1 - 1 && ((bt >> 3) & 1)) num -= 1.0f;
  if (k <= 0 && ((bt >> 4) & 1)) num -= 1.0f;
  if (k >= L.n2 - 1 && ((bt >> 5) & 1)) num -= 1.0f;
  return num;
}
// num_fluid from the solid bits
__device__ __forceinline__ float diag8(const Lv& L, int bt, int i, int j, int k, uint32_t mb) {
  return diag_free(L, bt, i, j, k) - (float)__popc(mb & 63u);
}

// One red-black half-sweep (as k_smooth).
template <class S>
__global__ void __launch_bounds__(BS) k_smooth8(Lv L, int bt, const S* __restrict__ r, S* __restrict__ z, int phase,
                                               const double* __restrict__ sc, int ignore_done) {
  if (!ignore_done && sc[S_DONE] != 0.0) return;
  const uint32_t t_ = blockIdx.x * blockDim.x + threadIdx.x;
  if (t_ >= (uint32_t)((int64_t)L.n0 * L.n1 * L.n2 >> 3)) return;
  int i, j, k0;
  chunk_ijk(L, t_, i, j, k0);
  const int64_t q0 = ((int64_t)i * L.n1 + j) * L.n2 + k0;
  const uint64_t M = nb8(L, q0);
  Nb8 N;
  load_nb(L, z, i, j, k0, q0, N);
  float rv[8], out[8];
  ld8(r + q0, rv);
#pragma unroll
  for (int c = 0; c < 8; ++c) {
    const int k = k0 + c;
    const uint32_t mb = nbc(M, c);
    out[c] = N.c[c];
    if (((i + j + k) & 1) == phase && !(mb & NB_SELF))
      out[c] = (rv[c] + nsum8(L, N, i, j, k0, c, mb)) / diag8(L, bt, i, j, k, mb);
  }
  st8(z + q0, out);
}

// Level-0 entry (as k_first0).  w, x: the warm start's decision, taken here
// in the first iteration (S_X0ACT, set by its finish): 1 r = w (the residual
// of the kept x0), 2 x = 0 (x0 dropped, r is the recentred b).
template <class S>
__global__ void __launch_bounds__(BS) k_first08(Lv L, int bt, float* __restrict__ r, S* __restrict__ rl,
                                               S* __restrict__ z, const double* __restrict__ sc, int ignore_done,
                                               const float* __restrict__ w, float* __restrict__ x) {
  if (!ignore_done && sc[S_DONE] != 0.0) return;
  const uint32_t t_ = blockIdx.x * blockDim.x + threadIdx.x;
  if (t_ >= (uint32_t)((int64_t)L.n0 * L.n1 * L.n2 >> 3)) return;
  int i, j, k0;
  chunk_ijk(L, t_, i, j, k0);
  const int64_t q0 = ((int64_t)i * L.n1 + j) * L.n2 + k0;
  const float mean = (float)sc[S_MEAN];
  const double act = sc[S_X0ACT];
  const uint64_t M = nb8(L, q0);
  float v[8], zv[8];
  if (act == 1.0) {
    ld8(w + q0, v);
  } else {
    ld8(r + q0, v);
    if (act == 2.0) {
      const float zero[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
      st8(x + q0, zero);
    }
  }
#pragma unroll
  for (int c = 0; c < 8; ++c) {
    const int k = k0 + c;
    const uint32_t mb = nbc(M, c);
    v[c] = v[c] - mean;
    zv[c] = 0.f;
    if (((i + j + k) & 1) == 0 && !(mb & NB_SELF)) zv[c] = v[c] / diag8(L, bt, i, j, k, mb);
  }
  st8(r + q0, v);
  if ((void*)rl != (void*)r) st8(rl + q0, v);
  st8(z + q0, zv);
}

// Prolongation + first black half-sweep (as k_prolong_black).
template <class S>
__global__ void __launch_bounds__(BS) k_prolong8(Lv F, Lv C, int bt, const S* __restrict__ rf, S* __restrict__ zf,
                                                const S* __restrict__ zc, const double* __restrict__ sc,
                                                int ignore_done) {
  if (!ignore_done && sc[S_DONE] != 0.0) return;
  const uint32_t t_ = blockIdx.x * blockDim.x + threadIdx.x;
  if (t_ >= (uint32_t)((int64_t)F.n0 * F.n1 * F.n2 >> 3)) return;
  int i, j, k0;
  chunk_ijk(F, t_, i, j, k0);
  const int64_t q0 = ((int64_t)i * F.n1 + j) * F.n2 + k0;
  const uint64_t M = nb8(F, q0);
  Nb8 N;
  load_nb(F, zf, i, j, k0, q0, N);
  // the coarse values under each neighbour: rows (i+-1)>>1, (j+-1)>>1, own row at k0/2-1 .. k0/2+4
  const int kc0 = k0 >> 1;
  auto crow = [&](int a, int b, float out[4]) {
    const int64_t qc = lid(C, a >> 1, b >> 1, kc0);
#pragma unroll
    for (int m = 0; m < 4; ++m) out[m] = ld(zc, qc + m);
  };
  float cip[4] = {0, 0, 0, 0}, cim[4] = {0, 0, 0, 0}, cjp[4] = {0, 0, 0, 0}, cjm[4] = {0, 0, 0, 0}, cc[4];
  if (i < F.n0 - 1) crow(i + 1, j, cip);
  if (i > 0) crow(i - 1, j, cim);
  if (j < F.n1 - 1) crow(i, j + 1, cjp);
  if (j > 0) crow(i, j - 1, cjm);
  crow(i, j, cc);
  const int64_t qc = lid(C, i >> 1, j >> 1, kc0);
  const float ckp = k0 + 8 < F.n2 ? ld(zc, qc + 4) : 0.f;
  const float ckm = k0 > 0 ? ld(zc, qc - 1) : 0.f;
  float rv[8], out[8];
  ld8(rf + q0, rv);
#pragma unroll
  for (int c = 0; c < 8; ++c) {
    const int k = k0 + c;
    const uint32_t mb = nbc(M, c);
    out[c] = N.c[c];
    if (((i + j + k) & 1) != 1 || (mb & NB_SELF)) continue;
    float ret = 0.f;
    if (i < F.n0 - 1 && !(mb & 1u)) ret += N.ip[c] + cip[c >> 1];
    if (i > 0 && !(mb & 2u)) ret += N.im[c] + cim[c >> 1];
    if (j < F.n1 - 1 && !(mb & 4u)) ret += N.jp[c] + cjp[c >> 1];
    if (j > 0 && !(mb & 8u)) ret += N.jm[c] + cjm[c >> 1];
    if (k < F.n2 - 1 && !(mb & 16u)) ret += (c < 7 ? N.c[c + 1] + cc[(c + 1) >> 1] : N.kp + ckp);
    if (k > 0 && !(mb & 32u)) ret += (c > 0 ? N.c[c - 1] + cc[(c - 1) >> 1] : N.km + ckm);
    out[c] = (rv[c] + ret) / diag8(F, bt, i, j, k, mb);
  }
  st8(zf + q0, out);
}

// Restriction + fill(0) + first red half-sweep of the coarse level (as
// k_restrict).  Four lanes per 4 coarse cells along k: lane (a, b) computes
// the residual of the 8 fine cells of fine row (2I + a, 2J + b); the block's
// residuals meet in LDS and one thread per coarse cell adds its 8 children in
// the order of k_restrict.
template <class S>
__global__ void __launch_bounds__(BS) k_restrict8(Lv F, Lv C, int bt, const S* __restrict__ rf, const S* __restrict__ zf,
                                                 S* __restrict__ rc, S* __restrict__ zc, const double* __restrict__ sc,
                                                 int ignore_done) {
  if (!ignore_done && sc[S_DONE] != 0.0) return;
  __shared__ float res[BS / 4][4][8];
  const uint32_t t_ = blockIdx.x * blockDim.x + threadIdx.x;
  const uint32_t per_row = (uint32_t)(C.n2 >> 2);
  const uint32_t nchunk = (uint32_t)C.n0 * (uint32_t)C.n1 * per_row;
  const uint32_t u = t_ >> 2;           // coarse chunk: 4 cells along k
  const int ab = (int)(t_ & 3u);        // fine row (a, b) here, coarse cell K0 + ab below
  const bool live = u < nchunk;
  const int lc = threadIdx.x >> 2;
  int I = 0, J = 0, K0 = 0;
  if (live) {
    const uint32_t row = u / per_row;
    K0 = 4 * (int)(u - row * per_row);
    J = (int)(row % (uint32_t)C.n1);
    I = (int)(row / (uint32_t)C.n1);
    const int a = ab >> 1, b = ab & 1;
    const int i = 2 * I + a, j = 2 * J + b, k0 = 2 * K0;
    const int64_t q0 = ((int64_t)i * F.n1 + j) * F.n2 + k0;
    const uint64_t M = nb8(F, q0);
    Nb8 N;
    load_nb(F, zf, i, j, k0, q0, N);
    float rv[8];
    ld8(rf + q0, rv);
#pragma unroll
    for (int c = 0; c < 8; ++c) {
      const uint32_t mb = nbc(M, c);
      // a solid fine cell adds +0 (k_restrict skips it: the same sum)
      res[lc][ab][c] = (mb & NB_SELF) ? 0.f
                                      : (rv[c] - (diag8(F, bt, i, j, k0 + c, mb) * N.c[c] - nsum8(F, N, i, j, k0, c, mb))) *
                                            1.0f / 2.0f;
    }
  }
  __syncthreads();
  if (!live) return;
  const int m = ab;
  float acc = 0.f;
#pragma unroll
  for (int c8 = 0; c8 < 8; ++c8) acc += res[lc][c8 >> 1][2 * m + (c8 & 1)];
  const int K = K0 + m;
  const int64_t q = ((int64_t)I * C.n1 + J) * C.n2 + K;
  st(rc, q, acc);
  const uint32_t mbc = C.nb ? (uint32_t)C.nb[q] : 0u;
  float zv = 0.f;
  if (((I + J + K) & 1) == 0 && !(mbc & NB_SELF)) zv = through<S>(acc) / diag8(C, bt, I, J, K, mbc);
  st(zc, q, zv);
}

// p = z + beta p on the fly, Ap, p.Ap (as k_p_Ap).
template <class S>
__global__ void __launch_bounds__(BS) k_p_Ap8(Lv L, int bt, const S* __restrict__ z, float* __restrict__ p,
                                             float* __restrict__ Ap, Red R) {
  EXIT_IF_DONE(R.sc);
  const float beta = (float)R.sc[S_BETA];
  double s = 0, sa = 0;
  const uint32_t t_ = blockIdx.x * blockDim.x + threadIdx.x;
  if (t_ < (uint32_t)((int64_t)L.n0 * L.n1 * L.n2 >> 3)) {
    int i, j, k0;
    chunk_ijk(L, t_, i, j, k0);
    const int64_t q0 = ((int64_t)i * L.n1 + j) * L.n2 + k0;
    const uint64_t M = nb8(L, q0);
    {
      Nb8 Z, P;
      load_nb(L, z, i, j, k0, q0, Z);
      load_nb(L, (const float*)p, i, j, k0, q0, P);
      Nb8 N;
#pragma unroll
      for (int c = 0; c < 8; ++c) {
        N.c[c] = Z.c[c] + beta * P.c[c];
        N.ip[c] = Z.ip[c] + beta * P.ip[c];
        N.im[c] = Z.im[c] + beta * P.im[c];
        N.jp[c] = Z.jp[c] + beta * P.jp[c];
        N.jm[c] = Z.jm[c] + beta * P.jm[c];
      }
      N.kp = Z.kp + beta * P.kp;
      N.km = Z.km + beta * P.km;
      float out[8];
#pragma unroll
      for (int c = 0; c < 8; ++c) {
        const uint32_t mb = nbc(M, c);
        const float pq = N.c[c];
        out[c] = 0.f;   // solid cells: Ap is never read there
        if (mb & NB_SELF) continue;
        out[c] = diag8(L, bt, i, j, k0 + c, mb) * pq - nsum8(L, N, i, j, k0, c, mb);
        s += (double)pq * (double)out[c];
        sa += (double)out[c];
      }
      st8(Ap + q0, out);
    }
  }
  block_partial(R, s, sa);
}

// k_residual_x0 on a chunked level (no solids: every cell is fluid).
__global__ void __launch_bounds__(BS) k_residual_x08(Lv L, int bt, const float* __restrict__ r,
                                                    const float* __restrict__ x, float* __restrict__ w, Red R) {
  const float mean = (float)R.sc[S_MEAN];
  double s = 0;
  const uint32_t t_ = blockIdx.x * blockDim.x + threadIdx.x;
  if (t_ < (uint32_t)((int64_t)L.n0 * L.n1 * L.n2 >> 3)) {
    int i, j, k0;
    chunk_ijk(L, t_, i, j, k0);
    const int64_t q0 = ((int64_t)i * L.n1 + j) * L.n2 + k0;
    const uint64_t M = nb8(L, q0);
    Nb8 X;
    load_nb(L, x, i, j, k0, q0, X);
    float rv[8], out[8];
    ld8(r + q0, rv);
#pragma unroll
    for (int c = 0; c < 8; ++c) {
      const uint32_t mb = nbc(M, c);
      const float ax = diag8(L, bt, i, j, k0 + c, mb) * X.c[c] - nsum8(L, X, i, j, k0, c, mb);
      out[c] = (rv[c] - mean) - ax;
      s += (double)out[c] * (double)out[c];
    }
    st8(w + q0, out);
  }
  block_partial(R, s, 0.0);
}

// p = z + beta p, x += alpha p, r -= alpha Ap, r.r and sum r (as k_update_xr).
template <class S>
// zn (may equal z, hence no __restrict__ on either): when not null, also the next
// V-cycle's fill(0) + first red half-sweep from the updated residual (the
// first level-0 pass of the next V-cycle, k_first08, is then skipped); rln,
// when not null, the level-0 copy of r in the levels' storage type (mixed
// mode; in dense mode r is level 0's r).  recentre: r also loses its mean, as
// k_first08 would subtract it: the mean of r - alpha Ap is predicted as
// mean(r) - alpha mean(Ap), both known (S_MEAN from the last update's sum,
// S_APMEAN from k_p_Ap8's), so the next update's sum measures what rounding
// left and the next prediction removes it too.
__global__ void __launch_bounds__(BS) k_update_xr8(Lv L, int bt, const S* z, float* __restrict__ x,
                                                  float* __restrict__ r, float* __restrict__ p,
                                                  const float* __restrict__ Ap, S* zn, S* __restrict__ rln, int recentre,
                                                  Red R) {
  EXIT_IF_DONE(R.sc);
  const float alpha = (float)R.sc[S_ALPHA];
  const float beta = (float)R.sc[S_BETA];
  const float mean = recentre ? (float)(R.sc[S_MEAN] - R.sc[S_ALPHA] * R.sc[S_APMEAN]) : 0.f;
  double s2 = 0, s1 = 0;
  const uint32_t t_ = blockIdx.x * blockDim.x + threadIdx.x;
  if (t_ < (uint32_t)((int64_t)L.n0 * L.n1 * L.n2 >> 3)) {
    const int64_t q0 = (int64_t)t_ * 8;
    float zv[8], xv[8], rv[8], pv[8], av[8];
    ld8(z + q0, zv);
    ld8(x + q0, xv);
    ld8(r + q0, rv);
    ld8(p + q0, pv);
    ld8(Ap + q0, av);
    const uint64_t M = nb8(L, q0);
#pragma unroll
    for (int c = 0; c < 8; ++c) {
      if (!(nbc(M, c) & NB_SELF)) {
        const float pn = zv[c] + beta * pv[c];
        pv[c] = pn;
        xv[c] += alpha * pn;
        rv[c] -= alpha * av[c];
        if (recentre) rv[c] -= mean;
        s2 += (double)rv[c] * (double)rv[c];
      }
      s1 += (double)rv[c];
    }
    st8(p + q0, pv);
    st8(x + q0, xv);
    st8(r + q0, rv);
    if (rln) st8(rln + q0, rv);
    if (zn) {
      int i, j, k0;
      chunk_ijk(L, t_, i, j, k0);
      float zo[8];
#pragma unroll
      for (int c = 0; c < 8; ++c) {
        const int k = k0 + c;
        const uint32_t mb = nbc(M, c);
        zo[c] = (((i + j + k) & 1) == 0 && !(mb & NB_SELF)) ? rv[c] / diag8(L, bt, i, j, k, mb) : 0.f;
      }
      st8(zn + q0, zo);
    }
  }
  block_partial(R, s2, s1);
}

// Last red half-sweep of level 0 + reduce(z, r) over fluid cells (as k_smooth_last_dot).
template <class S>
__global__ void __launch_bounds__(BS) k_last_dot8(Lv L, int bt, const S* __restrict__ rl, const float* __restrict__ r,
                                                 S* __restrict__ z, Red R, int ignore_done) {
  if (!ignore_done && R.sc[S_DONE] != 0.0) return;
  double s = 0;
  const uint32_t t_ = blockIdx.x * blockDim.x + threadIdx.x;
  if (t_ < (uint32_t)((int64_t)L.n0 * L.n1 * L.n2 >> 3)) {
    int i, j, k0;
    chunk_ijk(L, t_, i, j, k0);
    const int64_t q0 = ((int64_t)i * L.n1 + j) * L.n2 + k0;
    float rv[8];
    ld8(r + q0, rv);
    const uint64_t M = nb8(L, q0);
    {
      Nb8 N;
      load_nb(L, z, i, j, k0, q0, N);
      float rlv[8], out[8];
      ld8(rl + q0, rlv);
#pragma unroll
      for (int c = 0; c < 8; ++c) {
        const int k = k0 + c;
        const uint32_t mb = nbc(M, c);
        out[c] = N.c[c];
        if (mb & NB_SELF) continue;
        if (((i + j + k) & 1) == 0) out[c] = through<S>((rlv[c] + nsum8(L, N, i, j, k0, c, mb)) / diag8(L, bt, i, j, k, mb));
        s += (double)out[c] * (double)rv[c];
      }
      st8(z + q0, out);
    }
  }
  block_partial(R, s, 0.0);
}

// Bottom of the plain V-cycle: the 20 red/black pairs of sim/mgpcg.py:187-189
// (the first red half-sweep was fused into the restriction) in one workgroup,
// the level held in LDS; each thread decodes its (at most BOT_CPT) cells and
// their neighbour indices once.
constexpr int BOT_THREADS = 1024;
constexpr int BOT_CPT = 4;
template <class S>
__global__ void __launch_bounds__(BOT_THREADS) k_bottom_rbgs(Lv L, int bt, const S* __restrict__ r, S* __restrict__ z,
                                                            const double* __restrict__ sc, int ignore_done,
                                                            int pairs) {
  if (!ignore_done && sc[S_DONE] != 0.0) return;
  extern __shared__ float lds[];
  const int N = L.n0 * L.n1 * L.n2;
  float* zs = lds;
  int q_[BOT_CPT], nb[BOT_CPT][6], col[BOT_CPT];
  float rv[BOT_CPT], dg[BOT_CPT];
#pragma unroll
  for (int c = 0; c < BOT_CPT; ++c) {
    int q = threadIdx.x + c * BOT_THREADS;
    q_[c] = q;
    col[c] = -1;
    if (q >= N) continue;
    zs[q] = ld(z, q);
    int i, j, k;
    decode(L, q, i, j, k);
    bool sol = L.bm && L.bm[q] > 0;
    col[c] = sol ? -1 : ((i + j + k) & 1);
    rv[c] = ld(r, q);
    dg[c] = sol ? 1.f : num_fluid(L, bt, i, j, k);
    int cand[6][3] = {{i + 1, j, k}, {i - 1, j, k}, {i, j + 1, k}, {i, j - 1, k}, {i, j, k + 1}, {i, j, k - 1}};
    bool ok[6] = {i < L.n0 - 1, i > 0, j < L.n1 - 1, j > 0, k < L.n2 - 1, k > 0};
#pragma unroll
    for (int e = 0; e < 6; ++e)
      nb[c][e] = (ok[e] && !solid_at(L, cand[e][0], cand[e][1], cand[e][2])) ? (int)lid(L, cand[e][0], cand[e][1], cand[e][2])
                                                                             : -1;
  }
  __syncthreads();
  for (int s2 = 0; s2 < 2 * pairs - 1; ++s2) {
    int phase = (s2 & 1) == 0 ? 1 : 0;  // black, red, black, ...
#pragma unroll
    for (int c = 0; c < BOT_CPT; ++c) {
      if (col[c] != phase) continue;
      float ns = 0.f;
#pragma unroll
      for (int e = 0; e < 6; ++e)
        if (nb[c][e] >= 0) ns += zs[nb[c][e]];
      zs[q_[c]] = (rv[c] + ns) / dg[c];
    }
    __syncthreads();
  }
#pragma unroll
  for (int c = 0; c < BOT_CPT; ++c)
    if (q_[c] < N) st(z, q_[c], zs[q_[c]]);
}

// The coarse end of the V-cycle in one workgroup: levels lt..L-1 (each at most
// TAIL_MAX_CELLS cells) held in LDS, so their ~10 half-sweeps and transfers
// cost barriers instead of kernel launches.  Same smoother, transfers,
// storage type S and order of operations as the per-level kernels above.
//   part 0: everything (bottom = the 20 red-black pairs in LDS);
//   part 1: the way down to level L-1 (r[L-1] and every tail level written back);
//   part 2: the way up from a z[L-1] given in global memory (dense bottom).
// On entry level lt's r and z (first red half-sweep done) are in global
// memory; on exit level lt's z is.
constexpr int TAIL_THREADS = 1024;
#ifndef EDGE_TAIL_MAX_CELLS
#define EDGE_TAIL_MAX_CELLS 1024
#endif
constexpr int TAIL_MAX_CELLS = EDGE_TAIL_MAX_CELLS;   // larger levels are faster as row-chunk launches
struct Tail {
  int lt, L;
  Lv lv[MAXL];
  void* r[MAXL];   // global level storage (S)
  void* z[MAXL];
  int off[MAXL];   // LDS offset (elements of S) of level l's r; z follows
};

template <class S>
__device__ __forceinline__ void tail_smooth(const Lv& L, int bt, const S* r, S* z, int phase) {
  const int half = (L.n0 * L.n1 * L.n2) >> 1;
  for (int t = threadIdx.x; t < half; t += TAIL_THREADS) {
    int i, j, k;
    colour_cell(L, t, phase, i, j, k);
    int q = (int)lid(L, i, j, k);
    if (L.bm && L.bm[q] > 0) continue;
    float ns = neighbor_sum(L, i, j, k, [&](int a, int b, int c) { return ld(z, lid(L, a, b, c)); });
    st(z, q, (ld(r, q) + ns) / num_fluid(L, bt, i, j, k));
  }
  __syncthreads();
}

template <class S>
__global__ void __launch_bounds__(TAIL_THREADS) k_vtail(Tail T, int bt, int part, const double* __restrict__ sc,
                                                      int ignore_done) {
  if (!ignore_done && sc[S_DONE] != 0.0) return;
  extern __shared__ __align__(16) unsigned char tail_lds[];
  S* lds = (S*)tail_lds;
  auto R = [&](int l) { return lds + T.off[l]; };
  auto Z = [&](int l) { const Lv& L = T.lv[l]; return lds + T.off[l] + L.n0 * L.n1 * L.n2; };
  auto ncells = [&](int l) { const Lv& L = T.lv[l]; return L.n0 * L.n1 * L.n2; };
  const int last = T.L - 1;
  if (part != 2) {
    {
      const int n = ncells(T.lt);
      const S* gr = (const S*)T.r[T.lt];
      const S* gz = (const S*)T.z[T.lt];
      for (int q = threadIdx.x; q < n; q += TAIL_THREADS) {
        R(T.lt)[q] = gr[q];
        Z(T.lt)[q] = gz[q];
      }
      __syncthreads();
    }
    for (int l = T.lt; l < last; ++l) {
      const Lv& F = T.lv[l];
      const Lv& C = T.lv[l + 1];
      S* rf = R(l);
      S* zf = Z(l);
      tail_smooth(F, bt, rf, zf, 1);
      tail_smooth(F, bt, rf, zf, 0);
      tail_smooth(F, bt, rf, zf, 1);
      // restriction + fill(0) + first red half-sweep of level l+1 (as k_restrict)
      S* rc = R(l + 1);
      S* zc = Z(l + 1);
      const int nc = ncells(l + 1);
      for (int q = threadIdx.x; q < nc; q += TAIL_THREADS) {
        int I, J, K;
        decode(C, q, I, J, K);
        float acc = 0.f;
#pragma unroll
        for (int c = 0; c < 8; ++c) {
          int i = 2 * I + (c >> 2), j = 2 * J + ((c >> 1) & 1), k = 2 * K + (c & 1);
          int f = (int)lid(F, i, j, k);
          if (F.bm && F.bm[f] > 0) continue;
          float ns = neighbor_sum(F, i, j, k, [&](int a, int b, int cc) { return ld(zf, lid(F, a, b, cc)); });
          float res = ld(rf, f) - (num_fluid(F, bt, i, j, k) * ld(zf, f) - ns);
          acc += res * 1.0f / 2.0f;
        }
        st(rc, q, acc);
        float zv = 0.f;
        if (((I + J + K) & 1) == 0 && !(C.bm && C.bm[q] > 0)) zv = ld(rc, q) / num_fluid(C, bt, I, J, K);
        st(zc, q, zv);
      }
      __syncthreads();
    }
    if (part == 0) {
      // bottom: the remaining 39 half-sweeps of sim/mgpcg.py:187-189 (black first)
      for (int s2 = 0; s2 < 39; ++s2) tail_smooth(T.lv[last], bt, R(last), Z(last), (s2 & 1) == 0 ? 1 : 0);
    } else {
      // hand the tail levels to the dense bottom and the up part
      for (int l = T.lt; l <= last; ++l) {
        const int n = ncells(l);
        S* gr = (S*)T.r[l];
        S* gz = (S*)T.z[l];
        for (int q = threadIdx.x; q < n; q += TAIL_THREADS) {
          gr[q] = R(l)[q];
          gz[q] = Z(l)[q];
        }
      }
      return;
    }
  } else {
    for (int l = T.lt; l <= last; ++l) {
      const int n = ncells(l);
      const S* gr = (const S*)T.r[l];
      const S* gz = (const S*)T.z[l];
      for (int q = threadIdx.x; q < n; q += TAIL_THREADS) {
        R(l)[q] = gr[q];
        Z(l)[q] = gz[q];
      }
    }
    __syncthreads();
  }
  for (int l = last - 1; l >= T.lt; --l) {
    const Lv& F = T.lv[l];
    const Lv& C = T.lv[l + 1];
    S* rf = R(l);
    S* zf = Z(l);
    const S* zc = Z(l + 1);
    // prolongation + first black half-sweep (as k_prolong_black)
    const int half = ncells(l) >> 1;
    for (int t = threadIdx.x; t < half; t += TAIL_THREADS) {
      int i, j, k;
      colour_cell(F, t, 1, i, j, k);
      int q = (int)lid(F, i, j, k);
      if (F.bm && F.bm[q] > 0) continue;
      float ns = neighbor_sum(F, i, j, k, [&](int a, int b, int c) {
        return ld(zf, lid(F, a, b, c)) + ld(zc, lid(C, a >> 1, b >> 1, c >> 1));
      });
      st(zf, q, (ld(rf, q) + ns) / num_fluid(F, bt, i, j, k));
    }
    __syncthreads();
    tail_smooth(F, bt, rf, zf, 0);
    tail_smooth(F, bt, rf, zf, 1);
    tail_smooth(F, bt, rf, zf, 0);  // l > 0 in the tail
  }
  {
    const int n = ncells(T.lt);
    S* gz = (S*)T.z[T.lt];
    for (int q = threadIdx.x; q < n; q += TAIL_THREADS) gz[q] = Z(T.lt)[q];
  }
}

// ---------------------------------------------------------------------------
// Halo tiles for the coarse levels.  A level below the finest is a few
// microseconds of memory traffic, so the seven launches of its two V-cycle
// legs cost more than their work.  One workgroup per fine tile of 8 x 8 x TK
// cells loads the tile and a halo of HT_HALO cells into LDS and runs a whole
// leg there:
//   down: black, red, black half-sweeps, then the restriction, fill(0) and
//         first red half-sweep of the coarse level (as k_restrict8);
//   up:   prolongation + black, red, black, red half-sweeps (as k_prolong8
//         and three k_smooth8).
// Half-sweep s (1-based) updates the cells at distance >= s from the box edge;
// their neighbours are exact after s-1 half-sweeps, so after 3 (down) or 4
// (up) the tile itself is exact.  Each cell update is the row-chunk kernels'
// arithmetic (same neighbour order, z rounded through S after every update),
// so a leg gives the values of the per-level launches.  A leg reads z and
// writes another buffer (down: zl -> zt, up: zt -> zl): a tile's halo is its
// neighbours' interior, which must not change while the launch runs (a
// launch holds more tiles than fit on the GPU at once).  Loads, sweeps and
// stores are unrolled over compile-time extents so that each thread keeps
// several independent accesses in flight (one workgroup's latency is the
// leg's time).
// ---------------------------------------------------------------------------
constexpr int HT_HALO = 4;
constexpr int HT_T = 8;                       // tile extent along i and j
constexpr int HT_B = HT_T + 2 * HT_HALO;      // box extent along i and j
constexpr int HT_COLS = HT_B * HT_B;          // box columns (along k)
constexpr int HT_KS = 4;                      // threads per box column
constexpr int HT_THREADS = HT_COLS * HT_KS;
constexpr int HT_CB = HT_T / 2 + 4;           // coarse box extent along i and j (up leg)

struct HTile {
  int kp, ckp;    // LDS column strides (elements) of the fine and the coarse box
  int ntj, ntk;   // tiles along j and k
  int masked;     // solid bits staged (the level has solids)
};
template <int TK>
struct HTK {
  static constexpr int BK = TK + 2 * HT_HALO;   // box extent along k
  static constexpr int CK = TK / 2 + 4;          // coarse box extent along k
  static constexpr int SPAN = BK / HT_KS;        // box cells per thread along k
};

template <class S>
__device__ __forceinline__ S to_s(float v);
template <>
__device__ __forceinline__ float to_s<float>(float v) { return v; }
template <>
__device__ __forceinline__ __hip_bfloat16 to_s<__hip_bfloat16>(float v) { return __float2bfloat16(v); }

template <class S>
__device__ __forceinline__ float lds_f(const S* p, int q);
template <>
__device__ __forceinline__ float lds_f<float>(const float* p, int q) { return p[q]; }
template <>
__device__ __forceinline__ float lds_f<__hip_bfloat16>(const __hip_bfloat16* p, int q) {
  return __bfloat162float(p[q]);
}

// box origin of this workgroup's tile
template <int TK>
__device__ __forceinline__ void ht_origin(const HTile& H, int& o0, int& o1, int& o2) {
  int b = (int)blockIdx.x;
  const int tk_ = b % H.ntk;
  b /= H.ntk;
  const int tj_ = b % H.ntj, ti_ = b / H.ntj;
  o0 = ti_ * HT_T - HT_HALO;
  o1 = tj_ * HT_T - HT_HALO;
  o2 = tk_ * TK - HT_HALO;
}

// fine box <- z, r (and the solid bits); consecutive threads along k
template <class S, int TK>
__device__ __forceinline__ void ht_load(const Lv& L, const HTile& H, int o0, int o1, int o2, const S* __restrict__ z,
                                        const S* __restrict__ r, S* zs, S* rs, uint8_t* ms) {
  constexpr int BK = HTK<TK>::BK, N = HT_COLS * BK, PER = (N + HT_THREADS - 1) / HT_THREADS;
  S zv[PER], rv[PER];
  uint8_t mv[PER];
#pragma unroll
  for (int u = 0; u < PER; ++u) {
    const int t = threadIdx.x + u * HT_THREADS;
    const int col = t / BK, kk = t - col * BK;
    const int i = o0 + col / HT_B, j = o1 + col % HT_B, k = o2 + kk;
    const bool in = t < N && i >= 0 && i < L.n0 && j >= 0 && j < L.n1 && k >= 0 && k < L.n2;
    const int64_t g = in ? lid(L, i, j, k) : 0;   // outside: load cell 0 and drop it
    const S zg = z[g], rg = r[g];
    zv[u] = in ? zg : to_s<S>(0.f);
    rv[u] = in ? rg : to_s<S>(0.f);
    mv[u] = 0;
    if (H.masked) {
      const uint8_t mg = L.nb[g];
      mv[u] = in ? mg : (uint8_t)0;
    }
  }
#pragma unroll
  for (int u = 0; u < PER; ++u) {
    const int t = threadIdx.x + u * HT_THREADS;
    if (t >= N) continue;
    const int col = t / BK, kk = t - col * BK;
    const int q = col * H.kp + kk;
    zs[q] = zv[u];
    rs[q] = rv[u];
    if (H.masked) ms[q] = mv[u];
  }
}

// fine tile -> z
template <class S, int TK>
__device__ __forceinline__ void ht_store(const Lv& L, const HTile& H, int o0, int o1, int o2, const S* zs,
                                         S* __restrict__ z) {
  constexpr int N = HT_T * HT_T * TK, PER = (N + HT_THREADS - 1) / HT_THREADS;
#pragma unroll
  for (int u = 0; u < PER; ++u) {
    const int t = threadIdx.x + u * HT_THREADS;
    if (t >= N) continue;
    const int c = t / TK, kk = t - c * TK;
    const int bi = HT_HALO + c / HT_T, bj = HT_HALO + c % HT_T;
    z[lid(L, o0 + bi, o1 + bj, o2 + HT_HALO + kk)] = zs[(bi * HT_B + bj) * H.kp + HT_HALO + kk];
  }
}

// one half-sweep of colour `phase` over the cells at box distance >= dist; PRO
// adds the coarse correction of each neighbour (prolongation, as k_prolong8).
// All of a thread's new values are computed before any is stored (cells of
// one colour do not read each other).
template <class S, int TK, bool PRO>
__device__ __forceinline__ void ht_sweep(const Lv& L, int bt, const HTile& H, int o0, int o1, int o2, const S* rs,
                                         S* zs, const uint8_t* ms, const S* zcs, int phase, int dist) {
  constexpr int BK = HTK<TK>::BK, SPAN = HTK<TK>::SPAN, NU = (SPAN + 1) / 2;
  const int col = threadIdx.x % HT_COLS, part = threadIdx.x / HT_COLS;
  const int bi = col / HT_B, bj = col % HT_B;
  const int i = o0 + bi, j = o1 + bj;
  const int dc = min(min(bi, HT_B - 1 - bi), min(bj, HT_B - 1 - bj));
  if (dc >= dist && i >= 0 && i < L.n0 && j >= 0 && j < L.n1) {
    const int lo = max(part * SPAN, dist), hi = min((part + 1) * SPAN, BK - dist);
    const int first = lo + (((i + j + o2 + lo) & 1) != phase);
    const int SI = HT_B * H.kp, SJ = H.kp;
    const int c0 = (o0 + HT_HALO) / 2 - 2, c1 = (o1 + HT_HALO) / 2 - 2, c2 = (o2 + HT_HALO) / 2 - 2;
    const int CI = HT_CB * H.ckp, CJ = H.ckp;
    auto zc = [&](int a, int b, int c) {
      return lds_f(zcs, ((a >> 1) - c0) * CI + ((b >> 1) - c1) * CJ + ((c >> 1) - c2));
    };
    // Every read is issued unconditionally (the box holds every neighbour of a
    // cell at distance >= 1, and the padded column one more element) and
    // masked by selects: a term left out adds +0, which leaves a sum that is
    // never -0 unchanged.
    const bool ip = i < L.n0 - 1, im = i > 0, jp = j < L.n1 - 1, jm = j > 0;
    float nv[NU];
    bool ok[NU];
#pragma unroll
    for (int u = 0; u < NU; ++u) {
      const int bk = first + 2 * u;
      const int k = o2 + bk;
      const int q = col * H.kp + bk;
      const uint32_t mb = H.masked ? (uint32_t)ms[q] : 0u;
      ok[u] = bk < hi && k >= 0 && k < L.n2 && !(mb & NB_SELF);
      float v[6];
      v[0] = lds_f(zs, q + SI);
      v[1] = lds_f(zs, q - SI);
      v[2] = lds_f(zs, q + SJ);
      v[3] = lds_f(zs, q - SJ);
      v[4] = lds_f(zs, q + 1);
      v[5] = lds_f(zs, q - 1);
      if (PRO) {
        v[0] += zc(i + 1, j, k);
        v[1] += zc(i - 1, j, k);
        v[2] += zc(i, j + 1, k);
        v[3] += zc(i, j - 1, k);
        v[4] += zc(i, j, k + 1);
        v[5] += zc(i, j, k - 1);
      }
      const bool use[6] = {ip && !(mb & 1u), im && !(mb & 2u), jp && !(mb & 4u), jm && !(mb & 8u),
                           k < L.n2 - 1 && !(mb & 16u), k > 0 && !(mb & 32u)};
      float ret = 0.f;
#pragma unroll
      for (int e = 0; e < 6; ++e) ret += use[e] ? v[e] : 0.f;
      nv[u] = (lds_f(rs, q) + ret) / diag8(L, bt, i, j, k, mb);
    }
#pragma unroll
    for (int u = 0; u < NU; ++u)
      if (ok[u]) zs[col * H.kp + first + 2 * u] = to_s<S>(nv[u]);
  }
  __syncthreads();
}

template <class S, int TK>
__device__ __forceinline__ void ht_carve(const HTile& H, unsigned char* base, S*& zs, S*& rs, uint8_t*& ms, S*& zcs) {
  zs = (S*)base;
  rs = zs + HT_COLS * H.kp;
  zcs = rs + HT_COLS * H.kp;
  ms = (uint8_t*)(zcs + HT_CB * HT_CB * H.ckp);
}

// Down leg of level F (>= 1) into coarse level C.
template <class S, int TK>
__global__ void __launch_bounds__(HT_THREADS) k_down_tile(Lv F, Lv C, int bt, HTile H, const S* __restrict__ rf,
                                                         const S* __restrict__ zf, S* __restrict__ zout,
                                                         S* __restrict__ rc, S* __restrict__ zc,
                                                         const double* __restrict__ sc, int ignore_done) {
  if (!ignore_done && sc[S_DONE] != 0.0) return;
  extern __shared__ __align__(16) unsigned char ht_lds[];
  S *zs, *rs, *zcs;
  uint8_t* ms;
  ht_carve<S, TK>(H, ht_lds, zs, rs, ms, zcs);
  int o0, o1, o2;
  ht_origin<TK>(H, o0, o1, o2);
  ht_load<S, TK>(F, H, o0, o1, o2, zf, rf, zs, rs, ms);
  __syncthreads();
  ht_sweep<S, TK, false>(F, bt, H, o0, o1, o2, rs, zs, ms, zcs, 1, 1);
  ht_sweep<S, TK, false>(F, bt, H, o0, o1, o2, rs, zs, ms, zcs, 0, 2);
  ht_sweep<S, TK, false>(F, bt, H, o0, o1, o2, rs, zs, ms, zcs, 1, 3);
  ht_store<S, TK>(F, H, o0, o1, o2, zs, zout);
  // restriction of the tile's residual (its 4 x 4 x TK/2 coarse cells)
  constexpr int HK = TK / 2;
  const int SI = HT_B * H.kp, SJ = H.kp;
  const int t = threadIdx.x;
  if (t < (HT_T / 2) * (HT_T / 2) * HK) {
    const int a = t / ((HT_T / 2) * HK), b = (t / HK) % (HT_T / 2), c = t % HK;
    const int I = (o0 + HT_HALO) / 2 + a, J = (o1 + HT_HALO) / 2 + b, K = (o2 + HT_HALO) / 2 + c;
    float acc = 0.f;
#pragma unroll
    for (int c8 = 0; c8 < 8; ++c8) {
      const int bi = HT_HALO + 2 * a + (c8 >> 2), bj = HT_HALO + 2 * b + ((c8 >> 1) & 1), bk = HT_HALO + 2 * c + (c8 & 1);
      const int i = o0 + bi, j = o1 + bj, k = o2 + bk;
      const int q = (bi * HT_B + bj) * H.kp + bk;
      const uint32_t mb = H.masked ? (uint32_t)ms[q] : 0u;
      const float v[6] = {lds_f(zs, q + SI), lds_f(zs, q - SI), lds_f(zs, q + SJ),
                          lds_f(zs, q - SJ), lds_f(zs, q + 1), lds_f(zs, q - 1)};
      const bool use[6] = {i < F.n0 - 1 && !(mb & 1u), i > 0 && !(mb & 2u), j < F.n1 - 1 && !(mb & 4u),
                           j > 0 && !(mb & 8u), k < F.n2 - 1 && !(mb & 16u), k > 0 && !(mb & 32u)};
      float ns = 0.f;
#pragma unroll
      for (int e = 0; e < 6; ++e) ns += use[e] ? v[e] : 0.f;
      const float res = lds_f(rs, q) - (diag8(F, bt, i, j, k, mb) * lds_f(zs, q) - ns);
      acc += (mb & NB_SELF) ? 0.f : res * 1.0f / 2.0f;
    }
    const int64_t qc = lid(C, I, J, K);
    st(rc, qc, acc);
    const uint32_t mbc = C.nb ? (uint32_t)C.nb[qc] : 0u;
    float zv = 0.f;
    if (((I + J + K) & 1) == 0 && !(mbc & NB_SELF)) zv = through<S>(acc) / diag8(C, bt, I, J, K, mbc);
    st(zc, qc, zv);
  }
}

// Up leg of level F (>= 1) from the coarse level's z.
template <class S, int TK>
__global__ void __launch_bounds__(HT_THREADS) k_up_tile(Lv F, Lv C, int bt, HTile H, const S* __restrict__ rf,
                                                       const S* __restrict__ zf, S* __restrict__ zout,
                                                       const S* __restrict__ zc, const double* __restrict__ sc,
                                                       int ignore_done) {
  if (!ignore_done && sc[S_DONE] != 0.0) return;
  extern __shared__ __align__(16) unsigned char ht_lds[];
  S *zs, *rs, *zcs;
  uint8_t* ms;
  ht_carve<S, TK>(H, ht_lds, zs, rs, ms, zcs);
  int o0, o1, o2;
  ht_origin<TK>(H, o0, o1, o2);
  ht_load<S, TK>(F, H, o0, o1, o2, zf, rf, zs, rs, ms);
  {
    // coarse box: the parents of the fine box's cells
    constexpr int CK = HTK<TK>::CK, N = HT_CB * HT_CB * CK, PER = (N + HT_THREADS - 1) / HT_THREADS;
    const int c0 = (o0 + HT_HALO) / 2 - 2, c1 = (o1 + HT_HALO) / 2 - 2, c2 = (o2 + HT_HALO) / 2 - 2;
    S v[PER];
#pragma unroll
    for (int u = 0; u < PER; ++u) {
      const int t = threadIdx.x + u * HT_THREADS;
      const int col = t / CK, kk = t - col * CK;
      const int I = c0 + col / HT_CB, J = c1 + col % HT_CB, K = c2 + kk;
      const bool in = t < N && I >= 0 && I < C.n0 && J >= 0 && J < C.n1 && K >= 0 && K < C.n2;
      const S g = zc[in ? lid(C, I, J, K) : 0];
      v[u] = in ? g : to_s<S>(0.f);
    }
#pragma unroll
    for (int u = 0; u < PER; ++u) {
      const int t = threadIdx.x + u * HT_THREADS;
      if (t >= N) continue;
      const int col = t / CK, kk = t - col * CK;
      zcs[col * H.ckp + kk] = v[u];
    }
  }
  __syncthreads();
  ht_sweep<S, TK, true>(F, bt, H, o0, o1, o2, rs, zs, ms, zcs, 1, 1);
  ht_sweep<S, TK, false>(F, bt, H, o0, o1, o2, rs, zs, ms, zcs, 0, 2);
  ht_sweep<S, TK, false>(F, bt, H, o0, o1, o2, rs, zs, ms, zcs, 1, 3);
  ht_sweep<S, TK, false>(F, bt, H, o0, o1, o2, rs, zs, ms, zcs, 0, 4);
  ht_store<S, TK>(F, H, o0, o1, o2, zs, zout);
}

// Bottom of the mixed V-cycle: z = B^-1 r over the fluid cells of the coarsest
// level, B = A + c 1 1^T (SPD; equal to the pseudo-inverse on mean-free r), FP64.
template <class S>
__global__ void __launch_bounds__(256) k_bottom_dense(int nf, const int* __restrict__ fidx, const double* __restrict__ Binv,
                                                     const S* __restrict__ r, S* __restrict__ z, int64_t ncoarse,
                                                     const double* __restrict__ sc, int ignore_done) {
  if (!ignore_done && sc[S_DONE] != 0.0) return;
  int row = blockIdx.x;
  if (row >= nf) return;
  double s = 0;
  const double* Brow = Binv + (int64_t)row * nf;
  for (int c = threadIdx.x; c < nf; c += blockDim.x) s += Brow[c] * (double)ld(r, fidx[c]);
  s = wave_sum(s);
  __shared__ double sh[4];
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0;
    for (int w = 0; w < (int)(blockDim.x >> 6); ++w) t += sh[w];
    st(z, fidx[row], (float)t);
  }
}

// solid masks: level-0 mask from the user's int32 boundary_mask, then the
// coarser levels (sim/mgpcg_solid.py:87-97, every solve).
__global__ void k_mask0(int64_t N, const int* __restrict__ bm, int8_t* __restrict__ out) {
  int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q < N) out[q] = bm[q] > 0 ? 1 : 0;
}
__global__ void k_downsample(Lv F, Lv C, int8_t* __restrict__ out) {
  const int64_t NC = (int64_t)C.n0 * C.n1 * C.n2;
  int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= NC) return;
  int I, J, K;
  decode(C, q, I, J, K);
  int all = 1;
  for (int c = 0; c < 8; ++c)
    if (F.bm[lid(F, 2 * I + (c >> 2), 2 * J + ((c >> 1) & 1), 2 * K + (c & 1))] <= 0) all = 0;
  out[q] = (int8_t)all;
}

// the solid bits of every cell of a level (from its mask)
__global__ void k_nbmask(Lv L, uint8_t* __restrict__ out) {
  const int64_t N = (int64_t)L.n0 * L.n1 * L.n2;
  int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= N) return;
  int i, j, k;
  decode(L, q, i, j, k);
  uint32_t m = L.bm[q] > 0 ? NB_SELF : 0u;
  if (i < L.n0 - 1 && solid_at(L, i + 1, j, k)) m |= 1u;
  if (i > 0 && solid_at(L, i - 1, j, k)) m |= 2u;
  if (j < L.n1 - 1 && solid_at(L, i, j + 1, k)) m |= 4u;
  if (j > 0 && solid_at(L, i, j - 1, k)) m |= 8u;
  if (k < L.n2 - 1 && solid_at(L, i, j, k + 1)) m |= 16u;
  if (k > 0 && solid_at(L, i, j, k - 1)) m |= 32u;
  out[q] = (uint8_t)m;
}

__global__ void k_set_scalars(double* sc, int n, const double* vals) {
  int t = threadIdx.x;
  if (t < n) sc[t] = vals[t];
}
__global__ void k_set_slot(double* sc, int slot, double v) { sc[slot] = v; }

// A p for a given p (test hook; sim/mgpcg.py:119-124).
__global__ void __launch_bounds__(BS) k_apply_A(Lv L, int bt, const float* __restrict__ p, float* __restrict__ Ap) {
  const int64_t N = (int64_t)L.n0 * L.n1 * L.n2;
  int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= N) return;
  if (L.bm && L.bm[q] > 0) {
    Ap[q] = 0.f;
    return;
  }
  int i, j, k;
  decode(L, q, i, j, k);
  Ap[q] = num_fluid(L, bt, i, j, k) * p[q] - neighbor_sum(L, i, j, k, [&](int a, int b, int c) { return p[lid(L, a, b, c)]; });
}

template <class S>
__global__ void k_to_float(int64_t n, const S* __restrict__ a, float* __restrict__ o) {
  int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q < n) o[q] = ld(a, q);
}

// ---------------------------------------------------------------------------
// Solver object
// ---------------------------------------------------------------------------
struct Solver {
  int nx, ny, nz, L, mode, bt;
  int64_t N;
  Lv lv[MAXL];
  int64_t nl[MAXL];
  void* rl[MAXL];   // level residual (level 0 aliases r in plain mode)
  void* zl[MAXL];
  void* zt[MAXL];   // levels >= 1: z after the tile down leg (the up leg reads it, writes zl)
  int8_t* bm[MAXL];  // null when no solids
  uint8_t* nbm[MAXL];  // solid bits per cell (k_nbmask), null when no solids
  float *r, *x, *p, *Ap;
  double* partials;
  double* sc;
  double* sc_host;   // pinned
  bool solid = false;
  const int* user_bm = nullptr;
  const float* user_bvel = nullptr;
  // dense coarse solve
  double* Binv = nullptr;
  int* fidx = nullptr;
  int nf = 0;
  bool dense_ok = false;
  bool dense_tried = false;        // build_dense ran for dense_mask (it may have failed)
  std::vector<int8_t> dense_mask;  // coarse mask the inverse was last built (or tried) for
  size_t elem;
  // CG iteration batches captured as a HIP graph (replayed on the caller's stream)
  int use_graph = 0;
  hipStream_t cap = nullptr;
  hipGraphExec_t gexec[MAXB + 1] = {};  // [0]: the first iteration; [b]: a batch of b iterations
  int hist[4] = {0, 0, 0, 0};            // iterations of the last four solves, newest first (adaptive batches)
  // coarse levels from tail_from on go to one workgroup (0: never)
  int tail_from = 0;
  // halo-tile legs for the levels between the finest and the tail/bottom
  int use_tiles = 1;
  // marks the stop flag's copy in a batch's stream (run_cg's post-solve launches)
  hipEvent_t ev_check = nullptr;
};

// Launches that follow a solve (run_cg): guard = the scalars (skip unless the
// solve has stopped) or NULL (unconditional).
using PostSolve = std::function<void(hipStream_t, const double*)>;

// First level of the one-workgroup tail for the current setup (0 = no tail):
// the first level >= 1 small enough for LDS, used around a dense bottom (the
// tail replaces ~8 short launches by 2; with the red-black bottom the
// dedicated bottom kernel is faster than the tail's generic sweeps).
int tail_level(const Solver& s) {
  if (!s.tail_from) return 0;
  if (!(s.mode >= 1 && s.dense_ok) || s.tail_from == s.L - 1) return 0;
  return s.tail_from;
}

void drop_graph(Solver& s) {
  for (auto& g : s.gexec) {
    if (g) (void)hipGraphExecDestroy(g);
    g = nullptr;
  }
}

Red red(Solver& s) { return Red{s.partials, s.sc}; }

// the tail plan of the current level set (masks included)
Tail tail_plan(const Solver& s, size_t elem, int* lds_bytes) {
  Tail T;
  T.lt = tail_level(s);
  T.L = s.L;
  int off = 0;
  for (int l = 0; l < MAXL; ++l) {
    T.r[l] = nullptr;
    T.z[l] = nullptr;
    T.off[l] = 0;
  }
  for (int l = 0; l < s.L; ++l) {
    T.lv[l] = s.lv[l];
    T.r[l] = s.rl[l];
    T.z[l] = s.zl[l];
    if (l >= T.lt) {
      T.off[l] = off;
      off += 2 * (int)s.nl[l];
    }
  }
  if (lds_bytes) *lds_bytes = off * (int)elem;
  return T;
}

// the finishing launch of a reduction over `nb` blocks
void finish(Solver& s, hipStream_t st, int nb, Fin fin, int guard) {
  hipLaunchKernelGGL(k_finish, dim3(1), dim3(FIN_THREADS), 0, st, red(s), nb, (int)fin, guard);
}

template <class S>
struct Ops {
  // z_ready: level 0's z already holds the fill(0) + first red half-sweep
  static void vcycle(Solver& s, hipStream_t st, int ignore_done, bool z_ready = false);
};

int bottom_lds_bytes(const Solver& s) { return (int)(s.nl[s.L - 1] * sizeof(float)); }

int chunk_blocks(int64_t ncells) { return blocks_for(ncells / 8, BS); }

// LDS column stride: an odd number of 32-bit words between neighbouring columns
int ht_stride(int extent, size_t elem) {
  if (elem == 4) return extent | 1;
  return extent + ((2 - extent % 4) + 4) % 4;   // 2-byte elements: extent = 2 (mod 4)
}
constexpr int HT_MAX_LDS = 128 * 1024;

// Halo-tile plan of level l (between the finest and the coarsest): false when
// the level's extents do not split into 8 x 8 x tk tiles.
bool ht_plan(const Solver& s, int l, size_t elem, HTile& H, int& tk_out, int& blocks, int& lds) {
  if (!s.use_tiles || l < 1 || l >= s.L - 1 || !s.zt[l]) return false;
  const Lv& L = s.lv[l];
  if (L.n0 % HT_T || L.n1 % HT_T) return false;
  // the deepest tile that still gives one workgroup per CU, else the shallowest
  // (a leg's time is one workgroup's latency, which grows with the tile)
  int tk = 0;
  for (int c : {32, 16, 8}) {
    if (L.n2 % c) continue;
    tk = c;
    if ((int64_t)(L.n0 / HT_T) * (L.n1 / HT_T) * (L.n2 / c) >= 256) break;
  }
  if (!tk) return false;
  tk_out = tk;
  H.kp = ht_stride(tk + 2 * HT_HALO, elem);
  H.ckp = ht_stride(tk / 2 + 4, elem);
  H.ntj = L.n1 / HT_T;
  H.ntk = L.n2 / tk;
  H.masked = L.nb != nullptr;
  blocks = (L.n0 / HT_T) * H.ntj * H.ntk;
  lds = (int)(elem * (2 * HT_COLS * H.kp + HT_CB * HT_CB * H.ckp)) + (H.masked ? HT_COLS * H.kp : 0);
  return lds <= HT_MAX_LDS;
}

// one red-black half-sweep of level l
template <class S>
void smooth(Solver& s, hipStream_t st, int l, const S* r, S* z, int phase, int ig) {
  const double* sc = s.sc;
  if (chunked(s.lv[l]))
    hipLaunchKernelGGL(k_smooth8<S>, dim3(chunk_blocks(s.nl[l])), dim3(BS), 0, st, s.lv[l], s.bt, r, z, phase, sc, ig);
  else
    hipLaunchKernelGGL(k_smooth<S>, dim3(blocks_for(s.nl[l] / 2, BS)), dim3(BS), 0, st, s.lv[l], s.bt, r, z, phase, sc,
                       ig);
}

// level 0's last red half-sweep + z.r, then its finish action
template <class S>
void last_dot(Solver& s, hipStream_t st, Fin fin, int ig, int guard) {
  const Lv& L0 = s.lv[0];
  if (chunked(L0)) {
    hipLaunchKernelGGL(k_last_dot8<S>, dim3(chunk_blocks(s.N)), dim3(BS), 0, st, L0, s.bt, (const S*)s.rl[0], s.r,
                       (S*)s.zl[0], red(s), ig);
    if (fin != F_NONE) finish(s, st, chunk_blocks(s.N), fin, guard);
  } else {
    hipLaunchKernelGGL(k_smooth_last_dot<S>, dim3(red_blocks(s.N / 2)), dim3(BS), 0, st, L0, s.bt, (const S*)s.rl[0],
                       s.r, (S*)s.zl[0], red(s), ig);
    if (fin != F_NONE) finish(s, st, red_blocks(s.N / 2), fin, guard);
  }
}

template <class S>
void Ops<S>::vcycle(Solver& s, hipStream_t st, int ig, bool z_ready) {
  const int L = s.L;
  const double* sc = s.sc;
  S* r0 = (S*)s.rl[0];
  S* z0 = (S*)s.zl[0];
  if (z_ready) {
    // k_update_xr8 of the previous iteration wrote z (and, mixed, level 0's r)
  } else if (chunked(s.lv[0]))
    hipLaunchKernelGGL(k_first08<S>, dim3(chunk_blocks(s.N)), dim3(BS), 0, st, s.lv[0], s.bt, s.r, r0, z0, sc, ig,
                       (const float*)s.Ap, s.x);
  else
    hipLaunchKernelGGL(k_first0<S>, dim3(blocks_for(s.N, BS)), dim3(BS), 0, st, s.lv[0], s.bt, s.r, r0, z0, sc, ig);
  const int lt = tail_level(s);
  for (int l = 0; l < (lt ? lt : L - 1); ++l) {
    S* rl = (S*)s.rl[l];
    S* zl = (S*)s.zl[l];
    int64_t half = s.nl[l] / 2;
    (void)half;
    HTile H;
    int tk = 0, hb = 0, hl = 0;
    if (ht_plan(s, l, sizeof(S), H, tk, hb, hl)) {
      auto kern = tk == 32 ? k_down_tile<S, 32> : (tk == 16 ? k_down_tile<S, 16> : k_down_tile<S, 8>);
      hipLaunchKernelGGL(kern, dim3(hb), dim3(HT_THREADS), hl, st, s.lv[l], s.lv[l + 1], s.bt, H, rl, zl,
                         (S*)s.zt[l], (S*)s.rl[l + 1], (S*)s.zl[l + 1], sc, ig);
      continue;
    }
    // remaining pre-smoothing: black, red, black (sim/mgpcg.py:180-182)
    smooth<S>(s, st, l, rl, zl, 1, ig);
    smooth<S>(s, st, l, rl, zl, 0, ig);
    smooth<S>(s, st, l, rl, zl, 1, ig);
    if (chunked(s.lv[l]))
      hipLaunchKernelGGL(k_restrict8<S>, dim3(blocks_for(s.nl[l + 1], BS)), dim3(BS), 0, st, s.lv[l], s.lv[l + 1],
                         s.bt, rl, zl, (S*)s.rl[l + 1], (S*)s.zl[l + 1], sc, ig);
    else
      hipLaunchKernelGGL(k_restrict<S>, dim3(blocks_for(s.nl[l + 1], BS)), dim3(BS), 0, st, s.lv[l], s.lv[l + 1], s.bt,
                         rl, zl, (S*)s.rl[l + 1], (S*)s.zl[l + 1], sc, ig);
  }
  const bool dense = s.mode >= 1 && s.dense_ok;
  if (lt) {
    int lds = 0;
    Tail T = tail_plan(s, sizeof(S), &lds);
    if (!dense) {
      hipLaunchKernelGGL(k_vtail<S>, dim3(1), dim3(TAIL_THREADS), lds, st, T, s.bt, 0, sc, ig);
    } else {
      int l = L - 1;
      hipLaunchKernelGGL(k_vtail<S>, dim3(1), dim3(TAIL_THREADS), lds, st, T, s.bt, 1, sc, ig);
      hipLaunchKernelGGL(k_bottom_dense<S>, dim3(s.nf), dim3(256), 0, st, s.nf, s.fidx, s.Binv, (const S*)s.rl[l],
                         (S*)s.zl[l], s.nl[l], sc, ig);
      hipLaunchKernelGGL(k_vtail<S>, dim3(1), dim3(TAIL_THREADS), lds, st, T, s.bt, 2, sc, ig);
    }
  } else {
    int l = L - 1;
    if (dense) {
      hipLaunchKernelGGL(k_bottom_dense<S>, dim3(s.nf), dim3(256), 0, st, s.nf, s.fidx, s.Binv, (const S*)s.rl[l],
                         (S*)s.zl[l], s.nl[l], sc, ig);
    } else if (s.nl[l] <= BOT_THREADS * BOT_CPT) {
      hipLaunchKernelGGL(k_bottom_rbgs<S>, dim3(1), dim3(BOT_THREADS), bottom_lds_bytes(s), st, s.lv[l], s.bt,
                         (const S*)s.rl[l], (S*)s.zl[l], sc, ig, 20);
    } else {
      for (int it = 0; it < 39; ++it) smooth<S>(s, st, l, (const S*)s.rl[l], (S*)s.zl[l], (it & 1) == 0 ? 1 : 0, ig);
    }
  }
  for (int l = (lt ? lt : L - 1) - 1; l >= 0; --l) {
    S* rl = (S*)s.rl[l];
    S* zl = (S*)s.zl[l];
    int64_t half = s.nl[l] / 2;
    HTile H;
    int tk = 0, hb = 0, hl = 0;
    if (ht_plan(s, l, sizeof(S), H, tk, hb, hl)) {
      auto kern = tk == 32 ? k_up_tile<S, 32> : (tk == 16 ? k_up_tile<S, 16> : k_up_tile<S, 8>);
      hipLaunchKernelGGL(kern, dim3(hb), dim3(HT_THREADS), hl, st, s.lv[l], s.lv[l + 1], s.bt, H, rl,
                         (const S*)s.zt[l], zl, (const S*)s.zl[l + 1], sc, ig);
      continue;
    }
    if (chunked(s.lv[l]))
      hipLaunchKernelGGL(k_prolong8<S>, dim3(chunk_blocks(s.nl[l])), dim3(BS), 0, st, s.lv[l], s.lv[l + 1], s.bt, rl,
                         zl, (const S*)s.zl[l + 1], sc, ig);
    else
      hipLaunchKernelGGL(k_prolong_black<S>, dim3(blocks_for(half, BS)), dim3(BS), 0, st, s.lv[l], s.lv[l + 1], s.bt,
                         rl, zl, (const S*)s.zl[l + 1], sc, ig);
    smooth<S>(s, st, l, rl, zl, 0, ig);
    smooth<S>(s, st, l, rl, zl, 1, ig);
    if (l > 0) smooth<S>(s, st, l, rl, zl, 0, ig);
    // level 0's last red half-sweep is fused with z.r by the caller
  }
}


// Largest coarsest level (fluid cells) for the dense solve: the host Cholesky
// is O(n^3) (about a second at 2048) and reruns whenever the solid mask of the
// coarsest level changes.
constexpr int DENSE_MAX = 2048;

// Build B^-1 for the coarsest level on the host (FP64 Cholesky).  Returns
// false when B is not positive definite (e.g. a disconnected fluid region).
bool build_dense(Solver& s, const std::vector<int8_t>& cm) {
  const Lv& C = s.lv[s.L - 1];
  int n0 = C.n0, n1 = C.n1, n2 = C.n2;
  int nc = n0 * n1 * n2;
  std::vector<int> idx(nc, -1), fl;
  for (int q = 0; q < nc; ++q)
    if (!(s.solid && cm[q] > 0)) {
      idx[q] = (int)fl.size();
      fl.push_back(q);
    }
  int nf = (int)fl.size();
  if (nf == 0 || nf > DENSE_MAX) return false;
  std::vector<double> A((size_t)nf * nf, 0.0);
  bool all_neu = (s.bt == 63);
  auto solid = [&](int i, int j, int k) { return s.solid && cm[(i * n1 + j) * n2 + k] > 0; };
  for (int r = 0; r < nf; ++r) {
    int q = fl[r];
    int k = q % n2, j = (q / n2) % n1, i = q / (n1 * n2);
    int I[3] = {i, j, k}, n[3] = {n0, n1, n2};
    double num = 6;
    for (int a = 0; a < 3; ++a) {
      int lo[3] = {i, j, k}, hi[3] = {i, j, k};
      lo[a]--;
      hi[a]++;
      if (I[a] <= 0 && ((s.bt >> (2 * a)) & 1)) num -= 1;
      else if (I[a] > 0 && solid(lo[0], lo[1], lo[2])) num -= 1;
      if (I[a] >= n[a] - 1 && ((s.bt >> (2 * a + 1)) & 1)) num -= 1;
      else if (I[a] < n[a] - 1 && solid(hi[0], hi[1], hi[2])) num -= 1;
      if (I[a] > 0 && !solid(lo[0], lo[1], lo[2])) A[(size_t)r * nf + idx[(lo[0] * n1 + lo[1]) * n2 + lo[2]]] -= 1;
      if (I[a] < n[a] - 1 && !solid(hi[0], hi[1], hi[2]))
        A[(size_t)r * nf + idx[(hi[0] * n1 + hi[1]) * n2 + hi[2]]] -= 1;
    }
    A[(size_t)r * nf + r] += num;
  }
  if (all_neu) {
    double c = 1.0 / nf;
    for (auto& v : A) v += c;
  }
  // Cholesky A = L L^T (in place, lower)
  for (int j = 0; j < nf; ++j) {
    double d = A[(size_t)j * nf + j];
    for (int k = 0; k < j; ++k) d -= A[(size_t)j * nf + k] * A[(size_t)j * nf + k];
    if (!(d > 1e-10)) return false;
    d = std::sqrt(d);
    A[(size_t)j * nf + j] = d;
    for (int i = j + 1; i < nf; ++i) {
      double v = A[(size_t)i * nf + j];
      for (int k = 0; k < j; ++k) v -= A[(size_t)i * nf + k] * A[(size_t)j * nf + k];
      A[(size_t)i * nf + j] = v / d;
    }
  }
  // Linv (lower), then B^-1 = Linv^T Linv
  std::vector<double> Li((size_t)nf * nf, 0.0);
  for (int c = 0; c < nf; ++c) {
    Li[(size_t)c * nf + c] = 1.0 / A[(size_t)c * nf + c];
    for (int i = c + 1; i < nf; ++i) {
      double v = 0;
      for (int k = c; k < i; ++k) v -= A[(size_t)i * nf + k] * Li[(size_t)k * nf + c];
      Li[(size_t)i * nf + c] = v / A[(size_t)i * nf + i];
    }
  }
  std::vector<double> Bi((size_t)nf * nf, 0.0);
  for (int i = 0; i < nf; ++i)
    for (int j = 0; j <= i; ++j) {
      double v = 0;
      for (int k = i; k < nf; ++k) v += Li[(size_t)k * nf + i] * Li[(size_t)k * nf + j];
      Bi[(size_t)i * nf + j] = v;
      Bi[(size_t)j * nf + i] = v;
    }
  if (s.Binv) (void)hipFree(s.Binv);
  if (s.fidx) (void)hipFree(s.fidx);
  s.Binv = nullptr;
  s.fidx = nullptr;
  if (hipMalloc(&s.Binv, Bi.size() * sizeof(double)) != hipSuccess) return false;
  if (hipMalloc(&s.fidx, fl.size() * sizeof(int)) != hipSuccess) return false;
  if (hipMemcpy(s.Binv, Bi.data(), Bi.size() * sizeof(double), hipMemcpyHostToDevice) != hipSuccess) return false;
  if (hipMemcpy(s.fidx, fl.data(), fl.size() * sizeof(int), hipMemcpyHostToDevice) != hipSuccess) return false;
  s.nf = nf;
  s.dense_mask = cm;
  return true;
}

void free_solver(Solver* s) {
  if (!s) return;
  for (int l = 0; l < s->L; ++l) {
    if (s->rl[l] && s->rl[l] != (void*)s->r) (void)hipFree(s->rl[l]);
    if (s->zl[l]) (void)hipFree(s->zl[l]);
    if (s->zt[l]) (void)hipFree(s->zt[l]);
    if (s->bm[l]) (void)hipFree(s->bm[l]);
    if (s->nbm[l]) (void)hipFree(s->nbm[l]);
  }
  for (void* p : {(void*)s->r, (void*)s->x, (void*)s->p, (void*)s->Ap, (void*)s->partials,
                  (void*)s->sc, (void*)s->Binv, (void*)s->fidx})
    if (p) (void)hipFree(p);
  if (s->sc_host) (void)hipHostFree(s->sc_host);
  if (s->ev_check) (void)hipEventDestroy(s->ev_check);
  drop_graph(*s);
  if (s->cap) (void)hipStreamDestroy(s->cap);
  delete s;
}

int set_scalars(Solver& s, hipStream_t st, int max_iters, int stop_mode, double abs_tol, double rel_tol) {
  double v[S_COUNT];
  for (int i = 0; i < S_COUNT; ++i) v[i] = 0.0;
  v[S_MAXIT] = max_iters < 0 ? 1e18 : (double)max_iters;
  v[S_STOPMODE] = stop_mode;
  v[S_ABSTOL] = abs_tol;
  v[S_RELTOL] = rel_tol;
  v[S_SOLID] = s.solid ? 1.0 : 0.0;
  v[S_NCELL] = (double)s.N;
  v[S_NEU] = (s.bt == 63) ? 1.0 : 0.0;
  // The reference adds an absolute 1e-12 to the denominators of alpha and beta
  // (sim/mgpcg.py:236-239, 270-276); plain mode keeps it.  Once z.r falls to
  // ~1e-12 (a small right-hand side, or a good warm start) that constant
  // damps both coefficients and the residual stalls; the other modes divide
  // by the dot products themselves (0 only when the direction is 0).
  v[S_EPS] = (s.mode == 0) ? 1e-12 : 0.0;
  // pass through pinned host memory (stream-ordered copy, no memset)
  for (int i = 0; i < S_COUNT; ++i) s.sc_host[32 + i] = v[i];
  EDGE_CHECK_HIP(hipMemcpyAsync(s.sc, s.sc_host + 32, S_COUNT * sizeof(double), hipMemcpyHostToDevice, st));
  return EDGE_OK;
}

int prepare_masks(Solver& s, hipStream_t st) {
  if (!s.solid) return EDGE_OK;
  hipLaunchKernelGGL(k_mask0, dim3(blocks_for(s.N, BS)), dim3(BS), 0, st, s.N, s.user_bm, s.bm[0]);
  EDGE_CHECK_LAUNCH();
  for (int l = 1; l < s.L; ++l) {
    hipLaunchKernelGGL(k_downsample, dim3(blocks_for(s.nl[l], BS)), dim3(BS), 0, st, s.lv[l - 1], s.lv[l], s.bm[l]);
    EDGE_CHECK_LAUNCH();
  }
  for (int l = 0; l < s.L; ++l) {
    hipLaunchKernelGGL(k_nbmask, dim3(blocks_for(s.nl[l], BS)), dim3(BS), 0, st, s.lv[l], s.nbm[l]);
    EDGE_CHECK_LAUNCH();
  }
  if (s.mode >= 1) {
    // the dense coarse inverse depends on the coarse mask
    std::vector<int8_t> cm(s.nl[s.L - 1]);
    EDGE_CHECK_HIP(hipMemcpyAsync(cm.data(), s.bm[s.L - 1], cm.size(), hipMemcpyDeviceToHost, st));
    EDGE_CHECK_HIP(hipStreamSynchronize(st));
    // a mask whose build failed (too many coarse cells, or B not positive
    // definite) is remembered too, so an unchanged mask does not rerun the
    // O(n^3) host factorisation on every solve
    if (!s.dense_tried || cm != s.dense_mask) {
      s.dense_ok = build_dense(s, cm);
      s.dense_tried = true;
      s.dense_mask = cm;
      drop_graph(s);  // the bottom's matrix, size and the tail split may have changed
    }
  }
  return EDGE_OK;
}

// The update half of a CG iteration: Ap, alpha, x, r, r.r and the stop test
// (sim/mgpcg.py:234-259).
// The update of r also writes the next V-cycle's fill(0) + first red
// half-sweep (and, mixed, level 0's BF16 r), which saves k_first08's pass
// over r and z, on a chunked level 0 while the dense bottom exists (without it
// the red/black bottom needs the exactly recentred residual of k_first08):
// * no solids, mixed or dense mode: r recentred in the update by its predicted
//   mean (update_recentres; without any recentering the BF16 levels' residual
//   drifts and the 512x256x256 solve stalled at |r|/|b| = 0.85);
// * dense mode with solids: without the recenter (the dense coarsest solve
//   handles the constant; FP32 levels whose level-0 r is the CG residual).
inline bool update_recentres(const Solver& s) { return s.mode >= 1 && !s.solid; }
inline bool z_ready(const Solver& s) {
  return s.dense_ok && chunked(s.lv[0]) && (update_recentres(s) || s.mode == 2);
}

template <class S>
void cg_update(Solver& s, hipStream_t st) {
  Red R = red(s);
  const Lv& L0 = s.lv[0];
  if (chunked(L0)) {
    hipLaunchKernelGGL(k_p_Ap8<S>, dim3(chunk_blocks(s.N)), dim3(BS), 0, st, L0, s.bt, (const S*)s.zl[0], s.p, s.Ap, R);
    finish(s, st, chunk_blocks(s.N), F_ALPHA, 1);
    // the next V-cycle's first level-0 pass here (see z_ready)
    S* zn = z_ready(s) ? (S*)s.zl[0] : nullptr;
    S* rln = (zn && s.rl[0] != (void*)s.r) ? (S*)s.rl[0] : nullptr;
    hipLaunchKernelGGL(k_update_xr8<S>, dim3(chunk_blocks(s.N)), dim3(BS), 0, st, L0, s.bt, (const S*)s.zl[0], s.x,
                       s.r, s.p, s.Ap, zn, rln, (zn && update_recentres(s)) ? 1 : 0, R);
    finish(s, st, chunk_blocks(s.N), F_XR, 1);
  } else {
    hipLaunchKernelGGL(k_p_Ap<S>, dim3(red_blocks(s.N)), dim3(BS), 0, st, L0, s.bt, (const S*)s.zl[0], s.p, s.Ap, R);
    finish(s, st, red_blocks(s.N), F_ALPHA, 1);
    hipLaunchKernelGGL(k_update_xr<S>, dim3(red_blocks(s.N)), dim3(BS), 0, st, L0, (const S*)s.zl[0], s.x, s.r, s.p,
                       s.Ap, R);
    finish(s, st, red_blocks(s.N), F_XR, 1);
  }
}

// A CG iteration, rotated so that it ends with the stop test: z = M r, beta
// and p (sim/mgpcg.py:260-278 of the previous iteration), then the update.  A
// batch that ends at convergence then launches nothing after the test.
template <class S>
void cg_iteration(Solver& s, hipStream_t st) {
  Ops<S>::vcycle(s, st, 0, z_ready(s));
  last_dot<S>(s, st, F_ZR, 0, 1);
  cg_update<S>(s, st);
}

// The first iteration: the initial preconditioning, p = z and z.r
// (sim/mgpcg.py:216-228), then the update.
template <class S>
void cg_first(Solver& s, hipStream_t st) {
  Ops<S>::vcycle(s, st, 1);
  last_dot<S>(s, st, F_ZR_INIT, 1, 0);
  cg_update<S>(s, st);
}

// batch 0: the first iteration (cg_first); b >= 1: b iterations
template <class S>
int capture_batch(Solver& s, int batch) {
  if (s.gexec[batch]) return EDGE_OK;
  if (!s.cap) EDGE_CHECK_HIP(hipStreamCreateWithFlags(&s.cap, hipStreamNonBlocking));
  hipGraph_t g = nullptr;
  EDGE_CHECK_HIP(hipStreamBeginCapture(s.cap, hipStreamCaptureModeThreadLocal));
  if (batch == 0) cg_first<S>(s, s.cap);
  for (int b = 0; b < batch; ++b) cg_iteration<S>(s, s.cap);
  hipError_t le = hipGetLastError();
  hipError_t e = hipStreamEndCapture(s.cap, &g);
  if (le != hipSuccess || e != hipSuccess) {
    if (g) (void)hipGraphDestroy(g);
    set_error("edge_poisson: graph capture failed: %s", hipGetErrorString(le != hipSuccess ? le : e));
    return EDGE_EHIP;
  }
  e = hipGraphInstantiate(&s.gexec[batch], g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  if (e != hipSuccess) {
    s.gexec[batch] = nullptr;
    set_error("edge_poisson: graph instantiate failed: %s", hipGetErrorString(e));
    return EDGE_EHIP;
  }
  return EDGE_OK;
}

// CG iterations to run before the next host check.  check_every >= 1: fixed
// batches (at most MAXB).  0, adaptive: a projection step usually alternates
// two kinds of solve, so the first batch is predicted from the solves two and
// four back: their count when they agree, else one short of the smaller;
// later batches are single iterations.  A host check costs ~30 us, an
// iteration run after convergence ~100 us of launches (its kernels return at
// once, but each launch still takes its slot).
int next_batch(const Solver& s, int check_every, int done_iters) {
  if (check_every >= 1) return std::min(check_every, MAXB);
  if (done_iters > 0 || s.hist[1] <= 0) return 1;
  int b = (s.hist[3] > 0 && s.hist[3] == s.hist[1]) ? s.hist[1] : std::min(s.hist[1], s.hist[0] > 0 ? s.hist[0] : s.hist[1]) - 1;
  return std::max(1, std::min(b, MAXB));
}

template <class S>
int run_cg(Solver& s, hipStream_t st, int max_iters, int check_every, double* stats, const PostSolve* post) {
  int it = 0;
  int done = 0;
  const int maxit = max_iters < 0 ? (1 << 30) : max_iters;
  while (!done) {
    int batch = std::min(next_batch(s, check_every, it), std::max(1, maxit - it));
    // a batch: the first iteration's graph when it = 0, then the captured
    // graphs of sizes in the binary expansion of the remaining count
    if (s.use_graph) {
      int rest = batch;
      if (it == 0) {
        int rc = capture_batch<S>(s, 0);
        if (rc) return rc;
        EDGE_CHECK_HIP(hipGraphLaunch(s.gexec[0], st));
        --rest;
      }
      for (int b = MAXB; b >= 1; b >>= 1)
        if (rest & b) {
          int rc = capture_batch<S>(s, b);
          if (rc) return rc;
          EDGE_CHECK_HIP(hipGraphLaunch(s.gexec[b], st));
        }
    } else {
      for (int b = 0; b < batch; ++b) {
        if (it + b == 0)
          cg_first<S>(s, st);
        else
          cg_iteration<S>(s, st);
        EDGE_CHECK_LAUNCH();
      }
    }
    it += batch;
    EDGE_CHECK_HIP(hipMemcpyAsync(s.sc_host, s.sc, S_COUNT * sizeof(double), hipMemcpyDeviceToHost, st));
    if (post) {
      // the post-solve work goes in behind the flag's copy, guarded by the
      // flag, so the GPU runs it while the host turns the flag around
      if (!s.ev_check) EDGE_CHECK_HIP(hipEventCreateWithFlags(&s.ev_check, hipEventDisableTiming));
      EDGE_CHECK_HIP(hipEventRecord(s.ev_check, st));
      (*post)(st, s.sc);
      EDGE_CHECK_LAUNCH();
      EDGE_CHECK_HIP(hipEventSynchronize(s.ev_check));
    } else {
      EDGE_CHECK_HIP(hipStreamSynchronize(st));
    }
    done = s.sc_host[S_DONE] != 0.0 || it >= maxit;
  }
  // stopped on the host's count with the device flag unset (not expected: the
  // device stops at the same count): the guarded launches skipped, run them now
  if (post && s.sc_host[S_DONE] == 0.0) {
    (*post)(st, nullptr);
    EDGE_CHECK_LAUNCH();
  }
  for (int q = 3; q > 0; --q) s.hist[q] = s.hist[q - 1];
  s.hist[0] = (int)s.sc_host[S_ITERS];
  if (stats) {
    stats[0] = s.sc_host[S_ITERS];
    stats[1] = s.sc_host[S_RTR];
    stats[2] = s.sc_host[S_B2];
    stats[3] = s.sc_host[S_DSUM];
    stats[4] = s.sc_host[S_CORR] != 0.0 ? 1.0 : 0.0;
    stats[5] = s.sc_host[S_DONE] == 1.0 ? 1.0 : 0.0;
    stats[6] = s.sc_host[S_ZTR0];
    stats[7] = s.sc_host[S_R02];
    stats[8] = s.sc_host[S_USEX0];
  }
  return s.sc_host[S_DONE] == 1.0 ? EDGE_OK : EDGE_ENOCONV;
}

int run_cg_any(Solver& s, hipStream_t st, int max_iters, int check_every, double* stats,
               const PostSolve* post = nullptr) {
  if (s.mode == 1) return run_cg<__hip_bfloat16>(s, st, max_iters, check_every, stats, post);
  return run_cg<float>(s, st, max_iters, check_every, stats, post);
}

// Mixed and dense modes without solids: the recenter of r is left to the
// first level-0 pass (k_first08 / k_first0 subtract S_MEAN, the same
// arithmetic as k_recenter), ||b||^2 comes from k_init's sums (F_MEAN_B2).
inline bool fold_recenter(const Solver& s) { return s.mode >= 1 && !s.solid; }

// common start: r prepared by k_init (its finish: F_MEAN_B2 when
// fold_recenter(s), else F_MEAN); mean in S_MEAN.  Recentre (all-Neumann) and
// ||b||^2 unless folded, then the optional warm start: w = b - A x0, kept when
// ||w|| <= ||b||; r = w or x = 0 is applied by the first level-0 pass on a
// chunked level, here otherwise.
int start_solve(Solver& s, hipStream_t st, const float* x0, bool fold) {
  Red R = red(s);
  const Lv& L0 = s.lv[0];
  bool neu = (s.bt == 63);
  if (!fold) {
    if (!neu) hipLaunchKernelGGL(k_set_slot, dim3(1), dim3(1), 0, st, s.sc, (int)S_MEAN, 0.0);
    hipLaunchKernelGGL(k_recenter, dim3(red_blocks(s.N)), dim3(BS), 0, st, L0, s.r, R);
    finish(s, st, red_blocks(s.N), F_B2, 0);
    EDGE_CHECK_LAUNCH();
  }
  if (x0) {
    if (chunked(L0) && !s.solid) {
      hipLaunchKernelGGL(k_residual_x08, dim3(chunk_blocks(s.N)), dim3(BS), 0, st, L0, s.bt, s.r, s.x, s.Ap, R);
      finish(s, st, chunk_blocks(s.N), F_X0, 0);
    } else {
      hipLaunchKernelGGL(k_residual_x0, dim3(red_blocks(s.N)), dim3(BS), 0, st, L0, s.bt, s.r, s.x, s.Ap, R);
      finish(s, st, red_blocks(s.N), F_X0, 0);
    }
    if (!chunked(L0))   // k_first0 does not read S_X0ACT
      hipLaunchKernelGGL(k_apply_x0, dim3(blocks_for(s.N, BS)), dim3(BS), 0, st, s.N, s.r, s.Ap, s.x, s.sc);
    EDGE_CHECK_LAUNCH();
  }
  return EDGE_OK;
}

}  // namespace

extern "C" {

// Create a solver for an nx x ny x nz all-cell Poisson problem.
// mode 0 plain (FP32, reference V-cycle), 1 mixed (BF16 levels, dense FP64 bottom),
// 2 dense (FP32 levels, dense FP64 bottom).
// bt: host int[6] boundary_types (row-major [axis][side], 2 = wall), or NULL for all walls.
int edge_mg_create(int nx, int ny, int nz, int levels, int mode, const int* bt, void** handle) {
  clear_error();
  EDGE_REQUIRE(handle, "edge_mg_create: null handle");
  *handle = nullptr;
  EDGE_REQUIRE(nx >= 2 && ny >= 2 && nz >= 2, "edge_mg_create: bad dims");
  EDGE_REQUIRE(levels >= 2 && levels <= MAXL, "edge_mg_create: levels must be in 2..%d", MAXL);
  EDGE_REQUIRE(mode >= 0 && mode <= 2, "edge_mg_create: mode must be 0 (plain), 1 (mixed) or 2 (dense)");
  for (int l = 0; l < levels; ++l) {
    EDGE_REQUIRE(((nx >> l) << l) == nx && ((ny >> l) << l) == ny && ((nz >> l) << l) == nz,
                 "edge_mg_create: dims must be divisible by 2^(levels-1)");
    EDGE_REQUIRE((nz >> l) % 2 == 0 && (nx >> l) >= 1 && (ny >> l) >= 1,
                 "edge_mg_create: every level needs an even nz (got %d at level %d)", nz >> l, l);
  }
  Solver* s = new Solver();
  s->nx = nx;
  s->ny = ny;
  s->nz = nz;
  s->L = levels;
  s->mode = mode;
  s->bt = 0;
  for (int a = 0; a < 3; ++a)
    for (int sd = 0; sd < 2; ++sd)
      if (!bt || bt[2 * a + sd] == 2) s->bt |= 1 << (2 * a + sd);
  s->N = (int64_t)nx * ny * nz;
  s->elem = mode == 1 ? 2 : 4;
  std::memset(s->rl, 0, sizeof(s->rl));
  std::memset(s->zl, 0, sizeof(s->zl));
  std::memset(s->zt, 0, sizeof(s->zt));
  std::memset(s->bm, 0, sizeof(s->bm));
  std::memset(s->nbm, 0, sizeof(s->nbm));
  s->r = s->x = s->p = s->Ap = nullptr;
  s->partials = nullptr;
  s->sc = nullptr;
  s->sc_host = nullptr;
  auto fail = [&](const char* what) {
    set_error("edge_mg_create: %s failed", what);
    free_solver(s);
    return EDGE_ENOMEM;
  };
  for (int l = 0; l < levels; ++l) {
    s->lv[l] = make_lv(nx >> l, ny >> l, nz >> l);
    s->nl[l] = (int64_t)(nx >> l) * (ny >> l) * (nz >> l);
  }
  if (hipMalloc(&s->r, s->N * 4) != hipSuccess) return fail("hipMalloc r");
  if (hipMalloc(&s->x, s->N * 4) != hipSuccess) return fail("hipMalloc x");
  if (hipMalloc(&s->p, s->N * 4) != hipSuccess) return fail("hipMalloc p");
  if (hipMalloc(&s->Ap, s->N * 4) != hipSuccess) return fail("hipMalloc Ap");
  for (int l = 0; l < levels; ++l) {
    if (l == 0 && mode != 1) {
      s->rl[0] = s->r;
    } else if (hipMalloc(&s->rl[l], s->nl[l] * s->elem) != hipSuccess) {
      return fail("hipMalloc level r");
    }
    if (hipMalloc(&s->zl[l], s->nl[l] * s->elem) != hipSuccess) return fail("hipMalloc level z");
    if (l >= 1 && l < levels - 1 && hipMalloc(&s->zt[l], s->nl[l] * s->elem) != hipSuccess)
      return fail("hipMalloc level z (tiles)");
  }
  if (hipMalloc(&s->partials, (size_t)(red_blocks(s->N) + 1) * 2 * sizeof(double)) != hipSuccess)
    return fail("hipMalloc partials");
  if (hipMalloc(&s->sc, 64 * sizeof(double)) != hipSuccess) return fail("hipMalloc scalars");
  if (hipHostMalloc(&s->sc_host, 128 * sizeof(double)) != hipSuccess) return fail("hipHostMalloc");
  if (hipMemset(s->p, 0, s->N * 4) != hipSuccess) return fail("hipMemset");
  if (s->nl[levels - 1] * 3 * 4 > 160 * 1024 && mode == 0) {
    // bottom falls back to per-sweep launches; fine
  }
  if (mode >= 1) {
    std::vector<int8_t> cm(s->nl[levels - 1], 0);
    s->dense_ok = build_dense(*s, cm);
    s->dense_tried = true;
    s->dense_mask = cm;
  }
  for (int l = 1; l < levels; ++l)
    if (s->nl[l] <= TAIL_MAX_CELLS) {
      s->tail_from = l;
      break;
    }
  if (s->tail_from) {
    int bytes = 0;
    for (int l = s->tail_from; l < levels; ++l) bytes += 2 * (int)s->nl[l] * (int)s->elem;
    // The attribute belongs to the kernel, not the solver: only ever raise it, so
    // that a solver created later with a smaller tail cannot lower it below what
    // an earlier solver launches with.
    static std::mutex mu;
    static int tail_attr[2] = {0, 0};
    std::lock_guard<std::mutex> lock(mu);
    const int ty = mode == 1 ? 1 : 0;
    hipError_t e = hipSuccess;
    if (bytes > tail_attr[ty]) {
      e = ty ? hipFuncSetAttribute((const void*)k_vtail<__hip_bfloat16>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                   bytes)
             : hipFuncSetAttribute((const void*)k_vtail<float>, hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
      if (e == hipSuccess) tail_attr[ty] = bytes;
    }
    if (e != hipSuccess) s->tail_from = 0;  // per-level kernels instead
  }
  {
    const void* kerns[2][6] = {
        {(const void*)k_down_tile<float, 8>, (const void*)k_down_tile<float, 16>, (const void*)k_down_tile<float, 32>,
         (const void*)k_up_tile<float, 8>, (const void*)k_up_tile<float, 16>, (const void*)k_up_tile<float, 32>},
        {(const void*)k_down_tile<__hip_bfloat16, 8>, (const void*)k_down_tile<__hip_bfloat16, 16>,
         (const void*)k_down_tile<__hip_bfloat16, 32>, (const void*)k_up_tile<__hip_bfloat16, 8>,
         (const void*)k_up_tile<__hip_bfloat16, 16>, (const void*)k_up_tile<__hip_bfloat16, 32>}};
    bool ok = true;
    for (const void* k : kerns[mode == 1 ? 1 : 0])
      ok = ok && hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize, HT_MAX_LDS) == hipSuccess;
    if (!ok) s->use_tiles = 0;  // per-level kernels instead
  }
  if (hipDeviceSynchronize() != hipSuccess) return fail("hipDeviceSynchronize");
  *handle = s;
  return EDGE_OK;
}

int edge_mg_destroy(void* handle) {
  clear_error();
  if (handle) {
    (void)hipDeviceSynchronize();
    free_solver((Solver*)handle);
  }
  return EDGE_OK;
}

// Solids (sim/mgpcg_solid.py): bm int32 device mask (nx*ny*nz, >0 solid),
// bvel 3 planes of boundary velocity; both borrowed until the next call.
// NULL bm clears the mask.
int edge_mg_set_solid(void* handle, const int* bm, const float* bvel) {
  clear_error();
  EDGE_REQUIRE(handle, "edge_mg_set_solid: null handle");
  Solver& s = *(Solver*)handle;
  if (bm != s.user_bm || bvel != s.user_bvel || (bm != nullptr) != s.solid)
    drop_graph(s);  // captured launches hold the mask pointers
  if (!bm) {
    s.solid = false;
    for (int l = 0; l < s.L; ++l) s.lv[l].bm = nullptr, s.lv[l].nb = nullptr;
    s.user_bm = nullptr;
    s.user_bvel = nullptr;
    return EDGE_OK;
  }
  EDGE_REQUIRE(bvel, "edge_mg_set_solid: boundary velocity required with a mask");
  for (int l = 0; l < s.L; ++l) {
    if (!s.bm[l]) EDGE_CHECK_HIP(hipMalloc(&s.bm[l], s.nl[l]));
    if (!s.nbm[l]) EDGE_CHECK_HIP(hipMalloc(&s.nbm[l], s.nl[l]));
    s.lv[l].bm = s.bm[l];
    s.lv[l].nb = s.nbm[l];
  }
  s.solid = true;
  s.user_bm = bm;
  s.user_bvel = bvel;
  return EDGE_OK;
}

// MGPCG_3.Poisson / MGPCG_Solid_3.Poisson (sim/mgpcg.py:440-486,
// sim/mgpcg_solid.py:418-427): project u in place; p_out (optional) gets the
// pressure; f* (optional) get -(grad p) (Poisson_with_buffer).
// stop_mode 0: the reference criterion; 1: ||r|| <= rel_tol ||b||.
// x0 (optional): warm start, used only when ||b - A x0|| <= ||b||.
// check_every: CG iterations between host convergence checks (0 = adaptive).
// stats (host double[9]): iters, rTr, ||b||^2, fluid div sum, corrected,
// converged, z0.r0, ||b - A x0||^2, x0 used.
int edge_poisson(void* handle, float* ux, float* uy, float* uz, float* p_out, float* fx, float* fy, float* fz,
                 int max_iters, int stop_mode, double abs_tol, double rel_tol, const float* x0, int check_every,
                 double* stats, void* stream) {
  clear_error();
  EDGE_REQUIRE(handle && ux && uy && uz, "edge_poisson: null argument");
  EDGE_REQUIRE(check_every >= 0, "edge_poisson: check_every must be >= 0 (0 = adaptive)");
  EDGE_REQUIRE(stop_mode == 0 || stop_mode == 1, "edge_poisson: bad stop_mode");
  Solver& s = *(Solver*)handle;
  hipStream_t st = (hipStream_t)stream;
  Faces3 U{{ux, uy, uz}};
  int64_t nf = (int64_t)(s.nx + 1) * s.ny * s.nz + (int64_t)s.nx * (s.ny + 1) * s.nz + (int64_t)s.nx * s.ny * (s.nz + 1);
  int rc = set_scalars(s, st, max_iters, stop_mode, abs_tol, rel_tol);
  if (rc) return rc;
  rc = prepare_masks(s, st);
  if (rc) return rc;
  // apply_bc (sim/mgpcg.py:373-392) first.  Without solids it only zeroes the
  // box's wall faces, and everything after reads them through the divergence
  // (k_init, walls as 0) or overwrites them (k_subtract_grad writes 0 there):
  // the pass is skipped, with the same values.
  const int walls = s.solid ? 0 : s.bt;
  if (s.solid)
    hipLaunchKernelGGL(k_apply_bc, dim3(blocks_for(nf, BS)), dim3(BS), 0, st, make_grid(s.nx, s.ny, s.nz, 1.f), s.bt,
                       U, s.user_bm, s.user_bvel, nullptr);
  if (s.solid) {
    hipLaunchKernelGGL(k_div_sum, dim3(red_blocks(s.N)), dim3(BS), 0, st, make_grid(s.nx, s.ny, s.nz, 1.f), U, (const int8_t*)s.bm[0], red(s));
    finish(s, st, red_blocks(s.N), F_DIVSUM, 0);
  }
  hipLaunchKernelGGL(k_init, dim3(red_blocks(s.N)), dim3(BS), 0, st, make_grid(s.nx, s.ny, s.nz, 1.f), U, nullptr,
                     s.solid ? (const int8_t*)s.bm[0] : nullptr, s.r, s.x, s.p, x0, red(s), walls);
  finish(s, st, red_blocks(s.N), fold_recenter(s) ? F_MEAN_B2 : F_MEAN, 0);
  EDGE_CHECK_LAUNCH();
  rc = start_solve(s, st, x0, fold_recenter(s));
  if (rc) return rc;
  // get_result + subtract_grad_p + apply_bc, and p out.  Enqueued by run_cg
  // after every batch of iterations, before the host waits for the batch's
  // stop flag, each launch guarded by that flag (guard = s.sc): the batch that
  // ends the solve is followed by them at once, the others by launches that
  // return at their first instruction.  guard = NULL: unconditional.
  auto post = [&](hipStream_t pst, const double* guard) {
    auto al16 = [](const void* q) { return ((uintptr_t)q & 15u) == 0; };
    if (s.nz % 4 == 0 && al16(U.f[0]) && al16(U.f[1]) && (!fx || al16(fx)) && (!fy || al16(fy)) &&
        (int64_t)s.nx * s.ny * (s.nz + 4) < (int64_t)1 << 31) {
      const int64_t quads = ((int64_t)(s.nx + 1) * s.ny + (int64_t)s.nx * (s.ny + 1)) * (s.nz / 4) +
                            (int64_t)s.nx * s.ny * (s.nz + 1);
      hipLaunchKernelGGL(k_subtract_grad4, dim3(blocks_for(quads, BS)), dim3(BS), 0, pst,
                         make_grid(s.nx, s.ny, s.nz, 1.f), U, Faces3{{fx, fy, fz}}, (const float*)s.x, walls, guard);
    } else {
      hipLaunchKernelGGL(k_subtract_grad, dim3(blocks_for(nf, BS)), dim3(BS), 0, pst, make_grid(s.nx, s.ny, s.nz, 1.f),
                         U, Faces3{{fx, fy, fz}}, (const float*)s.x, walls, guard);
    }
    if (s.solid)
      hipLaunchKernelGGL(k_apply_bc, dim3(blocks_for(nf, BS)), dim3(BS), 0, pst, make_grid(s.nx, s.ny, s.nz, 1.f),
                         s.bt, U, s.user_bm, s.user_bvel, guard);
    if (p_out)
      hipLaunchKernelGGL(k_copy_done, dim3(blocks_for(s.N / 4 + 1, BS)), dim3(BS), 0, pst, s.N, (const float*)s.x,
                         p_out, guard);
  };
  PostSolve ps = post;
  int cg = run_cg_any(s, st, max_iters, check_every, stats, &ps);
  if (cg != EDGE_OK && cg != EDGE_ENOCONV) return cg;
  EDGE_CHECK_LAUNCH();
  if (cg == EDGE_ENOCONV) set_error("edge_poisson: no convergence within %d iterations", max_iters);
  return cg;
}

// The right-hand side edge_poisson solves for, without solving: u gets the
// boundary conditions (in place, as edge_poisson does first) and b the
// recentred -div u over fluid cells (sim/mgpcg.py:59-75, 169-175, 211-217).
int edge_mg_rhs(void* handle, float* ux, float* uy, float* uz, float* b, void* stream) {
  clear_error();
  EDGE_REQUIRE(handle && ux && uy && uz && b, "edge_mg_rhs: null argument");
  Solver& s = *(Solver*)handle;
  hipStream_t st = (hipStream_t)stream;
  Faces3 U{{ux, uy, uz}};
  int64_t nf = (int64_t)(s.nx + 1) * s.ny * s.nz + (int64_t)s.nx * (s.ny + 1) * s.nz + (int64_t)s.nx * s.ny * (s.nz + 1);
  int rc = set_scalars(s, st, 1, 1, 0.0, 0.0);
  if (rc) return rc;
  rc = prepare_masks(s, st);
  if (rc) return rc;
  const Grid g = make_grid(s.nx, s.ny, s.nz, 1.f);
  hipLaunchKernelGGL(k_apply_bc, dim3(blocks_for(nf, BS)), dim3(BS), 0, st, g, s.bt, U, s.solid ? s.user_bm : nullptr,
                     s.solid ? s.user_bvel : nullptr, nullptr);
  if (s.solid) {
    hipLaunchKernelGGL(k_div_sum, dim3(red_blocks(s.N)), dim3(BS), 0, st, g, U, (const int8_t*)s.bm[0], red(s));
    finish(s, st, red_blocks(s.N), F_DIVSUM, 0);
  }
  hipLaunchKernelGGL(k_init, dim3(red_blocks(s.N)), dim3(BS), 0, st, g, U, nullptr,
                     s.solid ? (const int8_t*)s.bm[0] : nullptr, s.r, s.x, s.p, nullptr, red(s), 0);
  finish(s, st, red_blocks(s.N), F_MEAN, 0);
  EDGE_CHECK_LAUNCH();
  rc = start_solve(s, st, nullptr, false);   // the recentred b is the result
  if (rc) return rc;
  EDGE_CHECK_HIP(hipMemcpyAsync(b, s.r, s.N * 4, hipMemcpyDeviceToDevice, st));
  return EDGE_OK;
}

// Solve A x = b (b recentred first when all walls), same options as edge_poisson.
int edge_mg_solve(void* handle, const float* b, float* x, int max_iters, int stop_mode, double abs_tol,
                  double rel_tol, const float* x0, int check_every, double* stats, void* stream) {
  clear_error();
  EDGE_REQUIRE(handle && b && x, "edge_mg_solve: null argument");
  EDGE_REQUIRE(check_every >= 0, "edge_mg_solve: check_every must be >= 0 (0 = adaptive)");
  Solver& s = *(Solver*)handle;
  hipStream_t st = (hipStream_t)stream;
  int rc = set_scalars(s, st, max_iters, stop_mode, abs_tol, rel_tol);
  if (rc) return rc;
  rc = prepare_masks(s, st);
  if (rc) return rc;
  hipLaunchKernelGGL(k_init, dim3(red_blocks(s.N)), dim3(BS), 0, st, make_grid(s.nx, s.ny, s.nz, 1.f), Faces3{}, b,
                     s.solid ? (const int8_t*)s.bm[0] : nullptr, s.r, s.x, s.p, x0, red(s), 0);
  finish(s, st, red_blocks(s.N), fold_recenter(s) ? F_MEAN_B2 : F_MEAN, 0);
  EDGE_CHECK_LAUNCH();
  rc = start_solve(s, st, x0, fold_recenter(s));
  if (rc) return rc;
  int cg = run_cg_any(s, st, max_iters, check_every, stats);
  if (cg != EDGE_OK && cg != EDGE_ENOCONV) return cg;
  EDGE_CHECK_HIP(hipMemcpyAsync(x, s.x, s.N * 4, hipMemcpyDeviceToDevice, st));
  if (cg == EDGE_ENOCONV) set_error("edge_mg_solve: no convergence within %d iterations", max_iters);
  return cg;
}

// One V-cycle z = M^-1 r (r used as given, not recentred) -- test hook.
int edge_mg_vcycle(void* handle, const float* r, float* z, void* stream) {
  clear_error();
  EDGE_REQUIRE(handle && r && z, "edge_mg_vcycle: null argument");
  Solver& s = *(Solver*)handle;
  hipStream_t st = (hipStream_t)stream;
  int rc = set_scalars(s, st, 1, 0, 0, 0);
  if (rc) return rc;
  rc = prepare_masks(s, st);
  if (rc) return rc;
  EDGE_CHECK_HIP(hipMemcpyAsync(s.r, r, s.N * 4, hipMemcpyDeviceToDevice, st));
  if (s.mode == 1) {
    Ops<__hip_bfloat16>::vcycle(s, st, 1);
    last_dot<__hip_bfloat16>(s, st, F_NONE, 1, 0);
    hipLaunchKernelGGL(k_to_float<__hip_bfloat16>, dim3(blocks_for(s.N, BS)), dim3(BS), 0, st, s.N,
                       (const __hip_bfloat16*)s.zl[0], z);
  } else {
    Ops<float>::vcycle(s, st, 1);
    last_dot<float>(s, st, F_NONE, 1, 0);
    EDGE_CHECK_HIP(hipMemcpyAsync(z, s.zl[0], s.N * 4, hipMemcpyDeviceToDevice, st));
  }
  EDGE_CHECK_LAUNCH();
  return EDGE_OK;
}

// Ap = A p (sim/mgpcg.py:119-124) -- test hook.
int edge_mg_apply_A(void* handle, const float* p, float* Ap, void* stream) {
  clear_error();
  EDGE_REQUIRE(handle && p && Ap, "edge_mg_apply_A: null argument");
  Solver& s = *(Solver*)handle;
  hipStream_t st = (hipStream_t)stream;
  int rc = prepare_masks(s, st);
  if (rc) return rc;
  hipLaunchKernelGGL(k_apply_A, dim3(blocks_for(s.N, BS)), dim3(BS), 0, st, s.lv[0], s.bt, p, Ap);
  EDGE_CHECK_LAUNCH();
  return EDGE_OK;
}

// Replay the CG iterations as a captured HIP graph (1) or launch them one by one (0).
int edge_mg_set_graph(void* handle, int enable) {
  clear_error();
  EDGE_REQUIRE(handle, "edge_mg_set_graph: null handle");
  ((Solver*)handle)->use_graph = enable ? 1 : 0;
  return EDGE_OK;
}

// Halo-tile legs on (default) or off for the levels between the finest and
// the coarsest; off runs every half-sweep as its own launch.
int edge_mg_set_tiles(void* handle, int enable) {
  clear_error();
  EDGE_REQUIRE(handle, "edge_mg_set_tiles: null handle");
  Solver& s = *(Solver*)handle;
  if ((enable ? 1 : 0) != s.use_tiles) drop_graph(s);
  s.use_tiles = enable ? 1 : 0;
  return EDGE_OK;
}

// 1 when the mixed solver's dense coarsest solve is active (0: red-black bottom).
int edge_mg_dense_bottom(void* handle) {
  if (!handle) return 0;
  return ((Solver*)handle)->dense_ok && ((Solver*)handle)->mode >= 1 ? 1 : 0;
}

}  // extern "C"
