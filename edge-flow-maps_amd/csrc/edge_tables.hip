// Hermite tables and finite-difference derivative builders.
//
// The three Hermite variants of sim/hermite.py are one tricubic Hermite with
// different derivative data:
//   garm-FD (sim/hermite.py:298-456, EDGE):  dT stored, d2T = calc_fd_3_from_grad_grad(dT)
//   full    (sim/hermite.py:459-607, ED8/GE): dT and d2T stored
//   FD      (sim/hermite.py:64-295, ED4):    dT = calc_fd_3_from_grad(T), d2T from that dT
// because value_3 of the garm variant is (sum of the dT central differences)/6
// * dx^2 = d2T * dx^3 with d2T as in sim/hermite.py:1258-1287, and the
// value_2* / value_3* terms of the FD variant are the dT / d2T of
// sim/hermite.py:1180-1255 times dx / dx^3.  The pack kernel writes the table
// that edge::hermite_eval reads.
#include "edge_common.h"

using namespace edge;

namespace {

// clamped fetch of plane c of a 3-plane field (util/taichi_utils.py:61-67)
__device__ __forceinline__ float smp(const float* __restrict__ f, const Grid& g, int c, int i, int j, int k) {
  i = max(0, min(i, g.nx - 1));
  j = max(0, min(j, g.ny - 1));
  k = max(0, min(k, g.nz - 1));
  return f[(int64_t)c * ncell(g) + cell_id(g, i, j, k)];
}

// sim/hermite.py:1258-1287 calc_fd_3_from_grad_grad for one component.
__device__ __forceinline__ float fd_xyz(const float* __restrict__ dTc, const Grid& g, int i, int j, int k) {
  return ((smp(dTc, g, 0, i + 1, j, k) - smp(dTc, g, 0, i - 1, j, k)) +
          (smp(dTc, g, 1, i, j + 1, k) - smp(dTc, g, 1, i, j - 1, k)) +
          (smp(dTc, g, 2, i, j, k + 1) - smp(dTc, g, 2, i, j, k - 1))) /
         6.0f / g.dx;
}

struct MapIn {
  const float* psi;
  const float* T[3];
  const float* dT[3];
  const float* d2T[3];  // full variant only
};

// mode 0: garm (d2T from dT); mode 1: full (d2T given)
__global__ void __launch_bounds__(256) k_pack_table(Grid g, MapIn m, float4* __restrict__ tab, int mode) {
  int64_t N = ncell(g);
  int64_t q = (int64_t)xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
  if (q >= N) return;
  int i, j, k;
  cell_ijk(g, q, i, j, k);
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    float4 lo, hi;
    lo.x = m.psi[c * N + q];
    lo.y = m.T[c][q];
    lo.z = m.T[c][N + q];
    lo.w = m.T[c][2 * N + q];
    hi.x = m.dT[c][q];
    hi.y = m.dT[c][N + q];
    hi.z = m.dT[c][2 * N + q];
    hi.w = (mode == 0) ? fd_xyz(m.dT[c], g, i, j, k) : m.d2T[c][q];
    tab[2 * c * N + q] = lo;          // the low plane of component c
    tab[(2 * c + 1) * N + q] = hi;    // the high plane
  }
}

// The high planes {dT_c, d2T_c} of a garm table whose low planes are written
// elsewhere (by k_march_edge with a new-table output): the words k_pack_table
// mode 0 stores there, from the same dT planes.
__global__ void __launch_bounds__(256) k_pack_hi(Grid g, MapIn m, float4* __restrict__ tab) {
  int64_t N = ncell(g);
  int64_t q = (int64_t)xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
  if (q >= N) return;
  int i, j, k;
  cell_ijk(g, q, i, j, k);
#pragma unroll
  for (int c = 0; c < 3; ++c)
    tab[(2 * c + 1) * N + q] =
        make_float4(m.dT[c][q], m.dT[c][N + q], m.dT[c][2 * N + q], fd_xyz(m.dT[c], g, i, j, k));
}

// sim/hermite.py:1180-1287: dT (want_dT) and d2T by central differences.
// Writes dT_c planes (3 per component) when want_dT, else only d2T.
__global__ void __launch_bounds__(256) k_fd_dT(Grid g, MapIn m, float* __restrict__ dTx, float* __restrict__ dTy,
                                                 float* __restrict__ dTz) {
  int64_t N = ncell(g);
  int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= N) return;
  int i, j, k;
  cell_ijk(g, q, i, j, k);
  float* out[3] = {dTx, dTy, dTz};
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    const float* Tc = m.T[c];
    out[c][q] = ((smp(Tc, g, 1, i, j, k + 1) - smp(Tc, g, 1, i, j, k - 1)) +
                 (smp(Tc, g, 2, i, j + 1, k) - smp(Tc, g, 2, i, j - 1, k))) / 4.0f / g.dx;
    out[c][N + q] = ((smp(Tc, g, 0, i, j, k + 1) - smp(Tc, g, 0, i, j, k - 1)) +
                     (smp(Tc, g, 2, i + 1, j, k) - smp(Tc, g, 2, i - 1, j, k))) / 4.0f / g.dx;
    out[c][2 * N + q] = ((smp(Tc, g, 0, i, j + 1, k) - smp(Tc, g, 0, i, j - 1, k)) +
                         (smp(Tc, g, 1, i + 1, j, k) - smp(Tc, g, 1, i - 1, j, k))) / 4.0f / g.dx;
  }
}

__global__ void __launch_bounds__(256) k_fd_d2T(Grid g, MapIn m, float* __restrict__ d2Tx, float* __restrict__ d2Ty,
                                                  float* __restrict__ d2Tz) {
  int64_t N = ncell(g);
  int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= N) return;
  int i, j, k;
  cell_ijk(g, q, i, j, k);
  float* out[3] = {d2Tx, d2Ty, d2Tz};
#pragma unroll
  for (int c = 0; c < 3; ++c) out[c][q] = fd_xyz(m.dT[c], g, i, j, k);
}

}  // namespace

// ---------------------------------------------------------------------------
// C-ABI
// ---------------------------------------------------------------------------
extern "C" {

// Build the Hermite table of a map.  kind 0: garm (psi, T, dT; d2T derived),
// kind 1: full (psi, T, dT, d2T).  tab: 3*N*8 floats (16-byte aligned).
int edge_pack_hermite_table(int kind, const float* psi, const float* Tx, const float* Ty, const float* Tz,
                            const float* dTx, const float* dTy, const float* dTz, const float* d2Tx,
                            const float* d2Ty, const float* d2Tz, float* tab, int nx, int ny, int nz, float dx,
                            void* stream) {
  clear_error();
  EDGE_REQUIRE(dims_ok(nx, ny, nz), "edge_pack_hermite_table: bad dims %d %d %d", nx, ny, nz);
  EDGE_REQUIRE(psi && Tx && Ty && Tz && dTx && dTy && dTz && tab, "edge_pack_hermite_table: null pointer");
  EDGE_REQUIRE(kind == 0 || (kind == 1 && d2Tx && d2Ty && d2Tz), "edge_pack_hermite_table: bad kind %d", kind);
  EDGE_REQUIRE(((uintptr_t)tab & 15) == 0, "edge_pack_hermite_table: table not 16-byte aligned");
  Grid g = make_grid(nx, ny, nz, dx);
  MapIn m{psi, {Tx, Ty, Tz}, {dTx, dTy, dTz}, {d2Tx, d2Ty, d2Tz}};
  int64_t N = ncell(g);
  hipLaunchKernelGGL(k_pack_table, dim3(blocks_for(N, 256)), dim3(256), 0, (hipStream_t)stream, g, m,
                     (float4*)tab, kind);
  EDGE_CHECK_LAUNCH();
  return EDGE_OK;
}

// The high planes of a garm table (dT and calc_fd_3_from_grad_grad of it,
// sim/hermite.py:1258-1287) whose low planes edge_march_edge_table wrote.
int edge_pack_table_hi(const float* dTx, const float* dTy, const float* dTz, float* tab, int nx, int ny, int nz,
                       float dx, void* stream) {
  clear_error();
  EDGE_REQUIRE(dims_ok(nx, ny, nz), "edge_pack_table_hi: bad dims");
  EDGE_REQUIRE(dTx && dTy && dTz && tab, "edge_pack_table_hi: null pointer");
  Grid g = make_grid(nx, ny, nz, dx);
  int64_t N = ncell(g);
  MapIn m{nullptr, {nullptr, nullptr, nullptr}, {dTx, dTy, dTz}, {nullptr, nullptr, nullptr}};
  hipLaunchKernelGGL(k_pack_hi, dim3(blocks_for(N, 256)), dim3(256), 0, (hipStream_t)stream, g, m, (float4*)tab);
  EDGE_CHECK_LAUNCH();
  return EDGE_OK;
}

// sim/hermite.py:1180-1255 calc_fd_3_from_grad (want_dT=1: dT from T, then d2T)
// and sim/hermite.py:1258-1287 calc_fd_3_from_grad_grad (want_dT=0: d2T from dT).
int edge_calc_fd3(int want_dT, const float* Tx, const float* Ty, const float* Tz, float* dTx, float* dTy,
                  float* dTz, float* d2Tx, float* d2Ty, float* d2Tz, int nx, int ny, int nz, float dx,
                  void* stream) {
  clear_error();
  EDGE_REQUIRE(dims_ok(nx, ny, nz), "edge_calc_fd3: bad dims");
  EDGE_REQUIRE(dTx && dTy && dTz && d2Tx && d2Ty && d2Tz, "edge_calc_fd3: null pointer");
  Grid g = make_grid(nx, ny, nz, dx);
  int64_t N = ncell(g);
  MapIn m{nullptr, {Tx, Ty, Tz}, {dTx, dTy, dTz}, {nullptr, nullptr, nullptr}};
  if (want_dT) {
    EDGE_REQUIRE(Tx && Ty && Tz, "edge_calc_fd3: null T");
    hipLaunchKernelGGL(k_fd_dT, dim3(blocks_for(N, 256)), dim3(256), 0, (hipStream_t)stream, g, m, dTx, dTy, dTz);
    EDGE_CHECK_LAUNCH();
  }
  hipLaunchKernelGGL(k_fd_d2T, dim3(blocks_for(N, 256)), dim3(256), 0, (hipStream_t)stream, g, m, d2Tx, d2Ty,
                     d2Tz);
  EDGE_CHECK_LAUNCH();
  return EDGE_OK;
}

}  // extern "C"
