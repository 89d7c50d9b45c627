// Device-side numerics shared by the EDGE kernels (gfx950, wave64).
//
// Interpolation follows util/taichi_utils.py and sim/hermite.py of the
// reference, re-expressed for the GPU:
//  * The quadratic B-spline (util/taichi_utils.py:230-319) is evaluated with
//    its 3 non-zero taps per axis (27 per component) instead of the reference's
//    64-tap loop; the per-tap weights are the reference's N_2 / dN_2 pieces,
//    selected by the fractional position.
//  * Every Hermite variant of sim/hermite.py is the tricubic Hermite of
//    sim/hermite.py:515-607 with per-corner data (psi_c, T_c, dT_c, d2T_c); the
//    garm-FD and FD-unrolled variants differ only in how dT/d2T are derived
//    (see edge_tables.hip), so one evaluator reads a packed per-cell table.
//  * dx is folded into the "g" basis functions, so the table holds the
//    reference's un-scaled derivative data.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define EDGE_DEV __device__ __forceinline__
// Keeps the instruction scheduler from interleaving the work on either side
// (bounds the gathers in flight, and so the registers they hold).
#ifndef EDGE_SCHED_FENCE
#define EDGE_SCHED_FENCE() __builtin_amdgcn_sched_barrier(0)
#endif
// Tuning knobs (compile-time): minimum waves per SIMD requested for the RK4
// march kernels (0 = the compiler's choice).
#ifndef EDGE_WPE_MARCH
#define EDGE_WPE_MARCH 0
#endif
#ifndef EDGE_WPE_ED8
#define EDGE_WPE_ED8 0
#endif
#ifndef EDGE_WPE_FACES
#define EDGE_WPE_FACES 0
#endif
#if EDGE_WPE_MARCH
#define EDGE_MARCH_ATTR __attribute__((amdgpu_waves_per_eu(EDGE_WPE_MARCH, 8)))
#else
#define EDGE_MARCH_ATTR
#endif
#if EDGE_WPE_ED8
#define EDGE_ED8_ATTR __attribute__((amdgpu_waves_per_eu(EDGE_WPE_ED8, 8)))
#else
#define EDGE_ED8_ATTR
#endif
#if EDGE_WPE_FACES
#define EDGE_FACES_ATTR __attribute__((amdgpu_waves_per_eu(EDGE_WPE_FACES, 8)))
#else
#define EDGE_FACES_ATTR
#endif

namespace edge {

// ---------------------------------------------------------------------------
// q / d for 0 <= q < 2^31 by a multiply-high (round-up magic number):
// q / d = (umulhi(q, m) + q) >> l with l = ceil(log2 d), m = 2^32 (2^l - d) / d + 1.
// ---------------------------------------------------------------------------
struct FDiv {
  uint32_t d, m, l;
};
inline FDiv make_fdiv(uint32_t d) {
  uint32_t l = 0;
  while ((1ull << l) < d) ++l;
  uint64_t m = ((1ull << 32) * ((1ull << l) - d)) / d + 1;
  return FDiv{d, (uint32_t)m, l};
}
EDGE_DEV uint32_t fdiv(uint32_t q, const FDiv& f) { return (__umulhi(q, f.m) + q) >> f.l; }

// ---------------------------------------------------------------------------
// Grid parameters passed by value to every kernel.
// ---------------------------------------------------------------------------
struct Grid {
  int nx, ny, nz;   // cells
  float dx;
  float idx;        // 1/dx (f32): positions enter the index space as p * idx, as the
                    // reference's fast-math `p / dx` compiles (no IEEE division per lookup)
  FDiv cz, cy;          // division by nz, ny (cell index decode)
  FDiv fz[3], fy[3];    // division by the k and j extents of face grid a
  uint32_t nf[3];       // face counts
  // Clamp upper bounds dim-2-1e-6 rounded to f32 (util/taichi_utils.py:262-264):
  // hi_c[a] for a cell-sized axis, hi_f[a] for the staggered (+1) axis.
  float hi_c[3], hi_f[3];
  // interp_1 clamp dim-1-1e-6 for cell fields (util/taichi_utils.py:179-181).
  float hi1_c[3];
};

__host__ __device__ __forceinline__ int64_t cell_id(const Grid& g, int i, int j, int k) {
  return ((int64_t)i * g.ny + j) * g.nz + k;
}
__host__ __device__ __forceinline__ int64_t ncell(const Grid& g) { return (int64_t)g.nx * g.ny * g.nz; }

// Face grid `a` has dims n + e_a.
EDGE_DEV void face_dims(const Grid& g, int a, int d[3]) {
  d[0] = g.nx + (a == 0);
  d[1] = g.ny + (a == 1);
  d[2] = g.nz + (a == 2);
}

// (i, j, k) of cell q (q < 2^31).
EDGE_DEV void cell_ijk(const Grid& g, int64_t q64, int& i, int& j, int& k) {
  uint32_t q = (uint32_t)q64;
  uint32_t row = fdiv(q, g.cz);
  k = (int)(q - row * (uint32_t)g.nz);
  uint32_t ii = fdiv(row, g.cy);
  j = (int)(row - ii * (uint32_t)g.ny);
  i = (int)ii;
}

// (i, j, k) of element t of face grid a.
EDGE_DEV void face_ijk(const Grid& g, int a, uint32_t t, int& i, int& j, int& k) {
  uint32_t row = fdiv(t, g.fz[a]);
  k = (int)(t - row * g.fz[a].d);
  uint32_t ii = fdiv(row, g.fy[a]);
  j = (int)(row - ii * g.fy[a].d);
  i = (int)ii;
}

// Thread t over the concatenated face grids x, y, z: axis a, index l in that
// grid, its size nf and (i, j, k).  False past the end.
EDGE_DEV bool face_split(const Grid& g, int64_t t64, int& a, int64_t& l, int64_t& nf, int& i, int& j, int& k) {
  uint32_t t = (uint32_t)t64;
  if (t64 >= (int64_t)g.nf[0] + g.nf[1] + g.nf[2]) return false;
  if (t < g.nf[0]) { a = 0; }
  else if (t < g.nf[0] + g.nf[1]) { a = 1; t -= g.nf[0]; }
  else { a = 2; t -= g.nf[0] + g.nf[1]; }
  l = t;
  nf = g.nf[a];
  face_ijk(g, a, t, i, j, k);
  return true;
}

// ---------------------------------------------------------------------------
// Quadratic B-spline, 3 taps per axis.
// ---------------------------------------------------------------------------
struct Axis3 {
  int i0;        // first tap (>= 0); taps i0, i0+1, i0+2 clamped to dmax
  float w[3];    // N_2 at the taps
  float dw[3];   // dN_2 at the taps
};

// u = p/dx - BL (already shifted); hi = dim-2-eps (>= 1: every grid has >= 3
// cells along an axis).
// The reference's 64-tap loop (util/taichi_utils.py:304-317) visits taps
// iu-1..iu+2 where iu = floor(clamped u); the weight of iu-1 vanishes when the
// fraction is >= 0.5 and that of iu+2 when it is <= 0.5.  The three taps with
// weight are therefore c-1, c, c+1 around the nearest node c = rint(s), and
// with t = s - c in [-0.5, 0.5] the weights are 0.5 (0.5 - t)^2, 0.75 - t^2,
// 0.5 (0.5 + t)^2 and their derivatives t - 0.5, -2t, t + 0.5.  Every step is
// exact in FP32 (s - c, 0.5 -+ t), so the weights carry the bits of the
// floor form 0.5 (1.5 - x)^2 ... with x = s - (c - 1); at a tie (t = +-0.5,
// where rint rounds to even) the taps shift by one and the tap dropped or
// added has weight 0.  The clamp is one v_med3 (the same as max(1, min(u, hi))
// for hi >= 1).
EDGE_DEV Axis3 bspline_axis(float u, float hi) {
  const float s = __builtin_amdgcn_fmed3f(u, 1.0f, hi);
  const float c = __builtin_rintf(s);
  const float t = s - c;
  Axis3 A;
  A.i0 = (int)(c - 1.0f);   // c - 1 exact; an integer -1 here splits the merged 3-tap loads
  const float h0 = 0.5f - t, h2 = 0.5f + t;
  A.w[0] = 0.5f * (h0 * h0);
  A.w[1] = 0.75f - t * t;
  A.w[2] = 0.5f * (h2 * h2);
  A.dw[0] = -h0;
  A.dw[1] = -2.0f * t;
  A.dw[2] = h2;
  return A;
}

// Buffer-resource gathers: the field base and size live in scalar registers,
// each lane supplies one 32-bit byte offset, the wave-uniform part of an
// address goes in the scalar offset and small constants in the immediate
// field.  Reads past `bytes` return 0 instead of faulting.
typedef __amdgpu_buffer_rsrc_t Rsrc;
EDGE_DEV Rsrc make_rsrc(const void* p, uint32_t bytes) {
  // dword 3 = 0x00020000: the raw-buffer format word of the gfx9 family (gfx950 included)
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(p), (short)0, (int)bytes, 0x00020000);
}
EDGE_DEV float bload(Rsrc r, uint32_t voff, uint32_t soff) {
  return __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(r, (int)voff, (int)soff, 0));
}
EDGE_DEV float4 bload4(Rsrc r, uint32_t voff, uint32_t soff) {
  auto v = __builtin_amdgcn_raw_buffer_load_b128(r, (int)voff, (int)soff, 0);
  return make_float4(__uint_as_float(v[0]), __uint_as_float(v[1]), __uint_as_float(v[2]), __uint_as_float(v[3]));
}

// Value and gradient (in index units) of the quadratic B-spline of one field
// with strides s0 (axis 0) and s1 (axis 1), in elements; A/B/C: axis weights.
// Every tap read is inside the grid: the position clamp to [1, dim-2-eps]
// puts the first tap in [0, dim-3] (bspline_axis), so the reference's sample()
// clamp (util/taichi_utils.py:61-67) never binds on a tap of non-zero weight.
template <bool GRAD>
EDGE_DEV float bspline_eval(Rsrc f, uint32_t s0, uint32_t s1, const Axis3& A, const Axis3& B, const Axis3& C,
                            float g[3]) {
  // Separable contraction, one axis at a time (k, then j, then i): 93 VALU
  // operations with the gradient instead of 117 for the 27 weight products.
  const uint32_t vb = 4u * ((uint32_t)A.i0 * s0 + (uint32_t)B.i0 * s1 + (uint32_t)C.i0);
  // Every sum is written as explicit fmas, so each term is rounded once
  // whatever the weights' expressions (left to itself the compiler moves the
  // -2 of dw[1] = -2t out of the product and rounds tap * t separately).
  float val = 0.f, gx = 0.f, gy = 0.f, gz = 0.f;
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    float t = 0.f, ty = 0.f, tz = 0.f;
#pragma unroll
    for (int b = 0; b < 3; ++b) {
      const uint32_t so = 4u * (a * s0 + b * s1);
      const float v0 = bload(f, vb, so), v1 = bload(f, vb + 4u, so), v2 = bload(f, vb + 8u, so);
      const float sk = __builtin_fmaf(v2, C.w[2], __builtin_fmaf(v1, C.w[1], v0 * C.w[0]));
      t = b ? __builtin_fmaf(sk, B.w[b], t) : sk * B.w[b];
      if (GRAD) {
        const float sdk = __builtin_fmaf(v2, C.dw[2], __builtin_fmaf(v1, C.dw[1], v0 * C.dw[0]));
        ty = b ? __builtin_fmaf(sk, B.dw[b], ty) : sk * B.dw[b];
        tz = b ? __builtin_fmaf(sdk, B.w[b], tz) : sdk * B.w[b];
      }
    }
    val = a ? __builtin_fmaf(t, A.w[a], val) : t * A.w[a];
    if (GRAD) {
      gx = a ? __builtin_fmaf(t, A.dw[a], gx) : t * A.dw[a];
      gy = a ? __builtin_fmaf(ty, A.w[a], gy) : ty * A.w[a];
      gz = a ? __builtin_fmaf(tz, A.w[a], gz) : tz * A.w[a];
    }
  }
  if (GRAD) {
    g[0] = gx;
    g[1] = gy;
    g[2] = gz;
  }
  return val;
}

// MAC velocity (run_edge.py:193-195): u_x (nx+1,ny,nz), u_y (nx,ny+1,nz), u_z (nx,ny,nz+1).
struct Mac {
  const float* __restrict__ ux;
  const float* __restrict__ uy;
  const float* __restrict__ uz;
};

// util/taichi_utils.py:322-341: interp_u_MAC(_grad).  Component a samples its
// face grid with BL = 0 on axis a and 0.5 elsewhere.  J rows = grad u_a (1/dx units).
template <bool GRAD>
EDGE_DEV void interp_u_mac(const Grid& g, const Mac& U, const float p[3], float out[3], float J[3][3]) {
  float q0 = p[0] * g.idx, q1 = p[1] * g.idx, q2 = p[2] * g.idx;
  // own (BL 0, staggered dims) and other (BL 0.5, cell dims) weights per axis
  Axis3 ow0 = bspline_axis(q0, g.hi_f[0]), ot0 = bspline_axis(q0 - 0.5f, g.hi_c[0]);
  Axis3 ow1 = bspline_axis(q1, g.hi_f[1]), ot1 = bspline_axis(q1 - 0.5f, g.hi_c[1]);
  Axis3 ow2 = bspline_axis(q2, g.hi_f[2]), ot2 = bspline_axis(q2 - 0.5f, g.hi_c[2]);
  float gr[3];
  const uint32_t nx = g.nx, ny = g.ny, nz = g.nz;
  out[0] = bspline_eval<GRAD>(make_rsrc(U.ux, 4u * (nx + 1) * ny * nz), ny * nz, nz, ow0, ot1, ot2, gr);
  if (GRAD) { float inv = g.idx; J[0][0] = gr[0] * inv; J[0][1] = gr[1] * inv; J[0][2] = gr[2] * inv; }
  out[1] = bspline_eval<GRAD>(make_rsrc(U.uy, 4u * nx * (ny + 1) * nz), (ny + 1) * nz, nz, ot0, ow1, ot2, gr);
  if (GRAD) { float inv = g.idx; J[1][0] = gr[0] * inv; J[1][1] = gr[1] * inv; J[1][2] = gr[2] * inv; }
  out[2] = bspline_eval<GRAD>(make_rsrc(U.uz, 4u * nx * ny * (nz + 1)), ny * (nz + 1), nz + 1, ot0, ot1, ow2, gr);
  if (GRAD) { float inv = g.idx; J[2][0] = gr[0] * inv; J[2][1] = gr[1] * inv; J[2][2] = gr[2] * inv; }
}

// Component a of interp_u_mac<false> alone (same weights and taps).
EDGE_DEV float interp_u_comp(const Grid& g, const Mac& U, const float p[3], int a) {
  const float q0 = p[0] * g.idx, q1 = p[1] * g.idx, q2 = p[2] * g.idx;
  const uint32_t nx = g.nx, ny = g.ny, nz = g.nz;
  if (a == 0)
    return bspline_eval<false>(make_rsrc(U.ux, 4u * (nx + 1) * ny * nz), ny * nz, nz, bspline_axis(q0, g.hi_f[0]),
                               bspline_axis(q1 - 0.5f, g.hi_c[1]), bspline_axis(q2 - 0.5f, g.hi_c[2]), nullptr);
  if (a == 1)
    return bspline_eval<false>(make_rsrc(U.uy, 4u * nx * (ny + 1) * nz), (ny + 1) * nz, nz,
                               bspline_axis(q0 - 0.5f, g.hi_c[0]), bspline_axis(q1, g.hi_f[1]),
                               bspline_axis(q2 - 0.5f, g.hi_c[2]), nullptr);
  return bspline_eval<false>(make_rsrc(U.uz, 4u * nx * ny * (nz + 1)), ny * (nz + 1), nz + 1,
                             bspline_axis(q0 - 0.5f, g.hi_c[0]), bspline_axis(q1 - 0.5f, g.hi_c[1]),
                             bspline_axis(q2, g.hi_f[2]), nullptr);
}

// Component A's grid of the MAC velocity: buffer, strides (elements) of axes 0
// and 1, and the clamp limit of each axis for that component's positions
// (its own axis is staggered).
template <int A>
struct MacComp {
  Rsrc f;
  uint32_t s0, s1;
  EDGE_DEV MacComp(const Grid& g, const Mac& U) {
    const uint32_t nx = g.nx, ny = g.ny, nz = g.nz;
    if (A == 0) { f = make_rsrc(U.ux, 4u * (nx + 1) * ny * nz); s0 = ny * nz; s1 = nz; }
    if (A == 1) { f = make_rsrc(U.uy, 4u * nx * (ny + 1) * nz); s0 = (ny + 1) * nz; s1 = nz; }
    if (A == 2) { f = make_rsrc(U.uz, 4u * nx * ny * (nz + 1)); s0 = ny * (nz + 1); s1 = nz + 1; }
  }
};

// The 27 taps of component A around face (i, j, k) of its own grid, held in
// registers.  For component A the face sits on a node in every axis (index
// position (i, j, k) after the BL shift), so every B-spline evaluation within
// half a cell of the face, after the same clamp, has its first taps at
// w0 = rint(clamp(i, j, k)) - 1: the semi-Lagrangian RK4 stages and foot of
// that face read these taps again and again.
template <int A>
struct OwnTaps {
  float v[9][3];
  int w0[3];
  EDGE_DEV OwnTaps(const Grid& g, const MacComp<A>& M, int i, int j, int k) {
    const int idx[3] = {i, j, k};
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      const float hi = d == A ? g.hi_f[d] : g.hi_c[d];
      w0[d] = (int)(__builtin_rintf(__builtin_amdgcn_fmed3f((float)idx[d], 1.0f, hi)) - 1.0f);
    }
    const uint32_t vb = 4u * ((uint32_t)w0[0] * M.s0 + (uint32_t)w0[1] * M.s1 + (uint32_t)w0[2]);
#pragma unroll
    for (int r = 0; r < 9; ++r) {
      const uint32_t so = 4u * ((r / 3) * M.s0 + (r % 3) * M.s1);
      v[r][0] = bload(M.f, vb, so);
      v[r][1] = bload(M.f, vb + 4u, so);
      v[r][2] = bload(M.f, vb + 8u, so);
    }
  }
};

// bspline_eval<false>'s contraction of taps already in registers (the same
// operations in the same order, so the same bits).
EDGE_DEV float bspline_contract(const float v[9][3], const Axis3& A, const Axis3& B, const Axis3& C) {
  float val = 0.f;
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    float t = 0.f;
#pragma unroll
    for (int b = 0; b < 3; ++b) {
      const float v0 = v[3 * a + b][0], v1 = v[3 * a + b][1], v2 = v[3 * a + b][2];
      const float sk = __builtin_fmaf(v2, C.w[2], __builtin_fmaf(v1, C.w[1], v0 * C.w[0]));
      t = b ? __builtin_fmaf(sk, B.w[b], t) : sk * B.w[b];
    }
    val = a ? __builtin_fmaf(t, A.w[a], val) : t * A.w[a];
  }
  return val;
}

// Component A at p from the held taps when p's taps are those, else gathered.
template <int A>
EDGE_DEV float own_eval(const MacComp<A>& M, const OwnTaps<A>& W, const Axis3& X, const Axis3& Y, const Axis3& Z) {
  if (X.i0 == W.w0[0] && Y.i0 == W.w0[1] && Z.i0 == W.w0[2]) return bspline_contract(W.v, X, Y, Z);
  return bspline_eval<false>(M.f, M.s0, M.s1, X, Y, Z, nullptr);
}

// 27 taps of one component at the first taps i0, held in registers.
struct Taps27 {
  float v[9][3];
  int i0[3];
};
EDGE_DEV void load_taps(Rsrc f, uint32_t s0, uint32_t s1, const Axis3& A, const Axis3& B, const Axis3& C,
                        Taps27& T) {
  T.i0[0] = A.i0;
  T.i0[1] = B.i0;
  T.i0[2] = C.i0;
  const uint32_t vb = 4u * ((uint32_t)A.i0 * s0 + (uint32_t)B.i0 * s1 + (uint32_t)C.i0);
#pragma unroll
  for (int r = 0; r < 9; ++r) {
    const uint32_t so = 4u * ((r / 3) * s0 + (r % 3) * s1);
    T.v[r][0] = bload(f, vb, so);
    T.v[r][1] = bload(f, vb + 4u, so);
    T.v[r][2] = bload(f, vb + 8u, so);
  }
}
// The value at axes A, B, C from held taps when they are the point's taps, else gathered.
EDGE_DEV float held_eval(Rsrc f, uint32_t s0, uint32_t s1, const Taps27& T, const Axis3& A, const Axis3& B,
                         const Axis3& C) {
  if (A.i0 == T.i0[0] && B.i0 == T.i0[1] && C.i0 == T.i0[2]) return bspline_contract(T.v, A, B, C);
  return bspline_eval<false>(f, s0, s1, A, B, C, nullptr);
}

// interp_u_mac<false> with component A's taps held (the same values).  HOLD
// for the two other components: 0 gathered, 1 their taps loaded into T1, T2
// (and used), 2 read from T1, T2 where they are the point's taps.
template <int A, int HOLD = 0>
EDGE_DEV void interp_u_mac_own(const Grid& g, const Mac& U, const float p[3], const MacComp<A>& M,
                               const OwnTaps<A>& W, float out[3], Taps27* T1 = nullptr, Taps27* T2 = nullptr) {
  float q0 = p[0] * g.idx, q1 = p[1] * g.idx, q2 = p[2] * g.idx;
  Axis3 ow0 = bspline_axis(q0, g.hi_f[0]), ot0 = bspline_axis(q0 - 0.5f, g.hi_c[0]);
  Axis3 ow1 = bspline_axis(q1, g.hi_f[1]), ot1 = bspline_axis(q1 - 0.5f, g.hi_c[1]);
  Axis3 ow2 = bspline_axis(q2, g.hi_f[2]), ot2 = bspline_axis(q2 - 0.5f, g.hi_c[2]);
  const uint32_t nx = g.nx, ny = g.ny, nz = g.nz;
  const Axis3* ax[3][3] = {{&ow0, &ot1, &ot2}, {&ot0, &ow1, &ot2}, {&ot0, &ot1, &ow2}};
  const Rsrc f[3] = {make_rsrc(U.ux, 4u * (nx + 1) * ny * nz), make_rsrc(U.uy, 4u * nx * (ny + 1) * nz),
                     make_rsrc(U.uz, 4u * nx * ny * (nz + 1))};
  const uint32_t s0[3] = {ny * nz, (ny + 1) * nz, ny * (nz + 1)}, s1[3] = {nz, nz, nz + 1};
#pragma unroll
  for (int m = 0; m < 3; ++m) {
    const Axis3 &X = *ax[m][0], &Y = *ax[m][1], &Z = *ax[m][2];
    if (m == A) {
      out[m] = own_eval<A>(M, W, X, Y, Z);
    } else {
      Taps27* T = m == (A + 1) % 3 ? T1 : T2;
      if (HOLD == 1) {
        load_taps(f[m], s0[m], s1[m], X, Y, Z, *T);
        out[m] = bspline_contract(T->v, X, Y, Z);
      } else if (HOLD == 2) {
        out[m] = held_eval(f[m], s0[m], s1[m], *T, X, Y, Z);
      } else {
        out[m] = bspline_eval<false>(f[m], s0[m], s1[m], X, Y, Z, nullptr);
      }
    }
  }
}

// interp_u_comp(g, U, p, A) with component A's taps held.
template <int A>
EDGE_DEV float interp_u_comp_own(const Grid& g, const float p[3], const MacComp<A>& M, const OwnTaps<A>& W) {
  const float q0 = p[0] * g.idx, q1 = p[1] * g.idx, q2 = p[2] * g.idx;
  if (A == 0)
    return own_eval<A>(M, W, bspline_axis(q0, g.hi_f[0]), bspline_axis(q1 - 0.5f, g.hi_c[1]),
                       bspline_axis(q2 - 0.5f, g.hi_c[2]));
  if (A == 1)
    return own_eval<A>(M, W, bspline_axis(q0 - 0.5f, g.hi_c[0]), bspline_axis(q1, g.hi_f[1]),
                       bspline_axis(q2 - 0.5f, g.hi_c[2]));
  return own_eval<A>(M, W, bspline_axis(q0 - 0.5f, g.hi_c[0]), bspline_axis(q1 - 0.5f, g.hi_c[1]),
                     bspline_axis(q2, g.hi_f[2]));
}

// ---------------------------------------------------------------------------
// Tricubic Hermite over a packed table.
//
// Table layout: per component c two planes of N 16-byte cells, the low plane
// {psi_c, T_c0, T_c1, T_c2} at float4 index (2c)*N + q and the high plane
// {dT_c0 (yz), dT_c1 (xz), dT_c2 (xy), d2T_c (xyz)} at (2c+1)*N + q
// (sim/hermite.py:610-620 "Hermite" row layout; dT as in sim/hermite.py:500-510).
// An evaluation that needs only the low coefficients (a face centre) reads
// only the low plane's cache lines.
// ---------------------------------------------------------------------------
struct HAxis {
  int i0;          // corner base (cells i0, i0+1)
  float F[2], G[2], dF[2], dG[2];  // basis at the two corners; G, dG pre-scaled by dx
};

// sim/hermite.py:4-41 for x = frac (corner 0) and x = frac - 1 (corner 1);
// the "g" functions carry the dx of value_1..3 (sim/hermite.py:500-510).
EDGE_DEV HAxis hermite_axis(float u, float hi, float dx) {
  float s = __builtin_amdgcn_fmed3f(u, 1.0f, hi);   // max(1, min(u, hi)), hi >= 1
  float iu = floorf(s);
  HAxis H;
  H.i0 = (int)iu;
  float a0 = s - iu;       // x >= 0 at corner 0
  float a1 = 1.0f - a0;    // |x| at corner 1 (x < 0 there)
  float a02 = a0 * a0, a12 = a1 * a1;
  H.F[0] = 2.0f * (a02 * a0) - 3.0f * a02 + 1.0f;
  H.F[1] = 2.0f * (a12 * a1) - 3.0f * a12 + 1.0f;
  H.dF[0] = 6.0f * a02 - 6.0f * a0;
  H.dF[1] = -(6.0f * a12 - 6.0f * a1);
  H.G[0] = ((a02 * a0) - 2.0f * a02 + a0) * dx;
  H.G[1] = -((a12 * a1) - 2.0f * a12 + a1) * dx;
  H.dG[0] = (3.0f * a02 - 4.0f * a0 + 1.0f) * dx;
  H.dG[1] = (3.0f * a12 - 4.0f * a1 + 1.0f) * dx;
  return H;
}

struct HPoint {
  HAxis X, Y, Z;
};

EDGE_DEV HPoint hermite_point(const Grid& g, const float p[3]) {
  HPoint P;
  P.X = hermite_axis(p[0] * g.idx - 0.5f, g.hi_c[0], g.dx);
  P.Y = hermite_axis(p[1] * g.idx - 0.5f, g.hi_c[1], g.dx);
  P.Z = hermite_axis(p[2] * g.idx - 0.5f, g.hi_c[2], g.dx);
  return P;
}

// Value (and gradient in index units, i.e. before the final 1/dx) of one
// component from its table.  Corner indices are always inside the grid: the
// clamp keeps the base in [1, dim-2].
// The 8 corners x 8 coefficients form one separable 4 x 4 x 4 tensor
// C[(i,tx)][(j,ty)][(k,tz)] (corner i/j/k, basis F or G per axis), contracted
// one axis at a time over all corners: z (value and derivative weights), then
// y, then x: 84 operations for the value, 192 with the gradient (288 corner by
// corner).
template <bool GRAD>
EDGE_DEV float hermite_eval(const Grid& g, Rsrc tab, const HPoint& P, float gr[3]) {
  const uint32_t s1 = (uint32_t)g.nz, s0 = (uint32_t)g.ny * s1;
  const uint32_t vb = 16u * ((uint32_t)P.X.i0 * s0 + (uint32_t)P.Y.i0 * s1 + (uint32_t)P.Z.i0);
  const uint32_t hib = 16u * (uint32_t)ncell(g);   // the high plane
  // per-axis weights of the combined index (corner, basis): [2 * corner + basis]
  const float Zw[4] = {P.Z.F[0], P.Z.G[0], P.Z.F[1], P.Z.G[1]};
  const float Yw[4] = {P.Y.F[0], P.Y.G[0], P.Y.F[1], P.Y.G[1]};
  const float Xw[4] = {P.X.F[0], P.X.G[0], P.X.F[1], P.X.G[1]};
  const float dZw[4] = {P.Z.dF[0], P.Z.dG[0], P.Z.dF[1], P.Z.dG[1]};
  const float dYw[4] = {P.Y.dF[0], P.Y.dG[0], P.Y.dF[1], P.Y.dG[1]};
  const float dXw[4] = {P.X.dF[0], P.X.dG[0], P.X.dF[1], P.X.dG[1]};
  float val = 0.f, gx = 0.f, gy = 0.f, gz = 0.f;
#pragma unroll
  for (int i = 0; i < 2; ++i) {
    // the four corners (i, j, k) of this x-layer: c[j][k][tx][ty][tz]
    float c[2][2][2][2][2];
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      const uint32_t so = 16u * (i * s0 + j * s1);
#pragma unroll
      for (int k = 0; k < 2; ++k) {
        const float4 lo = bload4(tab, vb + 16u * k, so), hi = bload4(tab, vb + 16u * k, so + hib);
        c[j][k][0][0][0] = lo.x;
        c[j][k][1][0][0] = lo.y;
        c[j][k][0][1][0] = lo.z;
        c[j][k][0][0][1] = lo.w;
        c[j][k][0][1][1] = hi.x;
        c[j][k][1][0][1] = hi.y;
        c[j][k][1][1][0] = hi.z;
        c[j][k][1][1][1] = hi.w;
      }
    }
#pragma unroll
    for (int tx = 0; tx < 2; ++tx) {
      float t = 0.f, ty = 0.f, tz = 0.f;
#pragma unroll
      for (int j = 0; j < 2; ++j) {
#pragma unroll
        for (int tyb = 0; tyb < 2; ++tyb) {
          // contract z over K = (k, tz)
          float sz = 0.f, dsz = 0.f;
#pragma unroll
          for (int k = 0; k < 2; ++k) {
#pragma unroll
            for (int tzb = 0; tzb < 2; ++tzb) {
              const float cv = c[j][k][tx][tyb][tzb];
              sz += cv * Zw[2 * k + tzb];
              if (GRAD) dsz += cv * dZw[2 * k + tzb];
            }
          }
          const int J = 2 * j + tyb;
          t += sz * Yw[J];
          if (GRAD) {
            ty += sz * dYw[J];
            tz += dsz * Yw[J];
          }
        }
      }
      const int I = 2 * i + tx;
      val += t * Xw[I];
      if (GRAD) {
        gx += t * dXw[I];
        gy += ty * Xw[I];
        gz += tz * Xw[I];
      }
    }
  }
  if (GRAD) {
    gr[0] = gx;
    gr[1] = gy;
    gr[2] = gz;
  }
  return val;
}

// Value and the derivative along axis D only (index units) of one component.
// The two other axes are contracted first and D last, so the derivative costs
// 2 more operations per corner (16 per corner instead of 36 for the full
// gradient).  Coefficient c[tx][ty][tz] (t = 0: F basis, 1: G basis) of a
// corner is the table row of hermite_eval.
// A corner whose weights on axis A1 or A2 are both exactly zero adds exact
// zeros and is skipped (neither read nor multiplied): at a face centre
// (fraction 0 on the two axes across the face) only the 2 corners along D
// remain, and the value is that of the 8-corner loop.
template <int D>
EDGE_DEV float hermite_eval_dir(const Grid& g, Rsrc tab, const HPoint& P, float& dval) {
  constexpr int A1 = D == 2 ? 0 : 2;   // contracted first
  constexpr int A2 = D == 1 ? 0 : 1;   // contracted second
  const uint32_t s1 = (uint32_t)g.nz, s0 = (uint32_t)g.ny * s1;
  const uint32_t vb = 16u * ((uint32_t)P.X.i0 * s0 + (uint32_t)P.Y.i0 * s1 + (uint32_t)P.Z.i0);
  const uint32_t hib = 16u * (uint32_t)ncell(g);   // the high plane
  const HAxis* ax[3] = {&P.X, &P.Y, &P.Z};
  const bool z1 = ax[A1]->F[1] == 0.f && ax[A1]->G[1] == 0.f;  // corner 1 of A1 weighs nothing
  const bool z2 = ax[A2]->F[1] == 0.f && ax[A2]->G[1] == 0.f;
  float val = 0.f, dv = 0.f;
#pragma unroll
  for (int i = 0; i < 2; ++i) {
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      const uint32_t so = 16u * (i * s0 + j * s1);
#pragma unroll
      for (int k = 0; k < 2; ++k) {
        const int cr[3] = {i, j, k};
        if ((cr[A1] == 1 && z1) || (cr[A2] == 1 && z2)) continue;
        const float w1f = ax[A1]->F[cr[A1]], w1g = ax[A1]->G[cr[A1]];
        const float w2f = ax[A2]->F[cr[A2]], w2g = ax[A2]->G[cr[A2]];
        // the upper four coefficients (mixed derivatives) all carry a derivative
        // along A1 or A2: with both G weights exactly zero (a face centre: the
        // point sits on the corner's A1 and A2 coordinates) they add +0, and
        // their 16 bytes are not fetched (the same value, half the bytes)
        const bool hi_needed = w1g != 0.f || w2g != 0.f;
        const float4 lo = bload4(tab, vb + 16u * k, so);
        float4 hi = make_float4(0.f, 0.f, 0.f, 0.f);
        if (hi_needed) hi = bload4(tab, vb + 16u * k, so + hib);
        float c[2][2][2];
        c[0][0][0] = lo.x;
        c[1][0][0] = lo.y;
        c[0][1][0] = lo.z;
        c[0][0][1] = lo.w;
        c[0][1][1] = hi.x;
        c[1][0][1] = hi.y;
        c[1][1][0] = hi.z;
        c[1][1][1] = hi.w;
        auto coef = [&](int t1, int t2, int td) {
          int t[3];
          t[A1] = t1;
          t[A2] = t2;
          t[D] = td;
          return c[t[0]][t[1]][t[2]];
        };
        float u[2];
#pragma unroll
        for (int td = 0; td < 2; ++td) {
          const float s0_ = coef(0, 0, td) * w1f + coef(1, 0, td) * w1g;
          const float s1_ = coef(0, 1, td) * w1f + coef(1, 1, td) * w1g;
          u[td] = s0_ * w2f + s1_ * w2g;
        }
        val += u[0] * ax[D]->F[cr[D]] + u[1] * ax[D]->G[cr[D]];
        dv += u[0] * ax[D]->dF[cr[D]] + u[1] * ax[D]->dG[cr[D]];
      }
    }
  }
  dval = dv;
  return val;
}

// All three components of the map at p and column D of grad psi (1/dx units).
template <int D, bool FENCE = false>
EDGE_DEV void map_eval_dir(const Grid& g, const float4* __restrict__ tab, const float p[3], float v[3],
                           float gcol[3]) {
  HPoint P = hermite_point(g, p);
  int64_t N = ncell(g);
  float inv = g.idx;
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    float d;
    v[c] = hermite_eval_dir<D>(g, make_rsrc(tab + 2 * N * c, (uint32_t)(32 * N)), P, d);
    gcol[c] = d * inv;
    if (FENCE) EDGE_SCHED_FENCE();
  }
}

// All three components of the map at p: value v and rows G (grad psi_c, 1/dx units).
// FENCE: evaluate the components one after another (fewer gathers in flight).
template <bool GRAD, bool FENCE = false>
EDGE_DEV void map_eval(const Grid& g, const float4* __restrict__ tab, const float p[3], float v[3], float G[3][3]) {
  HPoint P = hermite_point(g, p);
  int64_t N = ncell(g);
  float inv = g.idx;
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    float gr[3];
    v[c] = hermite_eval<GRAD>(g, make_rsrc(tab + 2 * N * c, (uint32_t)(32 * N)), P, gr);
    if (GRAD) {
      G[c][0] = gr[0] * inv;
      G[c][1] = gr[1] * inv;
      G[c][2] = gr[2] * inv;
    }
    if (FENCE) EDGE_SCHED_FENCE();
  }
}

// ---------------------------------------------------------------------------
// RK4 marches
// ---------------------------------------------------------------------------
// sim/hermite.py:806-860: RK4 of a point psi and its Jacobian.  On entry and
// exit T holds rows (grad psi_c); the columns C[b] = d psi / d x_b evolve as
// dC/dt = J(psi) C.
EDGE_DEV void rk4_psi_T(const Grid& g, const Mac& U, float psi[3], float T[3][3], float dt) {
  float C[3][3];
#pragma unroll
  for (int b = 0; b < 3; ++b)
#pragma unroll
    for (int a = 0; a < 3; ++a) C[b][a] = T[a][b];
  float ps[3] = {psi[0], psi[1], psi[2]};
  float Cs[3][3];
#pragma unroll
  for (int b = 0; b < 3; ++b)
#pragma unroll
    for (int a = 0; a < 3; ++a) Cs[b][a] = C[b][a];
  float su[3] = {0.f, 0.f, 0.f}, sC[3][3] = {{0.f, 0.f, 0.f}, {0.f, 0.f, 0.f}, {0.f, 0.f, 0.f}};
  const float half = 0.5f * dt;
#pragma unroll 1
  for (int st = 0; st < 4; ++st) {
    float u[3], J[3][3];
    interp_u_mac<true>(g, U, ps, u, J);
    float dC[3][3];
#pragma unroll
    for (int b = 0; b < 3; ++b)
#pragma unroll
      for (int a = 0; a < 3; ++a) dC[b][a] = J[a][0] * Cs[b][0] + J[a][1] * Cs[b][1] + J[a][2] * Cs[b][2];
    float wgt = (st == 0 || st == 3) ? 1.0f : 2.0f;
#pragma unroll
    for (int a = 0; a < 3; ++a) su[a] += wgt * u[a];
#pragma unroll
    for (int b = 0; b < 3; ++b)
#pragma unroll
      for (int a = 0; a < 3; ++a) sC[b][a] += wgt * dC[b][a];
    if (st < 3) {
      float cst = (st < 2) ? half : dt;
#pragma unroll
      for (int a = 0; a < 3; ++a) ps[a] = psi[a] + cst * u[a];
#pragma unroll
      for (int b = 0; b < 3; ++b)
#pragma unroll
        for (int a = 0; a < 3; ++a) Cs[b][a] = C[b][a] + cst * dC[b][a];
    }
  }
  float h = dt / 6.0f;
#pragma unroll
  for (int a = 0; a < 3; ++a) psi[a] = psi[a] + h * su[a];
#pragma unroll
  for (int b = 0; b < 3; ++b)
#pragma unroll
    for (int a = 0; a < 3; ++a) T[a][b] = C[b][a] + h * sC[b][a];
}

// sim/hermite.py:786-803: RK4 of a point only.
EDGE_DEV void rk4_psi(const Grid& g, const Mac& U, float psi[3], float dt) {
  float ps[3] = {psi[0], psi[1], psi[2]};
  float su[3] = {0.f, 0.f, 0.f};
  const float half = 0.5f * dt;
#pragma unroll 1
  for (int st = 0; st < 4; ++st) {
    float u[3];
    interp_u_mac<false>(g, U, ps, u, nullptr);
    float wgt = (st == 0 || st == 3) ? 1.0f : 2.0f;
#pragma unroll
    for (int a = 0; a < 3; ++a) su[a] += wgt * u[a];
    if (st < 3) {
      float cst = (st < 2) ? half : dt;
#pragma unroll
      for (int a = 0; a < 3; ++a) ps[a] = psi[a] + cst * u[a];
    }
  }
  float h = dt / 6.0f;
#pragma unroll
  for (int a = 0; a < 3; ++a) psi[a] = psi[a] + h * su[a];
}

// run_edge.py:336-357: RK4 of a face point and one column F.
EDGE_DEV void rk4_phi_F(const Grid& g, const Mac& U, float phi[3], float F[3], float dt) {
  float ps[3] = {phi[0], phi[1], phi[2]}, Fs[3] = {F[0], F[1], F[2]};
  float su[3] = {0.f, 0.f, 0.f}, sF[3] = {0.f, 0.f, 0.f};
  const float half = 0.5f * dt;
#pragma unroll 1
  for (int st = 0; st < 4; ++st) {
    float u[3], J[3][3];
    interp_u_mac<true>(g, U, ps, u, J);
    float dF[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) dF[a] = J[a][0] * Fs[0] + J[a][1] * Fs[1] + J[a][2] * Fs[2];
    float wgt = (st == 0 || st == 3) ? 1.0f : 2.0f;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      su[a] += wgt * u[a];
      sF[a] += wgt * dF[a];
    }
    if (st < 3) {
      float cst = (st < 2) ? half : dt;
#pragma unroll
      for (int a = 0; a < 3; ++a) {
        ps[a] = phi[a] + cst * u[a];
        Fs[a] = F[a] + cst * dF[a];
      }
    }
  }
  float h = dt / 6.0f;
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    phi[a] = phi[a] + h * su[a];
    F[a] = F[a] + h * sF[a];
  }
}

// Compose a marched point (psi, rows T) with the old map in `tab`
// (sim/hermite.py:959-980): psi <- psi_old(psi), T <- grad psi_old(psi) . T.
EDGE_DEV void compose(const Grid& g, const float4* __restrict__ tab, float psi[3], float T[3][3]) {
  float v[3], G[3][3];
  map_eval<true>(g, tab, psi, v, G);
  float Tn[3][3];
#pragma unroll
  for (int a = 0; a < 3; ++a)
#pragma unroll
    for (int b = 0; b < 3; ++b) Tn[a][b] = G[a][0] * T[0][b] + G[a][1] * T[1][b] + G[a][2] * T[2][b];
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    psi[a] = v[a];
#pragma unroll
    for (int b = 0; b < 3; ++b) T[a][b] = Tn[a][b];
  }
}

// Face coordinate (util/taichi_utils.py:29-44): axis a integral, others +0.5.
EDGE_DEV void face_coord(int a, int i, int j, int k, float dx, float X[3]) {
  X[0] = (a == 0 ? (float)i : (float)i + 0.5f) * dx;
  X[1] = (a == 1 ? (float)j : (float)j + 0.5f) * dx;
  X[2] = (a == 2 ? (float)k : (float)k + 0.5f) * dx;
}

}  // namespace edge
