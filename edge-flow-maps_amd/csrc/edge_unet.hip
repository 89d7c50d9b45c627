// Learned warm start for the projection: a 3-level 3D U-Net (16/32/64
// channels, 3x3x3 convolutions, BF16 activations and weights, FP32
// accumulation) that maps the projection's right-hand side b (and the fluid
// mask) to a pressure guess p0.  No counterpart in the reference; the solver
// uses p0 only if ||b - A p0|| <= ||b|| (x0 of edge_poisson).
//
// Layout: activations are channels-last BF16, act[voxel][C] with voxel =
// (i*n1 + j)*n2 + k and C a multiple of 8, so one lane's 8 consecutive
// channels are one 16-byte load.  Weights of a convolution: W[co][tap][ci]
// with tap = (di*3 + dj)*3 + dk (offsets di-1, dj-1, dk-1).
//
// Convolutions are implicit GEMMs on v_mfma_f32_16x16x32_bf16: M = 16 output
// voxels along k, N = 16 output channels, K = 27 taps x input channels in
// groups of 8.  Lane l supplies A[m = l%16][k = 8(l/16) .. +7] and
// B[k = 8(l/16) .. +7][n = l%16] and receives D[m = 4(l/16) + r][n = l%16].
// The data gradient is the same convolution with flipped, transposed weights.
#include <hip/hip_bf16.h>

#include <algorithm>
#include <cstring>

#include "edge_common.h"

using namespace edge;

namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef __hip_bfloat16 bf16;

__device__ __forceinline__ bf16x8 zero8() {
  bf16x8 z;
#pragma unroll
  for (int t = 0; t < 8; ++t) z[t] = (__bf16)0.0f;
  return z;
}
__device__ __forceinline__ bf16x8 load8(const bf16* p) { return *(const bf16x8*)p; }
__device__ __forceinline__ float bf2f(bf16 v) { return __bfloat162float(v); }
__device__ __forceinline__ bf16 f2bf(float v) { return __float2bfloat16(v); }

struct Dims {
  int n0, n1, n2;
};
__device__ __forceinline__ int vox(const Dims& d, int i, int j, int k) { return (i * d.n1 + j) * d.n2 + k; }
__device__ __forceinline__ bool inside(const Dims& d, int i, int j, int k) {
  return (unsigned)i < (unsigned)d.n0 && (unsigned)j < (unsigned)d.n1 && (unsigned)k < (unsigned)d.n2;
}

// ---------------------------------------------------------------------------
// 3x3x3 'same' convolution (zero padding).
//   input : x0 [N][C0] at this resolution, then (optional) x1 [N/8][C1] at half
//           resolution read through nearest upsampling: channels C0.. of the
//           concatenated input;
//   output: mode 0: y[N][Cout] = bf16(relu?(acc + bias));
//           mode 1 (data gradient): channels co < S0 added to g0[N][S0] (FP32),
//           channels co >= S0 atomically added to g1[N/8][Cout-S0] (FP32) at
//           the parent voxel (the transpose of the upsampling).
// One wave computes MT segments of 16 voxels (consecutive j rows) x Cout.
// ---------------------------------------------------------------------------
struct ConvArgs {
  const bf16* x0;
  const bf16* x1;
  int C0, C1;
  const bf16* w;
  const float* bias;
  bf16* y;
  float* g0;
  float* g1;
  int S0;
  int Cout;
  int relu;
  Dims d;
};

// j rows per wave (build knob; 4 measured 15% faster than 2 for this network)
#ifndef EDGE_CONV_MT
#define EDGE_CONV_MT 4
#endif
constexpr int MT = EDGE_CONV_MT;

template <int NT, int MODE>
__global__ void __launch_bounds__(256) k_conv3(ConvArgs a) {
  const int lane = threadIdx.x & 63;
  const int wave = blockIdx.x * 4 + (threadIdx.x >> 6);
  const Dims d = a.d;
  const int segs_k = d.n2 >> 4;
  const int rows_j = (d.n1 + MT - 1) / MT;
  const int nw = d.n0 * rows_j * segs_k;
  if (wave >= nw) return;
  const int ks = wave % segs_k;
  const int jr = (wave / segs_k) % rows_j;
  const int i = wave / (segs_k * rows_j);
  const int k0 = ks * 16;
  const int Cin = a.C0 + a.C1;
  const int CB = Cin >> 3;
  const int G = 27 * CB;
  const int steps = (G + 3) >> 2;
  const int m = lane & 15, kg = lane >> 4;
  const Dims dh{d.n0 >> 1, d.n1 >> 1, d.n2 >> 1};
  f32x4 acc[MT][NT];
#pragma unroll
  for (int s = 0; s < MT; ++s)
#pragma unroll
    for (int t = 0; t < NT; ++t) acc[s][t] = f32x4{0.f, 0.f, 0.f, 0.f};
  const int Kw = 27 * Cin;
  for (int st = 0; st < steps; ++st) {
    const int g = 4 * st + kg;
    bf16x8 b[NT];
    bf16x8 av[MT];
    if (g < G) {
      const int tap = g / CB, cb = g - tap * CB;
      const int di = tap / 9 - 1, dj = (tap / 3) % 3 - 1, dk = tap % 3 - 1;
      const int c = cb * 8;
#pragma unroll
      for (int t = 0; t < NT; ++t) b[t] = load8(a.w + (size_t)(16 * t + m) * Kw + 8 * g);
#pragma unroll
      for (int s = 0; s < MT; ++s) {
        const int j = jr * MT + s;
        const int ii = i + di, jj = j + dj, kk = k0 + m + dk;
        av[s] = zero8();
        if (j < d.n1 && inside(d, ii, jj, kk)) {
          if (c < a.C0)
            av[s] = load8(a.x0 + (size_t)vox(d, ii, jj, kk) * a.C0 + c);
          else
            av[s] = load8(a.x1 + (size_t)vox(dh, ii >> 1, jj >> 1, kk >> 1) * a.C1 + (c - a.C0));
        }
      }
    } else {
#pragma unroll
      for (int t = 0; t < NT; ++t) b[t] = zero8();
#pragma unroll
      for (int s = 0; s < MT; ++s) av[s] = zero8();
    }
#pragma unroll
    for (int s = 0; s < MT; ++s)
#pragma unroll
      for (int t = 0; t < NT; ++t) acc[s][t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av[s], b[t], acc[s][t], 0, 0, 0);
  }
  // epilogue: lane holds D[m = 4*kg + r][n = m_lane]
#pragma unroll
  for (int s = 0; s < MT; ++s) {
    const int j = jr * MT + s;
    if (j >= d.n1) continue;
#pragma unroll
    for (int t = 0; t < NT; ++t) {
      const int co = 16 * t + m;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int k = k0 + 4 * kg + r;
        const int v = vox(d, i, j, k);
        float val = acc[s][t][r];
        if (MODE == 0) {
          if (a.bias) val += a.bias[co];
          if (a.relu) val = fmaxf(val, 0.f);
          a.y[(size_t)v * a.Cout + co] = f2bf(val);
        } else {
          if (co < a.S0) {
            a.g0[(size_t)v * a.S0 + co] += val;
          } else {
            const int S1 = a.Cout - a.S0;
            atomicAdd(a.g1 + (size_t)vox(dh, i >> 1, j >> 1, k >> 1) * S1 + (co - a.S0), val);
          }
        }
      }
    }
  }
}

// ---------------------------------------------------------------------------
// Weight gradient: dW[co][tap][ci] += sum_v g[v][co] * x(v + off(tap))[ci],
// x the (concatenated, upsampled) input of the forward convolution.  MFMA with
// M = 16 output channels, N = 16 columns (tap, ci) of W, K = voxels in groups
// of 8 along k; each wave covers `per` groups of 32 voxels and adds its 16x16
// tile to the FP32 gradient with atomics.
// ---------------------------------------------------------------------------
struct WgradArgs {
  const bf16* g;   // [N][Cout] gradient w.r.t. the pre-activation
  const bf16* x0;
  const bf16* x1;
  int C0, C1, Cout;
  float* dw;       // [Cout][27*Cin]
  Dims d;
  int per;         // K-steps (32 voxels each) per wave
};

__global__ void __launch_bounds__(256) k_wgrad(WgradArgs a) {
  const int lane = threadIdx.x & 63;
  const int wave = blockIdx.x * 4 + (threadIdx.x >> 6);
  const Dims d = a.d;
  const Dims dh{d.n0 >> 1, d.n1 >> 1, d.n2 >> 1};
  const int Cin = a.C0 + a.C1;
  const int Kw = 27 * Cin;
  const int mt = a.Cout >> 4, nt = (Kw + 15) >> 4;
  const int N = d.n0 * d.n1 * d.n2;
  const int ksteps = N >> 5;
  const int splits = (ksteps + a.per - 1) / a.per;
  const int nwaves = mt * nt * splits;
  if (wave >= nwaves) return;
  const int sp = wave % splits;
  const int tile = wave / splits;
  const int tm = tile % mt, tn = tile / mt;
  const int m = lane & 15, kg = lane >> 4;
  const int co = 16 * tm + m;       // A row
  const int col = 16 * tn + m;      // B column
  const bool col_ok = col < Kw;
  const int tap = col_ok ? col / Cin : 0;
  const int ci = col_ok ? col - tap * Cin : 0;
  const int di = tap / 9 - 1, dj = (tap / 3) % 3 - 1, dk = tap % 3 - 1;
  f32x4 acc = {0.f, 0.f, 0.f, 0.f};
  const int s_end = min(ksteps, (sp + 1) * a.per);
  for (int st = sp * a.per; st < s_end; ++st) {
    // 8 consecutive voxels v0..v0+7 (one row: n2 is a multiple of 16)
    const int v0 = 32 * st + 8 * kg;
    const int k0 = v0 % d.n2, row = v0 / d.n2;
    const int j = row % d.n1, i = row / d.n1;
    bf16x8 av, bv;
#pragma unroll
    for (int t = 0; t < 8; ++t) {
      av[t] = (__bf16)bf2f(a.g[(size_t)(v0 + t) * a.Cout + co]);
      const int ii = i + di, jj = j + dj, kk = k0 + t + dk;
      float xv = 0.f;
      if (col_ok && inside(d, ii, jj, kk)) {
        if (ci < a.C0)
          xv = bf2f(a.x0[(size_t)vox(d, ii, jj, kk) * a.C0 + ci]);
        else
          xv = bf2f(a.x1[(size_t)vox(dh, ii >> 1, jj >> 1, kk >> 1) * a.C1 + (ci - a.C0)]);
      }
      bv[t] = (__bf16)xv;
    }
    acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av, bv, acc, 0, 0, 0);
  }
  if (!col_ok) return;
#pragma unroll
  for (int r = 0; r < 4; ++r) atomicAdd(a.dw + (size_t)(16 * tm + 4 * kg + r) * Kw + col, acc[r]);
}

// ---------------------------------------------------------------------------
// Small kernels
// ---------------------------------------------------------------------------
// 2x2x2 average pooling: y[N/8][C] from x[N][C] (d: fine dims).
__global__ void k_pool2(const bf16* __restrict__ x, bf16* __restrict__ y, Dims d, int C) {
  const Dims dh{d.n0 >> 1, d.n1 >> 1, d.n2 >> 1};
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t n = (int64_t)dh.n0 * dh.n1 * dh.n2 * C;
  if (t >= n) return;
  const int c = (int)(t % C);
  const int v = (int)(t / C);
  const int k = v % dh.n2, j = (v / dh.n2) % dh.n1, i = v / (dh.n2 * dh.n1);
  float s = 0.f;
#pragma unroll
  for (int q = 0; q < 8; ++q) s += bf2f(x[(size_t)vox(d, 2 * i + (q >> 2), 2 * j + ((q >> 1) & 1), 2 * k + (q & 1)) * C + c]);
  y[t] = f2bf(0.125f * s);
}

// gradient of the pooling: gx[v][c] += gy[parent][c] / 8
__global__ void k_pool2_bwd(const float* __restrict__ gy, float* __restrict__ gx, Dims d, int C) {
  const Dims dh{d.n0 >> 1, d.n1 >> 1, d.n2 >> 1};
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t n = (int64_t)d.n0 * d.n1 * d.n2 * C;
  if (t >= n) return;
  const int c = (int)(t % C);
  const int v = (int)(t / C);
  const int k = v % d.n2, j = (v / d.n2) % d.n1, i = v / (d.n2 * d.n1);
  gx[t] += 0.125f * gy[(size_t)vox(dh, i >> 1, j >> 1, k >> 1) * C + c];
}

// gradient through ReLU to BF16: g[t] = bf16(gy[t]) where y > 0, else 0 (y = the
// layer's post-activation output; relu = 0: no mask)
__global__ void k_relu_bwd(const float* __restrict__ gy, const bf16* __restrict__ y, bf16* __restrict__ g, int64_t n,
                           int relu) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  float v = gy[t];
  if (relu && !(bf2f(y[t]) > 0.f)) v = 0.f;
  g[t] = f2bf(v);
}

// bias gradient: db[c] += sum_v g[v][c]
__global__ void k_bias_grad(const bf16* __restrict__ g, float* __restrict__ db, int N, int C) {
  const int c = blockIdx.y;
  float s = 0.f;
  for (int v = blockIdx.x * blockDim.x + threadIdx.x; v < N; v += gridDim.x * blockDim.x) s += bf2f(g[(size_t)v * C + c]);
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
  if ((threadIdx.x & 63) == 0) atomicAdd(db + c, s);
}

// input features: x[v][0] = b[v] / scale, x[v][1] = fluid, x[v][2..7] = 0.
// scale = sqrt(mean b^2) (from sums[0]) or 1 when b = 0.
__global__ void k_input(const float* __restrict__ b, const int* __restrict__ bm, const double* __restrict__ sums,
                        int N, bf16* __restrict__ x, float* __restrict__ scale_out) {
  const int v = blockIdx.x * blockDim.x + threadIdx.x;
  double ms = sums[0] / (double)N;
  float sc = ms > 0.0 ? (float)sqrt(ms) : 1.0f;
  if (v == 0 && scale_out) *scale_out = sc;
  if (v >= N) return;
  bf16x8 o = zero8();
  o[0] = (__bf16)(b[v] / sc);
  o[1] = (__bf16)((bm && bm[v] > 0) ? 0.f : 1.f);
  *(bf16x8*)(x + (size_t)v * 8) = o;
}

// sums[0] += sum a^2, sums[1] += sum (b or a): wave sums, block sums in LDS,
// one FP64 atomic per block and value
__global__ void __launch_bounds__(256) k_sums(const float* __restrict__ a, const float* __restrict__ b, int N,
                                              double* __restrict__ sums) {
  double s0 = 0, s1 = 0;
  for (int v = blockIdx.x * blockDim.x + threadIdx.x; v < N; v += gridDim.x * blockDim.x) {
    double x = a[v];
    s0 += x * x;
    s1 += b ? (double)b[v] : x;
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    s0 += __shfl_xor(s0, off, 64);
    s1 += __shfl_xor(s1, off, 64);
  }
  __shared__ double sh[2][4];
  if ((threadIdx.x & 63) == 0) {
    sh[0][threadIdx.x >> 6] = s0;
    sh[1][threadIdx.x >> 6] = s1;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    atomicAdd(sums, sh[0][0] + sh[0][1] + sh[0][2] + sh[0][3]);
    atomicAdd(sums + 1, sh[1][0] + sh[1][1] + sh[1][2] + sh[1][3]);
  }
}

// head: p[v] = scale * (sum_c w[c] h[v][c] + bias), FP32; C a multiple of 8.
__global__ void k_head(const bf16* __restrict__ h, const float* __restrict__ w, const float* __restrict__ bias,
                       const float* __restrict__ scale, int N, int C, float* __restrict__ p) {
  const int v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= N) return;
  float s = bias[0];
  for (int c = 0; c < C; c += 8) {
    const bf16x8 hv = load8(h + (size_t)v * C + c);
#pragma unroll
    for (int t = 0; t < 8; ++t) s += w[c + t] * (float)hv[t];
  }
  p[v] = scale[0] * s;
}

// loss = mean((q - mean q)^2), q = (p - target) / scale (the pressure is defined up
// to a constant); dp[v] = 2/N (q - mean q) / scale is dloss/dp.  sums: [2] = sum q.
__global__ void k_loss_grad(const float* __restrict__ p, const float* __restrict__ target, const float* __restrict__ scale,
                            const double* __restrict__ sums, int N, float* __restrict__ dp, double* __restrict__ loss) {
  const int v = blockIdx.x * blockDim.x + threadIdx.x;
  const float sc = scale[0];
  const float mq = (float)(sums[2] / (double)N);
  double l = 0;
  if (v < N) {
    float q = (p[v] - target[v]) / sc - mq;
    dp[v] = 2.0f / (float)N * q / sc;
    l = (double)q * (double)q / (double)N;
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) l += __shfl_xor(l, off, 64);
  if ((threadIdx.x & 63) == 0) atomicAdd(loss, l);
}

__global__ void k_qsum(const float* __restrict__ p, const float* __restrict__ target, const float* __restrict__ scale,
                       int N, double* __restrict__ sums) {
  double s = 0;
  const float sc = scale[0];
  for (int v = blockIdx.x * blockDim.x + threadIdx.x; v < N; v += gridDim.x * blockDim.x) s += (double)((p[v] - target[v]) / sc);
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
  if ((threadIdx.x & 63) == 0) atomicAdd(sums + 2, s);
}

// head backward: gh[v][c] = dp[v] * scale * w[c]; dw[c] += sum_v dp scale h[v][c]; db += sum dp scale
__global__ void k_head_bwd(const bf16* __restrict__ h, const float* __restrict__ w, const float* __restrict__ scale,
                           const float* __restrict__ dp, int N, int C, float* __restrict__ gh, float* __restrict__ dw,
                           float* __restrict__ db) {
  const int v = blockIdx.x * blockDim.x + threadIdx.x;
  const float sc = scale[0];
  float d = v < N ? dp[v] * sc : 0.f;
  for (int c = 0; c < C; ++c) {
    float hv = v < N ? bf2f(h[(size_t)v * C + c]) : 0.f;
    if (v < N) gh[(size_t)v * C + c] = d * w[c];
    float s = d * hv;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
    if ((threadIdx.x & 63) == 0) atomicAdd(dw + c, s);
  }
  float s = d;
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
  if ((threadIdx.x & 63) == 0) atomicAdd(db, s);
}

// BF16 copies of a convolution's FP32 weights W[co][tap][ci]: the forward
// layout, and the data-gradient layout Wt[ci][26 - tap][co].
__global__ void k_pack_w(const float* __restrict__ w, int Cout, int Cin, bf16* __restrict__ wf, bf16* __restrict__ wt) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const int n = Cout * 27 * Cin;
  if (t >= n) return;
  const int ci = t % Cin, tap = (t / Cin) % 27, co = t / (27 * Cin);
  const bf16 v = f2bf(w[t]);
  wf[t] = v;
  if (wt) wt[((size_t)ci * 27 + (26 - tap)) * Cout + co] = v;
}

// Adam (decoupled from the model layout: one flat FP32 parameter vector).
__global__ void k_adam(float* __restrict__ p, const float* __restrict__ g, float* __restrict__ m1, float* __restrict__ m2,
                       int n, float lr, float b1, float b2, float eps, float c1, float c2) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  float gv = g[t];
  float a = b1 * m1[t] + (1.f - b1) * gv;
  float b = b2 * m2[t] + (1.f - b2) * gv * gv;
  m1[t] = a;
  m2[t] = b;
  p[t] -= lr * (a / c1) / (sqrtf(b / c2) + eps);
}

__global__ void k_zero_f(float* __restrict__ p, int64_t n) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) p[t] = 0.f;
}
__global__ void k_zero_d(double* __restrict__ p, int n) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) p[t] = 0.0;
}

int64_t nb(int64_t n, int bs = 256) { return (n + bs - 1) / bs; }

// one level's dims: 16-voxel segments along k; even dims where a half-resolution
// tensor is read or written
bool dims_conv(int n0, int n1, int n2, bool half) {
  if (n0 < 1 || n1 < 1 || n2 < 16 || n2 % 16 != 0 || (int64_t)n0 * n1 * n2 >= (1LL << 31)) return false;
  return !half || (n0 % 2 == 0 && n1 % 2 == 0);
}

int conv_launch(const ConvArgs& a, int mode, hipStream_t st) {
  const int rows_j = (a.d.n1 + MT - 1) / MT;
  const int64_t waves = (int64_t)a.d.n0 * rows_j * (a.d.n2 >> 4);
  dim3 grid((unsigned)nb(waves, 4)), block(256);
  if (a.Cout % 16 != 0) {
    set_error("edge_conv3d: output channels %d not a multiple of 16", a.Cout);
    return EDGE_EINVAL;
  }
  const int NT = a.Cout / 16;
#define EDGE_CONV_CASE(nt)                                                      \
  case nt:                                                                      \
    if (mode == 0) hipLaunchKernelGGL((k_conv3<nt, 0>), grid, block, 0, st, a); \
    else hipLaunchKernelGGL((k_conv3<nt, 1>), grid, block, 0, st, a);           \
    break;
  switch (NT) {
    EDGE_CONV_CASE(1)
    EDGE_CONV_CASE(2)
    EDGE_CONV_CASE(3)
    EDGE_CONV_CASE(4)
    EDGE_CONV_CASE(6)
    default:
      set_error("edge_conv3d: unsupported output channel count %d", a.Cout);
      return EDGE_EINVAL;
  }
#undef EDGE_CONV_CASE
  EDGE_CHECK_LAUNCH();
  return EDGE_OK;
}

}  // namespace

extern "C" {

// 3x3x3 convolution, forward: y = relu?(conv(concat(x0, up(x1))) + bias).
// x0 [N][C0], x1 [N/8][C1] or NULL (C1 = 0), w [Cout][27][C0+C1] BF16,
// bias [Cout] FP32 or NULL, y [N][Cout] BF16.  C0, C1 multiples of 8; Cout in
// {16, 32, 48, 64, 96}; n2 a multiple of 16 (n0, n1 even when x1 is given).
int edge_conv3d_bf16_fwd(const void* x0, int C0, const void* x1, int C1, const void* w, const float* bias, void* y,
                         int Cout, int relu, int n0, int n1, int n2, void* stream) {
  clear_error();
  EDGE_REQUIRE(dims_conv(n0, n1, n2, C1 > 0), "edge_conv3d_bf16_fwd: dims (%d, %d, %d) unsupported", n0, n1, n2);
  EDGE_REQUIRE(x0 && w && y && C0 > 0 && C0 % 8 == 0 && C1 >= 0 && C1 % 8 == 0 && (C1 == 0 || x1),
               "edge_conv3d_bf16_fwd: bad inputs");
  ConvArgs a{};
  a.x0 = (const bf16*)x0;
  a.x1 = (const bf16*)x1;
  a.C0 = C0;
  a.C1 = C1;
  a.w = (const bf16*)w;
  a.bias = bias;
  a.y = (bf16*)y;
  a.Cout = Cout;
  a.relu = relu;
  a.d = Dims{n0, n1, n2};
  return conv_launch(a, 0, (hipStream_t)stream);
}

// Data gradient: g0 += (conv^T g)[channels < S0], g1 (half resolution) += the
// remaining channels summed over each parent's 8 children.  g [N][Cg] BF16,
// wt [Cin][27][Cg] (the transposed, flipped weights of the forward layer whose
// input had Cin = S0 + S1 channels), g0 [N][S0], g1 [N/8][Cin-S0] FP32.
int edge_conv3d_bf16_bwd_data(const void* g, int Cg, const void* wt, int Cin, int S0, float* g0, float* g1, int n0,
                              int n1, int n2, void* stream) {
  clear_error();
  EDGE_REQUIRE(dims_conv(n0, n1, n2, S0 < Cin), "edge_conv3d_bf16_bwd_data: dims (%d, %d, %d) unsupported", n0, n1,
               n2);
  EDGE_REQUIRE(g && wt && g0 && Cg % 8 == 0 && S0 > 0 && S0 <= Cin && (S0 == Cin || g1),
               "edge_conv3d_bf16_bwd_data: bad inputs");
  ConvArgs a{};
  a.x0 = (const bf16*)g;
  a.C0 = Cg;
  a.C1 = 0;
  a.w = (const bf16*)wt;
  a.Cout = Cin;
  a.S0 = S0;
  a.g0 = g0;
  a.g1 = g1;
  a.d = Dims{n0, n1, n2};
  return conv_launch(a, 1, (hipStream_t)stream);
}

// Weight gradient: dw [Cout][27][C0+C1] (FP32) += sum_v g[v] x(v + tap).
int edge_conv3d_bf16_bwd_weight(const void* g, int Cout, const void* x0, int C0, const void* x1, int C1, float* dw,
                                int n0, int n1, int n2, void* stream) {
  clear_error();
  EDGE_REQUIRE(dims_conv(n0, n1, n2, C1 > 0), "edge_conv3d_bf16_bwd_weight: dims (%d, %d, %d) unsupported", n0, n1,
               n2);
  EDGE_REQUIRE(g && x0 && dw && Cout % 16 == 0 && C0 > 0 && (C1 == 0 || x1), "edge_conv3d_bf16_bwd_weight: bad inputs");
  WgradArgs a{};
  a.g = (const bf16*)g;
  a.x0 = (const bf16*)x0;
  a.x1 = (const bf16*)x1;
  a.C0 = C0;
  a.C1 = C1;
  a.Cout = Cout;
  a.dw = dw;
  a.d = Dims{n0, n1, n2};
  const int64_t N = (int64_t)n0 * n1 * n2;
  const int ksteps = (int)(N >> 5);
  const int Kw = 27 * (C0 + C1);
  const int tiles = (Cout / 16) * ((Kw + 15) / 16);
  // split K so that the launch has about 64k waves
  int splits = (int)std::max<int64_t>(1, std::min<int64_t>(ksteps, 65536 / std::max(1, tiles)));
  a.per = (ksteps + splits - 1) / splits;
  splits = (ksteps + a.per - 1) / a.per;
  const int64_t waves = (int64_t)tiles * splits;
  hipLaunchKernelGGL(k_wgrad, dim3((unsigned)nb(waves, 4)), dim3(256), 0, (hipStream_t)stream, a);
  EDGE_CHECK_LAUNCH();
  return EDGE_OK;
}

// 2x2x2 average pooling (x [N][C] -> y [N/8][C], BF16) and its gradient
// (gx [N][C] += gy[parent] / 8, FP32).  d: fine dims.
int edge_unet_pool(const void* x, void* y, int C, int n0, int n1, int n2, void* stream) {
  clear_error();
  EDGE_REQUIRE(x && y && C > 0 && n0 % 2 == 0 && n1 % 2 == 0 && n2 % 2 == 0, "edge_unet_pool: bad arguments");
  const int64_t n = (int64_t)n0 * n1 * n2 / 8 * C;
  hipLaunchKernelGGL(k_pool2, dim3((unsigned)nb(n)), dim3(256), 0, (hipStream_t)stream, (const bf16*)x, (bf16*)y,
                     Dims{n0, n1, n2}, C);
  EDGE_CHECK_LAUNCH();
  return EDGE_OK;
}
int edge_unet_pool_bwd(const float* gy, float* gx, int C, int n0, int n1, int n2, void* stream) {
  clear_error();
  EDGE_REQUIRE(gy && gx && C > 0 && n0 % 2 == 0 && n1 % 2 == 0 && n2 % 2 == 0, "edge_unet_pool_bwd: bad arguments");
  const int64_t n = (int64_t)n0 * n1 * n2 * C;
  hipLaunchKernelGGL(k_pool2_bwd, dim3((unsigned)nb(n)), dim3(256), 0, (hipStream_t)stream, gy, gx, Dims{n0, n1, n2}, C);
  EDGE_CHECK_LAUNCH();
  return EDGE_OK;
}

// g = bf16(gy masked by y > 0 when relu) and, if db, db[c] += sum_v g[v][c].
int edge_unet_relu_bwd(const float* gy, const void* y, void* g, float* db, int N, int C, int relu, void* stream) {
  clear_error();
  EDGE_REQUIRE(gy && g && (y || !relu) && N > 0 && C > 0, "edge_unet_relu_bwd: bad arguments");
  hipStream_t st = (hipStream_t)stream;
  const int64_t n = (int64_t)N * C;
  hipLaunchKernelGGL(k_relu_bwd, dim3((unsigned)nb(n)), dim3(256), 0, st, gy, (const bf16*)y, (bf16*)g, n, relu);
  if (db) hipLaunchKernelGGL(k_bias_grad, dim3(std::min<int64_t>(nb(N), 256), C), dim3(256), 0, st, (const bf16*)g, db, N, C);
  EDGE_CHECK_LAUNCH();
  return EDGE_OK;
}

// Network input x [N][8] (BF16) from b and the optional int32 solid mask;
// scale (one device float) = the RMS of b.  work: 4 device doubles.
int edge_unet_input(const float* b, const int* bm, int N, void* x, float* scale, double* work, void* stream) {
  clear_error();
  EDGE_REQUIRE(b && x && scale && work && N > 0, "edge_unet_input: bad arguments");
  hipStream_t st = (hipStream_t)stream;
  hipLaunchKernelGGL(k_zero_d, dim3(1), dim3(64), 0, st, work, 4);
  hipLaunchKernelGGL(k_sums, dim3((unsigned)std::min<int64_t>(nb(N), 1024)), dim3(256), 0, st, b, (const float*)nullptr,
                     N, work);
  hipLaunchKernelGGL(k_input, dim3((unsigned)nb(N)), dim3(256), 0, st, b, bm, (const double*)work, N, (bf16*)x, scale);
  EDGE_CHECK_LAUNCH();
  return EDGE_OK;
}

// p [N] = scale * (h . w + bias); h [N][C] BF16, w [C], bias [1] FP32.
int edge_unet_head(const void* h, const float* w, const float* bias, const float* scale, int N, int C, float* p,
                   void* stream) {
  clear_error();
  EDGE_REQUIRE(h && w && bias && scale && p && N > 0 && C > 0 && C % 8 == 0, "edge_unet_head: bad arguments");
  hipLaunchKernelGGL(k_head, dim3((unsigned)nb(N)), dim3(256), 0, (hipStream_t)stream, (const bf16*)h, w, bias, scale, N,
                     C, p);
  EDGE_CHECK_LAUNCH();
  return EDGE_OK;
}

// Loss of a prediction against a target pressure (both FP32 [N]), up to a
// constant and in units of scale: loss (work[3]) and dp = dloss/dp.  Then the
// head's gradients: gh [N][C] (FP32), dw [C] and db [1] accumulated.
int edge_unet_loss_head_bwd(const float* p, const float* target, const float* scale, const void* h, const float* w,
                            int N, int C, float* dp, float* gh, float* dw, float* db, double* work, void* stream) {
  clear_error();
  EDGE_REQUIRE(p && target && scale && h && w && dp && gh && dw && db && work && N > 0, "edge_unet_loss: bad arguments");
  hipStream_t st = (hipStream_t)stream;
  hipLaunchKernelGGL(k_zero_d, dim3(1), dim3(64), 0, st, work, 4);
  hipLaunchKernelGGL(k_qsum, dim3((unsigned)std::min<int64_t>(nb(N), 1024)), dim3(256), 0, st, p, target, scale, N, work);
  hipLaunchKernelGGL(k_loss_grad, dim3((unsigned)nb(N)), dim3(256), 0, st, p, target, scale, (const double*)work, N, dp,
                     work + 3);
  hipLaunchKernelGGL(k_head_bwd, dim3((unsigned)nb(N)), dim3(256), 0, st, (const bf16*)h, w, scale, (const float*)dp, N,
                     C, gh, dw, db);
  EDGE_CHECK_LAUNCH();
  return EDGE_OK;
}

// The head's gradients for a given dL/dp (FP32 [N]): gh [N][C], dw [C] and
// db [1] accumulated, as in edge_unet_loss_head_bwd without its loss.
int edge_unet_head_bwd(const float* dp, const float* scale, const void* h, const float* w, int N, int C, float* gh,
                       float* dw, float* db, void* stream) {
  clear_error();
  EDGE_REQUIRE(dp && scale && h && w && gh && dw && db && N > 0 && C > 0, "edge_unet_head_bwd: bad arguments");
  hipLaunchKernelGGL(k_head_bwd, dim3((unsigned)nb(N)), dim3(256), 0, (hipStream_t)stream, (const bf16*)h, w, scale,
                     dp, N, C, gh, dw, db);
  EDGE_CHECK_LAUNCH();
  return EDGE_OK;
}

// BF16 forward weights wf and (if wt) data-gradient weights of one layer.
int edge_unet_pack_weights(const float* w, int Cout, int Cin, void* wf, void* wt, void* stream) {
  clear_error();
  EDGE_REQUIRE(w && wf && Cout > 0 && Cin > 0, "edge_unet_pack_weights: bad arguments");
  const int n = Cout * 27 * Cin;
  hipLaunchKernelGGL(k_pack_w, dim3((unsigned)nb(n)), dim3(256), 0, (hipStream_t)stream, w, Cout, Cin, (bf16*)wf,
                     (bf16*)wt);
  EDGE_CHECK_LAUNCH();
  return EDGE_OK;
}

// One Adam step on n FP32 parameters (bias corrections c1 = 1 - b1^t, c2 = 1 - b2^t).
int edge_adam(float* p, const float* g, float* m1, float* m2, int n, float lr, float b1, float b2, float eps, float c1,
              float c2, void* stream) {
  clear_error();
  EDGE_REQUIRE(p && g && m1 && m2 && n >= 0, "edge_adam: bad arguments");
  if (n == 0) return EDGE_OK;
  hipLaunchKernelGGL(k_adam, dim3((unsigned)nb(n)), dim3(256), 0, (hipStream_t)stream, p, g, m1, m2, n, lr, b1, b2, eps,
                     c1, c2);
  EDGE_CHECK_LAUNCH();
  return EDGE_OK;
}

int edge_zero_f32(float* p, long n, void* stream) {
  clear_error();
  EDGE_REQUIRE(p && n >= 0, "edge_zero_f32: bad arguments");
  if (n == 0) return EDGE_OK;
  hipLaunchKernelGGL(k_zero_f, dim3((unsigned)nb(n)), dim3(256), 0, (hipStream_t)stream, p, (int64_t)n);
  EDGE_CHECK_LAUNCH();
  return EDGE_OK;
}

}  // extern "C"
