// Learned initial guess: a small MLP trained on the GPU with hand-written
// BF16 MFMA kernels, and the POD bases its outputs are coefficients in.
//
// MLP: nin -> nh (tanh) -> nh (tanh) -> nout (linear). Weights are kept as
// FP32 masters with BF16 copies; every product runs on the matrix cores as
// v_mfma_f32_16x16x32_bf16 (BF16 operands, FP32 accumulation): activations
// and gradients are rounded to BF16 as they are loaded into the MFMA, the
// results stay FP32. Training is full-batch Adam on a mean-squared loss.
//
// POD: for each right-hand side the basis is the leading left singular
// vectors of the snapshot matrix S (n x k, FP64), from a randomized range
// finder with power iterations and Cholesky-QR orthonormalisation: every
// product with S runs on the device (tall-skinny FP64 kernels), only r x r
// and r x k matrices (r <= 96) reach the host.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <random>

#include "dev_common.hpp"
#include "learn.hpp"

namespace bcm {

namespace {

typedef __bf16 v8bf __attribute__((ext_vector_type(8)));
typedef unsigned short v8u __attribute__((ext_vector_type(8)));
typedef float v4f __attribute__((ext_vector_type(4)));

// float -> bf16, round to nearest even (NaN kept quiet).
__device__ __forceinline__ unsigned short f2bf(float f) {
  uint32_t u = __float_as_uint(f);
  if ((u & 0x7fffffffu) > 0x7f800000u) return (unsigned short)((u >> 16) | 0x40u);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (unsigned short)(u >> 16);
}

__device__ __forceinline__ float ldf(const float* p, size_t i) { return p[i]; }
__device__ __forceinline__ float ldf(const unsigned short* p, size_t i) {
  return __uint_as_float((uint32_t)p[i] << 16);
}

// C(i,j) = epilogue(sum_k A(i,k) B(k,j)) with A(i,k) = A[i*ars + k*acs] and
// B(k,j) = B[k*brs + j*bcs] (strides express the transposes). One wave per
// 16 x 16 tile of C, K in steps of 32 (zero padded). Epilogue: + bias[j],
// then tanh (act == 1), then * (1 - D(i,j)^2) with D(i,j) = dmul[i*ldd + j].
template <class TA, class TB>
__global__ __launch_bounds__(256) void k_gemm_bf16(int M, int N, int K, const TA* __restrict__ A, int ars, int acs,
                                                   const TB* __restrict__ B, int brs, int bcs, float* __restrict__ C,
                                                   int ldc, const float* __restrict__ bias, int act,
                                                   const float* __restrict__ dmul, int ldd) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int tiles_n = (N + 15) >> 4;
  const int tiles = ((M + 15) >> 4) * tiles_n;
  const int tile = blockIdx.x * 4 + wave;
  if (tile >= tiles) return;  // wave-uniform: no block-level synchronisation below
  const int tm = tile / tiles_n, tn = tile - tm * tiles_n;
  const int r = lane & 15, g = lane >> 4;
  const int ia = tm * 16 + r;  // row of A this lane feeds
  const int jb = tn * 16 + r;  // column of B this lane feeds
  v4f acc = {0.f, 0.f, 0.f, 0.f};
  // split K: wave group blockIdx.y adds the K range [kz0, kz1) into its own
  // partial tile (C + blockIdx.y * M * ldc); no epilogue then
  const int kper = (((K + gridDim.y - 1) / gridDim.y) + 31) / 32 * 32;
  const int kz0 = blockIdx.y * kper, kz1 = min(K, kz0 + kper);
  C += (size_t)blockIdx.y * M * ldc;
  for (int k0 = kz0; k0 < kz1; k0 += 32) {
    v8u av, bv;
#pragma unroll
    for (int t = 0; t < 8; ++t) {
      const int k = k0 + 8 * g + t;
      av[t] = (ia < M && k < kz1) ? f2bf(ldf(A, (size_t)ia * ars + (size_t)k * acs)) : (unsigned short)0;
      bv[t] = (jb < N && k < kz1) ? f2bf(ldf(B, (size_t)k * brs + (size_t)jb * bcs)) : (unsigned short)0;
    }
    acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(v8bf, av), __builtin_bit_cast(v8bf, bv), acc,
                                                  0, 0, 0);
  }
#pragma unroll
  for (int t = 0; t < 4; ++t) {
    const int i = tm * 16 + 4 * g + t, j = tn * 16 + r;
    if (i < M && j < N) {
      float v = acc[t];
      if (bias) v += bias[j];
      if (act == 1) v = tanhf(v);
      if (dmul) {
        const float d = dmul[(size_t)i * ldd + j];
        v *= 1.0f - d * d;
      }
      C[(size_t)i * ldc + j] = v;
    }
  }
}

// Column sums db[j] = sum_i dz[i*ld + j] in two deterministic passes: block
// (tile, chunk) adds rows [chunk*rows_per, ...) of 64 columns (4 row phases
// per column, combined in a fixed order) into part[chunk*N + j]; k_colsum_fin
// adds the chunks in order.
constexpr int kColChunks = 32;
__global__ __launch_bounds__(256) void k_colsum(int M, int N, const float* __restrict__ dz, int ld,
                                                float* __restrict__ part) {
  __shared__ float sm[4][64];
  const int c = threadIdx.x & 63, ph = threadIdx.x >> 6;
  const int j = blockIdx.x * 64 + c;
  const int per = (M + kColChunks - 1) / kColChunks;
  const int i0 = blockIdx.y * per, i1 = min(M, i0 + per);
  float s = 0.f;
  if (j < N)
    for (int i = i0 + ph; i < i1; i += 4) s += dz[(size_t)i * ld + j];
  sm[ph][c] = s;
  __syncthreads();
  if (ph == 0 && j < N) part[(size_t)blockIdx.y * N + j] = ((sm[0][c] + sm[1][c]) + sm[2][c]) + sm[3][c];
}

__global__ void k_colsum_fin(int N, const float* __restrict__ part, float* __restrict__ db) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= N) return;
  float s = 0.f;
  for (int k = 0; k < kColChunks; ++k) s += part[(size_t)k * N + j];
  db[j] = s;
}

void colsum(hipStream_t s, int M, int N, const float* dz, int ld, float* db, float* part) {
  hipLaunchKernelGGL(k_colsum, dim3((N + 63) / 64, kColChunks), dim3(256), 0, s, M, N, dz, ld, part);
  hipLaunchKernelGGL(k_colsum_fin, dim3((N + 255) / 256), dim3(256), 0, s, N, part, db);
  BCM_HIP(hipGetLastError());
}

// dy = 2 (y - t) / (M N); per-block partials of sum (y - t)^2.
__global__ __launch_bounds__(256) void k_mse(int64_t total, const float* __restrict__ y, const float* __restrict__ t,
                                             float scale, float* __restrict__ dy, double* partials) {
  double acc[1] = {0.0};
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += stride) {
    const float d = y[i] - t[i];
    dy[i] = scale * d;
    acc[0] += (double)d * (double)d;
  }
  block_partials<1>(acc, partials);
}

__global__ __launch_bounds__(256) void k_sum_to(const double* partials, int nb, double scale, double* out) {
  const double s = sum_partials(partials, nb, 1, 0);
  if (threadIdx.x == 0) *out = s * scale;
}

// Adam step over all parameters; refreshes the bf16 copies of the weights.
__global__ void k_adam(int64_t n, float* __restrict__ w, const float* __restrict__ gr, float* __restrict__ m,
                       float* __restrict__ v, unsigned short* __restrict__ wb, const unsigned char* __restrict__ is_w,
                       const int64_t* __restrict__ wb_index, float lr, float b1, float b2, float eps, float c1,
                       float c2) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    const float g = gr[i];
    const float mi = b1 * m[i] + (1.f - b1) * g;
    const float vi = b2 * v[i] + (1.f - b2) * g * g;
    m[i] = mi;
    v[i] = vi;
    const float wi = w[i] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
    w[i] = wi;
    if (is_w[i]) wb[wb_index[i]] = f2bf(wi);
  }
}

__global__ void k_to_bf16(int64_t n, const float* __restrict__ w, unsigned short* __restrict__ wb) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) wb[i] = f2bf(w[i]);
}

// ---- FP64 tall-skinny products for the POD ----

// Partial C_s = X(rows of split s)^T Y(rows of split s): X n x p, Y n x q
// (column-major, leading dimension n); C tile 32 x 32 per block, 256
// threads x 4 outputs; rows staged through LDS in chunks of 64.
__global__ __launch_bounds__(256) void k_tsmm(int64_t n, int p, int q, const double* __restrict__ X,
                                              const double* __restrict__ Y, int64_t rows_per_split,
                                              double* __restrict__ part) {
  __shared__ double xs[64][33];
  __shared__ double ys[64][33];
  const int i0 = blockIdx.x * 32, j0 = blockIdx.y * 32;
  const int64_t r0 = (int64_t)blockIdx.z * rows_per_split;
  const int64_t r1 = std::min<int64_t>(n, r0 + rows_per_split);
  const int t = threadIdx.x;
  const int ci = t & 31, cj = (t >> 5) * 4;  // outputs (ci, cj..cj+3)
  double acc[4] = {0.0, 0.0, 0.0, 0.0};
  for (int64_t rb = r0; rb < r1; rb += 64) {
    for (int e = t; e < 64 * 32; e += 256) {
      const int rr = e & 63, cc = e >> 6;
      const int64_t row = rb + rr;
      xs[rr][cc] = (row < r1 && i0 + cc < p) ? X[(size_t)(i0 + cc) * n + row] : 0.0;
      ys[rr][cc] = (row < r1 && j0 + cc < q) ? Y[(size_t)(j0 + cc) * n + row] : 0.0;
    }
    __syncthreads();
#pragma unroll 8
    for (int rr = 0; rr < 64; ++rr) {
      const double xv = xs[rr][ci];
#pragma unroll
      for (int u = 0; u < 4; ++u) acc[u] += xv * ys[rr][cj + u];
    }
    __syncthreads();
  }
#pragma unroll
  for (int u = 0; u < 4; ++u) {
    const int i = i0 + ci, j = j0 + cj + u;
    if (i < p && j < q) part[((size_t)blockIdx.z * q + j) * p + i] = acc[u];
  }
}

// C = sum over splits of the partials (fixed order), column-major p x q.
__global__ void k_tsmm_sum(int p, int q, int splits, const double* __restrict__ part, double* __restrict__ C) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (int64_t)p * q) return;
  double s = 0.0;
  for (int z = 0; z < splits; ++z) s += part[(size_t)z * p * q + e];
  C[e] = s;
}

// Y(n x q) = X(n x p) C(p x q), column-major; 16 columns of Y per block row,
// C staged through LDS in chunks of 64 rows.
__global__ __launch_bounds__(256) void k_mm(int64_t n, int p, int q, const double* __restrict__ X,
                                            const double* __restrict__ Cm, double* __restrict__ Y, int64_t ldy) {
  __shared__ double cs[64][16];
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int j0 = blockIdx.y * 16;
  double acc[16];
#pragma unroll
  for (int u = 0; u < 16; ++u) acc[u] = 0.0;
  for (int k0 = 0; k0 < p; k0 += 64) {
    for (int e = threadIdx.x; e < 64 * 16; e += blockDim.x) {
      const int kk = e & 63, u = e >> 6;
      cs[kk][u] = (k0 + kk < p && j0 + u < q) ? Cm[(size_t)(j0 + u) * p + k0 + kk] : 0.0;
    }
    __syncthreads();
    if (i < n) {
      const int kn = std::min(64, p - k0);
      for (int kk = 0; kk < kn; ++kk) {
        const double xv = X[(size_t)(k0 + kk) * n + i];
#pragma unroll
        for (int u = 0; u < 16; ++u) acc[u] += xv * cs[kk][u];
      }
    }
    __syncthreads();
  }
  if (i < n) {
#pragma unroll
    for (int u = 0; u < 16; ++u)
      if (j0 + u < q) Y[(size_t)(j0 + u) * ldy + i] = acc[u];
  }
}

int grid1(int64_t n, int cap = 4096) { return (int)std::max<int64_t>(1, std::min<int64_t>((n + 255) / 256, cap)); }

// C(i,j) = sum over s of the split-K partial tiles P[s*M*N + i*N + j], in order.
__global__ void k_ksum(int M, int N, int ks, const float* __restrict__ P, float* __restrict__ C, int ldc) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (int64_t)M * N) return;
  float v = 0.f;
  for (int z = 0; z < ks; ++z) v += P[(size_t)z * M * N + e];
  const int i = (int)(e / N), j = (int)(e - (int64_t)i * N);
  C[(size_t)i * ldc + j] = v;
}

template <class TA, class TB>
void gemm(hipStream_t s, int M, int N, int K, const TA* A, int ars, int acs, const TB* B, int brs, int bcs, float* C,
          int ldc, const float* bias, int act, const float* dmul, int ldd, DevBuf* kscratch = nullptr) {
  const int tiles = ((M + 15) / 16) * ((N + 15) / 16);
  // few output tiles over a long K (the weight gradients): split K so that the
  // whole device works, then add the partial tiles in a fixed order
  int ks = 1;
  if (kscratch && !bias && !act && !dmul) {
    const int waves = (tiles + 3) / 4 * 4;
    ks = std::max(1, std::min({16, K / 128, (4 * 256 * 4) / std::max(1, waves)}));
  }
  if (ks == 1) {
    hipLaunchKernelGGL((k_gemm_bf16<TA, TB>), dim3((tiles + 3) / 4), dim3(256), 0, s, M, N, K, A, ars, acs, B, brs,
                       bcs, C, ldc, bias, act, dmul, ldd);
    BCM_HIP(hipGetLastError());
    return;
  }
  const size_t part = (size_t)M * N;
  if (kscratch->bytes < part * ks * 4) kscratch->alloc(part * ks * 4);
  float* P = kscratch->as<float>();
  hipLaunchKernelGGL((k_gemm_bf16<TA, TB>), dim3((tiles + 3) / 4, ks), dim3(256), 0, s, M, N, K, A, ars, acs, B, brs,
                     bcs, P, N, (const float*)nullptr, 0, (const float*)nullptr, 0);
  hipLaunchKernelGGL(k_ksum, dim3((unsigned)((part + 255) / 256)), dim3(256), 0, s, M, N, ks, P, C, ldc);
  BCM_HIP(hipGetLastError());
}

}  // namespace

// ---------------------------------------------------------------------------
// MLP

Mlp::Mlp(int nin_, int nh_, int nout_, uint64_t seed) : nin(nin_), nh(nh_), nout(nout_) {
  if (nin < 1 || nh < 1 || nout < 1 || nin > 4096 || nh > 4096 || nout > 4096)
    throw Error(kErrArg, "mlp: layer sizes must be in 1..4096");
  const int fan_in[3] = {nin, nh, nh}, fan_out[3] = {nh, nh, nout};
  int64_t off = 0, woff = 0;
  for (int l = 0; l < 3; ++l) {
    w_off[l] = off;
    off += (int64_t)fan_in[l] * fan_out[l];
    b_off[l] = off;
    off += fan_out[l];
    wb_off[l] = woff;
    woff += (int64_t)fan_in[l] * fan_out[l];
  }
  nparam = off;
  nweight = woff;
  std::vector<float> h(nparam, 0.0f);
  std::mt19937_64 g(seed);
  for (int l = 0; l < 3; ++l) {  // Glorot-uniform weights, zero biases
    const double lim = std::sqrt(6.0 / (fan_in[l] + fan_out[l]));
    std::uniform_real_distribution<double> u(-lim, lim);
    for (int64_t i = 0; i < (int64_t)fan_in[l] * fan_out[l]; ++i) h[w_off[l] + i] = (float)u(g);
  }
  params.alloc(nparam * 4);
  grads.alloc(nparam * 4);
  m1.alloc(nparam * 4);
  m2.alloc(nparam * 4);
  wbf.alloc(nweight * 2);
  std::vector<unsigned char> isw(nparam, 0);
  std::vector<int64_t> idx(nparam, 0);
  for (int l = 0; l < 3; ++l)
    for (int64_t i = 0; i < (int64_t)fan_in[l] * fan_out[l]; ++i) {
      isw[w_off[l] + i] = 1;
      idx[w_off[l] + i] = wb_off[l] + i;
    }
  is_w.upload(isw.data(), isw.size());
  wb_index.upload(idx.data(), idx.size() * 8);
  set_params(h.data());
}

void Mlp::set_params(const float* host) {
  BCM_HIP(hipMemcpy(params.p, host, nparam * 4, hipMemcpyHostToDevice));
  BCM_HIP(hipMemset(m1.p, 0, nparam * 4));
  BCM_HIP(hipMemset(m2.p, 0, nparam * 4));
  BCM_HIP(hipMemset(grads.p, 0, nparam * 4));
  for (int l = 0; l < 3; ++l) {
    const int64_t cnt = b_off[l] - w_off[l];
    hipLaunchKernelGGL(k_to_bf16, dim3(grid1(cnt)), dim3(256), 0, nullptr, cnt, params.as<float>() + w_off[l],
                       wbf.as<unsigned short>() + wb_off[l]);
  }
  BCM_HIP(hipDeviceSynchronize());
  step = 0;
}

void Mlp::ensure_batch(int M) {
  if (M <= cap_batch) return;
  h1.alloc((size_t)M * nh * 4);
  h2.alloc((size_t)M * nh * 4);
  y.alloc((size_t)M * nout * 4);
  dz1.alloc((size_t)M * nh * 4);
  dz2.alloc((size_t)M * nh * 4);
  dy.alloc((size_t)M * nout * 4);
  partials.alloc(4096 * 8 + 64);
  colpart.alloc((size_t)kColChunks * std::max(nh, nout) * 4);
  cap_batch = M;
}

void Mlp::forward(hipStream_t s, int M, const float* x, float* out) {
  ensure_batch(M);
  const float* P = params.as<float>();
  const unsigned short* W = wbf.as<unsigned short>();
  gemm(s, M, nh, nin, x, nin, 1, W + wb_off[0], nh, 1, h1.as<float>(), nh, P + b_off[0], 1, (const float*)nullptr, 0);
  gemm(s, M, nh, nh, h1.as<const float>(), nh, 1, W + wb_off[1], nh, 1, h2.as<float>(), nh, P + b_off[1], 1,
       (const float*)nullptr, 0);
  gemm(s, M, nout, nh, h2.as<const float>(), nh, 1, W + wb_off[2], nout, 1, out ? out : y.as<float>(), nout,
       P + b_off[2], 0, (const float*)nullptr, 0);
}

void Mlp::backward(hipStream_t s, int M, const float* x, const float* t, double* loss_dev) {
  forward(s, M, x, nullptr);
  float* G = grads.as<float>();
  const unsigned short* W = wbf.as<unsigned short>();
  const int64_t total = (int64_t)M * nout;
  const int nb = grid1(total, 4096);
  hipLaunchKernelGGL(k_mse, dim3(nb), dim3(256), 0, s, total, y.as<const float>(), t, 2.0f / (float)total,
                     dy.as<float>(), partials.as<double>());
  if (loss_dev)
    hipLaunchKernelGGL(k_sum_to, dim3(1), dim3(256), 0, s, partials.as<const double>(), nb, 1.0 / (double)total,
                       loss_dev);
  // layer 3: dW3 = H2^T dY, db3 = colsum(dY), dZ2 = (dY W3^T) * (1 - H2^2)
  gemm(s, nh, nout, M, h2.as<const float>(), 1, nh, dy.as<const float>(), nout, 1, G + w_off[2], nout,
       (const float*)nullptr, 0, (const float*)nullptr, 0, &kpart);
  colsum(s, M, nout, dy.as<const float>(), nout, G + b_off[2], colpart.as<float>());
  gemm(s, M, nh, nout, dy.as<const float>(), nout, 1, W + wb_off[2], 1, nout, dz2.as<float>(), nh,
       (const float*)nullptr, 0, h2.as<const float>(), nh);
  // layer 2
  gemm(s, nh, nh, M, h1.as<const float>(), 1, nh, dz2.as<const float>(), nh, 1, G + w_off[1], nh,
       (const float*)nullptr, 0, (const float*)nullptr, 0, &kpart);
  colsum(s, M, nh, dz2.as<const float>(), nh, G + b_off[1], colpart.as<float>());
  gemm(s, M, nh, nh, dz2.as<const float>(), nh, 1, W + wb_off[1], 1, nh, dz1.as<float>(), nh, (const float*)nullptr,
       0, h1.as<const float>(), nh);
  // layer 1
  gemm(s, nin, nh, M, x, 1, nin, dz1.as<const float>(), nh, 1, G + w_off[0], nh, (const float*)nullptr, 0,
       (const float*)nullptr, 0, &kpart);
  colsum(s, M, nh, dz1.as<const float>(), nh, G + b_off[0], colpart.as<float>());
}

void Mlp::adam(hipStream_t s, float lr) {
  ++step;
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  const float c1 = 1.0f - std::pow(b1, (float)step), c2 = 1.0f - std::pow(b2, (float)step);
  hipLaunchKernelGGL(k_adam, dim3(grid1(nparam)), dim3(256), 0, s, nparam, params.as<float>(), grads.as<float>(),
                     m1.as<float>(), m2.as<float>(), wbf.as<unsigned short>(), is_w.as<unsigned char>(),
                     wb_index.as<int64_t>(), lr, b1, b2, eps, c1, c2);
  BCM_HIP(hipGetLastError());
}

// ---------------------------------------------------------------------------
// FP64 products for the POD

void tsmm(hipStream_t s, int64_t n, int p, int q, const double* X, const double* Y, double* C, DevBuf& scratch) {
  const int splits = (int)std::max<int64_t>(1, std::min<int64_t>(64, n / 2048));
  const int64_t rows = (n + splits - 1) / splits;
  const size_t need = (size_t)splits * p * q * 8;
  if (scratch.bytes < need) scratch.alloc(need);
  hipLaunchKernelGGL(k_tsmm, dim3((p + 31) / 32, (q + 31) / 32, splits), dim3(256), 0, s, n, p, q, X, Y, rows,
                     scratch.as<double>());
  hipLaunchKernelGGL(k_tsmm_sum, dim3(((int64_t)p * q + 255) / 256), dim3(256), 0, s, p, q, splits,
                     scratch.as<const double>(), C);
  BCM_HIP(hipGetLastError());
}

void mm(hipStream_t s, int64_t n, int p, int q, const double* X, const double* Cm, double* Y, int64_t ldy) {
  hipLaunchKernelGGL(k_mm, dim3((unsigned)((n + 255) / 256), (q + 15) / 16), dim3(256), 0, s, n, p, q, X, Cm, Y, ldy);
  BCM_HIP(hipGetLastError());
}

}  // namespace bcm
