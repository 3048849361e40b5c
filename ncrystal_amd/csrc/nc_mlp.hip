// Small MLP (features -> 3 x 64 hidden -> 48 outputs) trained and evaluated
// on the GPU: BF16 matrix products on the CDNA4 matrix cores
// (v_mfma_f32_16x16x32_bf16, FP32 accumulation), FP32 elementwise work and
// FP32 Adam with FP32 master weights.
//
// The 48 outputs parameterise a mixture of 8 bivariate Gaussians over the
// unit square of the kinematic coordinates (w, kappa) of an S(alpha,beta)
// sample given the neutron energy (component k: logit, mean_w, mean_kappa,
// log sigma_w, log sigma_kappa, atanh-like correlation), trained by maximum
// likelihood on samples of the plain S(alpha,beta) sampler.  This model has
// no counterpart in NCrystal.
#include "nc_kernels.hh"

namespace ncamd {

typedef __attribute__((__vector_size__(8 * sizeof(short)))) short bf16x8_t;
typedef __attribute__((__vector_size__(4 * sizeof(float)))) float f32x4_t;

// FP32 -> BF16 with round-to-nearest-even (NaN kept quiet)
NCD unsigned short f32_to_bf16(float f) {
  const unsigned u = __float_as_uint(f);
  if ((u & 0x7fffffffu) > 0x7f800000u) return (unsigned short)((u >> 16) | 0x40u);
  const unsigned r = ((u >> 16) & 1u) + 0x7fffu;
  return (unsigned short)((u + r) >> 16);
}

// C[M x N] (+)= op(A)[M x K] * op(B)[K x N] (+ bias[N]) with FP32 storage,
// BF16 operands and FP32 accumulation.  One wavefront per 16 x 16 output
// tile, K in steps of 32; blockIdx.y splits K (accumulate mode only).
// Operand layout of v_mfma_f32_16x16x32_bf16: lane l holds A[row l%16][k0 +
// 8*(l/16) + j] and B[k0 + 8*(l/16) + j][col l%16], j = 0..7; the result
// lane l holds C[4*(l/16) + i][l%16], i = 0..3.
template <bool TA, bool TB, bool Accum>
__global__ void __launch_bounds__(256) k_gemm_bf16(const float* __restrict__ A, int lda, const float* __restrict__ B,
                                                   int ldb, float* __restrict__ C, int ldc, const float* __restrict__ bias,
                                                   int M, int N, int K, int kchunk) {
  const int lane = threadIdx.x & 63;
  const int wave = (int)(blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6));
  const int tn = (N + 15) >> 4;
  const int ntiles = ((M + 15) >> 4) * tn;
  if (wave >= ntiles) return;
  const int m0 = (wave / tn) << 4, n0 = (wave % tn) << 4;
  const int kb = (int)blockIdx.y * kchunk;
  const int ke = min(K, kb + kchunk);
  const int r = lane & 15, g = lane >> 4;
  f32x4_t acc = {0.f, 0.f, 0.f, 0.f};
  for (int k0 = kb; k0 < ke; k0 += 32) {
    bf16x8_t a, b;
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      const int k = k0 + 8 * g + j;
      const int row = m0 + r, col = n0 + r;
      float av = 0.f, bv = 0.f;
      if (k < ke && row < M) av = TA ? A[(size_t)k * lda + row] : A[(size_t)row * lda + k];
      if (k < ke && col < N) bv = TB ? B[(size_t)col * ldb + k] : B[(size_t)k * ldb + col];
      a[j] = (short)f32_to_bf16(av);
      b[j] = (short)f32_to_bf16(bv);
    }
    acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, acc, 0, 0, 0);
  }
  const int col = n0 + r;
  if (col >= N) return;
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const int row = m0 + 4 * g + i;
    if (row >= M) continue;
    if (Accum) {
      atomicAdd(&C[(size_t)row * ldc + col], acc[i]);
    } else {
      C[(size_t)row * ldc + col] = acc[i] + (bias ? bias[col] : 0.f);
    }
  }
}

void launch_gemm_bf16(const float* A, int lda, bool ta, const float* B, int ldb, bool tb, float* C, int ldc,
                      const float* bias, int M, int N, int K, bool accum, hipStream_t st) {
  const int ntiles = ((M + 15) / 16) * ((N + 15) / 16);
  const int blocks = (ntiles + 3) / 4;
  int ksplit = 1, kchunk = ((K + 31) / 32) * 32;
  if (accum) {  // split the reduction so that small outputs still fill the GPU
    ksplit = std::max(1, std::min(1024, (K + 1023) / 1024));
    kchunk = (((K + ksplit - 1) / ksplit + 31) / 32) * 32;
    ksplit = (K + kchunk - 1) / kchunk;
  }
  const dim3 grid((unsigned)blocks, (unsigned)ksplit);
#define NCAMD_GEMM(TA_, TB_, AC_)                                                                              \
  hipLaunchKernelGGL((k_gemm_bf16<TA_, TB_, AC_>), grid, dim3(256), 0, st, A, lda, B, ldb, C, ldc, bias, M, N, K, \
                     kchunk)
  if (accum) {
    if (ta && !tb) NCAMD_GEMM(true, false, true);
    else if (!ta && tb) NCAMD_GEMM(false, true, true);
    else if (ta && tb) NCAMD_GEMM(true, true, true);
    else NCAMD_GEMM(false, false, true);
  } else {
    if (ta && !tb) NCAMD_GEMM(true, false, false);
    else if (!ta && tb) NCAMD_GEMM(false, true, false);
    else if (ta && tb) NCAMD_GEMM(true, true, false);
    else NCAMD_GEMM(false, false, false);
  }
#undef NCAMD_GEMM
}

// ---- elementwise pieces
NCD float silu(float z) { return z / (1.f + __expf(-z)); }
NCD float silu_grad(float z) {
  const float s = 1.f / (1.f + __expf(-z));
  return s * (1.f + z * (1.f - s));
}

__global__ void k_silu(const float* __restrict__ z, float* __restrict__ h, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) h[i] = silu(z[i]);
}
__global__ void k_silu_back(const float* __restrict__ z, float* __restrict__ d, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) d[i] *= silu_grad(z[i]);
}
__global__ void k_fill_f32(float* __restrict__ p, int64_t n, float v) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}
// column sums of D [M x N] into out[N] (atomic)
__global__ void k_colsum(const float* __restrict__ D, int M, int N, float* __restrict__ out) {
  const int col = blockIdx.x;
  float s = 0.f;
  for (int r = threadIdx.x; r < M; r += blockDim.x) s += D[(size_t)r * N + col];
  __shared__ float red[256];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int o = blockDim.x / 2; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) atomicAdd(&out[col], red[0]);
}

// input features of an energy: x = log(E/elo)/log(ehi/elo), Fourier terms
NCD void mlp_features(double E, double lelo, double linv, float* f) {
  const float x = (float)((log(E) - lelo) * linv);
  f[0] = x;
  f[1] = __sinf(3.14159265f * x); f[2] = __cosf(3.14159265f * x);
  f[3] = __sinf(6.28318531f * x); f[4] = __cosf(6.28318531f * x);
  f[5] = __sinf(9.42477796f * x); f[6] = __cosf(9.42477796f * x);
  f[7] = x * x;
}
__global__ void k_mlp_features(const double* __restrict__ E, int n, double lelo, double linv, float* __restrict__ X) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) mlp_features(E[i], lelo, linv, X + (size_t)i * kMlpIn);
}

// Mixture negative log-likelihood and its gradient with respect to the 48
// outputs (scaled by 1/n).  Component k uses outputs 6k..6k+5:
//   logit, mu_w, mu_k, log sigma_w, log sigma_k, c  with rho = 0.95 tanh(c).
NCD float mix_nll_grad(const float* y, float w, float kap, float* dy) {
  float lp[kMlpComp];
  float zw[kMlpComp], zk[kMlpComp], sw[kMlpComp], sk[kMlpComp], rho[kMlpComp], om[kMlpComp];
  float lmax = -1e30f;
  for (int c = 0; c < kMlpComp; ++c) lmax = fmaxf(lmax, y[6 * c]);
  float se = 0.f;
  for (int c = 0; c < kMlpComp; ++c) se += __expf(y[6 * c] - lmax);
  const float lse = lmax + __logf(se);
  float best = -1e30f;
  for (int c = 0; c < kMlpComp; ++c) {
    const float* p = y + 6 * c;
    const float lsw = fminf(fmaxf(p[3], -9.f), 3.f), lsk = fminf(fmaxf(p[4], -9.f), 3.f);
    sw[c] = __expf(lsw); sk[c] = __expf(lsk);
    rho[c] = 0.95f * tanhf(p[5]);
    om[c] = 1.f - rho[c] * rho[c];
    zw[c] = (w - p[1]) / sw[c];
    zk[c] = (kap - p[2]) / sk[c];
    const float q = (zw[c] * zw[c] - 2.f * rho[c] * zw[c] * zk[c] + zk[c] * zk[c]) / om[c];
    lp[c] = (p[0] - lse) - 1.83787706f - lsw - lsk - 0.5f * __logf(om[c]) - 0.5f * q;
    best = fmaxf(best, lp[c]);
  }
  float tot = 0.f;
  for (int c = 0; c < kMlpComp; ++c) tot += __expf(lp[c] - best);
  const float ll = best + __logf(tot);
  // responsibilities and per-component gradients of -log L
  for (int c = 0; c < kMlpComp; ++c) {
    const float* p = y + 6 * c;
    float* d = dy + 6 * c;
    const float g = __expf(lp[c] - ll);                    // posterior weight
    const float pik = __expf(p[0] - lse);                  // prior weight
    d[0] = pik - g;                                        // d(-ll)/d logit
    const float a = (zw[c] - rho[c] * zk[c]) / om[c];      // -d log N / d zw
    const float b = (zk[c] - rho[c] * zw[c]) / om[c];
    d[1] = -g * a / sw[c];
    d[2] = -g * b / sk[c];
    const bool clw = p[3] < -9.f || p[3] > 3.f, clk = p[4] < -9.f || p[4] > 3.f;
    d[3] = clw ? 0.f : -g * (a * zw[c] - 1.f);
    d[4] = clk ? 0.f : -g * (b * zk[c] - 1.f);
    // d log N / d rho = rho/om + zw zk/om - rho q/om
    const float q = (zw[c] * zw[c] - 2.f * rho[c] * zw[c] * zk[c] + zk[c] * zk[c]) / om[c];
    const float dlr = (rho[c] + zw[c] * zk[c] - rho[c] * q) / om[c];
    const float th = tanhf(p[5]);
    d[5] = -g * dlr * 0.95f * (1.f - th * th);
  }
  return -ll;
}

__global__ void k_mix_loss(const float* __restrict__ Y, const float* __restrict__ T, int n, float* __restrict__ dY,
                           float* __restrict__ loss, float scale) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  float l = 0.f;
  if (i < n) {
    float d[kMlpOut];
    l = mix_nll_grad(Y + (size_t)i * kMlpOut, T[2 * i], T[2 * i + 1], d);
    for (int j = 0; j < kMlpOut; ++j) dY[(size_t)i * kMlpOut + j] = d[j] * scale;
  }
  __shared__ float red[256];
  red[threadIdx.x] = l;
  __syncthreads();
  for (int o = blockDim.x / 2; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) atomicAdd(loss, red[0] * scale);
}

// Adam on the FP32 master weights
__global__ void k_adam(float* __restrict__ w, const float* __restrict__ g, float* __restrict__ m, float* __restrict__ v,
                       int n, float lr, float b1, float b2, float c1, float c2, float eps) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i];
  const float mi = b1 * m[i] + (1.f - b1) * gi;
  const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  w[i] -= lr * (mi * c1) / (sqrtf(vi * c2) + eps);
}

// training targets: (alpha, beta) of the plain sampler at energy E -> (w, kappa)
// in the neutron's own kinematic coordinates, plus the energy features
__global__ void k_mlp_targets(const double* __restrict__ E, const double* __restrict__ a, const double* __restrict__ b,
                              int n, double kT, double bmax, double lelo, double linv, float* __restrict__ X,
                              float* __restrict__ T) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double eps = E[i] / kT;
  const double r = sqrt(dmax(0.0, b[i] + eps));
  const double wmax = sqrt(bmax + eps);
  const double alo = (r - sqrt(eps)) * (r - sqrt(eps)), ahi = (r + sqrt(eps)) * (r + sqrt(eps));
  double kap = ahi > alo ? (a[i] - alo) / (ahi - alo) : 0.5;
  kap = dclamp(kap, 0.0, 1.0);
  T[2 * i] = (float)(r / wmax);
  T[2 * i + 1] = (float)kap;
  mlp_features(E[i], lelo, linv, X + (size_t)i * kMlpIn);
}

static inline unsigned g1(int64_t n) { return (unsigned)((n + 255) / 256); }

void launch_silu(const float* z, float* h, int64_t n, hipStream_t st) {
  hipLaunchKernelGGL(k_silu, dim3(g1(n)), dim3(256), 0, st, z, h, n);
}
void launch_silu_back(const float* z, float* d, int64_t n, hipStream_t st) {
  hipLaunchKernelGGL(k_silu_back, dim3(g1(n)), dim3(256), 0, st, z, d, n);
}
void launch_fill_f32(float* p, int64_t n, float v, hipStream_t st) {
  hipLaunchKernelGGL(k_fill_f32, dim3(g1(n)), dim3(256), 0, st, p, n, v);
}
void launch_colsum(const float* D, int M, int N, float* out, hipStream_t st) {
  hipLaunchKernelGGL(k_colsum, dim3((unsigned)N), dim3(256), 0, st, D, M, N, out);
}
void launch_mlp_features(const double* E, int n, double lelo, double linv, float* X, hipStream_t st) {
  hipLaunchKernelGGL(k_mlp_features, dim3(g1(n)), dim3(256), 0, st, E, n, lelo, linv, X);
}
void launch_mix_loss(const float* Y, const float* T, int n, float* dY, float* loss, hipStream_t st) {
  hipLaunchKernelGGL(k_mix_loss, dim3(g1(n)), dim3(256), 0, st, Y, T, n, dY, loss, 1.f / (float)n);
}
void launch_adam(float* w, const float* g, float* m, float* v, int n, float lr, int step, hipStream_t st) {
  const float b1 = 0.9f, b2 = 0.999f;
  const float c1 = 1.f / (1.f - powf(b1, (float)step)), c2 = 1.f / (1.f - powf(b2, (float)step));
  hipLaunchKernelGGL(k_adam, dim3(g1(n)), dim3(256), 0, st, w, g, m, v, n, lr, b1, b2, c1, c2, 1e-8f);
}
void launch_mlp_targets(const double* E, const double* a, const double* b, int n, double kT, double bmax,
                        double lelo, double linv, float* X, float* T, hipStream_t st) {
  hipLaunchKernelGGL(k_mlp_targets, dim3(g1(n)), dim3(256), 0, st, E, a, b, n, kT, bmax, lelo, linv, X, T);
}

}  // namespace ncamd
