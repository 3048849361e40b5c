// Small MLP (features -> 3 x 64 hidden -> 48 outputs) trained and evaluated
// on the GPU: BF16 matrix products on the CDNA4 matrix cores
// (v_mfma_f32_16x16x32_bf16, FP32 accumulation), FP32 elementwise work and
// FP32 Adam with FP32 master weights.  A training step is one kernel
// (k_mlp_step): features and targets, forward, mixture likelihood, backward
// and, in the last workgroup to finish, Adam; inference uses the separate
// matrix-product and activation kernels.
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

// ---------------------------------------------------- fused training step
// One workgroup (four waves) takes kStepRows samples.  Everything of the
// step lives in LDS: the BF16 weights, the samples' features and targets,
// the pre-activations Z (FP32, for the SiLU derivative), the activations H
// (BF16, matrix-core operands), the outputs and their gradient dY, and the
// backpropagated dZ.  Forward: wave w owns samples 16w..16w+15 of the
// workgroup and computes all their columns.  Backward: the output tiles of
// each product are dealt round the waves.  Weight and bias gradients are
// summed over the workgroup's samples by the matrix cores and added into
// the step's gradient with device-scope atomics; the last workgroup to
// finish reads the gradient back at device scope and applies Adam (or,
// without update, leaves the gradient for the caller).  Operand layout of
// v_mfma_f32_16x16x32_bf16 as in k_gemm_bf16.
constexpr int kStepRows = 64;

struct MlpStepLds {
  unsigned short w1[kMlpIn * kMlpHidden];
  unsigned short w2[kMlpHidden * kMlpHidden];
  unsigned short w3[kMlpHidden * kMlpHidden];
  unsigned short w4[kMlpHidden * kMlpOut];
  float bias[3 * kMlpHidden + kMlpOut];
  float x[kStepRows * kMlpIn];
  float t[kStepRows * 2];
  float z[3][kStepRows * kMlpHidden];
  unsigned short h[3][kStepRows * kMlpHidden];
  float y[kStepRows * kMlpOut];
  float d[2][kStepRows * kMlpHidden];
  float red[4];
  int last;
};

NCD short bfs(float f) { return (short)f32_to_bf16(f); }

// Z = A W + b for the wave's 16 rows, K = 64 (A: BF16 activations), N = 16 nc
// columns; writes z (FP32) and h = silu(z) (BF16), or y when out != nullptr
NCD void step_layer_fwd(const unsigned short* __restrict__ A, const unsigned short* __restrict__ W, int N,
                        const float* __restrict__ bias, float* __restrict__ z, unsigned short* __restrict__ h,
                        float* __restrict__ out, int wv, int r, int g) {
  bf16x8_t a0, a1;
  for (int j = 0; j < 8; ++j) {
    a0[j] = (short)A[(16 * wv + r) * kMlpHidden + 8 * g + j];
    a1[j] = (short)A[(16 * wv + r) * kMlpHidden + 32 + 8 * g + j];
  }
  for (int c = 0; c < N / 16; ++c) {
    bf16x8_t b0, b1;
    for (int j = 0; j < 8; ++j) {
      b0[j] = (short)W[(8 * g + j) * N + 16 * c + r];
      b1[j] = (short)W[(32 + 8 * g + j) * N + 16 * c + r];
    }
    f32x4_t acc = {0.f, 0.f, 0.f, 0.f};
    acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a0, b0, acc, 0, 0, 0);
    acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a1, b1, acc, 0, 0, 0);
    const int col = 16 * c + r;
    for (int i = 0; i < 4; ++i) {
      const int row = 16 * wv + 4 * g + i;
      const float v = acc[i] + bias[col];
      if (out) {
        out[row * N + col] = v;
      } else {
        z[row * kMlpHidden + col] = v;
        h[row * kMlpHidden + col] = f32_to_bf16(silu(v));
      }
    }
  }
}

// gradients of a hidden layer (64 -> N outputs) from dZ (FP32, ld N):
// dW[in][out] += sum_s Hprev[s][in] dZ[s][out] (atomics into gw), db[out] +=
// sum_s dZ[s][out], and, when dprev != nullptr, dprev[s][in] = (sum_out
// dZ[s][out] W[in][out]) * silu'(zprev[s][in])
NCD void step_layer_bwd(const float* __restrict__ dZ, int N, const unsigned short* __restrict__ Hprev,
                        const unsigned short* __restrict__ W, const float* __restrict__ zprev,
                        float* __restrict__ dprev, float* __restrict__ gw, float* __restrict__ gb, int tid, int wv,
                        int r, int g) {
  // weight gradient: (64 / 16) x (N / 16) tiles over the waves, K = the 64 samples
  const int tn = N / 16, ntile = 4 * tn;
  for (int t = wv; t < ntile; t += 4) {
    const int ti = t / tn, tj = t % tn;
    f32x4_t acc = {0.f, 0.f, 0.f, 0.f};
    for (int ks = 0; ks < 2; ++ks) {
      bf16x8_t av, bv;
      for (int j = 0; j < 8; ++j) {
        const int s = 32 * ks + 8 * g + j;
        av[j] = (short)Hprev[s * kMlpHidden + 16 * ti + r];
        bv[j] = bfs(dZ[s * N + 16 * tj + r]);
      }
      acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av, bv, acc, 0, 0, 0);
    }
    for (int i = 0; i < 4; ++i) atomicAdd(&gw[(16 * ti + 4 * g + i) * N + 16 * tj + r], acc[i]);
  }
  if (tid < N) {
    float sum = 0.f;
    for (int s = 0; s < kStepRows; ++s) sum += dZ[s * N + tid];
    atomicAdd(&gb[tid], sum);
  }
  if (!dprev) return;
  // backpropagated gradient: the wave's 16 samples x 64 inputs, K = N
  bf16x8_t a[4];
  for (int ks = 0; ks < 2; ++ks)
    for (int j = 0; j < 8; ++j) {
      const int k = 32 * ks + 8 * g + j;
      a[ks][j] = k < N ? bfs(dZ[(16 * wv + r) * N + k]) : (short)0;
    }
  for (int c = 0; c < 4; ++c) {
    f32x4_t acc = {0.f, 0.f, 0.f, 0.f};
    for (int ks = 0; ks < 2; ++ks) {
      bf16x8_t bv;
      for (int j = 0; j < 8; ++j) {
        const int k = 32 * ks + 8 * g + j;
        bv[j] = k < N ? (short)W[(16 * c + r) * N + k] : (short)0;
      }
      acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[ks], bv, acc, 0, 0, 0);
    }
    const int col = 16 * c + r;
    for (int i = 0; i < 4; ++i) {
      const int row = 16 * wv + 4 * g + i;
      dprev[row * kMlpHidden + col] = acc[i] * silu_grad(zprev[row * kMlpHidden + col]);
    }
  }
}

struct MlpStepArgs {
  float* w;                 // FP32 master parameters
  float* grad;              // gradient accumulators (zero on entry)
  float* m;                 // Adam moments
  float* v;
  const double* E;          // samples (E, alpha, beta) of the plain sampler, or
  const double* sa;
  const double* sb;
  const float* X;           // ... precomputed features and targets (E null)
  const float* T;
  int n;
  double kT, bmax, lelo, linv;
  float scale;              // 1 / n
  float* loss;              // += mean loss of the step
  unsigned* counter;        // finished workgroups (zero on entry, reset at exit)
  int update;               // apply Adam (and clear the gradient)
  float lr, c1, c2;         // Adam step size and bias corrections
};

__global__ void __launch_bounds__(256) k_mlp_step(MlpStepArgs a) {
  __shared__ MlpStepLds L;
  const int tid = (int)threadIdx.x, lane = tid & 63, wv = tid >> 6, r = lane & 15, g = lane >> 4;
  // weights (BF16) and biases
  for (int i = tid; i < kMlpParams; i += 256) {
    const float x = a.w[i];
    if (i < kOffB1) L.w1[i] = f32_to_bf16(x);
    else if (i < kOffW2) L.bias[i - kOffB1] = x;
    else if (i < kOffB2) L.w2[i - kOffW2] = f32_to_bf16(x);
    else if (i < kOffW3) L.bias[kMlpHidden + i - kOffB2] = x;
    else if (i < kOffB3) L.w3[i - kOffW3] = f32_to_bf16(x);
    else if (i < kOffW4) L.bias[2 * kMlpHidden + i - kOffB3] = x;
    else if (i < kOffB4) L.w4[i - kOffW4] = f32_to_bf16(x);
    else L.bias[3 * kMlpHidden + i - kOffB4] = x;
  }
  // features and targets: (alpha, beta) at energy E -> (w, kappa) in the
  // neutron's kinematic coordinates
  const int s0 = (int)blockIdx.x * kStepRows;
  if (tid < kStepRows) {
    const int s = s0 + tid;
    float f[kMlpIn];
    float tw = 0.f, tk = 0.f;
    for (int j = 0; j < kMlpIn; ++j) f[j] = 0.f;
    if (s < a.n) {
      if (a.E) {
        const double E = a.E[s];
        const double eps = E / a.kT;
        const double rr = sqrt(dmax(0.0, a.sb[s] + eps));
        const double wmax = sqrt(a.bmax + eps);
        const double alo = (rr - sqrt(eps)) * (rr - sqrt(eps)), ahi = (rr + sqrt(eps)) * (rr + sqrt(eps));
        double kap = ahi > alo ? (a.sa[s] - alo) / (ahi - alo) : 0.5;
        kap = dclamp(kap, 0.0, 1.0);
        tw = (float)(rr / wmax);
        tk = (float)kap;
        mlp_features(E, a.lelo, a.linv, f);
      } else {
        for (int j = 0; j < kMlpIn; ++j) f[j] = a.X[(size_t)s * kMlpIn + j];
        tw = a.T[2 * (size_t)s];
        tk = a.T[2 * (size_t)s + 1];
      }
    }
    for (int j = 0; j < kMlpIn; ++j) L.x[tid * kMlpIn + j] = f[j];
    L.t[2 * tid] = tw;
    L.t[2 * tid + 1] = tk;
  }
  __syncthreads();
  // layer 1 (K = 8, padded to one 32-step)
  {
    bf16x8_t av;
    for (int j = 0; j < 8; ++j) av[j] = g == 0 ? bfs(L.x[(16 * wv + r) * kMlpIn + j]) : (short)0;
    for (int c = 0; c < 4; ++c) {
      bf16x8_t bv;
      for (int j = 0; j < 8; ++j) bv[j] = g == 0 ? (short)L.w1[j * kMlpHidden + 16 * c + r] : (short)0;
      f32x4_t acc = {0.f, 0.f, 0.f, 0.f};
      acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av, bv, acc, 0, 0, 0);
      const int col = 16 * c + r;
      for (int i = 0; i < 4; ++i) {
        const int row = 16 * wv + 4 * g + i;
        const float v = acc[i] + L.bias[col];
        L.z[0][row * kMlpHidden + col] = v;
        L.h[0][row * kMlpHidden + col] = f32_to_bf16(silu(v));
      }
    }
  }
  __syncthreads();
  step_layer_fwd(L.h[0], L.w2, kMlpHidden, L.bias + kMlpHidden, L.z[1], L.h[1], nullptr, wv, r, g);
  __syncthreads();
  step_layer_fwd(L.h[1], L.w3, kMlpHidden, L.bias + 2 * kMlpHidden, L.z[2], L.h[2], nullptr, wv, r, g);
  __syncthreads();
  step_layer_fwd(L.h[2], L.w4, kMlpOut, L.bias + 3 * kMlpHidden, nullptr, nullptr, L.y, wv, r, g);
  __syncthreads();
  // mixture likelihood: dY (scaled by 1/n) replaces Y; samples past n add nothing
  if (tid < kStepRows) {
    float d[kMlpOut];
    float l = 0.f;
    if (s0 + tid < a.n) {
      l = mix_nll_grad(L.y + tid * kMlpOut, L.t[2 * tid], L.t[2 * tid + 1], d) * a.scale;
      for (int j = 0; j < kMlpOut; ++j) d[j] *= a.scale;
    } else {
      for (int j = 0; j < kMlpOut; ++j) d[j] = 0.f;
    }
    for (int j = 0; j < kMlpOut; ++j) L.y[tid * kMlpOut + j] = d[j];
    for (int o = 32; o > 0; o >>= 1) l += __shfl_xor(l, o);
    if (tid == 0) atomicAdd(a.loss, l);
  }
  __syncthreads();
  step_layer_bwd(L.y, kMlpOut, L.h[2], L.w4, L.z[2], L.d[0], a.grad + kOffW4, a.grad + kOffB4, tid, wv, r, g);
  __syncthreads();
  step_layer_bwd(L.d[0], kMlpHidden, L.h[1], L.w3, L.z[1], L.d[1], a.grad + kOffW3, a.grad + kOffB3, tid, wv, r, g);
  __syncthreads();
  step_layer_bwd(L.d[1], kMlpHidden, L.h[0], L.w2, L.z[0], L.d[0], a.grad + kOffW2, a.grad + kOffB2, tid, wv, r, g);
  __syncthreads();
  // layer 1: dW1[in][out] = sum_s X[s][in] dZ1[s][out] (8 rows of the 16-row tile)
  {
    const float* dZ = L.d[0];
    f32x4_t acc = {0.f, 0.f, 0.f, 0.f};
    for (int ks = 0; ks < 2; ++ks) {
      bf16x8_t av, bv;
      for (int j = 0; j < 8; ++j) {
        const int s = 32 * ks + 8 * g + j;
        av[j] = r < kMlpIn ? bfs(L.x[s * kMlpIn + r]) : (short)0;
        bv[j] = bfs(dZ[s * kMlpHidden + 16 * wv + r]);
      }
      acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av, bv, acc, 0, 0, 0);
    }
    for (int i = 0; i < 4; ++i) {
      const int row = 4 * g + i;
      if (row < kMlpIn) atomicAdd(&a.grad[kOffW1 + row * kMlpHidden + 16 * wv + r], acc[i]);
    }
    if (tid < kMlpHidden) {
      float sum = 0.f;
      for (int s = 0; s < kStepRows; ++s) sum += dZ[s * kMlpHidden + tid];
      atomicAdd(&a.grad[kOffB1 + tid], sum);
    }
  }
  // the last workgroup: Adam over the complete gradient
  __threadfence();
  __syncthreads();
  if (tid == 0) L.last = atomicAdd(a.counter, 1u) == gridDim.x - 1 ? 1 : 0;
  __syncthreads();
  if (!L.last) return;
  __threadfence();
  if (a.update) {
    const float b1 = 0.9f, b2 = 0.999f;
    constexpr int kChunk = 8;
    for (int i0 = tid; i0 < kMlpParams; i0 += 256 * kChunk) {
      float gi[kChunk];
      for (int q = 0; q < kChunk; ++q) {
        const int i = i0 + 256 * q;
        gi[q] = i < kMlpParams ? __hip_atomic_load(&a.grad[i], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) : 0.f;
      }
      for (int q = 0; q < kChunk; ++q) {
        const int i = i0 + 256 * q;
        if (i >= kMlpParams) break;
        const float mi = b1 * a.m[i] + (1.f - b1) * gi[q];
        const float vi = b2 * a.v[i] + (1.f - b2) * gi[q] * gi[q];
        a.m[i] = mi;
        a.v[i] = vi;
        a.w[i] -= a.lr * (mi * a.c1) / (sqrtf(vi * a.c2) + 1e-8f);
        a.grad[i] = 0.f;
      }
    }
  }
  if (tid == 0) *a.counter = 0u;
}

void launch_mlp_step(float* w, float* grad, float* m, float* v, const double* E, const double* sa, const double* sb,
                     const float* X, const float* T, int n, double kT, double bmax, double lelo, double linv,
                     float* loss, unsigned* counter, bool update, float lr, int step, hipStream_t st) {
  MlpStepArgs a;
  a.w = w; a.grad = grad; a.m = m; a.v = v;
  a.E = E; a.sa = sa; a.sb = sb; a.X = X; a.T = T;
  a.n = n; a.kT = kT; a.bmax = bmax; a.lelo = lelo; a.linv = linv;
  a.scale = 1.f / (float)n;
  a.loss = loss; a.counter = counter;
  a.update = update ? 1 : 0;
  a.lr = lr;
  a.c1 = update ? 1.f / (1.f - powf(0.9f, (float)step)) : 1.f;
  a.c2 = update ? 1.f / (1.f - powf(0.999f, (float)step)) : 1.f;
  hipLaunchKernelGGL(k_mlp_step, dim3((unsigned)((n + kStepRows - 1) / kStepRows)), dim3(256), 0, st, a);
}

static inline unsigned g1(int64_t n) { return (unsigned)((n + 255) / 256); }

void launch_silu(const float* z, float* h, int64_t n, hipStream_t st) {
  hipLaunchKernelGGL(k_silu, dim3(g1(n)), dim3(256), 0, st, z, h, n);
}
void launch_mlp_features(const double* E, int n, double lelo, double linv, float* X, hipStream_t st) {
  hipLaunchKernelGGL(k_mlp_features, dim3(g1(n)), dim3(256), 0, st, E, n, lelo, linv, X);
}

}  // namespace ncamd
