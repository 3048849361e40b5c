// Kinematic-box envelope sampler of S(alpha,beta): table construction
// kernels and the sampling kernel (device code in nc_kb.hh).
#include "nc_kernels.hh"

namespace ncamd {

// per table cell: the largest of its four corner values of S (the
// log-linear/linear cell interpolation never exceeds it)
__global__ void k_kb_cellmax(SABView s, double* __restrict__ cmax) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int nca = s.na - 1;
  if (t >= (int64_t)nca * (s.nb - 1)) return;
  const int ib = (int)(t / nca), ia = (int)(t % nca);
  const double2* r0 = s.sl + (size_t)ib * s.na + ia;
  const double2* r1 = r0 + s.na;
  cmax[t] = dmax(dmax(r0[0].x, r0[1].x), dmax(r1[0].x, r1[1].x));
}

// lower index of the cell containing x (clamped into the grid)
NCD int kb_cell_clamped(const double* g, int n, double x) {
  const int up = upper_bound_d(g, 0, n, x);
  return up - 1 < 0 ? 0 : (up - 1 > n - 2 ? n - 2 : up - 1);
}

// upper bound of S over the (alpha, beta) image of the sub-box
// [w0,w1] x [k0,k1] of an interval's unit square: maximum cell-corner S over
// every table cell touched by the image's bounding rectangle
NCD double kb_smax_subbox(const SABView& s, const double* __restrict__ cmax, double eps, double seps, double wmax,
                          double w0, double w1, double k0, double k1) {
  const double r0 = w0 * wmax, r1 = w1 * wmax;
  double blo = r0 * r0 - eps, bhi = r1 * r1 - eps;
  // alpha(r, kappa) = r^2 + eps - 2 sqrt(eps) r (1 - 2 kappa): linear in
  // kappa, convex in r (minimum at r = sqrt(eps)(1 - 2 kappa))
  double alo = __builtin_huge_val(), ahi = -__builtin_huge_val();
  for (int c = 0; c < 2; ++c) {
    const double kap = c ? k1 : k0;
    const double f = 2.0 * seps * (1.0 - 2.0 * kap);
    const double a_r0 = r0 * r0 + eps - f * r0, a_r1 = r1 * r1 + eps - f * r1;
    ahi = dmax(ahi, dmax(a_r0, a_r1));
    double amin = dmin(a_r0, a_r1);
    const double rs = 0.5 * f;
    if (rs > r0 && rs < r1) amin = dmin(amin, eps - 0.25 * f * f);
    alo = dmin(alo, amin);
  }
  // margins for the rounding of the sampler's own (alpha, beta) arithmetic
  const double ma = 1e-9 * (1.0 + fabs(alo) + fabs(ahi)) + 1e-9 * eps;
  const double mb = 1e-9 * (1.0 + fabs(blo) + fabs(bhi)) + 1e-9 * eps;
  alo -= ma; ahi += ma; blo -= mb; bhi += mb;
  if (ahi < s.alpha[0] || alo > s.alpha[s.na - 1] || bhi < s.beta[0] || blo > s.beta[s.nb - 1]) return 0.0;
  const int ia0 = kb_cell_clamped(s.alpha, s.na, dmax(alo, s.alpha[0]));
  const int ia1 = kb_cell_clamped(s.alpha, s.na, dmin(ahi, s.alpha[s.na - 1]));
  const int ib0 = kb_cell_clamped(s.beta, s.nb, dmax(blo, s.beta[0]));
  const int ib1 = kb_cell_clamped(s.beta, s.nb, dmin(bhi, s.beta[s.nb - 1]));
  const int nca = s.na - 1;
  double smax = 0.0;
  for (int ib = ib0; ib <= ib1; ++ib)
    for (int ia = ia0; ia <= ia1; ++ia) smax = dmax(smax, cmax[(size_t)ib * nca + ia]);
  return smax;
}

// envelope of t = S * J over each (interval, box): one thread per box; the
// box is split into kSub x kSub sub-boxes, each bounded by its own maximum
// S times J at its upper w (J grows as w^2), and the box takes the largest
constexpr int kKBSub = 4;
__global__ void k_kb_envelope(SABView s, KBView k, const double* __restrict__ cmax, double* __restrict__ env) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (int64_t)k.nn * k.nbox) return;
  const int n = (int)(t / k.nbox), box = (int)(t % k.nbox);
  const int iw = box / k.bk, ik = box - iw * k.bk;
  const double4 par = k.node_par[n];
  const double eps = par.x, seps = par.y, wmax = par.z, jfac = par.w;
  double e = 0.0;
  for (int sw = 0; sw < kKBSub; ++sw) {
    const double w0 = (iw + (double)sw / kKBSub) / k.bw;
    const double w1 = (sw + 1 == kKBSub) ? (double)(iw + 1) / k.bw : (iw + (double)(sw + 1) / kKBSub) / k.bw;
    for (int sk = 0; sk < kKBSub; ++sk) {
      const double k0 = (ik + (double)sk / kKBSub) / k.bk;
      const double k1 = (sk + 1 == kKBSub) ? (double)(ik + 1) / k.bk : (ik + (double)(sk + 1) / kKBSub) / k.bk;
      e = dmax(e, kb_smax_subbox(s, cmax, eps, seps, wmax, w0, w1, k0, k1) * (jfac * w1 * w1));
    }
  }
  env[t] = e * (1.0 + 1e-9);
}

// S(alpha,beta)-only sampling with the envelope sampler; neutrons outside
// its energy range or without an accepted proposal go to the queue of the
// plain-sampler kernel, which restarts their Philox stream from draw 0.
__global__ void __launch_bounds__(256) k_kb_sample_iso(SABView s, KBView k, const double* __restrict__ ekin,
                                                       int64_t n, int64_t total, uint64_t seed, uint32_t stream,
                                                       double* __restrict__ oE, double* __restrict__ oMu,
                                                       unsigned long long* __restrict__ nerr,
                                                       int64_t* __restrict__ queue,
                                                       unsigned long long* __restrict__ qcount, int maxtries) {
  const int64_t gstride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t kk = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; kk < total; kk += gstride) {
    const int64_t i = (total == n) ? kk : (kk % n);
    const double E = ekin[i];
    Rng rng;
    rng.init(seed, stream, (uint64_t)kk);
    double a, b;
    if (kb_sample_ab(s, k, E, rng, a, b, maxtries) != kKBAccepted) {
      const unsigned long long slot = atomicAdd(qcount, 1ull);
      queue[slot] = kk;
      continue;
    }
    double Ef, mu;
    const int rc = sab_ab_to_emu(s, rng, E, a, b, Ef, mu);
    if (rc) { Ef = -1.0; mu = -999.0; atomicAdd(nerr, 1ull); }
    oE[kk] = Ef;
    oMu[kk] = mu;
  }
}

// raw (alpha, beta) of the envelope sampler (statistics tests); fallback
// neutrons get the plain sampler's values
__global__ void __launch_bounds__(256) k_kb_sample_ab(SABView s, KBView k, const double* __restrict__ ekin,
                                                      int64_t n, uint64_t seed, uint32_t stream,
                                                      double* __restrict__ oa, double* __restrict__ ob,
                                                      unsigned long long* __restrict__ nfallback, int maxtries) {
  const int64_t gstride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t kk = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; kk < n; kk += gstride) {
    Rng rng;
    rng.init(seed, stream, (uint64_t)kk);
    double a = -1.0, b = 0.0;
    if (kb_sample_ab(s, k, ekin[kk], rng, a, b, maxtries) != kKBAccepted) {
      atomicAdd(nfallback, 1ull);
      Rng r2;
      r2.init(seed, stream, (uint64_t)kk);
      if (sab_sample_ab<true>(s, ekin[kk], r2, a, b)) { a = -1.0; b = 0.0; }
    }
    oa[kk] = a;
    ob[kk] = b;
  }
}

// the same two kernels for the learned proposal
__global__ void __launch_bounds__(256) k_lp_sample_iso(SABView s, KBView k, LPView lp, const double* __restrict__ ekin,
                                                       int64_t n, int64_t total, uint64_t seed, uint32_t stream,
                                                       double* __restrict__ oE, double* __restrict__ oMu,
                                                       unsigned long long* __restrict__ nerr,
                                                       int64_t* __restrict__ queue,
                                                       unsigned long long* __restrict__ qcount, int maxtries) {
  const int64_t gstride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t kk = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; kk < total; kk += gstride) {
    const int64_t i = (total == n) ? kk : (kk % n);
    const double E = ekin[i];
    Rng rng;
    rng.init(seed, stream, (uint64_t)kk);
    double a, b;
    if (lp_sample_ab(s, k, lp, E, rng, a, b, maxtries) != kKBAccepted) {
      const unsigned long long slot = atomicAdd(qcount, 1ull);
      queue[slot] = kk;
      continue;
    }
    double Ef, mu;
    const int rc = sab_ab_to_emu(s, rng, E, a, b, Ef, mu);
    if (rc) { Ef = -1.0; mu = -999.0; atomicAdd(nerr, 1ull); }
    oE[kk] = Ef;
    oMu[kk] = mu;
  }
}
__global__ void __launch_bounds__(256) k_lp_sample_ab(SABView s, KBView k, LPView lp, const double* __restrict__ ekin,
                                                      int64_t n, uint64_t seed, uint32_t stream,
                                                      double* __restrict__ oa, double* __restrict__ ob,
                                                      unsigned long long* __restrict__ nfallback, int maxtries) {
  const int64_t gstride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t kk = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; kk < n; kk += gstride) {
    Rng rng;
    rng.init(seed, stream, (uint64_t)kk);
    double a = -1.0, b = 0.0;
    if (lp_sample_ab(s, k, lp, ekin[kk], rng, a, b, maxtries) != kKBAccepted) {
      atomicAdd(nfallback, 1ull);
      Rng r2;
      r2.init(seed, stream, (uint64_t)kk);
      if (sab_sample_ab<true>(s, ekin[kk], r2, a, b)) { a = -1.0; b = 0.0; }
    }
    oa[kk] = a;
    ob[kk] = b;
  }
}

static inline unsigned kb_grid(int64_t n) {
  int64_t g = (n + 255) / 256;
  const int64_t gmax = 256 * 64;
  if (g > gmax) g = gmax;
  if (g < 1) g = 1;
  return (unsigned)g;
}

void launch_kb_build(const SABView& s, const KBView& k, double* cmax, double* env, hipStream_t st) {
  const int64_t nc = (int64_t)(s.na - 1) * (s.nb - 1);
  hipLaunchKernelGGL(k_kb_cellmax, dim3((unsigned)((nc + 255) / 256)), dim3(256), 0, st, s, cmax);
  const int64_t nt = (int64_t)k.nn * k.nbox;
  hipLaunchKernelGGL(k_kb_envelope, dim3((unsigned)((nt + 255) / 256)), dim3(256), 0, st, s, k, cmax, env);
}

void launch_kb_sample_iso(const SABView& s, const KBView& k, const double* ekin, int64_t n, int64_t repeat,
                          uint64_t seed, uint32_t stream, double* oE, double* oMu, unsigned long long* nerr,
                          int64_t* queue, unsigned long long* qcount, int maxtries, hipStream_t st) {
  const int64_t total = n * repeat;
  if (total <= 0) return;
  launch_zero_u64(qcount, st);
  hipLaunchKernelGGL(k_kb_sample_iso, dim3(kb_grid(total)), dim3(256), 0, st, s, k, ekin, n, total, seed, stream, oE,
                     oMu, nerr, queue, qcount, maxtries);
  launch_sab_queue(s, ekin, n, total, seed, stream, oE, oMu, nerr, queue, qcount, st);
}

void launch_lp_sample_iso(const SABView& s, const KBView& k, const LPView& lp, const double* ekin, int64_t n,
                          int64_t repeat, uint64_t seed, uint32_t stream, double* oE, double* oMu,
                          unsigned long long* nerr, int64_t* queue, unsigned long long* qcount, int maxtries,
                          hipStream_t st) {
  const int64_t total = n * repeat;
  if (total <= 0) return;
  launch_zero_u64(qcount, st);
  hipLaunchKernelGGL(k_lp_sample_iso, dim3(kb_grid(total)), dim3(256), 0, st, s, k, lp, ekin, n, total, seed, stream,
                     oE, oMu, nerr, queue, qcount, maxtries);
  launch_sab_queue(s, ekin, n, total, seed, stream, oE, oMu, nerr, queue, qcount, st);
}
void launch_lp_sample_ab(const SABView& s, const KBView& k, const LPView& lp, const double* ekin, int64_t n,
                         uint64_t seed, uint32_t stream, double* oa, double* ob, unsigned long long* nfallback,
                         int maxtries, hipStream_t st) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_lp_sample_ab, dim3(kb_grid(n)), dim3(256), 0, st, s, k, lp, ekin, n, seed, stream, oa, ob,
                     nfallback, maxtries);
}

void launch_kb_sample_ab(const SABView& s, const KBView& k, const double* ekin, int64_t n, uint64_t seed,
                         uint32_t stream, double* oa, double* ob, unsigned long long* nfallback, int maxtries,
                         hipStream_t st) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_kb_sample_ab, dim3(kb_grid(n)), dim3(256), 0, st, s, k, ekin, n, seed, stream, oa, ob,
                     nfallback, maxtries);
}

}  // namespace ncamd
