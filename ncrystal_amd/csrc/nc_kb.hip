// Box-envelope sampler of S(alpha,beta): the kernel that bounds t over every
// box, and the sampling kernels (device code in nc_kb.hh).
#include "nc_kernels.hh"

namespace ncamd {

// S along table row ib at alpha inside cell ia (log-linear in alpha, linear
// where a corner value is zero), as the sampler's cell evaluation computes it
NCD double kb_row(const SABView& s, int ib, int ia, double a) {
  const double2* r = s.sl + (size_t)ib * s.na + ia;
  const double2 c0 = r[0], c1 = r[1];
  return kb_loglin(s.alpha[ia], c0.x, s.alpha[ia + 1], c1.x, a, c0.y, c1.y);
}

// lower index of the cell containing x (clamped into the grid)
NCD int kb_cell_clamped(const double* g, int n, double x) {
  const int up = upper_bound_d(g, 0, n, x);
  return up - 1 < 0 ? 0 : (up - 1 > n - 2 ? n - 2 : up - 1);
}

// Bounds of S over the rectangle [alo,ahi] x [blo,bhi] inside the table grid.
// On a piece (rectangle and one cell) S = (1 - tau) s0(alpha) + tau s1(alpha)
// with the two rows s0, s1 log-linear (or linear) in alpha: each row is
// convex and monotone on the cell, and S is linear in tau.  So S is convex in
// alpha and linear in beta, and its maximum over the piece is at one of the
// piece's corners; its minimum is at least min over the tau end points of
// (1 - tau) min(s0) + tau min(s1), with each row's minimum at an alpha end
// point of the piece.
NCD void kb_sbounds(const SABView& s, double alo, double ahi, double blo, double bhi, double& smax, double& smin) {
  const int ia0 = kb_cell_clamped(s.alpha, s.na, alo), ia1 = kb_cell_clamped(s.alpha, s.na, ahi);
  const int ib0 = kb_cell_clamped(s.beta, s.nb, blo), ib1 = kb_cell_clamped(s.beta, s.nb, bhi);
  double mx = 0.0, mn = __builtin_huge_val();
  for (int ib = ib0; ib <= ib1; ++ib) {
    const double b0 = s.beta[ib], b1 = s.beta[ib + 1];
    const double tA = (dmax(blo, b0) - b0) / (b1 - b0), tB = (dmin(bhi, b1) - b0) / (b1 - b0);
    for (int ia = ia0; ia <= ia1; ++ia) {
      const double aA = dmax(alo, s.alpha[ia]), aB = dmin(ahi, s.alpha[ia + 1]);
      const double r0A = kb_row(s, ib, ia, aA), r0B = kb_row(s, ib, ia, aB);
      const double r1A = kb_row(s, ib + 1, ia, aA), r1B = kb_row(s, ib + 1, ia, aB);
      mx = dmax(mx, dmax(dmax(r0A + (r1A - r0A) * tA, r0A + (r1A - r0A) * tB),
                         dmax(r0B + (r1B - r0B) * tA, r0B + (r1B - r0B) * tB)));
      const double m0 = dmin(r0A, r0B), m1 = dmin(r1A, r1B);
      mn = dmin(mn, dmin(m0 + (m1 - m0) * tA, m0 + (m1 - m0) * tB));
    }
  }
  smax = mx;
  smin = dmax(mn, 0.0);
}

// Bounds of S over the (alpha, beta) image of the sub-box [w0,w1] x [k0,k1]
// of an interval's unit square, through the image's bounding rectangle
// (widened by margins for the rounding of the sampler's own arithmetic).
// Points of the rectangle outside the table have S = 0.
NCD void kb_subbox_bounds(const SABView& s, double eps, double seps, double wmax, double w0, double w1, double k0,
                          double k1, double& smax, double& smin) {
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
  const double ma = 1e-9 * (1.0 + fabs(alo) + fabs(ahi)) + 1e-9 * eps;
  const double mb = 1e-9 * (1.0 + fabs(blo) + fabs(bhi)) + 1e-9 * eps;
  alo -= ma; ahi += ma; blo -= mb; bhi += mb;
  const double A0 = s.alpha[0], A1 = s.alpha[s.na - 1], B0 = s.beta[0], B1 = s.beta[s.nb - 1];
  if (ahi < A0 || alo > A1 || bhi < B0 || blo > B1) { smax = 0.0; smin = 0.0; return; }
  const bool inside = alo >= A0 && ahi <= A1 && blo >= B0 && bhi <= B1;
  kb_sbounds(s, dmax(alo, A0), dmin(ahi, A1), dmax(blo, B0), dmin(bhi, B1), smax, smin);
  if (!inside) smin = 0.0;
}

// Bounds of t = S * J over each box (boxes of every interval, possibly of
// several candidate grids: interval = (index / nbox) % nn).  The box is split
// into kKBSub x kKBSub sub-boxes; J = jfac w^2 is bounded by its values at
// the sub-box's w edges.
constexpr int kKBSub = 4;
__global__ void k_kb_bounds(SABView s, const double4* __restrict__ node_par, int nn, int nbox, int64_t total,
                            KBBox* __restrict__ box) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= total) return;
  const int n = (int)((t / nbox) % nn);
  const double4 par = node_par[n];
  const double eps = par.x, seps = par.y, wmax = par.z, jfac = par.w;
  KBBox b = box[t];
  const double w0 = b.w0, w1 = b.w1, k0 = b.k0, k1 = b.k1;
  double env = 0.0, tlow = __builtin_huge_val();
  for (int sw = 0; sw < kKBSub; ++sw) {
    const double sw0 = w0 + (w1 - w0) * ((double)sw / kKBSub);
    const double sw1 = (sw + 1 == kKBSub) ? w1 : w0 + (w1 - w0) * ((double)(sw + 1) / kKBSub);
    for (int sk = 0; sk < kKBSub; ++sk) {
      const double sk0 = k0 + (k1 - k0) * ((double)sk / kKBSub);
      const double sk1 = (sk + 1 == kKBSub) ? k1 : k0 + (k1 - k0) * ((double)(sk + 1) / kKBSub);
      double smax, smin;
      kb_subbox_bounds(s, eps, seps, wmax, sw0, sw1, sk0, sk1, smax, smin);
      env = dmax(env, smax * (jfac * sw1 * sw1));
      tlow = dmin(tlow, smin * (jfac * sw0 * sw0));
    }
  }
  b.env = env * (1.0 + 1e-9);
  b.tlow = tlow * (1.0 - 1e-9);
  box[t] = b;
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

static inline unsigned kb_grid(int64_t n) {
  int64_t g = (n + 255) / 256;
  const int64_t gmax = 256 * 64;
  if (g > gmax) g = gmax;
  if (g < 1) g = 1;
  return (unsigned)g;
}

void launch_kb_bounds(const SABView& s, const double4* node_par, int nn, int nbox, int64_t total, KBBox* box,
                      hipStream_t st) {
  if (total <= 0) return;
  hipLaunchKernelGGL(k_kb_bounds, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, st, s, node_par, nn, nbox,
                     total, box);
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

void launch_kb_sample_ab(const SABView& s, const KBView& k, const double* ekin, int64_t n, uint64_t seed,
                         uint32_t stream, double* oa, double* ob, unsigned long long* nfallback, int maxtries,
                         hipStream_t st) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_kb_sample_ab, dim3(kb_grid(n)), dim3(256), 0, st, s, k, ekin, n, seed, stream, oa, ob,
                     nfallback, maxtries);
}

}  // namespace ncamd
