// Exact rejection sampler of S(alpha,beta) in kinematic coordinates.
//
// Target: the density of (alpha, beta) that the scattering kernel defines at
// neutron energy E, t(alpha,beta) ~ S(alpha,beta) on the kinematically
// allowed region K(E) = {beta > -E/kT, alpha_-(beta) <= alpha <= alpha_+(beta)},
// with S interpolated inside each table cell as NCrystal's SABCellEval does
// (ncrystal_core/src/sab/NCSABEval.cc:617-632: log-linear in alpha, then
// linear in beta; ncrystal_core/include/NCrystal/internal/sab/NCSABUtils.hh:176-195
// for the log-linear form with linear fallback).
//
// Energies are split into intervals [E_n, E_n+1].  For an interval, with
// eps = E_n+1/kT, the coordinates
//   w     = sqrt(beta + eps) / sqrt(beta_max + eps)          in (0, 1]
//   kappa = (alpha - alpha_-(beta)) / (alpha_+(beta) - alpha_-(beta)) in [0, 1]
// map K(E_n+1) onto the unit square, and K(E) lies inside it for every E of
// the interval.  In (w, kappa) the target is t = S * J(w) with
// J = 8 sqrt(eps) wmax^3 w^2, and it is smooth up to the kinematic edges, so
// a piecewise-constant envelope on a grid of boxes fits it closely.  Each
// box's envelope value bounds t from above over the box (maximum of the
// cell-corner S over every table cell the box touches, times max J), the
// box is picked with probability proportional to envelope x area (alias
// table), a point is drawn uniformly in the box and accepted with
// probability t/envelope.  The accepted points follow t exactly.
#pragma once
#include "nc_dev.hh"

namespace ncamd {

struct KBView {
  int nn;                          // energy intervals
  int bw, bk, nbox;                // boxes per interval: bw (w) x bk (kappa)
  double elo, ehi, lr_inv;         // interval index guess: floor(log(E/elo) * lr_inv)
  const double* node_e;            // [nn+1] interval edges
  const double4* node_par;         // [nn]: (eps, sqrt(eps), wmax, 8 sqrt(eps) wmax^3)
  const double* prob;              // [nn*nbox] alias acceptance
  const int* alias;                // [nn*nbox] alias target
  const double* env;               // [nn*nbox] envelope of t per box
  // cell searches
  const int* a_bucket;             // IEEE-bit buckets of alpha: upper_bound(alpha, lower edge)
  long long a_base;
  int a_shift, a_nbkt;
  const int* b_guide;              // uniform buckets of beta: upper_bound(beta, beta0 + g/b_scale)
  int b_ng;
  double b_scale;
};

// cell of the table holding (alpha, beta); both inside the grid
NCD int kb_alpha_cell(const SABView& s, const KBView& k, double a) {
  if (!(a < s.alpha[s.na - 1])) return s.na - 2;
  const long long b = (long long)((unsigned long long)__double_as_longlong(a) >> k.a_shift) - k.a_base;
  const int lo = k.a_bucket[b < 0 ? 0 : b], hi = k.a_bucket[b + 1];
  const int up = upper_bound_d(s.alpha, lo, hi, a);
  return up - 1 < 0 ? 0 : (up - 1 > s.na - 2 ? s.na - 2 : up - 1);
}
NCD int kb_beta_cell(const SABView& s, const KBView& k, double b) {
  long long g = (long long)((b - s.beta[0]) * k.b_scale);
  g = g < 0 ? 0 : (g >= k.b_ng ? k.b_ng - 1 : g);
  const int lo = k.b_guide[g > 0 ? g - 1 : 0];
  const int hi = k.b_guide[g + 2 <= k.b_ng ? g + 2 : k.b_ng];
  const int up = upper_bound_d(s.beta, lo, hi, b);
  return up - 1 < 0 ? 0 : (up - 1 > s.nb - 2 ? s.nb - 2 : up - 1);
}

// interpolate_loglin_fallbacklinlin_fast (NCSABUtils.hh:176-195)
NCD double kb_loglin(double a, double fa, double b, double fb, double x, double lfa, double lfb) {
  const double bma = b - a;
  const double mid = 0.5 * (b + a);
  const bool lin = (fa * fb == 0.0);
  if (x < mid) {
    const double r = (x - a) / bma;
    return lin ? (fa + (fb - fa) * r) : exp(lfa + (lfb - lfa) * r);
  }
  const double sx = (b - x) / bma;
  return lin ? (fb + (fa - fb) * sx) : exp(lfb + (lfa - lfb) * sx);
}

// S(alpha,beta) inside the table grid (SABCellEval, alpha first)
NCD double kb_eval_s(const SABView& s, const KBView& k, double a, double b) {
  const int ia = kb_alpha_cell(s, k, a);
  const int ib = kb_beta_cell(s, k, b);
  const double a0 = s.alpha[ia], a1 = s.alpha[ia + 1];
  const double b0 = s.beta[ib], b1 = s.beta[ib + 1];
  const double2* r0 = s.sl + (size_t)ib * s.na + ia;
  const double2* r1 = r0 + s.na;
  const double2 c00 = r0[0], c10 = r0[1], c01 = r1[0], c11 = r1[1];
  const double s0 = kb_loglin(a0, c00.x, a1, c10.x, a, c00.y, c10.y);
  const double s1 = kb_loglin(a0, c01.x, a1, c11.x, a, c01.y, c11.y);
  return s0 + (s1 - s0) * ((b - b0) / (b1 - b0));
}

enum : int { kKBAccepted = 0, kKBFallback = 1 };

// One neutron: energy interval lookup, then up to maxtries proposals.
// Returns kKBFallback when E is outside the covered range or no proposal was
// accepted (the caller then uses the plain sampler from the first draw).
NCD int kb_sample_ab(const SABView& s, const KBView& k, double E, Rng& rng, double& a_out, double& b_out,
                     int maxtries) {
  if (!(E >= k.elo && E < k.ehi)) return kKBFallback;
  int n = (int)(log(E / k.elo) * k.lr_inv);
  n = n < 0 ? 0 : (n > k.nn - 1 ? k.nn - 1 : n);
  if (E < k.node_e[n]) --n;
  else if (!(E < k.node_e[n + 1])) ++n;
  if (n < 0 || n >= k.nn) return kKBFallback;
  const double4 par = k.node_par[n];
  const double eps = par.x, seps = par.y, wmax = par.z, jfac = par.w;
  const double ediv = E / s.kT;
  const double sed = sqrt(ediv);
  const double* pr = k.prob + (size_t)n * k.nbox;
  const int* al = k.alias + (size_t)n * k.nbox;
  const double* ev = k.env + (size_t)n * k.nbox;
  const double amin = s.alpha[0], amax = s.alpha[s.na - 1];
  const double bmin = s.beta[0], bmax = s.beta[s.nb - 1];
  for (int t = 0; t < maxtries; ++t) {
    // box from the alias table, then a uniform point inside it
    const double u0 = rng.gen() * k.nbox;
    int box = (int)u0;
    box = box >= k.nbox ? k.nbox - 1 : box;
    if (pr[box] < u0 - (double)box) box = al[box];
    const int iw = box / k.bk, ik = box - iw * k.bk;
    const double w = (iw + rng.gen()) / k.bw;
    const double kap = (ik + rng.gen()) / k.bk;
    const double u = rng.gen();
    const double r = w * wmax;
    const double beta = r * r - eps;
    const double alpha = r * r + eps - 2.0 * seps * r * (1.0 - 2.0 * kap);
    // kinematics at the neutron's own energy and the table range
    if (!(beta > -ediv) || beta < bmin || beta > bmax || alpha < amin || alpha > amax) continue;
    const double re = sqrt(ediv + beta);
    const double alo = (re - sed) * (re - sed), ahi = (re + sed) * (re + sed);
    if (alpha < alo || alpha > ahi) continue;
    const double tv = kb_eval_s(s, k, alpha, beta) * (jfac * w * w);
    if (u * ev[box] <= tv) {
      a_out = alpha;
      b_out = beta;
      return kKBAccepted;
    }
  }
  return kKBFallback;
}

// ---------------------------------------------------------------------------
// Learned proposal: per energy interval a mixture of 8 bivariate Gaussians
// over (w, kappa) evaluated by the MLP at the interval's upper energy, plus
// non-negative patches on a coarse box grid that restore the envelope
// property: with M the interval's scale,
//   M * (q_mlp(x) + patch(box(x))) >= t(x)  for every x of the unit square,
// which holds because each patch covers the largest deficit
// max(env_j / M - min_{box j} q_mlp, 0) over the fine boxes j inside it
// (env_j bounds t on fine box j; min_{box j} q_mlp is bounded from below by
// the sum over components of each component's minimum over the box, which is
// at a vertex because a Gaussian's superlevel sets are convex).  Proposals: the mixture with probability 1/Z, else a
// patch box by its mass (Z = 1 + total patch mass); points outside the unit
// square are rejected; a point is accepted with probability
// t / (M (q_mlp + patch)).  Accepted points follow t exactly.
constexpr int kLPComp = 8;
struct LPComp {
  double cpi;              // cumulative weight
  double mw, mk;           // means
  double l11, l21, l22;    // Cholesky factor of the covariance
  double iw, ik, r, lognorm;  // 1/sigma_w, 1/sigma_k, rho, log(pi_k / (2 pi sw sk sqrt(1-rho^2)))
  double om;               // 1 - rho^2
  double pad;
};
struct LPView {
  int nn, bwc, bkc, ncbox;
  const LPComp* comp;      // [nn*kLPComp]
  const double2* mz;       // [nn]: (M, Z)
  const double* patch;     // [nn*ncbox] patch density per coarse box
  const double* pprob;     // [nn*ncbox] alias acceptance over patch masses
  const int* palias;       // [nn*ncbox]
};

NCD double lp_qmix(const LPComp* c, double w, double k) {
  double q = 0.0;
#pragma unroll
  for (int j = 0; j < kLPComp; ++j) {
    const double zw = (w - c[j].mw) * c[j].iw, zk = (k - c[j].mk) * c[j].ik;
    const double qq = (zw * zw - 2.0 * c[j].r * zw * zk + zk * zk) / c[j].om;
    q += exp(c[j].lognorm - 0.5 * qq);
  }
  return q;
}

NCD int lp_sample_ab(const SABView& s, const KBView& k, const LPView& lp, double E, Rng& rng, double& a_out,
                     double& b_out, int maxtries) {
  if (!(E >= k.elo && E < k.ehi)) return kKBFallback;
  int n = (int)(log(E / k.elo) * k.lr_inv);
  n = n < 0 ? 0 : (n > k.nn - 1 ? k.nn - 1 : n);
  if (E < k.node_e[n]) --n;
  else if (!(E < k.node_e[n + 1])) ++n;
  if (n < 0 || n >= k.nn) return kKBFallback;
  const double2 mz = lp.mz[n];
  if (!(mz.x > 0.0)) return kKBFallback;
  const double4 par = k.node_par[n];
  const double eps = par.x, seps = par.y, wmax = par.z, jfac = par.w;
  const double ediv = E / s.kT;
  const double sed = sqrt(ediv);
  const LPComp* c = lp.comp + (size_t)n * kLPComp;
  const double* pt = lp.patch + (size_t)n * lp.ncbox;
  const double amin = s.alpha[0], amax = s.alpha[s.na - 1];
  const double bmin = s.beta[0], bmax = s.beta[s.nb - 1];
  for (int t = 0; t < maxtries; ++t) {
    double w, kap;
    if (rng.gen() * mz.y <= 1.0) {
      // mixture component, then a Gaussian point (Box-Muller)
      const double uc = rng.gen();
      int j = 0;
      while (j < kLPComp - 1 && uc > c[j].cpi) ++j;
      const double u1 = rng.gen(), u2 = rng.gen();
      const double rad = sqrt(-2.0 * log(u1));
      double sn, cs;
      sincos(6.283185307179586 * u2, &sn, &cs);
      const double z1 = rad * cs, z2 = rad * sn;
      w = c[j].mw + c[j].l11 * z1;
      kap = c[j].mk + c[j].l21 * z1 + c[j].l22 * z2;
      if (!(w > 0.0 && w <= 1.0 && kap >= 0.0 && kap <= 1.0)) continue;
    } else {
      const double u0 = rng.gen() * lp.ncbox;
      int box = (int)u0;
      box = box >= lp.ncbox ? lp.ncbox - 1 : box;
      if (lp.pprob[(size_t)n * lp.ncbox + box] < u0 - (double)box) box = lp.palias[(size_t)n * lp.ncbox + box];
      const int iw = box / lp.bkc, ik = box - iw * lp.bkc;
      w = (iw + rng.gen()) / lp.bwc;
      kap = (ik + rng.gen()) / lp.bkc;
    }
    const double u = rng.gen();
    int iw = (int)(w * lp.bwc), ik = (int)(kap * lp.bkc);
    iw = iw >= lp.bwc ? lp.bwc - 1 : iw;
    ik = ik >= lp.bkc ? lp.bkc - 1 : ik;
    const double qv = lp_qmix(c, w, kap) + pt[iw * lp.bkc + ik];
    const double r = w * wmax;
    const double beta = r * r - eps;
    const double alpha = r * r + eps - 2.0 * seps * r * (1.0 - 2.0 * kap);
    if (!(beta > -ediv) || beta < bmin || beta > bmax || alpha < amin || alpha > amax) continue;
    const double re = sqrt(ediv + beta);
    const double alo = (re - sed) * (re - sed), ahi = (re + sed) * (re + sed);
    if (alpha < alo || alpha > ahi) continue;
    const double tv = kb_eval_s(s, k, alpha, beta) * (jfac * w * w);
    if (u * mz.x * qv <= tv) {
      a_out = alpha;
      b_out = beta;
      return kKBAccepted;
    }
  }
  return kKBFallback;
}

}  // namespace ncamd
