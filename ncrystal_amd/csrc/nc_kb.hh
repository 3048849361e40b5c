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
// the interval.  In (w, kappa) the target is t = S * J with
// J = 8 sqrt(eps) wmax^3 w^2, and it is smooth up to the kinematic edges, so
// a piecewise-constant envelope on a grid of boxes fits it closely.
//
// The grid of an interval: w-edges, and inside each w-strip its own
// kappa-edges (uniform, or placed by the learned model).  Each box carries
// an upper bound `env` and a lower bound `tlow` of t over the box.  A box is
// picked with probability proportional to env x area (alias table), a point
// is drawn uniformly in it, and accepted if u * env <= t.  Since
// tlow <= t on the box, u * env <= tlow already decides acceptance without
// evaluating the table (squeeze); the decision is the same either way.  The
// accepted points follow t exactly.

#pragma once
#include "nc_dev.hh"

namespace ncamd {

// one box: its geometry (float edges shared with the neighbouring boxes, so
// the boxes tile the square exactly) and the bounds of t over it
struct KBBox {
  float w0, w1, k0, k1;
  double env, tlow;
};
struct KBAlias {
  double prob;  // acceptance of the slot's own box
  int alias;    // box taken otherwise
  int pad;
};

struct KBView {
  int nn;                          // energy intervals
  int bw, bk, nbox;                // boxes per interval: bw w-strips x bk boxes per strip
  int squeeze;                     // 1: accept on u * env <= tlow without evaluating the table
  double elo, ehi, lr_inv;         // interval index guess: floor(log(E/elo) * lr_inv)
  const double* node_e;            // [nn+1] interval edges
  const double4* node_par;         // [nn]: (eps, sqrt(eps), wmax, 8 sqrt(eps) wmax^3)
  const KBAlias* alias;            // [nn*nbox] alias slots
  const KBBox* box;                // [nn*nbox] boxes, strip-major
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
  const KBAlias* al = k.alias + (size_t)n * k.nbox;
  const KBBox* bx = k.box + (size_t)n * k.nbox;
  const double amin = s.alpha[0], amax = s.alpha[s.na - 1];
  const double bmin = s.beta[0], bmax = s.beta[s.nb - 1];
  for (int t = 0; t < maxtries; ++t) {
    // box from the alias table, then a uniform point inside it
    const double u0 = rng.gen() * k.nbox;
    int slot = (int)u0;
    slot = slot >= k.nbox ? k.nbox - 1 : slot;
    const KBAlias as = al[slot];
    const int box = (as.prob < u0 - (double)slot) ? as.alias : slot;
    const KBBox b = bx[box];
    const double w0 = b.w0, w1 = b.w1, k0 = b.k0, k1 = b.k1;
    const double w = w0 + (w1 - w0) * rng.gen();
    const double kap = k0 + (k1 - k0) * rng.gen();
    const double ue = rng.gen() * b.env;
    const double r = w * wmax;
    const double beta = r * r - eps;
    const double alpha = r * r + eps - 2.0 * seps * r * (1.0 - 2.0 * kap);
    // kinematics at the neutron's own energy and the table range
    if (!(beta > -ediv) || beta < bmin || beta > bmax || alpha < amin || alpha > amax) continue;
    const double re = sqrt(ediv + beta);
    const double alo = (re - sed) * (re - sed), ahi = (re + sed) * (re + sed);
    if (alpha < alo || alpha > ahi) continue;
    if ((k.squeeze && ue <= b.tlow) || ue <= kb_eval_s(s, k, alpha, beta) * (jfac * w * w)) {
      a_out = alpha;
      b_out = beta;
      return kKBAccepted;
    }
  }
  return kKBFallback;
}

}  // namespace ncamd
