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
//
// Random numbers per proposal: one Philox block of the neutron's stream.
// Words 0-1 form a 53-bit uniform that picks the alias slot (u * nbox) and,
// through the residual of that pick (uniform and independent of the box
// chosen), the position in w; word 2 gives the position in kappa and word 3
// the acceptance test, as 32-bit uniforms (j + 1/2) / 2^32.  Every lane of a
// wave thus computes exactly one block per proposal.

#pragma once
#include "nc_dev.hh"

namespace ncamd {

// one box: its geometry (float edges shared with the neighbouring boxes, so
// the boxes tile the square exactly), the bounds of t over it (rounded
// outwards to float) and the ranges of table cells its (alpha, beta) range
// covers (first | last << 16), inside which the cell searches run
struct KBBox {
  float w0, w1, k0, k1;
  float env, tlow;
  uint32_t arange, brange;
};
// alias slot j with box j stored in it (one 64-byte record): acceptance of
// the slot's own box with the reciprocals that turn the pick's residual into
// a uniform number, the box taken otherwise, and box j itself.  A proposal
// reads one record, and a second one only when the alias is taken.
struct alignas(64) KBSlot {
  double prob, inv_p, inv_q;
  KBBox box;
  int alias, pad;
};
static_assert(sizeof(KBSlot) == 64, "one slot record per 64 bytes");

struct KBView {
  int nn;                          // energy intervals
  int bw, bk, nbox;                // boxes per interval: bw w-strips x bk boxes per strip
  int squeeze;                     // 1: accept on u * env <= tlow without evaluating the table
  double elo, ehi, lr_inv;         // interval index guess: floor(log(E/elo) * lr_inv)
  double inv_elo, inv_kT;          // 1/elo (for the guess), 1/kT (E/kT of a neutron as E * inv_kT)
  const double* node_e;            // [nn+1] interval edges
  const double4* node_par;         // [nn]: (eps, sqrt(eps), wmax, 8 sqrt(eps) wmax^3)
  const KBSlot* slot;              // [nn*nbox] alias slots with their boxes, strip-major
};

// interpolate_loglin_fallbacklinlin_fast (NCSABUtils.hh:176-195), with the
// reciprocal of the cell width given (one division per cell, not per
// interpolation; the envelope's bounds carry float margins far above the
// rounding this changes)
NCD double kb_loglin(double a, double fa, double b, double fb, double x, double lfa, double lfb, double inv_bma) {
  const double mid = 0.5 * (b + a);
  const bool lin = (fa * fb == 0.0);
  if (x < mid) {
    const double r = (x - a) * inv_bma;
    return lin ? (fa + (fb - fa) * r) : exp(lfa + (lfb - lfa) * r);
  }
  const double sx = (b - x) * inv_bma;
  return lin ? (fb + (fa - fb) * sx) : exp(lfb + (lfa - lfb) * sx);
}

// cell holding x among the cells [lo, hi] of a grid: largest i in [lo, hi]
// with g[i] <= x (lo if there is none), by binary search
NCD int kb_cell_in(const double* __restrict__ g, uint32_t range, double x) {
  const int lo = (int)(range & 0xFFFFu), hi = (int)(range >> 16);
  return upper_bound_d(g, lo + 1, hi + 1, x) - 1;
}

// S(alpha,beta) inside the table grid (SABCellEval, alpha first), with the
// cell searches restricted to the box's cell ranges
NCD double kb_eval_s(const SABView& s, double a, double b, uint32_t arange, uint32_t brange) {
  const int ia = kb_cell_in(s.alpha, arange, a);
  const int ib = kb_cell_in(s.beta, brange, b);
  const double a0 = s.alpha[ia], a1 = s.alpha[ia + 1];
  const double b0 = s.beta[ib], b1 = s.beta[ib + 1];
  const double2* r0 = s.sl + (size_t)ib * s.na + ia;
  const double2* r1 = r0 + s.na;
  const double2 c00 = r0[0], c10 = r0[1], c01 = r1[0], c11 = r1[1];
  const double inv_da = 1.0 / (a1 - a0);
  const double s0 = kb_loglin(a0, c00.x, a1, c10.x, a, c00.y, c10.y, inv_da);
  const double s1 = kb_loglin(a0, c01.x, a1, c11.x, a, c01.y, c11.y, inv_da);
  return s0 + (s1 - s0) * ((b - b0) / (b1 - b0));
}

// Energy interval of E, -1 outside [elo, ehi).  The FP32 logarithm gives the
// index to within one; the edge comparisons make it exact.
NCD int kb_interval(const KBView& k, double E) {
  if (!(E >= k.elo && E < k.ehi)) return -1;
  int n = (int)(__logf((float)(E * k.inv_elo)) * (float)k.lr_inv);
  n = n < 0 ? 0 : (n > k.nn - 1 ? k.nn - 1 : n);
  while (n > 0 && E < k.node_e[n]) --n;
  while (n < k.nn - 1 && !(E < k.node_e[n + 1])) ++n;
  return n;
}

// One proposal for a neutron of energy E (ediv = E/kT) in interval iv,
// from one Philox block (r0..r3); true if accepted, with (alpha, beta).  The
// interval's constants are read here (a small cached table), so that a lane
// carries only the neutron's index, energy, interval and draw counter
// between proposals.
NCD bool kb_propose(const SABView& s, const KBView& k, int iv, double ediv, uint32_t r0, uint32_t r1, uint32_t r2,
                    uint32_t r3, double& a_out, double& b_out) {
  const double4 par = k.node_par[iv];
  const double eps = par.x, seps = par.y, wmax = par.z, jfac = par.w;
  const KBSlot* sl = k.slot + (size_t)iv * k.nbox;
  // alias slot; the residual of the pick, rescaled, is the position in w
  const double u = Rng::to01(((uint64_t)r1 << 32) | r0) * k.nbox;
  int slot = (int)u;
  slot = slot >= k.nbox ? k.nbox - 1 : slot;
  const double f = u - (double)slot;
  const KBSlot& rec = sl[slot];
  const double prob = rec.prob;
  double uw;
  KBBox b;
  if (f < prob) {
    uw = f * rec.inv_p;
    b = rec.box;
  } else {
    uw = (f - prob) * rec.inv_q;
    b = sl[rec.alias].box;
  }
  const double w0 = b.w0, w1 = b.w1, k0 = b.k0;
  const double w = w0 + (w1 - w0) * uw;
  const double kap = k0 + ((double)b.k1 - k0) * (((double)r2 + 0.5) * 0x1.0p-32);
  const double ue = (((double)r3 + 0.5) * 0x1.0p-32) * (double)b.env;
  const double r = w * wmax;
  const double beta = r * r - eps;
  const double alpha = r * r + eps - 2.0 * seps * r * (1.0 - 2.0 * kap);
  // kinematics at the neutron's own energy: beta > -E/kT and
  // (alpha - 2E/kT - beta)^2 <= 4 (E/kT) (E/kT + beta), i.e. alpha between
  // the kinematic limits (sqrt(E/kT + beta) -+ sqrt(E/kT))^2; table range
  const double ef = ediv + beta;
  const double d = alpha - ediv - ef;
  if (!(ef > 0.0) || d * d > 4.0 * ediv * ef) return false;
  if (beta < s.beta[0] || beta > s.beta[s.nb - 1] || alpha < s.alpha[0] || alpha > s.alpha[s.na - 1]) return false;
  if ((k.squeeze && ue <= (double)b.tlow) || ue <= kb_eval_s(s, alpha, beta, b.arange, b.brange) * (jfac * w * w)) {
    a_out = alpha;
    b_out = beta;
    return true;
  }
  return false;
}

}  // namespace ncamd
