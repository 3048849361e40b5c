// Free-gas (alpha,beta) sampling on the device: the S(alpha,beta) high-energy
// extender.  Algorithms: ncrystal_core/src/phys_utils/NCFreeGasUtils.cc:86-930,
// ncrystal_core/include/NCrystal/internal/phys_utils/NCFreeGasUtils.hh:132-146,
// ncrystal_core/src/utils/NCMath.cc:338-417 (erfcdiff, erfc_rescaled),
// ncrystal_core/src/utils/NCRandUtils.cc:224-380 (randExpDivSqrt).
// The erfc bound table (1103 entries, bin width 0.01 on [-2,9]) is built on
// the host with the C library's erfc and uploaded once.
#pragma once

namespace ncamd {

constexpr int kErfcTabLen = 1103;

NCD double erfcdiff_notaylor(double a, double b) {
  if (b < 0) { const double t = -b; b = -a; a = t; }
  const double erfca = a > 27.3 ? 0.0 : erfc(a);
  if (b > a + 4.0 && (a >= 4 || (a < 0.0 && b > 6.0))) return erfca;
  const double erfcb = b > 27.3 ? 0.0 : erfc(b);
  return erfca - erfcb;
}
NCD double erfcdiff(double a, double b) {
  if (dmax(fabs(a), fabs(b)) < 0.32) {
    constexpr double c1 = -2.0 * kInvSqrtPi, c3 = 2.0 * kInvSqrtPi / 3.0, c5 = -0.2 * kInvSqrtPi,
                     c7 = kInvSqrtPi / 21.0, c9 = -kInvSqrtPi / 108.0, c11 = kInvSqrtPi / 660.0,
                     c13 = -kInvSqrtPi / 4680.0, c15 = kInvSqrtPi / 37800.0;
    const double a2 = a * a, b2 = b * b;
    const double a3 = a * a2 * (c3 + a2 * (c5 + a2 * (c7 + (a2 * (c9 + a2 * (c11 + a2 * (c13 + a2 * c15)))))));
    const double b3 = b * b2 * (c3 + b2 * (c5 + b2 * (c7 + (b2 * (c9 + b2 * (c11 + b2 * (c13 + b2 * c15)))))));
    return c1 * (a - b) + (a3 - b3);
  }
  return a > b ? -erfcdiff_notaylor(b, a) : erfcdiff_notaylor(a, b);
}
NCD double erfc_rescaled(double x, double b) {
  if (b < -745.1) return 0.0;
  if ((x < 23.0 && fabs(b) < 700) || x < 5) return exp(b) * erfc(x);
  const double bxx = b - x * x;
  if (bxx < -745.1) return 0.0;
  constexpr double c3 = -0.5, c5 = 0.75, c7 = -1.875, c9 = 6.5625, c11 = -29.53125;
  const double y = 1 / x, y2 = y * y;
  return kInvSqrtPi * exp(bxx) * (y + y2 * (c3 + y2 * (c5 + y2 * (c7 + y2 * (c9 + y2 * c11)))));
}
NCD void erfc_bounds(const double* tab, double x, double& lb, double& ub) {
  constexpr double bw = (9.0 - (-2.0)) / (kErfcTabLen - 3);
  constexpr double ibw = (kErfcTabLen - 3) / (9.0 - (-2.0));
  constexpr int nb_inf = kErfcTabLen - 1;
  constexpr double vlow = -2.0 - 0.5 * bw, vhigh = 9.0 + 0.5 * bw;
  const double xs = dclamp(x, vlow, vhigh);
  const double xr = (xs - (-2.0));
  int bi = (int)(1.0 + xr * ibw);
  if (bi > nb_inf) bi = nb_inf;
  if (bi < 0) bi = 0;
  lb = tab[bi + 1] * 0.99999999;
  ub = tab[bi] * 1.00000001;
}

struct FGEval {
  double beta, normfact, expmbeta, k11, k12, k21, k22;
  NCD FGEval(double c, double invA, double sqrtAc, double b, double nf) {
    beta = b; normfact = nf; expmbeta = -1.0;
    const double eps = b / c;
    const double sq1pe = sqrt(1 + eps);
    const double S = (b < 0.0 ? -1.0 : 1.0);
    const double sqrtepsprime = (eps >= 0.0 ? 1.0 : sq1pe);
    const double sqrtgp = sqrt(2.0 + eps + 2.0 * sq1pe);
    const double SP = 0.5 * (S + invA), SM = 0.5 * (S - invA);
    const double ia_se = invA * sqrtepsprime, mS_se = -S * sqrtepsprime;
    const double SPsgp = sqrtgp * SP, SMsgp = sqrtgp * SM;
    k11 = sqrtAc * (-ia_se + SPsgp);
    k12 = sqrtAc * (mS_se + SPsgp);
    k21 = sqrtAc * (mS_se + SMsgp);
    k22 = sqrtAc * (ia_se + SMsgp);
  }
  NCD void evalExpMBeta() { if (expmbeta < 0) expmbeta = beta < -700.0 ? 0.0 : exp(-beta); }
  NCD double exact() {
    const double t1 = erfcdiff(k11, k12);
    evalExpMBeta();
    if (!expmbeta) return normfact * t1;
    const double t2 = erfcdiff(k21, k22);
    return normfact * (t1 + t2 * expmbeta);
  }
  NCD void bounds(const double* tab, double& lb, double& ub) {
    double l11, u11, l12, u12, l21, u21, l22, u22;
    erfc_bounds(tab, k11, l11, u11); erfc_bounds(tab, k12, l12, u12);
    const double t1l = l11 - u12, t1u = u11 - l12;
    erfc_bounds(tab, k21, l21, u21); erfc_bounds(tab, k22, l22, u22);
    const double t2l = l21 - u22, t2u = u21 - l22;
    if (t2u > 0.0) {
      evalExpMBeta();
      lb = normfact * (t1l + t2l * expmbeta);
      ub = normfact * (t1u + t2u * expmbeta);
      return;
    }
    lb = normfact * t1l; ub = normfact * t1u;
  }
};

NCD double rand_exp_div_sqrt(Rng& rng, double c, double a, double b) {
  const double A = c * a;
  if (A > 0.1) {
    const double U = c * (b - a);
    const double invA = 1.0 / A;
    const double m_c1 = -1.0, m_c2 = expm1(-1.0 * (U - 0));
    while (true) {
      const double ug = 0 + m_c1 * log(1.0 + rng.gen() * m_c2);
      const double R = rng.gen();
      if ((1.0 + ug * invA) * R * R < 1.0) return dclamp((ug + A) / c, a, b);
    }
  }
  constexpr double Ulim = 16.1180956509583;
  const double U = dmin(c * (b - a), Ulim);
  const double B = U + A;
  if (!(B > A)) return a;
  const double sA = sqrt(A), sB = sqrt(B), d = sB - sA, tsA = 2 * sA;
  double ug;
  while (true) {
    const double T = rng.gen() * d;
    ug = T * (T + tsA);
    const double Ra = rng.gen();
    if (ug < 2.0) {
      constexpr double c1 = -1.0, c2 = 1.0 / 2.0, c3 = -1.0 / 6.0, c4 = 1.0 / 24.0, c5 = -1.0 / 120.0, c6 = 1.0 / 720.0;
      const double t6 = 1.0 + ug * (c1 + ug * (c2 + ug * (c3 + ug * (c4 + ug * (c5 + ug * c6)))));
      if (Ra > t6) continue;
      if (Ra + 0.020221 < t6) break;
    } else {
      if (Ra > 0.135335283236614) continue;
      if (ug > 4.0 && Ra > 0.0183156388887343) continue;
    }
    if (Ra < exp(-ug)) break;
  }
  return dclamp((ug + A) / c, a, b);
}

NCD double fe_eval(double x, double xmax, double c) {
  const double ea = (x - xmax) / (x * xmax) - c * (x - xmax);
  if (ea >= 706.0) return 1.0;
  return ea < -745.1 ? 0.0 : exp(ea) * sqrt(xmax / x);
}

// randExpMInvXMCXDivSqrtX (NCFreeGasUtils.cc:240-490)
NCD double rand_expminvx(Rng& rng, double c, double xm, double xp) {
  if (xp == xm) return xm;
  const double sqrtc = sqrt(c);
  const double invsqrtc = 1 / sqrtc;
  const double xpeak = (c > 1e-5 ? (c > 1e200 ? invsqrtc : (sqrt(16.0 * c + 1.0) - 1.0) / (4.0 * c))
                                 : (2.0 - c * (8.0 - c * (64.0 - c * (640.0 - c * 7168.0)))));
  if (xpeak == 0.0) return xm > 0.0 ? xm : dmin(kDblMin, xp);
  const double xmax = (xm > xpeak ? xm : dmin(xp, xpeak));
  if (!(xmax > 0.0)) return xm;
  const double xlarge = dmax(5.0 / sqrt(c), 2 * xpeak);
  const double xsmall = dmin(0.2 / sqrt(c), 0.5 * xpeak);
  if (xp > xlarge) xp = dmin(xp, dmax(xm, xlarge) + 15.0 / c);
  if (xm < xsmall) {
    const double xsm = dmin(xp, xsmall);
    xm = dmax(xm, xsm / (1 + 30.0 * xsm));
  }
  if ((xm = dmax(kDblMin, dmax(kDblMin / xp, xm))) >= xp) return xp;
  constexpr double fcut = 1e-9;
  if (xp < xpeak) {
    while (true) {
      const double xm_new = xp - 0.01 * (xp - xm);
      const double fv = fe_eval(xm_new, xmax, c);
      if (fv >= fcut) break;
      xm = xm_new;
    }
  }
  double pflat, xswitch, area_right = -1.0;
  if (xm >= xlarge) { pflat = 0.0; xswitch = xm; }
  else if (c > 25 || xp <= xlarge) { pflat = 1.0; xswitch = xp; }
  else {
    xswitch = xlarge;
    const double area_left = (xswitch - xm);
    const double B = c * xmax + 1 / xmax - 1 / xp;
    area_right = (erfc_rescaled(sqrtc * sqrt(xswitch), B) - erfc_rescaled(sqrtc * sqrt(xp), B)) * sqrt(kPi * (xmax / c));
    pflat = area_left / (area_left + area_right);
  }
  bool always_left = (pflat > 1.0 - fcut);
  bool single_side = always_left || (pflat < fcut);
  if (!single_side && fe_eval(xswitch, xmax, c) < fcut * 1.1) {
    pflat = 1.0; area_right = 0.0; xp = xswitch; always_left = true; single_side = false;
  }
  while (true) {
    const bool do_flat = single_side ? always_left : (rng.gen() < pflat);
    if (do_flat) {
      double dx = xswitch - xm;
      const double xthr_up = xswitch, xthr_low = xm;
      constexpr double fth = 0.05;
      const double xgen = xm + rng.gen() * dx;
      const double Ra = rng.gen();
      if (!in_interval(xthr_low, xthr_up, xgen) && Ra > fth) continue;
      const double fv = fe_eval(xgen, xmax, c);
      if (fv < fth) {
        if (fv < fcut) {
          if (xgen < xmax) xm = xgen;
          else xswitch = xgen;
          dx = xswitch - xm;
          if (!single_side) {
            const double area_left = dx;
            pflat = area_left / (area_left + area_right);
            always_left = (pflat > 1.0 - fcut);
            single_side = always_left || (pflat < fcut);
          }
          continue;
        }
      }
      if (Ra <= fv) return xgen;
    } else {
      const double xgen = rand_exp_div_sqrt(rng, c, xswitch, xp);
      if (rng.gen() < exp((xgen - xp) / (xgen * xp))) return xgen;
    }
  }
}

struct FGSampler {
  double c, kT, sqrtAc, invA, Adiv4, normfact, c_real;
  NCD FGSampler(double E, double T, double mass) {
    kT = kBoltzmann * T;
    c = dmin(1e14, dmax(1e-10, E / kT));
    sqrtAc = sqrt(mass * c / kNeutronMassAmu);
    const double A = (1.0 / kNeutronMassAmu) * mass;
    invA = 1.0 / A;
    Adiv4 = 0.25 * A;
    normfact = 0.5 / erf(sqrt(c * invA));
    c_real = E / kT;
  }
  // FreeGasSampler::sampleBeta (NCFreeGasUtils.cc:530-848)
  NCD double sample_beta(Rng& rng, const double* tab) const {
    if (c_real > 1e4) {
      const double A = 1.0 / invA, A2 = A * A;
      const double thr = 1e4 * dmin(1000.0 * A, A2 * A2 * A2);
      if (c_real > thr) {
        const double q = (1.0 - invA) / (1.0 + invA);
        const double elossmax = c_real * (1.0 - q * q);
        return -elossmax * rng.gen();
      }
    }
    constexpr double Tlim = 2.0, Tlim_k1 = 0.135335283236612691893999494972484403407, Tlim_k2 = 274. / 315.;
    constexpr double fcut = 1e-6;
    double aa = dmax(-c_real, -c), bb = 13.815510557964274;
    if (invA <= 1.0 / 10.0) {
      if (c > 10.1) {
        while (true) {
          const double an = aa * 0.2;
          if (an > -1e-99) break;
          FGEval h(c, invA, sqrtAc, an, normfact);
          if (h.exact() > fcut) break;
          aa = an;
        }
      }
      while (true) {
        const double bn = bb * 0.25;
        if (bn < 1e-99) break;
        FGEval h(c, invA, sqrtAc, bn, normfact);
        double lb, ub;
        h.bounds(tab, lb, ub);
        if (ub > fcut) break;
        bb = bn;
      }
    }
    if (!(bb > aa)) return aa;
    double a, b, pdown, pnotclose;
    double es_a = 0, es_c1 = 0, es_c2 = 0;
    auto setab = [&](double A_, double B_) {
      a = A_; b = B_;
      const double adown = -a;
      double afar, aclose;
      if (b <= Tlim) {
        afar = 0.0;
        aclose = b * (1.0 + b * (-1. / 2. + b * (1. / 6. + b * (-1. / 24. + b * (1. / 120. + b * (-1. / 720. + b * (1. / 5040.)))))));
      } else {
        afar = Tlim_k1 - exp(-b);
        aclose = Tlim_k2;
      }
      const double iat = 1.0 / (adown + aclose + afar);
      pdown = adown * iat;
      pnotclose = (afar + adown) * iat;
      es_a = es_c1 = es_c2 = 0.0;
    };
    setab(aa, bb);
    constexpr double fth = 0.1;
    double afth = a, bfth = b;
    while (true) {
      double beta, fov;
      const double Rs = rng.gen();
      if (Rs < pdown) {
        beta = rng.gen() * a;
        fov = 1.0;
      } else if (Rs < pnotclose) {
        if (!(es_c1 < 0.0)) { es_a = Tlim; es_c1 = -1.0; es_c2 = expm1(-1.0 * (b - Tlim)); }
        beta = es_a + es_c1 * log(1.0 + rng.gen() * es_c2);
        fov = exp(-beta);
      } else {
        const double bmax = dmin(b, Tlim);
        while (true) {
          beta = rng.gen() * bmax;
          const double R0 = rng.gen();
          constexpr double kcheap = 19. / 45.;
          if (R0 > 1.0 - kcheap * beta) continue;
          fov = 1.0 + beta * (-1. + beta * (1. / 2. + beta * (-1. / 6. + beta * (1. / 24. + beta * (-1. / 120. + beta * (1. / 720.))))));
          if (R0 < fov) break;
        }
      }
      const double facc = rng.gen() * fov;
      if (facc > fth && !in_interval(afth, bfth, beta)) continue;
      FGEval h(c, invA, sqrtAc, beta, normfact);
      bool need_exact = true;
      double fval = 0.0;
      if (beta > 0) {
        double lb, ub;
        h.bounds(tab, lb, ub);
        if (facc <= lb) return beta;
        fval = ub;
        if (facc > ub) need_exact = false;
      }
      if (need_exact) {
        fval = h.exact();
        if (facc < fval) return beta;
      }
      if (fval < fcut) {
        if (beta < 0) setab(beta, b); else setab(a, beta);
        continue;
      }
      if (fval < fth) {
        if (beta < 0) afth = dmax(afth, beta);
        else bfth = dmin(bfth, beta);
      }
    }
  }
  // FreeGasSampler::sampleAlpha (NCFreeGasUtils.cc:851-930)
  NCD double sample_alpha(double beta, Rng& rng) const {
    if (c_real < c || mu_iso_at_beta(beta, c)) {
      const Lims l = alpha_limits(c_real, beta);
      const double al = l.lo + rng.gen() * (l.hi - l.lo);
      return dclamp(al, l.lo, l.hi);
    }
    beta = dmax(-c, beta);
    const Lims l = alpha_limits(c, beta);
    const double am = l.lo, ap = l.hi;
    if (am == ap) return am;
    const double bsq = beta * beta;
    const double t = bsq * Adiv4;
    const double cc = 0.0625 * bsq;
    if (dmin(t, cc) < 1e-5) {
      const double fourA = Adiv4 * 16.0;
      const double inv4A = 1.0 / fourA;
      const double xxm = am * inv4A, xxp = ap * inv4A;
      while (true) {
        const double xx = rand_exp_div_sqrt(rng, 1.0, xxm, xxp);
        const double al = xx * fourA;
        if (al < am || al > ap) continue;
        if (al * ap * (-log(rng.gen())) >= t * (ap - al)) return al;
      }
    }
    const double invt = 1.0 / t;
    const double x = rand_expminvx(rng, cc, am * invt, ap * invt);
    return dclamp(x * t, am, ap);
  }
};

// FreeGas process (ncrystal_core/src/freegas/NCFreeGas.cc:25-77): temperature,
// target mass, free cross section, a^2 = ca E, and the erfc bound table
struct FGView {
  double T, mass, sigma_free, ca;
  const double* erfc_tab;
};
// FreeGasXSProvider::crossSection (NCFreeGasUtils.hh:122-125)
NCD double fg_xs(const FGView& g, double E) { return g.sigma_free * freegas_xs_shape(g.ca * E); }
// AbsOOV::crossSectionIsotropic (ncrystal_core/src/absoov/NCAbsOOV.cc:39-43)
NCD double absoov_xs(double c, double E) {
  const double sE = sqrt(E);
  return sE ? c / sE : __builtin_huge_val();
}

// FreeGasSampler::sampleDeltaEMu (NCFreeGasUtils.hh:146-160) with
// convertAlphaBetaToDeltaEMu (NCKinUtils.cc:25-58) at the sampler's clamped
// energy c kT, then FreeGas::sampleScatterIsotropic's E' = max(0, E + dE).
// Energies that are not finite and >= 0 are refused (the rejection loops
// need a finite c).
NCD int fg_sample_demu(const FGView& g, Rng& rng, double E, double& Ef, double& mu) {
  if (!(E >= 0.0 && E < __builtin_huge_val())) return kErrKinematics;
  const FGSampler f(E, g.T, g.mass);
  const double b = f.sample_beta(rng, g.erfc_tab);
  double dE, m;
  if (b <= -f.c || mu_iso_at_beta(b, f.c)) {
    dE = b * f.kT;
    m = rng.gen() * 2.0 - 1.0;
  } else {
    const double a = f.sample_alpha(b, rng);
    const double ek = f.c * f.kT;
    dE = b * f.kT;
    const double ekf = ek + dE;
    const double denom = 2.0 * sqrt(ek * ekf);
    if (!denom) return kErrKinematics;
    StableSum sm;
    sm.add(ek); sm.add(ekf); sm.add(-a * f.kT);
    m = dclamp(sm.sum() / denom, -1.0, 1.0);
  }
  Ef = dmax(0.0, E + dE);
  mu = m;
  return kOK;
}

NCD void freegas_sample_alpha_beta(double E, double T, double mass, const double* tab, Rng& rng,
                                   double& alpha, double& beta) {
  const FGSampler f(E, T, mass);
  const double b = f.sample_beta(rng, tab);
  if (b < -f.c || mu_iso_at_beta(b, f.c)) {
    const Lims l = alpha_limits(f.c_real, b);
    const double al = l.lo + rng.gen() * (l.hi - l.lo);
    alpha = dclamp(al, l.lo, l.hi);
    beta = b;
    return;
  }
  alpha = f.sample_alpha(b, rng);
  beta = b;
}

}  // namespace ncamd
