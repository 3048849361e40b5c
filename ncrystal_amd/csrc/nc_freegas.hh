// Free-gas scattering on the device, drawn as a two-body collision.
//
// NCrystal's free-gas model (cross section
// ncrystal_core/src/phys_utils/NCFreeGasUtils.cc:64-84, scattering function
// ncrystal_core/include/NCrystal/internal/phys_utils/NCFreeGasUtils.hh:36-61)
// is the kernel of elastic collisions of the neutron with atoms of mass A
// (in neutron masses) whose velocities follow the Maxwell distribution at
// temperature T, with a constant cross section and isotropic scattering in
// the centre-of-mass frame.  The device samples that collision directly
// rather than the (alpha, beta) marginals:
//
//  1. the target velocity, from the Maxwellian weighted by the relative
//     speed |v - V|.  In reduced speeds x = sqrt(A) |V|, y = sqrt(A) |v|
//     (neutron speed u = sqrt(E/kT), target density ~ w^2 exp(-A w^2)) the
//     bound |v - V| <= |v| + |V| splits the envelope (x + y) x^2 e^{-x^2}
//     into x^3 e^{-x^2} (mass 1/2, x^2 ~ Gamma(2)) and y x^2 e^{-x^2}
//     (mass y sqrt(pi)/4, x^2 ~ Gamma(3/2)); the angle cosine is uniform
//     and the draw is kept with probability |v - V| / (|v| + |V|), which is
//     at least about 2/3 on average at every energy;
//  2. an isotropic direction Omega of the outgoing relative velocity.
//
// With c the centre-of-mass velocity, r = v - V and g = A |r| / (1 + A), the
// outgoing neutron velocity is v' = c + g Omega and
//   alpha = |v' - v|^2       = 4 g^2 t              (1 - cos chi = 2 t),
//   beta  = |v'|^2 - |v|^2   = 2 g c.(Omega - r^)
// are evaluated as these differences, so small energy and momentum
// transfers keep their relative precision.  Every random number comes from
// the neutron's own Philox stream, in the order written here (the CPU
// checker draws in the same order).
#pragma once

namespace ncamd {

constexpr double kSqrtPi = 1.772453850905516027298167483341145182797549456;
constexpr int kFGMaxTries = 1000;  // acceptance is >= ~2/3 per try

struct FGDraw {
  double alpha, beta, mu;
};

// One collision of a neutron of energy c kT (c >= 0 and finite) with a gas of
// mass ratio A.  Returns kOK or kErrSamplerLoop.
NCD int fg_collide(double c, double A, Rng& rng, FGDraw& o) {
  const double u = sqrt(c);
  const double sA = sqrt(A);
  const double y = sA * u;
  const double p_x3 = 2.0 / (2.0 + kSqrtPi * y);
  double x, s, rel;
  int tries = 0;
  while (true) {
    if (++tries > kFGMaxTries) return kErrSamplerLoop;
    double x2;
    if (rng.gen() < p_x3) {
      const double r1 = rng.gen();
      const double r2 = rng.gen();
      x2 = -log(r1 * r2);
    } else {
      const double r1 = rng.gen();
      const double r2 = rng.gen();
      const double cq = cos(kPiHalf * rng.gen());
      x2 = -log(r1) - log(r2) * (cq * cq);
    }
    x = sqrt(x2);
    s = rng.gen();  // 1 - cos(angle between v and V) = 2 s
    const double d = x - y;
    rel = sqrt(d * d + 4.0 * x * y * s);
    if (rng.gen() * (x + y) < rel) break;
  }
  // velocities in units of sqrt(kT / m_n) per axis convention E = kT |v|^2;
  // v along z, V in the x-z plane
  const double iA = 1.0 / sA;
  const double w = x * iA;
  const double rn = rel * iA;                    // |r|
  const double wx = w * (2.0 * sqrt(s * (1.0 - s)));
  const double rx = -wx;
  const double rz = (u - w) + 2.0 * w * s;       // u - w cos
  const double q = A / (1.0 + A);
  const double g = q * rn;
  const double irn = 1.0 / rn;
  const double c_r = u * rz * irn - g;           // c . r^
  const double c_e = u * wx * irn;               // c . e1, e1 = (rz, 0, -rx)/|r|
  const double t = rng.gen();
  const double sinchi = 2.0 * sqrt(t * (1.0 - t));
  const double cphi = cos(k2Pi * rng.gen());
  o.alpha = 4.0 * g * g * t;
  o.beta = 2.0 * g * (sinchi * cphi * c_e - 2.0 * t * c_r);
  // outgoing direction cosine from the components of v' = c + g Omega
  const double coschi = 1.0 - 2.0 * t;
  const double cx = q * wx, cz = u - q * rz;
  const double ox = coschi * rx + sinchi * cphi * rz;
  const double oz = coschi * rz - sinchi * cphi * rx;
  const double vx = cx + g * ox * irn;
  const double vz = cz + g * oz * irn;
  const double gs = g * sinchi;
  const double vy2 = gs * gs * (1.0 - cphi * cphi);
  const double nrm = sqrt(vx * vx + vy2 + vz * vz);
  o.mu = nrm > 0.0 ? dclamp(vz / nrm, -1.0, 1.0) : 1.0;
  return kOK;
}

// FreeGas process (ncrystal_core/src/freegas/NCFreeGas.cc:25-77): kT, mass
// ratio A, free cross section, a^2 = ca E.
struct FGView {
  double kT, A, sigma_free, ca;
};
// FreeGasXSProvider::crossSection (NCFreeGasUtils.hh:122-125)
NCD double fg_xs(const FGView& g, double E) { return g.sigma_free * freegas_xs_shape(g.ca * E); }
// AbsOOV::crossSectionIsotropic (ncrystal_core/src/absoov/NCAbsOOV.cc:39-43)
NCD double absoov_xs(double c, double E) {
  const double sE = sqrt(E);
  return sE ? c / sE : __builtin_huge_val();
}

// FreeGas::sampleScatterIsotropic (NCFreeGas.cc:70-77): E' = max(0, E + dE)
// with dE = beta kT.  Energies that are not finite and >= 0 are refused.
NCD int fg_sample_demu(const FGView& g, Rng& rng, double E, double& Ef, double& mu) {
  if (!(E >= 0.0 && E < __builtin_huge_val())) return kErrKinematics;
  FGDraw d;
  const int rc = fg_collide(E / g.kT, g.A, rng, d);
  if (rc != kOK) return rc;
  Ef = dmax(0.0, E + d.beta * g.kT);
  mu = d.mu;
  return kOK;
}

}  // namespace ncamd
