// Device side of the mosaic single-crystal Bragg process (SCBragg): one lane
// per neutron scans the lab-frame demi-normals of every reflection family
// with 2d > lambda, evaluates the Gaussian-mosaic circle integral for the
// normals within the truncated mosaic cone, and for sampling picks one
// contribution by its cumulative weight and draws the outgoing direction on
// the circle of valid mosaic normals.
//
// Reference algorithms (NCrystal v4.4.2):
//  - cache/families/sampling: ncrystal_core/src/scbragg/NCSCBragg.cc:225-323
//  - GaussMos::calcCrossSections, genScat, InteractionPars::set:
//    ncrystal_core/src/phys_utils/NCGaussMos.cc:29-36,147-280 and the inline
//    calcRawCrossSectionValue of ncrystal_core/include/NCrystal/internal/phys_utils/NCGaussMos.hh
//  - GaussOnSphere::circleIntegral (fast spline path) and circleIntegralSlow,
//    genPointOnCircle: ncrystal_core/include/NCrystal/internal/phys_utils/NCGaussOnSphere.hh:194-211,
//    ncrystal_core/src/phys_utils/NCGaussOnSphere.cc:58-140,380-508
//  - cubic spline lookup: ncrystal_core/include/NCrystal/internal/utils/NCSpline.hh:118-160
//  - CosSinGridGen: ncrystal_core/include/NCrystal/internal/utils/NCMath.hh:489-524
//  - randPointOnUnitCircle: ncrystal_core/src/utils/NCRandUtils.cc:98-111
//  - rotateToFrame: ncrystal_core/src/utils/NCRotMatrix.cc:84-147
//  - pickRandIdxByWeight: ncrystal_core/src/utils/NCRandUtils.cc:198-222
//
// Lanes of a wavefront scan the demi-normals together: the normal index is
// wave-uniform (the records are scalar loads shared by the 64 lanes) and a
// lane whose families are exhausted is masked.  The rare backscattering
// geometry that needs the Romberg circle integral is left to a second kernel
// (AllowSlow=true) through a queue, so the main kernel carries no Romberg
// state.
#pragma once
#include "nc_dev.hh"
#include "nc_math.hh"

namespace ncamd {

struct SplineView {
  double a, invdelta;
  int n;
  const double2* __restrict__ yy2;  // (y_i, y''_i) pairs
};

struct SCBView {
  int nfam;
  int nnormals;
  const double* __restrict__ inv2d;     // [nfam] 1/(2d), non-decreasing
  const double* __restrict__ xsf;       // [nfam] |F|^2/(V0 n) * 0.5
  const int* __restrict__ off;          // [nfam+1] demi-normal offsets
  // one record per demi-normal in scan order: lab-frame unit normal and the
  // family's 1/(2d) rounded to 1e-15 (GaussMos cache rounding)
  const double4* __restrict__ nrec;     // [nnormals + kScbPad]
  // the same records in FP32 for the pre-filter of the scan
  const float4* __restrict__ nrec32;    // [nnormals + kScbPad]
  const double2* __restrict__ nfx;      // [nnormals + kScbPad] (family xsf, family index) of each demi-normal
  double threshold;                     // ekin below which nothing reflects
  double cta, k1, k2, numacc;           // GaussOnSphere: cos(truncation), fast-path bounds, Romberg accuracy
  float band32;                         // pre-filter half-width: truncation angle + rounding margins
  SplineView lt_evalcosx, lt_sofcosd;
};

// scan group size; the record arrays carry this many padding entries so
// that a group never reads past the end
constexpr int kScbGroup = 8;
constexpr int kScbPad = kScbGroup;

NCD double scb_round(double x) { return floor(x * 1e15 + 0.5) * 1e-15; }
NCD double gm_round(double x) { return floor(dmax(x, 1e-15) * 1e15 + 0.5) * 1e-15; }

NCD double slt_eval(const SplineView& t, double x) {
  const double xx = (x - t.a) * t.invdelta;
  const int nm2 = t.n - 2;
  const int idx = xx < 0.0 ? 0 : (xx < (double)nm2 ? (int)xx : nm2);
  const double b = xx - (double)idx;
  const double a = 1.0 - b;
  const double2 p0 = t.yy2[idx], p1 = t.yy2[idx + 1];
  double tmp = a * p0.x;
  double tmp2 = (a * a * a - a) * p0.y;
  tmp += b * p1.x;
  tmp2 += (b * b * b - b) * p1.y;
  return tmp + 0.166666666666666666666666666666666666666666666666666667 * tmp2;
}
NCD double gos_evalcosx_inrange(const SCBView& v, double cosx) {
  return dmax(0.0, slt_eval(v.lt_evalcosx, cosx));
}
NCD double gos_evalcosx(const SCBView& v, double cosx) {
  return cosx >= v.lt_evalcosx.a ? gos_evalcosx_inrange(v, cosx) : 0.0;
}

// cos/sin of offset + i*delta, i = 0..n-1, by rotation with periodic exact
// recomputation (every 128 steps)
struct CSGrid {
  double c, s, cd, sd, phimax, negdelta;
  unsigned left, recalc;
  NCD void init(unsigned n, double offset, double delta) {
    left = n - 1;
    recalc = ((128u - 1) + (n / 128u) * 128u) - n;
    phimax = offset + (n - 1) * delta;
    negdelta = -delta;
    sincos_0pi32(offset, c, s);
    sincos_mpi8pi8(delta, cd, sd);
  }
  NCD bool step() {
    if (!left) return false;
    --left;
    if ((left + recalc) % 128u) {
      const double cn = c * cd - s * sd;
      s = c * sd + s * cd;
      c = cn;
    } else {
      const double ang = phimax + negdelta * left;
      c = cos(ang);
      s = sin(ang);
    }
    return true;
  }
};

struct GCI {
  double sasg, cacg, acc;
};
NCD double gci_f(const SCBView& v, const GCI& g, double c) { return gos_evalcosx_inrange(v, g.sasg * c + g.cacg); }
NCD bool gci_accept(const GCI& g, unsigned level, double prev, double est) {
  if (fabs(prev - est) <= g.acc * fabs(est)) return true;
  if (level < 11) return false;
  return true;
}
NCD double gci_sum(const SCBView& v, const GCI& g, unsigned n, double off, double delta) {
  CSGrid gr;
  gr.init(n, off, delta);
  double sum = 0.0;
  do { sum += gci_f(v, g, gr.c); } while (gr.step());
  return sum;
}
// Romberg integration of the mosaic density along the circle, t in [a,b]
__device__ __noinline__ double gci_integrate(const SCBView& v, const GCI& g, double a, double b) {
  double h = (b - a);
  double fv[17];
  {
    CSGrid gr;
    gr.init(17, a, h * 0.0625);
    unsigned i = 0;
    do { fv[i++] = gci_f(v, g, gr.c); } while (gr.step());
  }
  h *= 0.5;
  const double R00 = (fv[0] + fv[16]) * h;
  const double R10 = h * fv[8] + 0.5 * R00;
  const double R11 = (4. / 3.) * R10 + (-1. / 3.) * R00;
  h *= 0.5;
  const double R20 = h * (fv[4] + fv[12]) + 0.5 * R10;
  const double R21 = (4. / 3.) * R20 + (-1. / 3.) * R10;
  const double R22 = (16. / 15.) * R21 + (-1. / 15.) * R11;
  h *= 0.5;
  const double R30 = h * ((fv[2] + fv[6]) + (fv[10] + fv[14])) + 0.5 * R20;
  const double R31 = (4. / 3.) * R30 + (-1. / 3.) * R20;
  const double R32 = (16. / 15.) * R31 + (-1. / 15.) * R21;
  const double R33 = (64. / 63.) * R32 + (-1. / 63.) * R22;
  h *= 0.5;
  const double R40 = h * (((fv[1] + fv[3]) + (fv[5] + fv[7])) + ((fv[9] + fv[11]) + (fv[13] + fv[15]))) + 0.5 * R30;
  const double R41 = (4. / 3.) * R40 + (-1. / 3.) * R30;
  const double R42 = (16. / 15.) * R41 + (-1. / 15.) * R31;
  const double R43 = (64. / 63.) * R42 + (-1. / 63.) * R32;
  const double R44 = (256. / 255.) * R43 + (-1. / 255.) * R33;
  if (gci_accept(g, 4, R33, R44)) return R44;
  const double c5 = gci_sum(v, g, 16, a + h * 0.5, h);
  h *= 0.5;
  const double R50 = h * c5 + 0.5 * R40;
  const double R51 = (4. / 3.) * R50 + (-1. / 3.) * R40;
  const double R52 = (16. / 15.) * R51 + (-1. / 15.) * R41;
  const double R53 = (64. / 63.) * R52 + (-1. / 63.) * R42;
  const double R54 = (256. / 255.) * R53 + (-1. / 255.) * R43;
  const double R55 = (1024. / 1023.) * R54 + (-1. / 1023.) * R44;
  if (gci_accept(g, 5, R44, R55)) return R55;
  constexpr unsigned maxlevel = 16;
  double rows[2][maxlevel];
  int ip = 0;
  rows[0][0] = R50; rows[0][1] = R51; rows[0][2] = R52;
  rows[0][3] = R53; rows[0][4] = R54; rows[0][5] = R55;
  unsigned nj = 16;
  for (unsigned i = 6; i < maxlevel; ++i) {
    const double hh = h;
    h *= 0.5;
    nj *= 2;
    const double cc = gci_sum(v, g, nj, a + h, hh);
    double* row_prev = rows[ip];
    double* row = rows[ip ^ 1];
    row[0] = h * cc + 0.5 * row_prev[0];
    double n_k = 1.;
    for (unsigned j = 0; j < i; ++j) {
      n_k *= 4.0;
      row[j + 1] = (n_k * row[j] - row_prev[j]) / (n_k - 1.0);
    }
    if (gci_accept(g, i, row_prev[i - 1], row[i])) return row[i];
    ip ^= 1;
  }
  return rows[ip][maxlevel - 1];
}

NCD double gos_circle_slow(const SCBView& v, double cg, double sg, double ca, double sa) {
  const double sasg = sa * sg;
  const double cacg = ca * cg;
  const double cd = cacg + sasg;
  if (cd <= v.cta) return 0.0;
  if (sasg < 1e-14) return k2Pi * sa * gos_evalcosx(v, ca);
  const double cos_tmax = (v.cta - cacg) / sasg;
  const double tmax = (cos_tmax <= -1.0 ? kPi : acos(dmin(1.0, cos_tmax)));
  if (tmax <= 1e-12) return 0.0;
  constexpr double kArcSec = 0.00000484813681109535993589914102357947975956353302;
  double intacc = v.numacc;
  if (tmax < 10 * kArcSec) {
    intacc = dmax(intacc, 1e-6);
    if (tmax < kArcSec) {
      intacc = dmax(intacc, 1e-5);
      if (tmax < 0.1 * kArcSec) intacc = dmax(intacc, 1e-4);
    }
  }
  GCI g{sasg, cacg, intacc};
  return 2.0 * sa * gci_integrate(v, g, 0.0, tmax);
}

// true if the fast spline evaluation applies (NCGaussOnSphere.hh:194-211)
NCD bool gos_circle_fast_ok(const SCBView& v, double cg, double sg, double ca, double sa, double& val) {
  const double sasg = sa * sg;
  const double cacg = ca * cg;
  const double cd = cacg + sasg;
  if (cd > v.cta && sasg >= 1e-14 && v.k2 > v.k1 * sasg + cacg) {
    val = slt_eval(v.lt_sofcosd, cd) * sqrt(sa / sg);
    return true;
  }
  return false;
}

// randPointOnUnitCircle: rejection in the square, 2 draws per try
NCD void rand_point_on_circle(Rng& rng, double& c, double& s) {
  double a, b, m2;
  do {
    a = -1.0 + rng.gen() * 2.0;
    b = -1.0 + rng.gen() * 2.0;
    m2 = a * a + b * b;
  } while (!in_interval(0.001, 1.0, m2));
  const double m = 1.0 / sqrt(m2);
  c = a * m;
  s = b * m;
}

NCD bool gos_gen_point(const SCBView& v, Rng& rng, double cg, double sg, double ca, double sa, double& ct,
                       double& st) {
  const double sasg = sa * sg;
  const double cacg = ca * cg;
  const double cd = cacg + sasg;
  if (cd <= v.cta) return false;
  if (sasg < 1e-14) {
    if (sa < 1e-7) return false;
    rand_point_on_circle(rng, ct, st);
    return true;
  }
  const double cos_tmax = (v.cta - cacg) / sasg;
  if (cos_tmax >= 1.0) return false;
  const double tmax = (cos_tmax <= -1.0 ? kPi : acos(cos_tmax));
  const double densitymax = gos_evalcosx_inrange(v, cd) * 1.00000001;
  int triesleft = 1001;
  while (--triesleft) {
    ct = cos_mpipi(rng.gen() * tmax);
    const double cd_at_t = sasg * ct + cacg;
    const double d = gos_evalcosx_inrange(v, cd_at_t);
    if (d > densitymax * rng.gen()) break;
  }
  if (triesleft <= 0) return false;
  st = sqrt(1.0 - ct * ct);
  st = (rng.gen() > 0.5 ? st : -st);
  return true;
}

struct V3 {
  double x, y, z;
};
NCD double v3dot(const V3& a, const V3& b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
NCD V3 v3cross(const V3& a, const V3& o) {
  return V3{a.y * o.z - a.z * o.y, a.z * o.x - a.x * o.z, a.x * o.y - a.y * o.x};
}
// Vector::normalise: unchanged when |v|^2 is within 2 ulp of 1
NCD V3 v3normalise(V3 v) {
  const double m2 = v.x * v.x + v.y * v.y + v.z * v.z;
  constexpr double eps = 2.220446049250313080847263336181640625e-16;
  if (m2 >= 1.0 - 2.0 * eps && m2 <= 1.0 + 2.0 * eps) return v;
  const double ff = 1.0 / sqrt(m2);
  v.x *= ff; v.y *= ff; v.z *= ff;
  return v;
}
NCD V3 phirot(double cp, double sp, const V3& v, const V3& axis) {
  const V3 ac = v3cross(axis, v);
  const double ad = v3dot(axis, v);
  V3 r{v.x * cp, v.y * cp, v.z * cp};
  r.x += ac.x * sp; r.y += ac.y * sp; r.z += ac.z * sp;
  const double g2 = ad * (1.0 - cp);
  r.x += axis.x * g2; r.y += axis.y * g2; r.z += axis.z * g2;
  return r;
}
// maps v from the frame (c1, c2, b) built from unit vectors a, b with
// sin/cos of their angle to the lab frame
NCD V3 rotate_to_frame(double sinab, double cosab, const V3& a, const V3& b, V3 v, Rng& rng) {
  if (fabs(sinab) < 1e-10) {
    const double bz = b.z;
    const double cp = bz, sp = -sqrt(1.0 - bz * bz);
    V3 axis{b.y, -b.x, 0.};
    const double m2 = axis.x * axis.x + axis.y * axis.y + axis.z * axis.z;
    if (m2 > 1e-12) {
      const double f = 1.0 / sqrt(m2);
      axis.x *= f; axis.y *= f; axis.z *= f;
      v = phirot(cp, sp, v, axis);
    } else if (b.z < 0.0) {
      v.z *= -1.0;
    }
    double rc, rs;
    rand_point_on_circle(rng, rc, rs);
    v = phirot(rc, rs, v, b);
    return v3normalise(v);
  }
  const double s = 1.0 / sinab;
  V3 c1 = b;
  c1.x *= (-cosab); c1.y *= (-cosab); c1.z *= (-cosab);
  c1.x += a.x; c1.y += a.y; c1.z += a.z;
  c1.x *= s; c1.y *= s; c1.z *= s;
  V3 c2 = v3cross(b, a);
  c2.x *= s; c2.y *= s; c2.z *= s;
  const V3 r{v.x * c1.x + v.y * c2.x + v.z * b.x, v.x * c1.y + v.y * c2.y + v.z * b.y,
             v.x * c1.z + v.y * c2.z + v.z * b.z};
  return v3normalise(r);
}

// GaussMos::genScat for the chosen (normal, inv2d) contribution
NCD V3 gm_genscat(const SCBView& v, Rng& rng, const V3& normal, double inv2d_rounded, double wl_raw, const V3& indir) {
  const double wl = gm_round(wl_raw);
  const double stb = wl * inv2d_rounded;
  if (stb == 0.) return indir;
  const double ca = stb;
  const double sa = sqrt(1.0 - ca * ca);
  const double cg = dclamp(-(v3dot(indir, normal)), -1.0, 1.0);
  const double sg = sqrt(1.0 - cg * cg);
  double ct, st;
  if (!gos_gen_point(v, rng, cg, sg, ca, sa, ct, st)) return indir;
  const double s2a = 2 * sa * ca;
  const double c2a = ca * ca - sa * sa;
  const V3 w{s2a * ct, s2a * st, c2a};
  const V3 mb{-indir.x, -indir.y, -indir.z};
  return rotate_to_frame(sg, cg, normal, mb, w, rng);
}

// ---------------------------------------------------------------------------
// wave-uniform helpers
NCD int wave_max_i(int x) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    const int y = __shfl_xor(x, o, 64);
    x = x > y ? x : y;
  }
  return __builtin_amdgcn_readfirstlane(x);
}
NCD double wave_sum_d(double x) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o, 64);
  return x;
}
NCD double wave_min_d(double x) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) x = dmin(x, __shfl_xor(x, o, 64));
  return x;
}
NCD float wave_min_f(float x) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) x = fminf(x, __shfl_xor(x, o, 64));
  return x;
}
NCD float wave_max_f(float x) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) x = fmaxf(x, __shfl_xor(x, o, 64));
  return x;
}
NCD float readlane_f(float x, int l) { return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(x), l)); }


// Cross sections of the anti-normal side (xs0) and normal side (xs1) of
// demi-normal record r (family cross-section factor xsf) for a neutron with
// rounded wavelength w (wl3 = w^3) and unit direction u; 0 where a side does
// not reflect (NCSCBragg.cc:263-300 with GaussMos::calcCrossSections).  The
// reference evaluates Q and cos(theta) lazily once per family; both depend
// on the family alone (sin(theta) = w / 2d), so evaluating them per
// demi-normal gives the same values.  Returns false, without AllowSlow, if a
// side needs the Romberg circle integral.
//
// A demi-normal usually reflects on one side only, and lanes of a wave
// differ in which: the sides that pass the cone test are evaluated by one
// loop, so that a wave runs the circle integral once per candidate rather
// than once per side.
template <bool AllowSlow>
NCD bool scb_candidate(const SCBView& v, const double4& r, double xsf, double w, double wl3, double ux, double uy,
                       double uz, double& xs0, double& xs1) {
  xs0 = 0.0;
  xs1 = 0.0;
  const double cta = v.cta;
  const double dot = r.x * ux + r.y * uy + r.z * uz;
  const double sinpt = w * r.w;
  const double cptsq = 1 - sinpt * sinpt;
  const double sdotcptsq = (1.0 - dot * dot) * cptsq;
  const double ds = dot * sinpt;
  const double A0 = dmax(0.0, cta - fabs(ds));
  if (sdotcptsq <= A0 * A0) return true;
  const double A1 = dmax(0.0, cta - ds), A2 = dmax(0.0, cta + ds);
  unsigned sides = (sdotcptsq > A1 * A1 ? 1u : 0u) | (sdotcptsq > A2 * A2 ? 2u : 0u);
  if (!sides) return true;
  const double cospt = sqrt(cptsq);
  const double tmp2 = cospt * sinpt;
  const double Qprime = tmp2 > 0 ? wl3 / tmp2 : ((sinpt > 0.5 && xsf) ? -2.0 : 0.0);
  const double Q = Qprime > 0. ? Qprime * xsf : -1.0;
  if (!(Q > 0.)) {
    const double xs = Qprime ? __builtin_huge_val() : 0.0;
    if (sides & 1u) xs0 = xs;
    if (sides & 2u) xs1 = xs;
    return true;
  }
#pragma nounroll
  while (sides) {
    const int side = __builtin_ctz(sides);
    sides &= sides - 1u;
    const double cc = dclamp(side == 0 ? dot : -dot, -1.0, 1.0);
    const double sc2 = sqrt(1.0 - cc * cc);
    double ci;
    if (!gos_circle_fast_ok(v, cc, sc2, sinpt, cospt, ci)) {
      if constexpr (AllowSlow) ci = gos_circle_slow(v, cc, sc2, sinpt, cospt);
      else return false;
    }
    if (side == 0) xs0 = Q * ci;
    else xs1 = Q * ci;
  }
  return true;
}

// Candidate lists: a wave keeps, per lane, the demi-normals that passed the
// pre-filter in LDS (entry t of lane l at list[t * 64 + l], as an offset
// from the list's base index) and evaluates them afterwards in lock step,
// lane by lane in scan order.  Evaluating each candidate as the scan meets
// it would run the evaluation once per group in which any lane of the wave
// has a candidate, i.e. for nearly every group.
constexpr int kScbListCap = 64;                      // entries per lane
constexpr int kScbListWave = kScbListCap * 64;       // uint16 entries per wave
constexpr int kScbListSpan = 65536;                  // index range of one list
// accumulation of the contributions in scan order (GaussMos::calcCrossSections
// cumulative per family on top of the previous families' total)
struct ScbAcc {
  int cur_f;
  double commul_last, xsoffset, xssum;
};

// One lane's state during a scan of the reflection families.
struct ScbLane {
  double wl;       // wavelength (from the rounded energy)
  V3 u;            // unit direction
  int jlim;        // demi-normals of the families with 2d > wl (0: lane takes no part)
  int count;       // number of non-zero contributions
  double total;    // cumulative xs after the last contribution
  bool slow;       // a contribution needs the Romberg circle integral
  // selection result (Select=true)
  bool found;
  V3 sel_normal;
  double sel_inv2d_r;
};

// Evaluates the first n entries of the lane's list (indices base + entry) in
// order and accumulates them; all lanes of the wave call this together.
// Select=false: count and cumulative total.  Select=true: stop at the
// contribution picked by pickRandIdxByWeight for (n_select, target).
template <bool Select, bool AllowSlow>
NCD void scb_eval_list(const SCBView& v, ScbLane& L, ScbAcc& a, bool& alive, const uint16_t* __restrict__ list,
                       int base, int n, double w, double wl3, int n_select, double target) {
  const int lane = (int)__lane_id();
  const int tmax = wave_max_i(alive ? n : 0);
  for (int t = 0; t < tmax; ++t) {
    if (!(alive && t < n)) continue;
    const int jj = base + (int)list[t * 64 + lane];
    const double4 r = v.nrec[jj];
    const double2 fx = v.nfx[jj];
    const int f = (int)fx.y;
    if (f != a.cur_f) {
      a.cur_f = f;
      a.xsoffset = a.commul_last;
      a.xssum = 0.0;
    }
    double xs2[2];
    if (!scb_candidate<AllowSlow>(v, r, fx.x, w, wl3, L.u.x, L.u.y, L.u.z, xs2[0], xs2[1])) {
      L.slow = true;
      alive = false;
      continue;
    }
    for (int side = 0; side < 2; ++side) {
      const double xs = xs2[side];
      if (!xs) continue;
      const double cm = a.xsoffset + (a.xssum += xs);
      a.commul_last = cm;
      if constexpr (Select) {
        const bool hit = n_select < 5 ? (cm > target) : !(cm < target);
        if (hit || L.count == n_select - 1) {
          L.found = true;
          L.sel_inv2d_r = r.w;
          L.sel_normal = side == 0 ? V3{-r.x, -r.y, -r.z} : V3{r.x, r.y, r.z};
          alive = false;
          break;
        }
      }
      L.count++;
    }
  }
}

// FP32 pre-filter of demi-normal record r (FP32 normal, rounded 1/(2d)) for
// a lane with FP32 direction f and wavelength fw: contracted FP32 arithmetic,
// the band covers its rounding
NCD bool scb_prefilter(float rx, float ry, float rz, float rw, float fx, float fy, float fz, float fw,
                       float band32) {
  const float d32 = __fmaf_rn(rz, fz, __fmaf_rn(ry, fy, rx * fx));
  const float s32 = fw * rw;
  return fabsf(fabsf(d32) - s32) < band32;
}

// wave-level filter applies while the directions of the wave's lanes lie
// within this chord of their mean direction (beyond it too few demi-normals
// are ruled out to pay for the second test)
constexpr double kScbCoarseChord = 0.2;

// Scan of all (family, demi-normal, side) contributions in the reference's
// order (families by decreasing d, normals in family order, anti-normal side
// first), NCSCBragg.cc:263-300.  Select as for scb_eval_list.
//
// A normal can only contribute if the angle phi between the neutron
// direction and its plane is within the truncation angle of the Bragg angle
// theta; since |sin(phi) - sin(theta)| <= |phi - theta|, the FP32 test
// | |n.u| - sin(theta) | < band32 (band32 = truncation angle plus margins
// for FP32 rounding, about 1e-5) is a superset of the reference's cone test.
// Candidates go to the lane's list and are evaluated with the FP64 record
// and the reference's test and arithmetic unchanged whenever a list could
// overflow and at the end.  Per-family state (cumulative offset) is
// (re)initialised at the first candidate of each family, which leaves every
// value as the reference's family loop computes it.
//
// Two scans give the same lists:
//  - coarse: when the wave's directions lie within a chord rho of their mean
//    m, lane t first tests demi-normal jb + t against the whole wave:
//    |n.u - n.m| <= |u - m| <= rho for every lane, and sin(theta) lies
//    between fw_lo/(2d) and fw_hi/(2d) (FP32 products are monotone in fw),
//    so a normal with | |n.m| - s | >= rho + band32 for all s in that range
//    (plus 1e-5 for the FP32 dot products) fails every lane's pre-filter.
//    The normals left are broadcast (readlane) and pre-filtered per lane.
//  - dense: every lane pre-filters every normal, in groups of kScbGroup with
//    a wave-uniform index (scalar loads of the FP32 records).
// `complete` is set (wave-uniform) when the list held every candidate of the
// scan, so that a second pass can evaluate it without scanning again.  All
// lanes of the wavefront must call this together.
template <bool Select, bool AllowSlow>
NCD void scb_walk(const SCBView& v, ScbLane& L, int n_select, double target, uint16_t* __restrict__ list,
                  int& nlist, bool& complete) {
  L.count = 0;
  L.total = 0.0;
  L.slow = false;
  L.found = false;
  bool alive = L.jlim > 0;
  const int jmax = wave_max_i(L.jlim);
  const int lane = (int)__lane_id();
  const double w = gm_round(L.wl);
  const double wl3 = w * w * w;
  const float fx = (float)L.u.x, fy = (float)L.u.y, fz = (float)L.u.z, fw = (float)w;
  const float band32 = v.band32;
  ScbAcc a{-1, 0.0, 0.0, 0.0};
  int n = 0, base = 0;
  complete = true;
  // bounds of the wave: mean direction, chord, wavelength range
  const double mx0 = wave_sum_d(alive ? L.u.x : 0.0), my0 = wave_sum_d(alive ? L.u.y : 0.0),
               mz0 = wave_sum_d(alive ? L.u.z : 0.0);
  const double mn = sqrt(mx0 * mx0 + my0 * my0 + mz0 * mz0);
  const double mx = mx0 / mn, my = my0 / mn, mz = mz0 / mn;
  const double cmin = wave_min_d(alive ? L.u.x * mx + L.u.y * my + L.u.z * mz : 1.0);
  const double rho = sqrt(dmax(0.0, 2.0 - 2.0 * cmin));
  // (a NaN in the bounds, e.g. from directions that cancel, selects the dense scan)
  if (jmax > 0 && mn > 0.0 && rho < kScbCoarseChord) {
    const float R = (float)rho + band32 + 1e-5f;
    const float gx = (float)mx, gy = (float)my, gz = (float)mz;
    const float fwlo = wave_min_f(alive ? fw : __builtin_huge_valf());
    const float fwhi = wave_max_f(alive ? fw : -__builtin_huge_valf());
    for (int jb = 0; jb < jmax; jb += 64) {
      const int jl = jb + lane;
      float4 r = make_float4(0.f, 0.f, 0.f, 0.f);
      bool pass = false;
      if (jl < jmax) {
        r = v.nrec32[jl];
        const float D = fabsf(__fmaf_rn(r.z, gz, __fmaf_rn(r.y, gy, r.x * gx)));
        pass = (fwlo * r.w < D + R) && (fwhi * r.w > D - R);
      }
      uint64_t M = __ballot(pass);
      if (!M) continue;
      if (__any(n + __popcll(M) > kScbListCap) || jb + 64 - base > kScbListSpan) {
        scb_eval_list<Select, AllowSlow>(v, L, a, alive, list, base, n, w, wl3, n_select, target);
        n = 0;
        base = jb;
        complete = false;
        if (!__any(alive)) break;
      }
      while (M) {
        const int t = __builtin_ctzll(M);
        M &= M - 1ull;
        const float rx = readlane_f(r.x, t), ry = readlane_f(r.y, t), rz = readlane_f(r.z, t),
                    rw = readlane_f(r.w, t);
        const int j = jb + t;
        if (alive && j < L.jlim && scb_prefilter(rx, ry, rz, rw, fx, fy, fz, fw, band32)) {
          list[n * 64 + lane] = (uint16_t)(j - base);
          ++n;
        }
      }
    }
  } else {
    for (int jb = 0; jb < jmax; jb += kScbGroup) {
      if (__any(n > kScbListCap - kScbGroup) || jb + kScbGroup - base > kScbListSpan) {
        scb_eval_list<Select, AllowSlow>(v, L, a, alive, list, base, n, w, wl3, n_select, target);
        n = 0;
        base = jb;
        complete = false;
        if (!__any(alive)) break;
      }
      unsigned cand = 0u;
#pragma unroll
      for (int q = 0; q < kScbGroup; ++q) {
        const float4 r = v.nrec32[jb + q];
        cand |= (scb_prefilter(r.x, r.y, r.z, r.w, fx, fy, fz, fw, band32) ? 1u : 0u) << q;
      }
      const int left = L.jlim - jb;
      if (left < kScbGroup) cand &= left > 0 ? (1u << left) - 1u : 0u;
      if (!alive) cand = 0u;
      while (cand) {
        const int q = __builtin_ctz(cand);
        cand &= cand - 1u;
        list[n * 64 + lane] = (uint16_t)(jb + q - base);
        ++n;
      }
    }
  }
  scb_eval_list<Select, AllowSlow>(v, L, a, alive, list, base, n, w, wl3, n_select, target);
  nlist = n;
  L.total = a.commul_last;
}

// Second pass over a complete list (scb_walk set `complete`): the same
// contributions as scb_walk<Select> without scanning again.
template <bool AllowSlow>
NCD void scb_select_list(const SCBView& v, ScbLane& L, int n_select, double target, const uint16_t* __restrict__ list,
                         int nlist) {
  L.count = 0;
  L.slow = false;
  L.found = false;
  bool alive = L.jlim > 0;
  const double w = gm_round(L.wl);
  const double wl3 = w * w * w;
  ScbAcc a{-1, 0.0, 0.0, 0.0};
  scb_eval_list<true, AllowSlow>(v, L, a, alive, list, 0, alive ? nlist : 0, w, wl3, n_select, target);
}

// per-lane set-up of a neutron (NCSCBragg.cc:233-259): returns false when
// the neutron cannot reflect (below threshold, zero wavelength).  jlim =
// demi-normals of the families with 1/(2d) below (1-2eps)/wl.
NCD bool scb_lane_setup(const SCBView& v, double E, double dx, double dy, double dz, ScbLane& L) {
  L.jlim = 0;
  if (!(E > v.threshold)) return false;
  const double ekin = scb_round(E);
  const double wl = ekin ? sqrt(kWl2Ekin / ekin) : __builtin_huge_val();
  if (wl == 0) return false;
  L.wl = wl;
  L.u = v3normalise(V3{dx, dy, dz});
  const double inv2dcut = (1.0 - 2 * 2.220446049250313080847263336181640625e-16) / wl;
  L.jlim = v.off[lower_bound_d(v.inv2d, 0, v.nfam, inv2dcut)];
  return true;
}

}  // namespace ncamd
