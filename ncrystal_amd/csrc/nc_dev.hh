// Device-side building blocks of the per-neutron physics for gfx950.
//
// One lane evaluates one neutron.  All arithmetic on the reference-parity
// path is FP64; the library is compiled with -ffp-contract=off so that the
// operation order written here is the operation order executed.
//
// Reference algorithms (NCrystal v4.4.2, paths relative to its repository):
//  - PowderBragg: ncrystal_core/src/powderbragg/NCPowderBragg.cc:152-216
//  - ElInc: ncrystal_core/src/phys_utils/NCElIncXS.cc:34-190
//  - SAB xs: ncrystal_core/src/sab/NCSABXSProvider.cc:54-95
//  - SAB sampling: ncrystal_core/src/sab/NCSABSampler.cc:59-236,
//    ncrystal_core/src/sab/NCSABSamplerModels.cc:48-234,
//    ncrystal_core/src/utils/NCPointwiseDist.cc:76-107
//  - kinematics: ncrystal_core/include/NCrystal/internal/phys_utils/NCKinUtils.hh:64-121,
//    ncrystal_core/src/phys_utils/NCKinUtils.cc:25-58
//  - free gas: cross section and scattering function of
//    ncrystal_core/src/phys_utils/NCFreeGasUtils.cc:64-84 (sampled as a
//    two-body collision, nc_freegas.hh)
//  - RNG to (0,1]: ncrystal_core/include/NCrystal/core/NCDefs.hh:1308-1329
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

#define NCD __device__ __forceinline__
#define NCHD __host__ __device__ __forceinline__

namespace ncamd {

constexpr double kBoltzmann = 8.6173303e-5;
constexpr double kNeutronMassAmu = 1.00866491588;
constexpr double kDalton2eVc2 = 931494095.17;       // [eV/c^2]
constexpr double kSpeedOfLightAas = 299792458e10;   // [Aa/s]
constexpr double kPi = 3.1415926535897932384626433832795028841971694;
constexpr double k2Pi = 6.2831853071795864769252867665590057683943388;
constexpr double kPiHalf = 1.5707963267948966192313216916397514420985847;
constexpr double kPiSq = 9.86960440108935861883449099987615113531369941;
constexpr double kInvSqrtPi = 0.564189583547756286948079451560772585844050629;
constexpr double kWl2Ekin = 0.081804209605330899;
constexpr double kEkin2WlsqInv = 12.22430978582345950656;
constexpr double kDblMin = 2.2250738585072014e-308;

NCHD double dmin(double a, double b) { return a < b ? a : b; }
NCHD double dmax(double a, double b) { return a > b ? a : b; }
NCHD double dclamp(double v, double lo, double hi) { return dmin(dmax(v, lo), hi); }
NCHD bool in_interval(double a, double b, double x) { return (a <= x) & (x <= b); }

// Neumaier compensated sum (NCMath.hh:526-536)
struct StableSum {
  double s = 0.0, c = 0.0;
  NCHD void add(double x) {
    double t = s + x;
    c += fabs(s) >= fabs(x) ? (s - t) + x : (x - t) + s;
    s = t;
  }
  NCHD double sum() const { return s + c; }
};

// ---------------------------------------------------------------------------
// Counter-based RNG: Philox4x32-10.  Stream of a neutron = (key=seed,
// counter=(block, stream, idx_lo, idx_hi)); draw 2j = low 64 bits of block j,
// draw 2j+1 = high 64 bits.

// One Philox4x32 round.  Each product is formed as one 32x32 -> 64-bit
// multiply (v_mad_u64_u32), not as separate low and high halves.
NCD void philox_round(uint32_t& x0, uint32_t& x1, uint32_t& x2, uint32_t& x3, uint32_t q0, uint32_t q1) {
  const uint64_t p0 = (uint64_t)0xD2511F53u * x0;
  const uint64_t p1 = (uint64_t)0xCD9E8D57u * x2;
  x0 = (uint32_t)(p1 >> 32) ^ x1 ^ q0;
  x1 = (uint32_t)p1;
  x2 = (uint32_t)(p0 >> 32) ^ x3 ^ q1;
  x3 = (uint32_t)p0;
}

struct Rng {
  uint32_t k0, k1, c0, c1, c2, c3;
  uint64_t b1;
  int nleft;
  NCD void init(uint64_t seed, uint32_t stream, uint64_t idx) {
    k0 = (uint32_t)seed; k1 = (uint32_t)(seed >> 32);
    c0 = 0; c1 = stream; c2 = (uint32_t)idx; c3 = (uint32_t)(idx >> 32);
    nleft = 0;
  }
  NCD static double to01(uint64_t x) {
    const double r1 = (double)(x >> 11) * 0x1.0p-53;
    const double r2 = (double)(x & 0x7FF) * 0x1.0p-64;
    return (1.0 - r1) - r2;
  }
  // one Philox block as four 32-bit words; drops a cached half block, so
  // that every caller consumes whole blocks in step
  NCD void gen4(uint32_t& w0, uint32_t& w1, uint32_t& w2, uint32_t& w3) {
    uint32_t x0 = c0, x1 = c1, x2 = c2, x3 = c3, q0 = k0, q1 = k1;
#pragma unroll
    for (int r = 0; r < 10; ++r) {
      if (r) { q0 += 0x9E3779B9u; q1 += 0xBB67AE85u; }
      philox_round(x0, x1, x2, x3, q0, q1);
    }
    ++c0;
    nleft = 0;
    w0 = x0; w1 = x1; w2 = x2; w3 = x3;
  }
  NCD double gen() {
    if (nleft) { nleft = 0; return to01(b1); }
    uint32_t x0 = c0, x1 = c1, x2 = c2, x3 = c3, q0 = k0, q1 = k1;
#pragma unroll
    for (int r = 0; r < 10; ++r) {
      if (r) { q0 += 0x9E3779B9u; q1 += 0xBB67AE85u; }
      philox_round(x0, x1, x2, x3, q0, q1);
    }
    ++c0;
    b1 = ((uint64_t)x3 << 32) | x2;
    nleft = 1;
    return to01(((uint64_t)x1 << 32) | x0);
  }
};

// ---------------------------------------------------------------------------
// Position of this thread's entry in a queue that the threads of the
// workgroup with `want` append to: one atomic per workgroup (appends to one
// counter serialise, so a launch that appends per wave spends most of its
// time there).  Every thread of the workgroup must call it: it synchronises
// the workgroup.  lds: NW + 1 workgroup-shared entries, NW = waves per
// workgroup.
template <int NW>
NCD unsigned long long block_append(unsigned long long* counter, bool want, unsigned long long* lds) {
  const unsigned long long m = __ballot(want);
  const int lane = (int)__lane_id(), w = (int)(threadIdx.x >> 6);
  if (lane == 0) lds[w] = (unsigned long long)__popcll(m);
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long tot = 0;
    for (int j = 0; j < NW; ++j) {
      const unsigned long long c = lds[j];
      lds[j] = tot;
      tot += c;
    }
    lds[NW] = tot ? atomicAdd(counter, tot) : 0ull;
  }
  __syncthreads();
  const unsigned long long below = lane ? (m & (~0ull >> (64 - lane))) : 0ull;
  const unsigned long long pos = lds[NW] + lds[w] + (unsigned long long)__popcll(below);
  __syncthreads();  // lds is free for the next call
  return pos;
}

// ---------------------------------------------------------------------------
// Position of this lane's entry in a queue that the active lanes with `want`
// append to: one atomic per wave (the first such lane adds their count) and
// each lane's rank among them.  Every active lane of the wave must call it
// with the same counter.  (Written out because the compiler does not merge
// atomics whose address it sees as a per-lane value.)
NCD unsigned long long wave_append(unsigned long long* counter, bool want) {
  const unsigned long long m = __ballot(want);
  if (!m) return 0ull;
  const int lane = (int)__lane_id();
  const int leader = __ffsll((long long)m) - 1;
  unsigned long long base = 0ull;
  if (lane == leader) base = atomicAdd(counter, (unsigned long long)__popcll(m));
  base = __shfl(base, leader, 64);
  const unsigned long long below = lane ? (m & (~0ull >> (64 - lane))) : 0ull;
  return base + (unsigned long long)__popcll(below);
}

// ---------------------------------------------------------------------------
// searches (std::upper_bound / std::lower_bound semantics)
NCD int upper_bound_d(const double* __restrict__ v, int lo, int hi, double x) {
  int n = hi - lo;
  while (n > 0) {
    int h = n >> 1, m = lo + h;
    if (!(x < v[m])) { lo = m + 1; n -= h + 1; } else n = h;
  }
  return lo;
}
NCD int lower_bound_d(const double* __restrict__ v, int lo, int hi, double x) {
  int n = hi - lo;
  while (n > 0) {
    int h = n >> 1, m = lo + h;
    if (v[m] < x) { lo = m + 1; n -= h + 1; } else n = h;
  }
  return lo;
}
NCD int lower_bound_f(const float* __restrict__ v, int lo, int hi, float x) {
  int n = hi - lo;
  while (n > 0) {
    int h = n >> 1, m = lo + h;
    if (v[m] < x) { lo = m + 1; n -= h + 1; } else n = h;
  }
  return lo;
}

// pickRandIdxByWeight (NCRandUtils.cc:198-222, NCRandUtils.hh:130-133)
NCD int pick_idx(Rng& rng, const double* commul, int n) {
  if (n == 1) return 0;
  const double r = rng.gen();
  if (n < 5) {
    const double rc = commul[n - 1] * r;
    for (int i = 0; i < n; ++i) if (commul[i] > rc) return i;
    return n - 1;
  }
  int it = lower_bound_d(commul, 0, n, commul[n - 1] * r);
  return it < n - 1 ? it : n - 1;
}

// ---------------------------------------------------------------------------
// PowderBragg
struct PBView {
  const double* e2d;    // [n] 2d-spacing thresholds in energy units, descending d
  const double* fdm;    // [n] cumulative d*|F|^2*mult*c (FP64)
  const float* fdm32;   // [n] same, FP32 (compensated build), may be null
  int n;
  double threshold;     // e2d[0]
  // search guides (exact: every bracket contains the plain binary-search answer)
  const int* e_bucket;  // [e_nbkt+1] upper_bound(e2d, bucket lower edge), IEEE-bit buckets of E
  long long e_base;     // bits(e2d[0]) >> e_shift
  int e_shift, e_nbkt;
  const int* c_guide;   // [c_n+1] lower_bound(fdm, g / c_scale), last entry n
  int c_n;
  double c_scale;       // c_n / fdm[n-1]
};
// index of the last plane with 2d >= wavelength (E >= threshold assumed)
NCD int pb_last_valid(const PBView& pb, double E) {
  if (!(E < pb.e2d[pb.n - 1])) return pb.n - 1;
  const long long b = (long long)((unsigned long long)__double_as_longlong(E) >> pb.e_shift) - pb.e_base;
  const int lo = max(pb.e_bucket[b], 1), hi = pb.e_bucket[b + 1];
  return upper_bound_d(pb.e2d, lo, hi, E) - 1;
}
// lower_bound over fdm[0..idx) of a target in [0, fdm[idx]]
NCD int pb_pick_plane(const PBView& pb, int idx, double target) {
  long long g = (long long)(target * pb.c_scale);
  g = g < 0 ? 0 : (g >= pb.c_n ? pb.c_n - 1 : g);
  const int lo = pb.c_guide[g > 0 ? g - 1 : 0];
  int hi = pb.c_guide[g + 2 <= pb.c_n ? g + 2 : pb.c_n];
  hi = hi < idx ? hi : idx;
  return lower_bound_d(pb.fdm, lo < hi ? lo : hi, hi, target);
}
NCD bool pb_active(const PBView& pb, double E) {
  return pb.n > 0 && !(E < pb.threshold) && isfinite(E);
}
NCD double pb_xs(const PBView& pb, double E) {
  if (!pb_active(pb, E)) return 0.0;
  const int idx = pb_last_valid(pb, E);
  const double inv = 1.0 / E;
  return pb.fdm[idx] * inv;
}
NCD double pb_sample_mu(const PBView& pb, Rng& rng, double E) {
  if (!pb_active(pb, E)) return 1.0;
  const int idx = pb_last_valid(pb, E);
  const double u = rng.gen();
  const int i = pb_pick_plane(pb, idx, u * pb.fdm[idx]);
  const double s2 = pb.e2d[i] / E;
  return 1.0 - 2.0 * s2;
}

// ---------------------------------------------------------------------------
// Incoherent elastic (Debye-Waller), up to kMaxElInc species
constexpr int kMaxElInc = 16;
struct ElIncView {
  int n;
  double msd[kMaxElInc];
  double bixs[kMaxElInc];  // bound incoherent xs * scale
};
NCD double eval_1mexpmtdivt(double t) {
  if (t < 0.01) return (1 + t * (-0.5 + t * 0.16666666666666666666666666666666666666666667 * (1. - 0.25 * t)));
  if (t > 24.0) return 1.0 / t;
  t = -t;
  return expm1(t) / t;
}
NCD double exp_smallarg_approx(double x) {
  return 1.0 + x * (1 + x * (0.5 + x * (0.16666666666666666666666666666666666667 + x * (0.04166666666666666666666666666666666667
         + x * (0.00833333333333333333333333333333333333 + x * (0.00138888888888888888888888888888888889
         + x * 0.00019841269841269841269841269841269841))))));
}
NCD double elinc_xs(const ElIncView& ei, double E, double* commul) {
  constexpr double kkk = 16.0 * kPiSq * kEkin2WlsqInv;
  const double e = kkk * E;
  double xs = 0.0;
  for (int i = 0; i < ei.n; ++i) {
    xs += ei.bixs[i] * eval_1mexpmtdivt(ei.msd[i] * e);
    if (commul) commul[i] = xs;
  }
  return xs;
}
NCD double elinc_mu_mono(Rng& rng, double E, double msd) {
  constexpr double kkk = 8.0 * kPiSq * kEkin2WlsqInv;
  const double a = (kkk * E) * msd;
  if (a < 0.01) {
    const double maxval = exp_smallarg_approx(a);
    while (true) {
      const double mu = rng.gen() * 2.0 - 1.0;
      if (rng.gen() * maxval < exp_smallarg_approx(a * mu)) return mu;
    }
  }
  return dclamp(log1p(rng.gen() * expm1(2.0 * a)) / a - 1.0, -1.0, 1.0);
}
NCD double elinc_sample_mu(const ElIncView& ei, Rng& rng, double E) {
  if (ei.n == 1) return elinc_mu_mono(rng, E, ei.msd[0]);
  double commul[kMaxElInc];
  elinc_xs(ei, E, commul);
  const int i = pick_idx(rng, commul, ei.n);
  return elinc_mu_mono(rng, E, ei.msd[i]);
}

// ---------------------------------------------------------------------------
// kinematics
struct Lims { double lo, hi; };
NCHD Lims alpha_limits(double e, double beta) {
  const double kk = e + beta;
  if (!(kk >= 0.0)) return {1.0, -1.0};
  const double a = kk + e;
  const double b = 2.0 * sqrt(e * kk);
  double am;
  if (fabs(beta) < 0.01 * e) {
    const double x = beta / e;
    constexpr double c9 = -715. / 32768., c8 = 429. / 16384., c7 = -33. / 1024., c6 = 21. / 512.,
                     c5 = -7. / 128., c4 = 5. / 64., c3 = -1. / 8., c2 = 1. / 4.;
    am = beta * x * (c2 + x * (c3 + x * (c4 + x * (c5 + x * (c6 + x * (c7 + x * (c8 + x * c9)))))));
  } else {
    am = dmax(0.0, a - b);
  }
  return {am, a + b};
}
NCHD bool mu_iso_at_beta(double beta, double e) { return beta <= e * (-1.0 + 1e-14); }

// ---------------------------------------------------------------------------
// S(alpha,beta) utilities (NCSABUtils.hh:176-300)
NCHD double interp_loglin_fast(double a, double fa, double b, double fb, double x,
                               double logfa, double logfb) {
  const double bma = b - a;
  const double mid = 0.5 * (b + a);
  const bool linlin = (fa * fb == 0.0);
  if (x < mid) {
    const double r = (x - a) / bma;
    return linlin ? (fa + (fb - fa) * r) : exp(logfa + (logfb - logfa) * r);
  } else {
    const double s = (b - x) / bma;
    return linlin ? (fb + (fa - fb) * s) : exp(logfb + (logfa - logfb) * s);
  }
}
NCHD double integrate_alpha_interval_fast(double a1, double s1, double a2, double s2,
                                          double logs1, double logs2) {
  const double da = a2 - a1;
  const double ps = s1 + s2;
  const double ds = s2 - s1;
  if (dmin(s1, s2) < 1e-300) return 0.5 * da * ps;
  if (fabs(ds) > 0.006 * ps) return da * ds / (logs2 - logs1);
  const double y = ds / ps;
  constexpr double c1 = 0.166666666666666666666666666666666666666666667;
  constexpr double c2 = 0.0444444444444444444444444444444444444444444444;
  constexpr double c3 = 0.0232804232804232804232804232804232804232804233;
  const double ysq = y * y;
  return da * ps * (0.5 - ysq * (c1 + ysq * (c2 + ysq * c3)));
}
NCD double sample_loglin_fast(double a, double fa, double b, double fb, double r,
                              double logfa, double logfb) {
  double df = fb - fa;
  if (fa * fb * df != 0.0) {
    const double a_sub_b = a - b;
    const double lfab = logfb - logfa;
    if (a_sub_b * lfab != 0.0)
      return a_sub_b * log(fa * exp(a * lfab / a_sub_b) / (fa + r * df)) / lfab;
    df = 0.0;
  }
  if (!df) return a + r * (b - a);
  const double x = (b - a) * sqrt(r);
  return fa ? b - x : a + x;
}

// Free-gas cross-section shape sigma/sigma_free as a function of
// a^2 = A E / kT (NCFreeGasUtils.cc:64-84 evaluates the same function):
// (1 + 1/(2 a^2)) erf(a) + exp(-a^2) / (a sqrt(pi)).  Both terms are
// positive, so the closed form keeps full relative precision from a -> 0
// (where it tends to 2/(a sqrt(pi))) to a -> inf (1 + 1/(2 a^2)).
NCHD double freegas_xs_shape(double a2) {
  const double a = sqrt(a2);
  if (a == 0.0) return __builtin_huge_val();
  const double inva = 1.0 / a;
  return (1.0 + 0.5 / a2) * erf(a) + kInvSqrtPi * exp(-a2) * inva;
}

// ---------------------------------------------------------------------------
// S(alpha,beta) device tables.  Per energy-grid point the Alg1 sampler data
// live at fixed strides: beta distribution at ie*(nb+1), alpha-sampler infos
// at ie*nb.
struct AInfo {
  double f_alpha, f_sval, f_logsval;
  double b_alpha, b_sval, b_logsval;
  double prob_front, prob_notback;
  int f_idx, b_idx;
};
static_assert(sizeof(AInfo) == 72, "AInfo layout");

struct SABView {
  int na, nb, ne;
  const double* alpha;    // [na]
  const double* beta;     // [nb]
  const double2* sl;      // [nb*na] (S, log S)
  const double* cumul;    // [nb*na] cumulative alpha integrals per beta row
  const double* egrid;    // [ne]
  const double* xs;       // [ne]
  const int4* shdr;       // [ne] (kind, ibeta_low, nvals, ninfos)
  const double* sfirst;   // [ne] first-bin kinematic endpoint value
  const double* scdf;     // [ne*(nb+1)]
  const double2* sxy;     // [ne*(nb+1)] (x, normalised y)
  const AInfo* sinfo;     // [ne*nb]
  double kT, T, mass, fg_A, sigma_free, fg_ca, k_ext, k1, k2;  // fg_A: mass / neutron mass
  double egrid_margin;
  // Search acceleration.  Every table stores, per bucket, the search result
  // at the bucket's lower edge, so the bracket [table[b], table[b+1]] always
  // contains the result of the full binary search: identical answers, fewer
  // dependent loads.
  const int* eg_bucket;            // [eg_nbkt+1]: upper_bound(egrid, lower edge of IEEE-bit bucket b)
  int eg_nbkt, eg_shift;
  long long eg_base;               // bucket b = (bits(E) >> eg_shift) - eg_base
  const unsigned short* cdf_guide; // [ne][kCdfGuide+1]: lower_bound(cdf, g/kCdfGuide)
  const unsigned short* row_guide; // [nb][kRowGuide+1]: upper_bound(cumul row, g/row_scale)
  const double* row_scale;         // [nb]: kRowGuide / cumul[row][na-1] (0: row without area)
};
constexpr int kCdfGuide = 256;  // power of two: p*kCdfGuide is exact
constexpr int kRowGuide = 64;

NCD double sab_fg_xs(const SABView& s, double E) { return s.sigma_free * freegas_xs_shape(s.fg_ca * E); }

// upper_bound(egrid, E) through the IEEE-bit bucket table
NCD int egrid_upper_bound(const SABView& s, double E) {
  if (!(E >= s.egrid[0])) return (E < s.egrid[0]) ? 0 : s.ne;  // below grid / NaN
  if (E >= s.egrid[s.ne - 1]) return s.ne;
  const long long b = (long long)((unsigned long long)__double_as_longlong(E) >> s.eg_shift) - s.eg_base;
  const int lo = s.eg_bucket[b], hi = s.eg_bucket[b + 1];
  return upper_bound_d(s.egrid, lo, hi, E);
}

NCD double sab_xs(const SABView& s, double E) {
  const int up = egrid_upper_bound(s, E);
  if (up == s.ne) return s.k_ext / E + sab_fg_xs(s, E);
  if (up == 0) return E > 0.0 ? sqrt(s.egrid[0] / E) * s.xs[0] : __builtin_huge_val();
  const int lo = up - 1;
  const double x0 = s.xs[lo], x1 = s.xs[up];
  const double e0 = s.egrid[lo], e1 = s.egrid[up];
  return x0 + (x1 - x0) * (E - e0) / (e1 - e0);
}

// PointwiseDist::percentileWithIndex
NCD double pw_percentile(const double* __restrict__ cdf, const double2* __restrict__ xy,
                         const unsigned short* __restrict__ guide, int n, double p, int& idx) {
  if (p == 1.) { idx = n - 2; return xy[n - 1].x; }
  const int g = (int)(p * kCdfGuide);  // exact scaling; p in (0,1) -> g in [0, kCdfGuide-1]
  int lb = lower_bound_d(cdf, guide[g], guide[g + 1], p);
  int i = lb < n - 1 ? lb : n - 1;
  if (i < 1) i = 1;
  const double2 p0 = xy[i - 1], p1 = xy[i];
  const double dx = p1.x - p0.x;
  const double c = (p - cdf[i - 1]);
  const double a = p0.y;
  const double d = p1.y - a;
  double zdx;
  if (!a) {
    zdx = d > 0.0 ? sqrt((2.0 * c * dx) / d) : 0.5 * dx;
  } else {
    const double e = d * c / (dx * a * a);
    if (fabs(e) > 1e-7) zdx = (sqrt(1.0 + 2.0 * e) - 1.0) * dx * a / d;
    else zdx = (1 + 0.5 * e * (e - 1.0)) * c / a;
  }
  idx = i - 1;
  return dclamp(p0.x + zdx, p0.x, p1.x);
}

NCD double alg1_sample_alpha(const SABView& s, int ie, int ibl, int ibeta, double rp) {
  const AInfo info = s.sinfo[(size_t)ie * s.nb + (ibeta - ibl)];
  const size_t row = (size_t)s.na * ibeta;
  const double* __restrict__ cumul = s.cumul + row;
  const double2* __restrict__ sl = s.sl + row;
  const double* __restrict__ ag = s.alpha;
  if (rp <= info.prob_front) {
    if (info.prob_front == 2.0) {
      const double da = info.b_alpha - info.f_alpha;
      return info.f_alpha + rp * da;
    } else if (info.prob_front == 1.0) {
      return sample_loglin_fast(info.f_alpha, info.f_sval, info.b_alpha, info.b_sval, rp,
                                info.f_logsval, info.b_logsval);
    } else {
      const double p2 = dclamp(rp / info.prob_front, kDblMin, 1.0);
      const int ia = info.f_idx;
      const double2 q = sl[ia];
      return sample_loglin_fast(info.f_alpha, info.f_sval, ag[ia], q.x, p2, info.f_logsval, q.y);
    }
  } else if (rp <= info.prob_notback) {
    const double p2 = dclamp((rp - info.prob_front) / (info.prob_notback - info.prob_front), 0.0, 1.0);
    const int il = info.f_idx, iu = info.b_idx;
    const double cl = cumul[il], cu = cumul[iu];
    const double sel = cl + p2 * (cu - cl);
    int e;
    const double rs = s.row_scale[ibeta];
    if (rs > 0.0) {
      // the global upper_bound of sel lies in the guide bracket (one bucket of
      // margin either side absorbs rounding of sel*rs); clamping it to iu+1
      // equals the bounded search of the reference (cumul[il] <= sel)
      int g = (int)(sel * rs);
      g = g < 0 ? 0 : (g > kRowGuide ? kRowGuide : g);
      const unsigned short* rg = s.row_guide + (size_t)ibeta * (kRowGuide + 1);
      const int lo = rg[g > 0 ? g - 1 : 0];
      const int hi = rg[g + 2 <= kRowGuide ? g + 2 : kRowGuide];
      e = upper_bound_d(cumul, lo, hi, sel);
      if (e > iu + 1) e = iu + 1;
    } else {
      e = upper_bound_d(cumul, il, iu + 1, sel);
    }
    if (e > iu) return ag[iu];
    if (e <= il) return ag[il];
    const int el = e - 1;
    const double c0 = cumul[el], c1 = cumul[e];
    const double binArea = c1 - c0;
    const double rr = dclamp((sel - c0) / binArea, kDblMin, 1.0);
    const double2 q0 = sl[el], q1 = sl[e];
    return sample_loglin_fast(ag[el], q0.x, ag[e], q1.x, rr, q0.y, q1.y);
  } else {
    const double p2 = dclamp((rp - info.prob_notback) / (1.0 - info.prob_notback), kDblMin, 1.0);
    const int ib = info.b_idx;
    const double2 q = sl[ib];
    return sample_loglin_fast(ag[ib], q.x, info.b_alpha, info.b_sval, p2, q.y, info.b_logsval);
  }
}

// status codes written into the per-call error counter
enum : int { kOK = 0, kErrAlg1Loop = -1, kErrSamplerLoop = -2, kErrBadSampler = -3,
             kErrHighEDiscard = -4, kErrKinematics = -5 };

// SABSamplerAtE_Alg1::sampleAlphaBeta (with the index -1 read of the
// reference replaced by a rejection; see the oracle for the same choice)
NCD int alg1_sample(const SABView& s, int ie, double ediv, Rng& rng, double& a_out, double& b_out) {
  const int4 hdr = s.shdr[ie];
  const int ibl = hdr.y, nv = hdr.z;
  const double firstbin = s.sfirst[ie];
  const double* __restrict__ cdf = s.scdf + (size_t)ie * (s.nb + 1);
  const double2* __restrict__ xy = s.sxy + (size_t)ie * (s.nb + 1);
  const unsigned short* __restrict__ guide = s.cdf_guide + (size_t)ie * (kCdfGuide + 1);
  const double* __restrict__ bg = s.beta;
  for (int it = 0; it < 100; ++it) {
    int ibs;
    double beta = pw_percentile(cdf, xy, guide, nv, rng.gen(), ibs);
    if (ibs == 0 && firstbin <= 0.0) {
      const double b0 = firstbin;
      const double b1 = xy[1].x;
      if (b1 < -ediv) continue;
      const double db = b1 - b0;
      double alphaval = 0.0;
      for (int iii = 0; iii < 30; ++iii) {
        beta = dmax(firstbin, b0 + db * rng.gen());
        if (beta < -ediv) break;
        alphaval = alg1_sample_alpha(s, ie, ibl, ibl, rng.gen());
        const Lims l = alpha_limits(-firstbin, beta);
        if (in_interval(l.lo, l.hi, alphaval)) break;
        if (iii == 29) { alphaval = 0.5 * (l.lo + l.hi); break; }
      }
      if (beta < -ediv) continue;
      const Lims l = alpha_limits(ediv, beta);
      if (in_interval(l.lo, l.hi, alphaval)) { a_out = alphaval; b_out = beta; return kOK; }
      continue;
    }
    if (beta <= dmax(-ediv, bg[0])) continue;
    const double rp = rng.gen();
    const int ibeta = ibl + ibs;
    if (ibeta < ibl + 1) continue;
    const double bl = bg[ibeta - 1];
    const double al = alg1_sample_alpha(s, ie, ibl, ibeta - 1, rp);
    const double bh = bg[ibeta];
    const double ah = alg1_sample_alpha(s, ie, ibl, ibeta, rp);
    const double alpha = al + (ah - al) * (beta - bl) / (bh - bl);
    const Lims l = alpha_limits(ediv, beta);
    if (in_interval(l.lo, l.hi, alpha)) { a_out = alpha; b_out = beta; return kOK; }
  }
  return kErrAlg1Loop;
}

}  // namespace ncamd

#include "nc_freegas.hh"

namespace ncamd {

// SABSampler::sampleHighE (NCSABSampler.cc:59-156); a_out=-1 requests the
// table sampler at Emax.
NCD int sab_sample_highe(const SABView& s, double E, Rng& rng, double& a_out, double& b_out) {
  const double emax = s.egrid[s.ne - 1];
  const double exm = E * sab_fg_xs(s, E);
  const double P_inside = s.k1 / ((s.k1 - s.k2) + exm);
  const double P_ext_inside = s.k2 / exm;
  const double P_discard = (P_ext_inside >= P_inside ? (1.0 - P_inside / P_ext_inside) : 0.0);
  if (P_discard > 0.95) return kErrHighEDiscard;
  if (P_ext_inside < P_inside) {
    const double aa = 1.0 - P_ext_inside;
    const double P_extra = aa > 1e-10 ? (P_inside - P_ext_inside) / aa : 1.0;
    if (rng.gen() < P_extra) { a_out = -1.0; b_out = 0.0; return kOK; }
  }
  const double emax_div_kt = emax / s.kT;
  for (int it = 0; it < 10000; ++it) {  // a discard happens with probability < 0.95
    FGDraw d;
    const int rc = fg_collide(E / s.kT, s.fg_A, rng, d);
    if (rc != kOK) return rc;
    const double al = d.alpha, be = d.beta;
    if (be <= -emax_div_kt) { a_out = al; b_out = be; return kOK; }
    const Lims l = alpha_limits(emax_div_kt, be);
    if (!in_interval(l.lo, l.hi, al)) { a_out = al; b_out = be; return kOK; }
    if (P_discard && rng.gen() < P_discard) continue;
    a_out = -1.0; b_out = 0.0; return kOK;
  }
  return kErrSamplerLoop;
}

// SABSampler::sampleAlphaBeta (NCSABSampler.cc:158-227)
template <bool WithHighE = true>
NCD int sab_sample_ab(const SABView& s, double E, Rng& rng, double& a_out, double& b_out) {
  int up = egrid_upper_bound(s, E);
  bool ultra = false;
  const double ultra_e = s.egrid[0];
  int isamp;
  if (up == s.ne) {
    if constexpr (!WithHighE) return kErrBadSampler;  // caller routes E >= Emax elsewhere
    double a, b;
    const int rc = sab_sample_highe(s, E, rng, a, b);
    if (rc) return rc;
    if (a >= 0.0) { a_out = a; b_out = b; return kOK; }
    E = s.egrid[s.ne - 1];
    isamp = s.ne - 1;
  } else if (up == 0) {
    isamp = 0;
    ultra = (E < ultra_e);
  } else {
    while (up + 1 != s.ne && E * s.egrid_margin > s.egrid[up]) ++up;
    isamp = up;
  }
  const int kind = s.shdr[isamp].x;
  const double ediv = E / s.kT;
  const double sdiv = ultra ? ultra_e / s.kT : ediv;
  for (int loop = 0; loop < 100; ++loop) {
    double alpha, beta;
    if (kind == 0) { alpha = 0.0; beta = 0.0; }
    else if (kind != 1) return kErrBadSampler;
    else {
      const int rc = alg1_sample(s, isamp, sdiv, rng, alpha, beta);
      if (rc) return rc;
    }
    if (beta < -ediv) continue;
    const Lims l = alpha_limits(ediv, beta);
    if (in_interval(l.lo, l.hi, alpha)) { a_out = alpha; b_out = beta; return kOK; }
    if (ultra) { a_out = l.lo + rng.gen() * (l.hi - l.lo); b_out = beta; return kOK; }
  }
  return kErrSamplerLoop;
}

// SABSampler::sampleDeltaEMu + SABScatter::sampleScatterIsotropic:
// (alpha, beta) -> (E', mu) (ncrystal_core/src/phys_utils/NCKinUtils.cc:25-58)
NCD int sab_ab_to_emu(const SABView& s, Rng& rng, double E, double a, double b, double& Ef, double& mu) {
  const double kT = s.kT;
  double dE, m;
  if (mu_iso_at_beta(b, E / kT)) {
    dE = b * kT;
    m = rng.gen() * 2.0 - 1.0;
  } else {
    dE = b * kT;
    const double ef = E + dE;
    const double denom = 2.0 * sqrt(E * ef);
    if (!denom) return kErrKinematics;
    StableSum sm;
    sm.add(E); sm.add(ef); sm.add(-a * kT);
    m = dclamp(sm.sum() / denom, -1.0, 1.0);
  }
  Ef = dmax(0.0, E + dE);
  mu = m;
  return kOK;
}
template <bool WithHighE = true>
NCD int sab_sample_iso(const SABView& s, Rng& rng, double E, double& Ef, double& mu) {
  double a, b;
  const int rc = sab_sample_ab<WithHighE>(s, E, rng, a, b);
  if (rc) return rc;
  return sab_ab_to_emu(s, rng, E, a, b, Ef, mu);
}

// randDirectionGivenScatterMu (NCRandUtils.cc:51-96)
NCD void rand_dir_given_mu(Rng& rng, double mu, double ix, double iy, double iz,
                           double& ox, double& oy, double& oz) {
  const double m2 = ix * ix + iy * iy + iz * iz;
  const double invm = (fabs(m2 - 1.0) < 1e-14 ? 1.0 : 1.0 / sqrt(m2));
  double u0 = ix * invm, u1 = iy * invm, u2 = iz * invm;
  double t0 = 0, t1 = 0, t2 = 0, tm2 = 0.0;
  do {
    const double x0 = 2.0 * rng.gen() - 1.0;
    const double x1 = 2.0 * rng.gen() - 1.0;
    const double s = x0 * x0 + x1 * x1;
    if (s < 1.0) {
      const double t = 2.0 * sqrt(1.0 - s);
      const double a0 = x0 * t, a1 = x1 * t, a2 = 1.0 - 2.0 * s;
      t0 = a1 * u2 - a2 * u1; t1 = a2 * u0 - a0 * u2; t2 = a0 * u1 - a1 * u0;
      tm2 = t0 * t0 + t1 * t1 + t2 * t2;
    }
  } while (tm2 < 0.001);
  u0 *= mu; u1 *= mu; u2 *= mu;
  const double f = sqrt((1 - mu * mu) / tm2);
  t0 *= f; t1 *= f; t2 *= f;
  ox = u0 + t0; oy = u1 + t1; oz = u2 + t2;
}

}  // namespace ncamd
