// VDOS -> S(alpha,beta): NCrystal's expansion of a phonon density of states
// into a scattering kernel (createScatteringKernel,
// ncrystal_core/src/vdos/NCVDOSToScatKnl.cc:787-944), for the GPU.
//
// Host (init-time, sequential): regularisation of the VDOS grid
// (NCVDOSEval.cc:466-640), normalisation, gamma0 and the MSD (VDOSEval,
// NCVDOSEval.cc:147-445), the phonon-order functions G_n as the chain
// G_n = G_{n-n/2} (x) G_{n/2} of FFT convolutions with truncation and
// thinning (NCVDOSGn.cc:133-492), the number of orders and the alpha/beta
// grids (NCVDOSToScatKnl.cc:332-785).
//
// GPU: the Sjolander sum S(alpha_i, beta_j) = sum_n A[n][i] G[n][j] over the
// orders (NCVDOSToScatKnl.cc:99-297), with A[n][i] = x^n exp(-x)/n! kT
// (x = alpha_i 2 m kT msd / hbar^2) and G[n][j] = G_n(beta_j kT), is a
// product of two order-major tables: two operand kernels fill them and an
// FP64 matrix-core kernel (v_mfma_f64_16x16x4f64) forms the product for the
// rows beta <= 0; the rows beta > 0 follow by detailed balance,
// S(alpha, +beta) = S(alpha, -beta) exp(-beta), in the same kernel.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <exception>
#include <memory>
#include <thread>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <queue>
#include <stdexcept>
#include <string>
#include <vector>

#include "nc_vdos.hh"

namespace ncamd {
namespace {

constexpr double kKB = 8.6173303e-5;
constexpr double kNMassAmu = 1.00866491588;
constexpr double kC = 299792458e10;
constexpr double kDalton2eVc2 = 931494095.17;
constexpr double kHbar = 4.135667662e-15 * 0.159154943091895335768883763372514362034459645740;
constexpr double kNMassEvc2 = kNMassAmu * kDalton2eVc2 / (kC * kC);
constexpr double kInvSqrt2Pi = 0.398942280401432677939946059934381868475858631;
constexpr double kInvE = 0.367879441171442321595523770161460867445811131;
constexpr double k2Pi = 6.2831853071795864769252867665590057683943388;

[[noreturn]] void fail(const char* type, const std::string& msg) { throw VdosError(type, msg); }

#define VDOS_HIP(call)                                                                          \
  do {                                                                                          \
    hipError_t e_ = (call);                                                                     \
    if (e_ != hipSuccess) fail("HIPError", std::string(#call) + ": " + hipGetErrorString(e_));  \
  } while (0)

struct Sum {
  double s = 0.0, c = 0.0;
  void add(double x) {
    const double t = s + x;
    c += std::fabs(s) >= std::fabs(x) ? (s - t) + x : (x - t) + s;
    s = t;
  }
  double sum() const { return s + c; }
};

std::vector<double> linspace(double a, double b, unsigned n) {
  std::vector<double> v(n);
  const unsigned nm1 = n - 1;
  const double dx = (b - a) / nm1;
  for (unsigned i = 0; i < nm1; ++i) v[i] = a + i * dx;
  v[nm1] = b;
  return v;
}
std::vector<double> geomspace(double a, double b, unsigned n) {
  std::vector<double> v(n);
  double start = std::log10(a);
  const double stop = std::log10(b);
  const double interval = (stop - start) / (n - 1);
  for (unsigned i = 0; i < n; ++i) { v[i] = std::pow(10.0, start); start += interval; }
  v[n - 1] = std::pow(10.0, stop);
  v[0] = a;
  v[n - 1] = b;
  return v;
}

// Romberg integration with 17 (level 4) or 33 (level 5) points, the
// integrateRomberg17/33 of NCMath.hh:580-604
template <class F>
double romberg(const F& f, double a, double b, bool lvl5) {
  double h = (b - a);
  double fv[17];
  const double delta = h * 0.0625;
  for (int i = 0; i < 17; ++i) fv[i] = f(a + delta * (double)i);
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
  if (!lvl5) return R44;
  double c5 = 0.0;
  const double off = a + h * 0.5;
  for (double i = 0; i < 16.0; ++i) c5 += f(off + h * i);
  h *= 0.5;
  const double R50 = h * c5 + 0.5 * R40;
  const double R51 = (4. / 3.) * R50 + (-1. / 3.) * R40;
  const double R52 = (16. / 15.) * R51 + (-1. / 15.) * R41;
  const double R53 = (64. / 63.) * R52 + (-1. / 63.) * R42;
  const double R54 = (256. / 255.) * R53 + (-1. / 255.) * R43;
  return (1024. / 1023.) * R54 + (-1. / 1023.) * R44;
}

// ---------------------------------------------------------- regularisation
bool is_linspaced(const std::vector<double>& g, double tol) {
  const size_t n = g.size();
  if (n == 2) return true;
  const double bw = (g.back() - g.front()) / (n - 1.0);
  const double eps = tol * bw;
  for (size_t i = 0; i < n; ++i)
    if (std::fabs((g.front() + (double)i * bw) - g[i]) > eps) return false;
  return true;
}
double check_regular(double emin, double emax, unsigned long npts, double tol) {
  const double bwa = (emax - emin) / (npts - 1);
  const double nb = emin / bwa;
  if (nb < 0.99 || std::fabs(nb - std::round(nb)) > tol) return 0.0;
  const unsigned long nbi = (unsigned long)(nb + 0.5);
  const double bw = emin / nbi;
  return emin + (npts - 1) * bw;
}

}  // namespace

RegularVdos vdos_regularise(const std::vector<double>& egrid, const std::vector<double>& dens) {
  const size_t n = dens.size();
  if (n < 5) fail("BadInput", "Too few points in VDOS density.");
  if (!(egrid.size() == 2 || egrid.size() == n))
    fail("BadInput", "VDOS egrid must have two points, or the same number of points as the density.");
  for (size_t i = 1; i < egrid.size(); ++i)
    if (!(egrid[i] > egrid[i - 1])) fail("BadInput", "VDOS egrid must be increasing.");
  if (egrid.front() < 1e-5)
    fail("BadInput", "VDOS energy range can not be specified for values less than 1e-5eV = 0.01meV");
  for (double d : dens)
    if (!(d >= 0.0) || !std::isfinite(d)) fail("BadInput", "VDOS density values must be finite and >= 0.");
  const double tol = 1e-6;
  RegularVdos r;
  if (is_linspaced(egrid, tol)) {
    const double ec = check_regular(egrid.front(), egrid.back(), (unsigned long)n, tol);
    if (ec) {
      r.emin = egrid.front();
      r.emax = ec;
      r.density = dens;
      return r;
    }
  }
  const double emin = egrid.front(), oldEmax = egrid.back();
  const double d = oldEmax - emin, ratio = d / emin;
  double k = std::max(1.0, std::round(2000.0 / ratio));
  double best_eps = HUGE_VAL, best_k = 0.0;
  while (true) {
    const double m = std::floor(ratio * k);
    if (m < 1.0) { k += 1.0; continue; }
    const double bw = emin / k;
    const double eps = std::max(0.0, d - (m * bw));
    if (eps == 0.0 || best_k == 0.0 || (best_k * std::floor(ratio * k) > k * std::floor(ratio * best_k))) {
      best_eps = eps;
      best_k = k;
    }
    double t = 1e-6;
    if (m > 5000) {
      t = 1e-5;
      if (m > 10000) {
        t = 1e-4;
        if (m > 15000) t = m > 19000 ? 1e-2 : 1e-3;
      }
    }
    if (best_eps < d * t) break;
    if (m >= 20000) fail("BadInput", "Could not regularise input energy grid. Are the energy ranges highly unusual?");
    k += 1.0;
  }
  const double nbw = emin / best_k;
  unsigned npts = (unsigned)(std::floor(ratio * best_k) + 0.5) + 1;
  double nemax = emin + nbw * (npts - 1);
  if (nemax < oldEmax) {
    ++npts;
    nemax = emin + nbw * (npts - 1);
  }
  const std::vector<double> ex = egrid.size() == 2 ? linspace(egrid[0], egrid[1], (unsigned)n) : egrid;
  const double inv_bw = egrid.size() == 2 ? ((n - 1.0) / (egrid[1] - egrid[0])) : 0.0;
  r.density.resize(npts);
  size_t it = 0;
  const size_t last = n - 1;
  for (unsigned i = 0; i < npts; ++i) {
    const double ev = (i + 1 == npts) ? nemax : emin + nbw * i;
    while (it != last && ev >= ex[it + 1]) ++it;
    if (ev == ex[it]) { r.density[i] = dens[it]; continue; }
    if (it == last) {
      r.density[i] = ev > ex[it] ? 0.0 : dens[last];
    } else {
      const double y0 = dens[it], y1 = dens[it + 1], x0 = ex[it], x1 = ex[it + 1];
      Sum s;
      s.add(y0 * x1);
      s.add(-y0 * ev);
      s.add(ev * y1);
      s.add(-x0 * y1);
      r.density[i] = inv_bw ? s.sum() * inv_bw : s.sum() / (x1 - x0);
    }
  }
  r.emin = emin;
  r.emax = nemax;
  return r;
}

void vdos_debye(double debye_temp, double* emin, double* emax, std::vector<double>& dens) {
  const double ed = kKB * debye_temp;
  const std::vector<double> eg = linspace(0.5 * ed, ed, 20);
  const double scale = 1.0 / (ed * ed);
  dens.resize(20);
  for (int i = 0; i < 20; ++i) dens[i] = eg[i] * eg[i] * scale;
  *emin = eg.front();
  *emax = eg.back();
}

namespace {

// ---------------------------------------------------------------- VDOSEval
double safe_xcothx(double x) {
  if (x < 0.1) {
    const double y = x * x;
    return 1.0 + y * (1. / 3 + y * (-1. / 45 + y * (2. / 945 + y * (-1. / 4725 + y * (2. / 93555 +
           y * (-1382. / 638512875. + y * (4. / 18243225.)))))));
  }
  return x / std::tanh(x);
}
double safe_x3cothx(double x) {
  if (x < 0.1) {
    const double y = x * x;
    return y * (1.0 + y * (1. / 3 + y * (-1. / 45 + y * (2. / 945 + y * (-1. / 4725 + y * (2. / 93555 +
           y * (-1382. / 638512875 + y * (4. / 18243225 + y * (-3617. / 162820783125 +
           y * (87734. / 38979295480125))))))))));
  }
  return x * x * x / std::tanh(x);
}
double safe_xdivsinhx(double x) {
  if (x < 0.07) {
    const double y = x * x;
    return 1.0 + y * (-1. / 6. + y * (7. / 360. + y * (-31. / 15120. + y * (127. / 604800.))));
  }
  return x / std::sinh(x);
}
double safe_rec_expm1(double x) {
  if (x <= 700) return 1.0 / std::expm1(x);
  const double emx = std::exp(-x);
  return emx / (1.0 - emx);
}

struct Eval {
  std::vector<double> d;
  double emin, emax, kT, T, mass, bw, ibw, k, integral;
  unsigned long nptsext;

  Eval(const RegularVdos& v, double temperature, double mass_amu) {
    const int n = (int)v.density.size();
    if (!(mass_amu > 0.5 && mass_amu < 2000.0)) fail("BadInput", "element mass out of range");
    if (!(temperature > 0.0 && temperature <= 1e6)) fail("BadInput", "temperature out of range");
    if (v.emin < 1e-5)
      fail("BadInput", "VDOS energy range should not be specified for values less than 1e-5eV = 0.01meV");
    if (!check_regular(v.emin, v.emax, (unsigned long)n, 1e-6)) fail("BadInput", "Received non-regularised VDOS.");
    d = v.density;
    emin = v.emin;
    emax = v.emax;
    T = temperature;
    kT = kKB * temperature;
    mass = mass_amu;
    bw = (emax - emin) / (n - 1);
    nptsext = (unsigned long)n + (unsigned long)(emin / bw + 0.5);
    k = d[0] / (emin * emin);
    ibw = 1.0 / bw;
    Sum s;
    const double onethird = 1.0 / 3.0;
    s.add(onethird * d[0] * emin);
    bins([](double) { return 1.0; }, s);
    integral = s.sum();
    if (!(integral > 0.0)) fail("BadInput", "VDOS integral is not positive");
    const double sf = 1.0 / integral;
    for (double& x : d) x *= sf;
    k *= sf;
  }
  template <class F>
  void bins(const F& f, Sum& sum) const {
    const unsigned nbins = (unsigned)d.size() - 1;
    for (unsigned ib = 0; ib < nbins; ++ib) {
      const double d0 = d[ib], d1 = d[ib + 1];
      const double e0 = emin + bw * ib;
      const double e1 = (ib + 1 == nbins ? emax : emin + bw * (ib + 1));
      const double A = (d1 - d0) * ibw;
      const double B = d0 - e0 * A;
      sum.add(romberg([&](double e) { return f(e) * (A * e + B); }, e0, e1, false));
    }
  }
  double eval(double e) const {
    if (e <= emin) return k * e * e;
    double rel = std::min(std::max((e - emin) * ibw, -0.5), (double)d.size() + 0.5);
    const int ib = (int)rel;
    if (ib >= (int)d.size() - 1) return 0.0;
    rel = std::min(std::max(rel - ib, 0.0), 1.0);
    return (1.0 - rel) * d[ib] + rel * d[ib + 1];
  }
  double gamma0() const {
    const double twokT = 2.0 * kT, inv2kT = 1.0 / twokT;
    Sum s;
    s.add(k * (twokT * twokT) * romberg([](double x) { return safe_xcothx(x); }, 0.0, emin * inv2kT, true));
    bins([inv2kT](double e) { return 1.0 / (e * std::tanh(e * inv2kT)); }, s);
    return emax * s.sum();
  }
  double teff() const {
    const double twokT = 2.0 * kT, inv2kT = 1.0 / twokT;
    Sum s;
    s.add(k * (twokT * twokT * twokT) * romberg([](double x) { return safe_x3cothx(x); }, 0.0, emin * inv2kT, true));
    bins([inv2kT](double e) { return safe_xcothx(e * inv2kT); }, s);
    return T * s.sum();
  }
  double msd(double g0) const {
    const double convfact = 0.5 * (kHbar * kHbar) * (kC * kC) / kDalton2eVc2;
    return convfact * g0 / (mass * emax);
  }
  double g1sym(double e, double g0) const {
    if (e < 0) e = -e;
    const double u = e / (2 * kT);
    if (e <= emin) return (k * kT * emax / g0) * safe_xdivsinhx(u);
    return eval(e) * emax / (e * 2.0 * g0 * std::sinh(u));
  }
  void g1pair(double e, double g0, double& neg, double& pos) const {
    if (e < 200.0 * kT) {
      const double g = g1sym(e, g0);
      if (!g) { neg = pos = 0.0; return; }
      const double dbf = std::exp(e / (2 * kT));
      neg = g * dbf;
      pos = g / dbf;
      return;
    }
    const double kkk = eval(e) * emax / (e * g0);
    if (!kkk) { neg = pos = 0.0; return; }
    const double edkt = e / kT;
    neg = -kkk / std::expm1(-edkt);
    pos = kkk * safe_rec_expm1(+edkt);
  }
};

// --------------------------------------------------------- FFT convolution
// y = dt |IFFT(FFT(a1) FFT(a2))| over the next power of two >= n1 + n2 - 1
// (FastConvolve::convolve, NCFastConvolve.cc:224-268): radix-2, decimation
// in time, twiddles exp(+-2 pi i j / N) from libm.
// Device radix-2 FFT for the large convolutions (the low orders, whose
// spectra have tens of thousands of points).  Every butterfly performs the
// host loop's multiplications, additions and subtractions in the same order
// with contraction into fused multiply-adds disabled, on the same twiddles,
// and the bit-reversal permutation moves values without arithmetic, so the
// device transform is bit-identical to the host one.  Layout of the work
// buffer: re1 | im1 | re2 | im2, n values each.
__global__ void k_fft_load(const double* __restrict__ in, int m1, int m2, int n, int lg, double* __restrict__ buf) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  const int r = (int)(__builtin_bitreverse32((unsigned)j) >> (32 - lg));
  buf[r] = j < m1 ? in[j] : 0.0;
  buf[n + r] = 0.0;
  buf[2 * n + r] = j < m2 ? in[m1 + j] : 0.0;
  buf[3 * n + r] = 0.0;
}

// bit-reversed copy of (re1, im1) into (re2, im2): the inverse transform's
// permutation
__global__ void k_fft_perm(double* __restrict__ buf, int n, int lg) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  const int r = (int)(__builtin_bitreverse32((unsigned)j) >> (32 - lg));
  buf[2 * n + r] = buf[j];
  buf[3 * n + r] = buf[n + j];
}

// one stage (blocks of 2*half points) on array pair blockIdx.y (re at
// buf + 2*y*n, im after it), or on pair `pair` when gridDim.y is 1
__global__ void k_fft_stage(double* __restrict__ buf, int n, int lhalf, int step, const double* __restrict__ wc,
                            const double* __restrict__ ws, int forward, int pair) {
#pragma clang fp contract(off)
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (n >> 1)) return;
  const int y = gridDim.y > 1 ? (int)blockIdx.y : pair;
  double* re = buf + (size_t)2 * y * n;
  double* im = re + n;
  const int half = 1 << lhalf;
  const int j = t & (half - 1);
  const int p = ((t >> lhalf) << (lhalf + 1)) + j, q = p + half;
  const double c = wc[j * step];
  const double sn = forward ? ws[j * step] : -ws[j * step];
  const double rq = re[q], iq = im[q], rp = re[p], ip = im[p];
  const double tr = rq * c - iq * sn;
  const double ti = rq * sn + iq * c;
  re[q] = rp - tr;
  im[q] = ip - ti;
  re[p] = rp + tr;
  im[p] = ip + ti;
}

__global__ void k_fft_mul(double* __restrict__ buf, int n) {
#pragma clang fp contract(off)
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  const double a = buf[j], bb = buf[n + j], c = buf[2 * n + j], d = buf[3 * n + j];
  buf[j] = a * c - bb * d;
  buf[n + j] = a * d + bb * c;
}

// FFT convolution of two sampled spectra (the reference's FFT-based
// convolution as used by NCVDOSGn.cc): zero-padded to a power of two, two
// forward transforms, pointwise product, inverse transform, magnitude times
// dt / n.  Transforms of kDevFFTMin points or more run on the device.
struct Convolver {
  static constexpr int kDevFFTMin = 1 << 15;
  hipStream_t st = nullptr;
  int logn = -1;
  std::vector<double> wc, ws, b;
  // device workspace: twiddles (wc | ws), work buffer, inputs
  double *d_tw = nullptr, *d_buf = nullptr, *d_in = nullptr;
  int d_logn = -1;
  size_t d_in_cap = 0, d_buf_cap = 0;
  Convolver() = default;
  Convolver(const Convolver&) = delete;
  Convolver& operator=(const Convolver&) = delete;
  ~Convolver() {
    if (d_tw) (void)hipFree(d_tw);
    if (d_buf) (void)hipFree(d_buf);
    if (d_in) (void)hipFree(d_in);
  }
  void fft(double* re, double* im, int lg, bool forward) const {
    const int n = 1 << lg;
    for (int j = 1, i = 0; j < n; ++j) {
      int bit = n >> 1;
      for (; i & bit; bit >>= 1) i ^= bit;
      i ^= bit;
      if (j < i) { std::swap(re[i], re[j]); std::swap(im[i], im[j]); }
    }
    for (int len = 2; len <= n; len <<= 1) {
      const int half = len >> 1, step = n / len;
      for (int s = 0; s < n; s += len)
        for (int j = 0; j < half; ++j) {
          const double c = wc[j * step];
          const double sn = forward ? ws[j * step] : -ws[j * step];
          const int p = s + j, q = p + half;
          const double tr = re[q] * c - im[q] * sn;
          const double ti = re[q] * sn + im[q] * c;
          re[q] = re[p] - tr;
          im[q] = im[p] - ti;
          re[p] = re[p] + tr;
          im[p] = im[p] + ti;
        }
    }
  }
  void twiddles(int lg) {
    if (lg == logn) return;
    const int n = 1 << lg;
    logn = lg;
    wc.resize(n / 2 + 1);
    ws.resize(n / 2 + 1);
    // separate libm cos and sin calls (through volatile pointers, which no
    // compiler fuses into sincos: that routine can differ from them in the
    // last bit, and the truncation edges of the G_n depend on such bits)
    double (*volatile fcos)(double) = ::cos;
    double (*volatile fsin)(double) = ::sin;
    for (int j = 0; j <= n / 2; ++j) {
      const double ang = k2Pi * (double)j / (double)n;
      wc[j] = fcos(ang);
      ws[j] = fsin(ang);
    }
  }
  // transformed product on the device; leaves re, im of the inverse
  // transform's first m points in b[0, m) and b[n, n + m)
  void device_product(const std::vector<double>& a1, const std::vector<double>& a2, int lg) {
    const int n = 1 << lg, half = n / 2;
    const size_t nin = a1.size() + a2.size();
    if (nin > d_in_cap) {
      if (d_in) (void)hipFree(d_in);
      d_in = nullptr;
      VDOS_HIP(hipMalloc(&d_in, sizeof(double) * nin));
      d_in_cap = nin;
    }
    if ((size_t)4 * n > d_buf_cap) {
      if (d_buf) (void)hipFree(d_buf);
      d_buf = nullptr;
      VDOS_HIP(hipMalloc(&d_buf, sizeof(double) * 4 * (size_t)n));
      d_buf_cap = (size_t)4 * n;
    }
    if (d_logn != lg) {
      if (d_tw) (void)hipFree(d_tw);
      d_tw = nullptr;
      VDOS_HIP(hipMalloc(&d_tw, sizeof(double) * 2 * (size_t)(half + 1)));
      VDOS_HIP(hipMemcpyAsync(d_tw, wc.data(), sizeof(double) * (half + 1), hipMemcpyHostToDevice, st));
      VDOS_HIP(hipMemcpyAsync(d_tw + half + 1, ws.data(), sizeof(double) * (half + 1), hipMemcpyHostToDevice, st));
      d_logn = lg;
    }
    const int m1 = (int)a1.size(), m2 = (int)a2.size();
    VDOS_HIP(hipMemcpyAsync(d_in, a1.data(), sizeof(double) * m1, hipMemcpyHostToDevice, st));
    VDOS_HIP(hipMemcpyAsync(d_in + m1, a2.data(), sizeof(double) * m2, hipMemcpyHostToDevice, st));
    const double* dwc = d_tw;
    const double* dws = d_tw + half + 1;
    const unsigned gn = (unsigned)((n + 255) / 256), gh = (unsigned)((half + 255) / 256);
    hipLaunchKernelGGL(k_fft_load, dim3(gn), dim3(256), 0, st, d_in, m1, m2, n, lg, d_buf);
    for (int l = 0; l < lg; ++l)
      hipLaunchKernelGGL(k_fft_stage, dim3(gh, 2), dim3(256), 0, st, d_buf, n, l, n >> (l + 1), dwc, dws, 1, 0);
    hipLaunchKernelGGL(k_fft_mul, dim3(gn), dim3(256), 0, st, d_buf, n);
    hipLaunchKernelGGL(k_fft_perm, dim3(gn), dim3(256), 0, st, d_buf, n, lg);
    for (int l = 0; l < lg; ++l)
      hipLaunchKernelGGL(k_fft_stage, dim3(gh, 1), dim3(256), 0, st, d_buf, n, l, n >> (l + 1), dwc, dws, 0, 1);
    VDOS_HIP(hipGetLastError());
    const int m = m1 + m2 - 1;
    b.resize((size_t)n * 2);
    VDOS_HIP(hipMemcpyAsync(b.data(), d_buf + 2 * (size_t)n, sizeof(double) * m, hipMemcpyDeviceToHost, st));
    VDOS_HIP(hipMemcpyAsync(b.data() + n, d_buf + 3 * (size_t)n, sizeof(double) * m, hipMemcpyDeviceToHost, st));
    VDOS_HIP(hipStreamSynchronize(st));
  }
  std::vector<double> convolve(const std::vector<double>& a1, const std::vector<double>& a2, double dt) {
    const int m = (int)(a1.size() + a2.size() - 1);
    int lg = 0;
    while ((1 << lg) < m) ++lg;
    const int n = 1 << lg;
    twiddles(lg);
    double *r1, *i1;
    if (n >= kDevFFTMin) {
      device_product(a1, a2, lg);
      r1 = b.data();
      i1 = r1 + n;
    } else {
      b.assign((size_t)n * 4, 0.0);
      r1 = b.data();
      i1 = r1 + n;
      double *r2 = r1 + 2 * n, *i2 = r1 + 3 * n;
      std::copy(a1.begin(), a1.end(), r1);
      std::copy(a2.begin(), a2.end(), r2);
      fft(r1, i1, lg, true);
      fft(r2, i2, lg, true);
      for (int j = 0; j < n; ++j) {
        const double a = r1[j], bb = i1[j], c = r2[j], d = i2[j];
        r1[j] = a * c - bb * d;
        i1[j] = a * d + bb * c;
      }
      fft(r1, i1, lg, false);
    }
    std::vector<double> y(m);
    const double kf = dt / n;
    for (int j = 0; j < m; ++j) y[j] = std::sqrt(r1[j] * r1[j] + i1[j] * i1[j]) * kf;
    return y;
  }
};

// ------------------------------------------------------------------ VDOSGn
struct Gn {
  std::vector<double> s;
  double lo, bw, hi, smax;
  unsigned long thin;
  Gn(std::vector<double>&& spec, double lower, double binw, unsigned long tf)
      : s(std::move(spec)), lo(lower), bw(binw), thin(tf) {
    hi = lo + (s.size() - 1) * bw;
    double area = 0.0;
    for (double v : s) area += v;
    area *= bw;
    const double inv = 1.0 / area;
    smax = -HUGE_VAL;
    for (double& v : s) { v *= inv; smax = v > smax ? v : smax; }
  }
  void erange(double rel, double& a, double& b) const {
    const double thr = rel * smax;
    a = lo;
    b = hi;
    for (size_t i = 0; i < s.size(); ++i)
      if (s[i] >= thr) { a = lo + i * bw; break; }
    for (size_t i = s.size(); i > 0; --i)
      if (s[i - 1] >= thr) { b = std::min(b, lo + (i - 1) * bw); break; }
  }
};

std::vector<double> thin_vec(const std::vector<double>& v, unsigned long tf) {
  std::vector<double> o;
  o.reserve((v.size() + tf - 1) / tf);
  for (size_t i = 0; i < v.size(); i += tf) o.push_back(v[i]);
  return o;
}

struct GnSet {
  std::vector<Gn> g;
  double kT;
  unsigned cap = 0;
  Convolver cv;
  explicit GnSet(const Eval& ev) : kT(ev.kT) {
    unsigned long nbins = ev.nptsext - 1;
    const unsigned long thicken = (unsigned long)std::ceil(400.0 / nbins);
    nbins *= thicken;
    const std::vector<double> eg = linspace(0.0, ev.emax, (unsigned)(nbins + 1));
    const double bw = eg.back() / nbins;
    std::vector<double> spec(eg.size() * 2 - 1, 0.0);
    const double g0 = ev.gamma0();
    for (size_t i = 0; i < eg.size(); ++i) {
      double neg, pos;
      ev.g1pair(eg[i], g0, neg, pos);
      spec[nbins + i] = pos;
      spec[nbins - i] = neg;
    }
    double lower = -ev.emax;
    size_t first = 0, last = spec.size() - 1;
    while (first != last && !(spec[first] > 0.0) && !(spec[first + 1] > 0.0)) ++first;
    while (last != 0 && !(spec[last] > 0.0) && !(spec[last - 1] > 0.0)) --last;
    if (first >= last || last - first < 3) fail("CalcError", "Too few non-zero pts in G1 spectrum.");
    if (first != 0 || last != spec.size() - 1) {
      lower += first * bw;
      spec = std::vector<double>(spec.begin() + first, spec.begin() + last + 1);
    }
    g.emplace_back(std::move(spec), lower, bw, 1);
  }
  // order n = G_{n - n/2} convolved with G_{n/2}, thinned as the reference
  // does (NCVDOSGn.cc): a function of the two lower orders only
  Gn order_from(unsigned order, Convolver& conv) const {
    const unsigned o2 = order / 2, o1 = order - o2;
    const Gn& p1 = g[o1 - 1];
    const Gn& p2 = g[o2 - 1];
    unsigned long t1 = p1.thin, t2 = p2.thin;
    double dt;
    std::vector<double> tmp;
    const std::vector<double>* in1 = &p1.s;
    const std::vector<double>* in2 = &p2.s;
    if (t1 == t2) {
      dt = p1.bw;
    } else {
      dt = std::max(p1.bw, p2.bw);
      if (t1 > t2) {
        tmp = thin_vec(*in2, t1 / t2);
        in2 = &tmp;
        t2 *= t1 / t2;
      } else {
        tmp = thin_vec(*in1, t2 / t1);
        in1 = &tmp;
        t1 *= t2 / t1;
      }
    }
    double start = p1.lo + p2.lo;
    std::vector<double> y = conv.convolve(*in1, *in2, dt);
    unsigned long xthin = 1;
    if (order >= 5u) {
      const double mx = *std::max_element(y.begin(), y.end());
      const double cut = 1e-14 * mx;
      size_t ifront = 0, iback = y.size() - 1;
      for (; ifront < iback; ++ifront)
        if (y[ifront] > cut) break;
      for (; iback > ifront; --iback)
        if (y[iback] > cut) break;
      if (iback > ifront) y = std::vector<double>(y.begin() + ifront, y.begin() + iback + 1);
      start += ifront * dt;
      if (y.size() > 1000) {
        while (y.size() > 1000ul * xthin) xthin *= 2;
        if (xthin >= 8 && order <= 10u) xthin /= 2;
        y = thin_vec(y, xthin);
        dt *= xthin;
      }
    }
    return Gn(std::move(y), start, dt, t1 * xthin);
  }
  // orders up to target (and, when cap is set, ahead up to cap).  The orders
  // c+1..2c depend only on orders <= c, so each such generation is computed
  // on host threads; every order comes from the same operations in the same
  // sequence as a one-thread growth, so the results do not depend on the
  // thread count
  void grow(unsigned target) {
    while (g.size() < target) {
      const unsigned c = (unsigned)g.size();
      const unsigned hi = std::min(2 * c, std::max(target, cap));
      const unsigned cnt = hi - c;
      unsigned nthr = std::min<unsigned>(cnt, std::min(16u, std::max(1u, std::thread::hardware_concurrency())));
      if (cnt < 3) nthr = 1;
      if (nthr == 1) {
        for (unsigned o = c + 1; o <= hi; ++o) g.push_back(order_from(o, cv));
        continue;
      }
      std::vector<std::unique_ptr<Gn>> res(cnt);
      std::atomic<unsigned> next{0};
      std::vector<std::exception_ptr> errs(nthr);
      auto work = [&](unsigned t) {
        try {
          Convolver conv;
          conv.st = cv.st;
          for (unsigned i = next++; i < cnt; i = next++) res[i].reset(new Gn(order_from(c + 1 + i, conv)));
        } catch (...) {
          errs[t] = std::current_exception();
        }
      };
      std::vector<std::thread> pool;
      for (unsigned t = 1; t < nthr; ++t) {
        try {
          pool.emplace_back(work, t);
        } catch (...) {
          break;  // fewer threads: the others take the remaining orders
        }
      }
      work(0);
      for (auto& th : pool) th.join();
      for (auto& e : errs)
        if (e) std::rethrow_exception(e);
      for (auto& r : res) g.push_back(std::move(*r));
    }
  }
};

// ------------------------------------------------- orders, limits and grids
double alpha2x(double kT, double msd) {
  const double fact = (2.0 * kNMassEvc2 / (kHbar * kHbar));
  return fact * kT * msd;
}
// findRoot (ncrystal_core/src/utils/NCMath.cc:291-332) of y exp(-y) = k
double root_yexp(double k, double a, double b, double acc) {
  double fa = a * std::exp(-a) - k, fb = b * std::exp(-b) - k;
  if (!(b > a) || ((!fa) * fb < 0.0)) fail("CalcError", "root finding requires b>a and f(a)*f(b)<0.");
  acc *= 0.5;
  unsigned i = 60;
  while (--i) {
    double c = (a * fb - b * fa) / (fb - fa);
    if (b - a < acc) return c;
    const double kk = 0.15 * (b - a);
    c = std::max(a + kk, std::min(b - kk, c));
    const double fc = c * std::exp(-c) - k;
    if (!fc) return c;
    if (fa * fc < 0) { b = c; fb = fc; } else { a = c; fa = fc; }
  }
  fail("CalcError", "Root search failed to converge!");
}
void range_xnexpmx(unsigned n, double eps, double& lo, double& hi) {
  const double fn = (double)n;
  const double k = kInvE * std::pow(eps, 1.0 / fn);
  lo = fn * root_yexp(k, 0.0, 1.0, 1e-13);
  hi = fn * root_yexp(k, 1.0, 700.0, 1e-13);
}
bool within_alphaplus(double c, double alpha, double beta) {
  const double cpb = c + beta;
  if (cpb < 0.0) return false;
  const double t = 0.5 * (alpha - beta) - c;
  return t <= 0.0 || c * cpb >= t * t;
}
void extreme_point(double E, double a0, double a1, double b0, double b1, double& ea, double& eb) {
  if (b1 <= -E) { ea = eb = -1.0; return; }
  const double apb1 = 2 * E + b1 + 2 * std::sqrt(E * (E + b1));
  if (apb1 <= a0) { ea = eb = -1.0; return; }
  b0 = std::max(b0, -E);
  const double apb0 = 2 * E + b0 + 2 * std::sqrt(E * (E + b0));
  if (apb0 >= a1) { ea = a1; eb = b0; return; }
  a1 = std::min(a1, apb1);
  if (apb0 < a0) b0 = a0 - 2.0 * std::sqrt(E * a0);
  ea = a1;
  eb = b0;
}

std::vector<double> alpha_grid(double kT, double msd, double alphaMax, unsigned npts) {
  const double x2a = 1.0 / alpha2x(kT, msd);
  const double alphaMin = x2a * 1e-50, alphaMin2 = x2a * 1e-10, g1 = x2a * 1.0, g15 = x2a * 15.0;
  const unsigned n0 = (unsigned)(npts * 0.15 + 0.5), n1 = (unsigned)(npts * 0.29 + 0.5),
                 n2 = (unsigned)(npts * 0.23 + 0.5), n3 = npts - (n1 + n2 + n0), n123 = n1 + n2 + n3;
  const double upscat = (n0 < 10 ? 6.0 : (n0 > 50 ? 14.0 : 10.0));
  const std::vector<double> r0 = linspace(std::min(1e-3, alphaMax * 0.01), std::min(upscat, alphaMax * 0.99), n0);
  std::vector<double> g;
  if (alphaMax <= alphaMin * 100.0) {
    g = linspace(alphaMax * 0.001, alphaMax, n123);
  } else {
    g.push_back(alphaMin);
    if (alphaMax <= g1 * 10.0) {
      const auto v = linspace(alphaMin2, alphaMax, n123 - 1);
      g.insert(g.end(), v.begin(), v.end());
    } else {
      const auto v = linspace(alphaMin2, g1, n1 - 1);
      g.insert(g.end(), v.begin(), v.end());
      if (alphaMax < 2.0 * g15) {
        const auto ls = linspace(g1, alphaMax, n2 + n3 + 2);
        g.insert(g.end(), ls.begin() + 1, ls.end() - 1);
      } else {
        const auto ls = linspace(g1, g15, n2 + 2);
        g.insert(g.end(), ls.begin() + 1, ls.end() - 1);
        const auto gs = geomspace(g15, alphaMax, n3);
        g.insert(g.end(), gs.begin(), gs.end());
      }
    }
  }
  // merge the region-0 points, each placed halfway between its neighbours
  std::vector<std::pair<double, bool>> m;
  for (double v : g) m.emplace_back(v, false);
  for (double v : r0) m.emplace_back(v, true);
  std::stable_sort(m.begin(), m.end());
  for (size_t i = 1; i + 1 < m.size(); ++i)
    if (m[i].second) m[i].first = 0.5 * (m[i - 1].first + m[i + 1].first);
  std::vector<double> out;
  for (const auto& e : m) out.push_back(e.first);
  return out;
}

// reducePtsInDistribution (ncrystal_core/src/utils/NCMath.cc:419-505):
// drop the interior point whose removal changes the curve least (area times
// log-area squared) until targetN remain; equal importances in insertion order
void reduce_pts(std::vector<double>& x, std::vector<double>& y, size_t target) {
  const size_t n = x.size();
  if (target >= n) return;
  const double inv = 1.0 / *std::max_element(y.begin(), y.end());
  std::vector<double> lny(n);
  std::vector<long> prv(n), nxt(n);
  std::vector<unsigned long> cur(n, 0);
  std::vector<char> alive(n, 1);
  for (size_t i = 0; i < n; ++i) {
    lny[i] = std::log(std::max(1e-20, y[i] * inv));
    prv[i] = (long)i - 1;
    nxt[i] = (long)i + 1;
  }
  auto imp = [&](size_t i) {
    const size_t a = (size_t)prv[i], b = (size_t)nxt[i];
    const double ac = std::fabs(x[a] * (y[i] - y[b]) + x[i] * (y[b] - y[a]) + x[b] * (y[a] - y[i]));
    const double lc = std::fabs(x[a] * (lny[i] - lny[b]) + x[i] * (lny[b] - lny[a]) + x[b] * (lny[a] - lny[i]));
    return ac * lc * lc;
  };
  struct Ent { double imp; unsigned long seq; size_t idx; };
  auto cmp = [](const Ent& p, const Ent& q) { return p.imp > q.imp || (p.imp == q.imp && p.seq > q.seq); };
  std::priority_queue<Ent, std::vector<Ent>, decltype(cmp)> pq(cmp);
  unsigned long seq = 0;
  for (size_t i = 1; i + 1 < n; ++i) {
    cur[i] = seq;
    pq.push(Ent{imp(i), seq++, i});
  }
  size_t remaining = n;
  while (remaining > target) {
    const Ent e = pq.top();
    pq.pop();
    if (!alive[e.idx] || cur[e.idx] != e.seq) continue;
    const size_t i = e.idx, a = (size_t)prv[i], b = (size_t)nxt[i];
    alive[i] = 0;
    --remaining;
    nxt[a] = (long)b;
    prv[b] = (long)a;
    if (a != 0) { cur[a] = seq; pq.push(Ent{imp(a), seq++, a}); }
    if (b != n - 1) { cur[b] = seq; pq.push(Ent{imp(b), seq++, b}); }
  }
  size_t j = 0;
  for (size_t i = 0; i < n; ++i)
    if (alive[i]) { x[j] = x[i]; y[j] = y[i]; ++j; }
  x.resize(j);
  y.resize(j);
}

std::vector<double> beta_grid(const GnSet& gs, double betaMax, unsigned lux) {
  const double invkT = 1.0 / gs.kT;
  const Gn& g1 = gs.g[0];
  const double G1 = std::fabs(g1.lo * invkT);
  double G3 = std::fabs(gs.g[std::min<size_t>(3, gs.g.size()) - 1].lo * invkT);
  if (G1 == G3) G3 = G1 * 1.0001;
  betaMax = std::max(betaMax, G1 * 1.01);
  G3 = std::min(G3, betaMax * 0.9999);
  const double D = std::min(betaMax * 0.9999, std::max(G3, std::max<int>(2, (int)(lux - 1u)) * 10.0));
  const unsigned ntotal = 100u * (1u << lux);
  std::vector<double> ev = linspace(g1.lo * invkT, g1.hi * invkT, (unsigned)g1.s.size());
  std::vector<double> rs = g1.s;
  const double epsilon = -0.1 * g1.bw;
  while (ev.back() > epsilon) ev.pop_back();
  rs.resize(ev.size());
  const unsigned n1_max = (unsigned)(ev.size() * 1.25 + 6.5) & ~1u;
  const unsigned n0 = 1;
  double L1 = 2 * G1, L2 = 2 * (D - G1);
  const double L3 = betaMax - D;
  L1 *= 4;
  L2 *= 2;
  const double Lnorm = 1.0 / (L1 + L2 + L3);
  double f1 = std::max(0.2, L1 * Lnorm), f2 = L2 * Lnorm;
  if (f1 + f2 > 0.99) {
    const double tmp = 0.99 / (f1 + f2);
    f1 *= tmp;
    f2 *= tmp;
  }
  unsigned n1 = std::min<unsigned>(n1_max, (unsigned)(f1 * ntotal + 0.5) / 2);
  unsigned n2 = (unsigned)(f2 * ntotal + 0.5) / 2;
  n1 = std::max<unsigned>(15, n1);
  n2 = std::max<unsigned>(1, n2);
  unsigned n012;
  while (true) {
    n012 = 2 * (n1 + n2) + n0;
    if (n012 >= ntotal - 1) {
      if (n2 > n1) --n2; else --n1;
    } else {
      break;
    }
  }
  const unsigned n3 = ntotal - n012;
  std::vector<double> grid;
  {
    const auto v = linspace(-betaMax, -D, n3 + 1);
    grid.insert(grid.end(), v.begin(), v.end() - 1);
  }
  const size_t idx_r2 = grid.size();
  {
    const auto v = linspace(-D, -G1, n2 + 1);
    grid.insert(grid.end(), v.begin(), v.end() - 1);
  }
  {
    unsigned near0 = std::max<unsigned>(5, (unsigned)(n1 * 0.2 + 0.5));
    unsigned nspec = n1 - near0;
    if (nspec >= ev.size()) {
      nspec = (unsigned)ev.size();
      near0 = n1 - nspec;
    } else {
      unsigned ngaps = (nspec > 30 && ev.size() - nspec > 10) ? std::max<unsigned>(5, (unsigned)(nspec * 0.1 + 0.5)) : 0;
      nspec -= ngaps;
      const double g1bw = ev[1] - ev[0];
      reduce_pts(ev, rs, nspec);
      if (ngaps > 0) {
        struct Gap { double b0, b1; unsigned n; };
        auto less = [](const Gap& s, const Gap& o) {
          const double a = (s.b1 - s.b0) / (s.n + 1), b = (o.b1 - o.b0) / (o.n + 1);
          if (std::fabs(a - b) <= (0.5 * 1e-13) * (std::fabs(a) + std::fabs(b)) + 1e-13) return s.b0 > o.b0;
          return a > b;
        };
        std::vector<Gap> gaps;
        const double dmn = 1.5 * g1bw;
        for (size_t i = 0; i + 1 < ev.size(); ++i)
          if (ev[i + 1] - ev[i] > dmn) gaps.push_back(Gap{ev[i], ev[i + 1], 0});
        while (ngaps > 0 && !gaps.empty()) {
          std::stable_sort(gaps.begin(), gaps.end(), less);
          gaps.front().n += 1;
          --ngaps;
        }
        for (const auto& gp : gaps) {
          if (!gp.n) continue;
          const double bw = (gp.b1 - gp.b0) / (gp.n + 1.0);
          for (unsigned q = 0; q < gp.n; ++q) ev.push_back(gp.b0 + (q + 1) * bw);
        }
        std::sort(ev.begin(), ev.end());
      }
      if (ngaps > 0) near0 += ngaps;
    }
    grid.insert(grid.end(), ev.begin(), ev.end());
    const double last = grid.back();
    auto v = geomspace(std::min(1e-50, -0.001 * last), -last * 0.1, near0);
    std::reverse(v.begin(), v.end());
    for (double e : v) grid.push_back(-e);
  }
  const size_t idx_r1back = grid.size() - 1;
  grid.push_back(0.0);
  for (size_t i = idx_r1back; i >= idx_r2; --i) {
    grid.push_back(-grid[i]);
    if (i == 0) break;
  }
  return grid;
}

double stirling9(double x) {
  const double c1 = 1. / 12., c2 = 1. / 288., c3 = -139 / 51840., c4 = -571. / 2488320.,
               c5 = 163879. / 209018880., c6 = 5246819. / 75246796800., c7 = -534703531. / 902961561600.,
               c8 = -4483131259. / 86684309913600., c9 = 432261921612371. / 514904800886784000.;
  return 1.0 + x * (c1 + x * (c2 + x * (c3 + x * (c4 + x * (c5 + x * (c6 + x * (c7 + x * (c8 + x * c9))))))));
}

// ---------------------------------------------------------------- GPU fill

struct GnDev {
  long long off;  // first value in the concatenated spectra
  int n;
  double lo, hi, invbw;
};

// G[k][j] = G_{k+1}(beta_j kT) for the rows beta_j <= 0 (interpolation of
// VDOSGnData::interpolateDensity, NCVDOSGn.cc:95-107)
__global__ void k_vdos_gtab(const GnDev* __restrict__ gd, const double* __restrict__ spec,
                            const double* __restrict__ beta, double kT, int K, int M, int Kp,
                            double* __restrict__ G) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long long)Kp * M) return;
  const int k = (int)(t / M), j = (int)(t % M);
  double v = 0.0;
  if (k < K) {
    const GnDev g = gd[k];
    const double e = beta[j] * kT;
    if (g.lo <= e && e <= g.hi) {
      const double a = (e - g.lo) * g.invbw;
      const double fa = floor(a);
      long long idx = (long long)fa;
      if (idx > g.n - 2) idx = g.n - 2;
      const double f = a - fa;
      const double* s = spec + g.off + idx;
      v = s[0] * (1.0 - f) + f * s[1];
    }
  }
  G[t] = v;
}

// A[k][i] = x^n exp(-x)/n! kT for n = k+1, x = x_i: by recursion
// f(x,n) = f(x,n-1) x/n below order 16 (half of exp(-x) applied at the end),
// from Stirling's series for n! above (NCVDOSToScatKnl.cc:149-188).  Thread
// (i, c) fills orders 16c+1..16c+16 of alpha i: chunk 0 runs the recursion,
// the others are independent per order.
__global__ void k_vdos_afac(const double* __restrict__ x, const double* __restrict__ emh,
                            const double* __restrict__ lx, const double* __restrict__ ofact,
                            const double* __restrict__ ologn, double kT, int K, int N, int Kp,
                            double* __restrict__ A) {
  const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x);
  if (i >= N) return;
  const int kb = (int)blockIdx.y * 16;
  const int ke = kb + 16 < Kp ? kb + 16 : Kp;
  const double xi = x[i], eh = emh[i], li = lx[i];
  double fxn = eh;
  for (int k = kb; k < ke; ++k) {
    const int n = k + 1;
    double af = 0.0;
    if (k >= K) {
      af = 0.0;
    } else if (n < 16) {
      fxn *= xi * (1.0 / n);
      af = fxn * eh * kT;
    } else {
      const double fact = ofact[k];
      af = fact ? fact * exp(n * (li - ologn[k] + 1.0) - xi) : 0.0;
    }
    A[(size_t)k * N + i] = af;
  }
}

// S[j][i] = sum_k G[k][j] A[k][i] for the rows beta_j <= 0 on the FP64 matrix
// cores: each wave one 16x16 output tile, four tiles (32x32) per workgroup,
// K in steps of 4 (v_mfma_f64_16x16x4f64: lane l holds A-operand element
// (row l%16, k l/16), B-operand element (k l/16, column l%16), and in its
// output register r the element (row 4r + l/16, column l%16): the rows of the
// FP64 instruction interleave, unlike the 16-bit ones).  Rows with a mirror at +|beta| also
// write S(alpha,+|beta|) = S(alpha,-|beta|) exp(-|beta|).
typedef double vdos_f64x4 __attribute__((ext_vector_type(4)));
__global__ void __launch_bounds__(256) k_vdos_fill(const double* __restrict__ G, const double* __restrict__ A,
                                                   const double* __restrict__ expb, int M, int N, int Kp, int iz,
                                                   int iflip, double* __restrict__ S) {
  const int lane = (int)(threadIdx.x & 63), wave = (int)(threadIdx.x >> 6);
  const int j0 = (int)blockIdx.y * 32 + (wave >> 1) * 16;
  const int i0 = (int)blockIdx.x * 32 + (wave & 1) * 16;
  const int lr = lane & 15, lk = lane >> 4;
  const int jj = j0 + lr, ii = i0 + lr;
  vdos_f64x4 acc = {0.0, 0.0, 0.0, 0.0};
  for (int k0 = 0; k0 < Kp; k0 += 4) {
    const int k = k0 + lk;
    const double a = jj < M ? G[(size_t)k * M + jj] : 0.0;
    const double b = ii < N ? A[(size_t)k * N + ii] : 0.0;
    acc = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, acc, 0, 0, 0);
  }
  const int col = i0 + lr;
  if (col >= N) return;
  for (int r = 0; r < 4; ++r) {
    const int row = j0 + 4 * r + lk;
    if (row >= M) continue;
    const double v = acc[r];
    S[(size_t)row * N + col] = v;
    if (row >= iflip && row != iz) S[(size_t)(iz + (iz - row)) * N + col] = v * expb[row];
  }
}

template <class T>
struct DevVec {
  T* p = nullptr;
  explicit DevVec(size_t n) { VDOS_HIP(hipMalloc(&p, sizeof(T) * (n ? n : 1))); }
  ~DevVec() { if (p) (void)hipFree(p); }
  DevVec(const DevVec&) = delete;
  DevVec& operator=(const DevVec&) = delete;
};
template <class T>
void upload(DevVec<T>& d, const std::vector<T>& h, hipStream_t st) {
  if (!h.empty()) VDOS_HIP(hipMemcpyAsync(d.p, h.data(), sizeof(T) * h.size(), hipMemcpyHostToDevice, st));
}

}  // namespace

VdosAnalysis vdos_analyse(const RegularVdos& v, double temperature, double mass) {
  const Eval ev(v, temperature, mass);
  VdosAnalysis a;
  a.gamma0 = ev.gamma0();
  a.msd = ev.msd(a.gamma0);
  a.teff = ev.teff();
  a.integral = ev.integral;
  return a;
}

std::vector<double> vdos_gn(const RegularVdos& v, double temperature, double mass, unsigned order, double& lo,
                            double& binwidth, hipStream_t st) {
  if (order < 1 || order > 100000) fail("BadInput", "order must be in 1..100000");
  const Eval ev(v, temperature, mass);
  GnSet gs(ev);
  gs.cv.st = st;
  gs.grow(order);
  const Gn& g = gs.g[order - 1];
  lo = g.lo;
  binwidth = g.bw;
  return g.s;
}

VdosKernel vdos_kernel(const RegularVdos& v, double temperature, double mass, unsigned lux, double target_emax_req,
                       hipStream_t st) {
  if (lux > 5) fail("BadInput", "vdoslux must be in 0..5");
  if (!(target_emax_req >= 0.0)) fail("BadInput", "target Emax must be >= 0");
  const double lux2emax[6] = {0.5, 1.0, 3.0, 5.0, 8.0, 12.0};
  double targetEmax = target_emax_req > 0.0 ? target_emax_req : lux2emax[lux];
  const Eval ev(v, temperature, mass);
  const double kT = ev.kT, invkT = 1.0 / kT;
  const double gamma0 = ev.gamma0();
  const double msd = ev.msd(gamma0);
  double tE = targetEmax * invkT;
  unsigned maxo = 4;
  GnSet gs(ev);
  gs.cv.st = st;
  gs.grow(maxo);
  unsigned order_limit = 1000;
  if (target_emax_req > 0.0 || lux == 5) order_limit *= 10;
  if (lux == 0) order_limit /= 10;
  // orders are grown a generation ahead (at most up to the first order past
  // the limit); the ones beyond the final count are dropped below
  gs.cap = order_limit + 1;
  const double emax_lowest = target_emax_req > 0.0 ? target_emax_req : 1e-15;
  const double rel = std::pow(10.0, -(3.0 + 2.0 * lux));
  const double x2a = 1.0 / alpha2x(kT, msd);
  auto ranges = [&](unsigned n, double& a0, double& a1, double& b0, double& b1) {
    gs.g[n - 1].erange(rel, b0, b1);
    b0 *= invkT;
    b1 *= invkT;
    range_xnexpmx(n, rel, a0, a1);
    a0 *= x2a;
    a1 *= x2a;
  };
  while (true) {
    gs.grow(maxo);
    double a0, a1, b0, b1;
    ranges(maxo, a0, a1, b0, b1);
    if (within_alphaplus(tE, a0, b1)) ++maxo;
    else break;
    if (maxo > order_limit) {
      double red = targetEmax;
      do {
        red *= 0.99;
        if (red < emax_lowest) {
          char buf[512];
          snprintf(buf, sizeof buf,
                   "VDOS expansion too slow - can not reach E=%geV after %u phonon convolutions (likely causes: "
                   "either the target energy value is too high, vdoslux too low, the temperature too high, or the "
                   "VDOS is very unusual).", emax_lowest, order_limit);
          fail("CalcError", buf);
        }
      } while (within_alphaplus(red * invkT, a0, b1));
      tE = red * invkT;
      targetEmax = red;
      break;
    }
  }
  gs.grow(maxo);
  gs.g.erase(gs.g.begin() + maxo, gs.g.end());
  double betaMin = 0.0, alphaMax = 0.0;
  for (unsigned o = 1; o <= maxo; ++o) {
    double a0, a1, b0, b1, ea, eb;
    ranges(o, a0, a1, b0, b1);
    extreme_point(tE, a0, a1, b0, b1, ea, eb);
    alphaMax = std::max(alphaMax, ea);
    betaMin = std::min(betaMin, eb);
  }
  if (!(betaMin < 0.0 && alphaMax > 0.0)) fail("CalcError", "no kinematically accessible region");
  VdosKernel out;
  out.beta = beta_grid(gs, -betaMin * 1.01, lux);
  out.alpha = alpha_grid(kT, msd, alphaMax * 1.01, (unsigned)(out.beta.size() / 2));
  const int na = (int)out.alpha.size(), nb = (int)out.beta.size();
  out.norders = (int)gs.g.size();
  out.suggested_emax = targetEmax;
  out.msd = msd;
  out.gamma0 = gamma0;
  // the orders summed: all of them, or, above 16 orders, the blocks of
  // norders / (norders / 16) orders that the reference adds up (orders past
  // the last whole block are not included, NCVDOSToScatKnl.cc:299-360)
  const int nord = out.norders;
  const int njobs = nord <= 16 ? 1 : nord / 16;
  const int K = njobs == 1 ? nord : njobs * (nord / njobs);
  out.orders_summed = K;
  // rows beta <= 0 and the mirror map
  const auto& bg = out.beta;
  auto closest = [&](double val) {
    const size_t lo = (size_t)(std::lower_bound(bg.begin(), bg.end(), val) - bg.begin());
    if (lo == 0) return (size_t)0;
    if (lo == bg.size()) return bg.size() - 1;
    return std::fabs(bg[lo] - val) < std::fabs(bg[lo - 1] - val) ? lo : lo - 1;
  };
  const size_t iz = closest(0.0), iflip = closest(-bg.back());
  if (bg[iz] != 0.0 || bg[iflip] != -bg.back()) fail("LogicError", "beta grid without its mirror points");
  const int M = (int)iz + 1;
  const int Kp = (K + 3) / 4 * 4;
  // operands: per-order spectra, per-alpha x, exp(-x/2), log(x), per-order
  // Stirling factors (host libm), then the tables on the GPU
  std::vector<GnDev> gd(K);
  std::vector<double> spec;
  for (int k = 0; k < K; ++k) {
    const Gn& g = gs.g[k];
    gd[k] = GnDev{(long long)spec.size(), (int)g.s.size(), g.lo, g.hi, 1.0 / g.bw};
    spec.insert(spec.end(), g.s.begin(), g.s.end());
  }
  const double a2x = alpha2x(kT, msd);
  std::vector<double> x(na), emh(na), lx(na), ofact(K, 0.0), ologn(K, 0.0), expb(M);
  for (int i = 0; i < na; ++i) {
    x[i] = out.alpha[i] * a2x;
    emh[i] = std::exp(-0.5 * x[i]);
    lx[i] = nord >= 16 ? std::log(x[i]) : 0.0;
  }
  for (int k = 15; k < K; ++k) {
    const unsigned n = (unsigned)k + 1;
    const double invn = 1.0 / n;
    const double gn = stirling9(invn);
    ofact[k] = kT * kInvSqrt2Pi / (std::sqrt(n) * gn);
    ologn[k] = std::log(n);
  }
  for (int j = 0; j < M; ++j) expb[j] = std::exp(bg[j]);
  DevVec<GnDev> d_gd(K);
  DevVec<double> d_spec(spec.size()), d_beta(M), d_x(na), d_emh(na), d_lx(na), d_of(K), d_ol(K), d_expb(M),
      d_G((size_t)Kp * M), d_A((size_t)Kp * na), d_S((size_t)na * nb);
  upload(d_gd, gd, st);
  upload(d_spec, spec, st);
  VDOS_HIP(hipMemcpyAsync(d_beta.p, bg.data(), sizeof(double) * M, hipMemcpyHostToDevice, st));
  upload(d_x, x, st);
  upload(d_emh, emh, st);
  upload(d_lx, lx, st);
  upload(d_of, ofact, st);
  upload(d_ol, ologn, st);
  upload(d_expb, expb, st);
  VDOS_HIP(hipMemsetAsync(d_S.p, 0, sizeof(double) * (size_t)na * nb, st));
  // device time of the three kernels (operand tables and the order sum)
  hipEvent_t ev0, ev1;
  VDOS_HIP(hipEventCreate(&ev0));
  VDOS_HIP(hipEventCreate(&ev1));
  VDOS_HIP(hipEventRecord(ev0, st));
  const long long ng = (long long)Kp * M;
  hipLaunchKernelGGL(k_vdos_gtab, dim3((unsigned)((ng + 255) / 256)), dim3(256), 0, st, d_gd.p, d_spec.p, d_beta.p, kT,
                     K, M, Kp, d_G.p);
  hipLaunchKernelGGL(k_vdos_afac, dim3((unsigned)((na + 63) / 64), (unsigned)((Kp + 15) / 16)), dim3(64), 0, st, d_x.p,
                     d_emh.p, d_lx.p, d_of.p, d_ol.p, kT, K, na, Kp, d_A.p);
  hipLaunchKernelGGL(k_vdos_fill, dim3((unsigned)((na + 31) / 32), (unsigned)((M + 31) / 32)), dim3(256), 0, st, d_G.p,
                     d_A.p, d_expb.p, M, na, Kp, (int)iz, (int)iflip, d_S.p);
  VDOS_HIP(hipGetLastError());
  VDOS_HIP(hipEventRecord(ev1, st));
  out.sab.resize((size_t)na * nb);
  VDOS_HIP(hipMemcpyAsync(out.sab.data(), d_S.p, sizeof(double) * out.sab.size(), hipMemcpyDeviceToHost, st));
  VDOS_HIP(hipStreamSynchronize(st));
  float ms = 0.0f;
  VDOS_HIP(hipEventElapsedTime(&ms, ev0, ev1));
  (void)hipEventDestroy(ev0);
  (void)hipEventDestroy(ev1);
  out.gpu_ms = ms;
  out.rows_summed = M;
  return out;
}

}  // namespace ncamd
