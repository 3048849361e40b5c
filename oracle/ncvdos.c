/*
 * ncvdos.c -- TEST INFRASTRUCTURE ONLY (the CPU checker; never linked by the
 * product library).  Plain-C restatement of NCrystal v4.4.2's expansion of a
 * phonon density of states (VDOS) into a scattering kernel S(alpha,beta).
 * Citations are reference file:line (relative to the reference repo):
 *  - regulariseVDOSGrid, checkIsRegularVDOSGrid:
 *    ncrystal_core/src/vdos/NCVDOSEval.cc:449-640
 *  - VDOSEval (normalisation, gamma0, effective temperature, msd, G1):
 *    ncrystal_core/src/vdos/NCVDOSEval.cc:33-445
 *  - VDOSGn (G1 on the thickened grid, G_n = G_{n-n/2} (x) G_{n/2},
 *    truncation and thinning): ncrystal_core/src/vdos/NCVDOSGn.cc:48-492
 *  - FastConvolve::convolve: ncrystal_core/src/utils/NCFastConvolve.cc:224-268
 *    (the transform itself is a radix-2 FFT of this file's own, twiddles from
 *    libm, so its rounding differs from the reference's table-driven one)
 *  - createScatteringKernel and its helpers (phonon orders, alpha/beta
 *    limits, grids, Sjolander II.28 fill):
 *    ncrystal_core/src/vdos/NCVDOSToScatKnl.cc:38-944
 *  - reducePtsInDistribution: ncrystal_core/src/utils/NCMath.cc:419-505
 *  - createVDOSDebye: ncrystal_core/src/dyninfoutils/NCDynInfoUtils.cc:377-396
 * Pins: tests/test_vdos_pins.py (reference logs analysevdos.log, vdosgn.log,
 * vdosregularise.log, vdos2knl.log; the Debye-model MSD; sum rules).
 */
#include "ncoracle.h"
#include <math.h>
#include <float.h>
#include <stdlib.h>
#include <string.h>
#include <stdio.h>

/* NCDefs.hh:79-128, NCMath.hh:34-49 */
#define V_KB 8.6173303e-5
#define V_NMASS_AMU 1.00866491588
#define V_C 299792458e10
#define V_DALTON2EVC2 931494095.17
#define V_PLANCK 4.135667662e-15
#define V_INV2PI 0.159154943091895335768883763372514362034459645740
#define V_HBAR (V_PLANCK * V_INV2PI)
#define V_NMASS_EVC2 (V_NMASS_AMU * V_DALTON2EVC2 / (V_C * V_C))
#define V_INVSQRT2PI 0.398942280401432677939946059934381868475858631
#define V_INVE 0.367879441171442321595523770161460867445811131
#define V_2PI 6.2831853071795864769252867665590057683943388

/* ------------------------------------------------------------------ */
/* small helpers */
typedef struct { double s, c; } vsum_t;
static inline void vs_add(vsum_t* ss, double x) {
  const double t = ss->s + x;
  ss->c += fabs(ss->s) >= fabs(x) ? (ss->s - t) + x : (x - t) + ss->s;
  ss->s = t;
}
static inline double vs_sum(const vsum_t* ss) { return ss->s + ss->c; }
static inline double vmax(double a, double b) { return a > b ? a : b; }
static inline double vmin(double a, double b) { return a < b ? a : b; }
static inline double vclamp(double v, double lo, double hi) { return vmin(vmax(v, lo), hi); }

static const char* g_vdos_err = NULL;
const char* orc_vdos_lasterror(void) { return g_vdos_err; }
#define VFAIL(msg) do { g_vdos_err = (msg); goto fail; } while (0)

/* linspace / geomspace (NCMath.cc:44-81) */
static void v_linspace(double a, double b, unsigned n, double* out) {
  const unsigned nm1 = n - 1;
  const double dx = (b - a) / nm1;
  for (unsigned i = 0; i < nm1; ++i) out[i] = a + i * dx;
  out[nm1] = b;
}
static void v_geomspace(double a, double b, unsigned n, double* out) {
  double start = log10(a);
  const double stop = log10(b);
  const double interval = (stop - start) / (n - 1);
  for (unsigned i = 0; i < n; ++i) { out[i] = pow(10.0, start); start += interval; }
  out[n - 1] = pow(10.0, stop);
  out[0] = a;
  out[n - 1] = b;
}

/* Romberg with 17 and 33 points (NCMath.hh:580-604 over NCRomberg.cc) */
typedef double (*vfunc_t)(double, const void*);
static double romberg_lvl(vfunc_t f, const void* ctx, double a, double b, int lvl5) {
  double h = (b - a);
  double fv[17];
  const double delta = h * 0.0625;
  for (int i = 0; i < 17; ++i) fv[i] = f(a + delta * (double)i, ctx);
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
  for (double i = 0; i < 16.0; ++i) c5 += f(off + h * i, ctx);
  h *= 0.5;
  const double R50 = h * c5 + 0.5 * R40;
  const double R51 = (4. / 3.) * R50 + (-1. / 3.) * R40;
  const double R52 = (16. / 15.) * R51 + (-1. / 15.) * R41;
  const double R53 = (64. / 63.) * R52 + (-1. / 63.) * R42;
  const double R54 = (256. / 255.) * R53 + (-1. / 255.) * R43;
  return (1024. / 1023.) * R54 + (-1. / 1023.) * R44;
}

/* ------------------------------------------------------------------ */
/* regularisation (NCVDOSEval.cc:449-640) */
static int v_is_linspaced(const double* g, int n, double tol) {
  if (n == 2) return 1;
  const double bw = (g[n - 1] - g[0]) / (n - 1.0);
  const double eps = tol * bw;
  for (int i = 0; i < n; ++i)
    if (fabs((g[0] + (double)i * bw) - g[i]) > eps) return 0;
  return 1;
}
static double v_check_regular_pair(double emin, double emax, unsigned long npts, double tol) {
  const double bwa = (emax - emin) / (npts - 1);
  const double nb = emin / bwa;
  if (nb < 0.99 || fabs(nb - round(nb)) > tol) return 0.0;
  const unsigned long nbi = (unsigned long)(nb + 0.5);
  const double bw = emin / nbi;
  return emin + (npts - 1) * bw;
}

int orc_vdos_regularise(const double* egrid, int negrid, const double* density, int n, double* emin_out,
                        double* emax_out, double* dens_out, int cap) {
  g_vdos_err = NULL;
  if (n <= 2 || !(negrid == 2 || negrid == n)) VFAIL("bad VDOS grid sizes");
  if (egrid[0] < 1e-5) VFAIL("VDOS energy range can not be specified for values less than 1e-5eV = 0.01meV");
  for (int i = 1; i < negrid; ++i)
    if (!(egrid[i] > egrid[i - 1])) VFAIL("VDOS egrid not increasing");
  const double tol = 1e-6;
  if (v_is_linspaced(egrid, negrid, tol)) {
    const double ec = v_check_regular_pair(egrid[0], egrid[negrid - 1], (unsigned long)n, tol);
    if (ec) {
      if (cap >= n) {
        *emin_out = egrid[0];
        *emax_out = ec;
        memcpy(dens_out, density, sizeof(double) * (size_t)n);
      }
      return n;
    }
  }
  const double emin = egrid[0];
  const double oldEmax = egrid[negrid - 1];
  const double d = oldEmax - emin;
  const double r = d / emin;
  const double kBegin = 2000.0 / r;
  double k = vmax(1.0, round(kBegin));
  double best_eps = INFINITY, best_k = 0.0;
  while (1) {
    const double m = floor(r * k);
    if (m < 1.0) { k += 1.0; continue; }
    const double bw = emin / k;
    double eps = d - (m * bw);
    eps = vmax(0.0, eps);
    const double kbest = best_k;
    if (eps == 0.0 || kbest == 0.0 || (kbest * floor(r * k) > k * floor(r * kbest))) {
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
    if (m >= 20000) VFAIL("Could not regularise input energy grid. Are the energy ranges highly unusual?");
    k += 1.0;
  }
  const double nbw = emin / best_k;
  const double mm = floor(r * best_k);
  unsigned npts = (unsigned)(mm + 0.5) + 1;
  double nemax = emin + nbw * (npts - 1);
  if (nemax < oldEmax) {
    ++npts;
    nemax = emin + nbw * (npts - 1);
  }
  if (cap < (int)npts) return (int)npts;
  double* ex = (double*)malloc(sizeof(double) * (size_t)n);
  if (negrid == 2) v_linspace(egrid[0], egrid[1], (unsigned)n, ex);
  else memcpy(ex, egrid, sizeof(double) * (size_t)n);
  const double inv_bw = negrid == 2 ? ((n - 1.0) / (egrid[1] - egrid[0])) : 0.0;
  int it = 0;
  const int last = n - 1;
  for (unsigned i = 0; i < npts; ++i) {
    const double ev = (i + 1 == npts) ? nemax : emin + nbw * i;
    while (it != last && ev >= ex[it + 1]) ++it;
    if (ev == ex[it]) { dens_out[i] = density[it]; continue; }
    if (it == last) {
      dens_out[i] = ev > ex[it] ? 0.0 : density[last];
    } else {
      const double y0 = density[it], y1 = density[it + 1], x0 = ex[it], x1 = ex[it + 1];
      vsum_t s = {0.0, 0.0};
      vs_add(&s, y0 * x1);
      vs_add(&s, -y0 * ev);
      vs_add(&s, ev * y1);
      vs_add(&s, -x0 * y1);
      dens_out[i] = inv_bw ? vs_sum(&s) * inv_bw : vs_sum(&s) / (x1 - x0);
    }
  }
  free(ex);
  *emin_out = emin;
  *emax_out = nemax;
  return (int)npts;
fail:
  return -1;
}

/* createVDOSDebye (NCDynInfoUtils.cc:377-396): 20 points on
   [E_D/2, E_D], density E^2/E_D^2 */
void orc_vdos_debye(double debye_temp, double* emin, double* emax, double* density20) {
  const double ed = V_KB * debye_temp;
  double eg[20];
  v_linspace(0.5 * ed, ed, 20, eg);
  const double scale = 1.0 / (ed * ed);
  for (int i = 0; i < 20; ++i) density20[i] = eg[i] * eg[i] * scale;
  *emin = eg[0];
  *emax = eg[19];
}

/* ------------------------------------------------------------------ */
/* VDOSEval (NCVDOSEval.cc:147-445) */
typedef struct {
  double* d;
  int n;
  double emin, emax, kT, T, mass, bw, ibw, k, integral;
  unsigned long nptsext;
} veval_t;

typedef struct { const veval_t* ve; double A, B; vfunc_t f; const void* fctx; } vbin_t;
static double vbin_integrand(double e, const void* c) {
  const vbin_t* b = (const vbin_t*)c;
  return b->f(e, b->fctx) * (b->A * e + b->B);
}
static void v_integrate_bins(const veval_t* ve, vfunc_t f, const void* fctx, vsum_t* sum) {
  const unsigned nbins = (unsigned)ve->n - 1;
  for (unsigned ib = 0; ib < nbins; ++ib) {
    const double d0 = ve->d[ib], d1 = ve->d[ib + 1];
    const double e0 = ve->emin + ve->bw * ib;
    const double e1 = (ib + 1 == nbins ? ve->emax : ve->emin + ve->bw * (ib + 1));
    vbin_t b;
    b.ve = ve;
    b.A = (d1 - d0) * ve->ibw;
    b.B = d0 - e0 * b.A;
    b.f = f;
    b.fctx = fctx;
    vs_add(sum, romberg_lvl(vbin_integrand, &b, e0, e1, 0));
  }
}
static double v_one(double e, const void* c) { (void)e; (void)c; return 1.0; }

static double safe_xcothx(double x) {
  if (x < 0.1) {
    const double y = x * x;
    return 1.0 + y * (1. / 3 + y * (-1. / 45 + y * (2. / 945 + y * (-1. / 4725 + y * (2. / 93555 +
           y * (-1382. / 638512875. + y * (4. / 18243225.)))))));
  }
  return x / tanh(x);
}
static double safe_x3cothx(double x) {
  if (x < 0.1) {
    const double y = x * x;
    return y * (1.0 + y * (1. / 3 + y * (-1. / 45 + y * (2. / 945 + y * (-1. / 4725 + y * (2. / 93555 +
           y * (-1382. / 638512875 + y * (4. / 18243225 + y * (-3617. / 162820783125 +
           y * (87734. / 38979295480125))))))))));
  }
  return x * x * x / tanh(x);
}
static double safe_xdivsinhx(double x) {
  if (x < 0.07) {
    const double y = x * x;
    return 1.0 + y * (-1. / 6. + y * (7. / 360. + y * (-31. / 15120. + y * (127. / 604800.))));
  }
  return x / sinh(x);
}
static double safe_rec_expm1(double x) {
  if (x <= 700) return 1.0 / expm1(x);
  const double emx = exp(-x);
  return emx / (1.0 - emx);
}

static int veval_init(veval_t* ve, double emin, double emax, const double* density, int n, double T,
                      double mass) {
  memset(ve, 0, sizeof(*ve));
  if (!(mass > 0.5 && mass < 2000.0) || n < 2 || !(emin >= 0.0 && emax > emin)) {
    g_vdos_err = "bad VDOS input";
    return -1;
  }
  if (emin < 1e-5) {
    g_vdos_err = "VDOS energy range should not be specified for values less than 1e-5eV = 0.01meV";
    return -1;
  }
  if (!v_check_regular_pair(emin, emax, (unsigned long)n, 1e-6)) {
    g_vdos_err = "Received non-regularised VDOS";
    return -1;
  }
  ve->d = (double*)malloc(sizeof(double) * (size_t)n);
  memcpy(ve->d, density, sizeof(double) * (size_t)n);
  ve->n = n;
  ve->emin = emin;
  ve->emax = emax;  /* the reference keeps the given emax outside its verbose mode */
  ve->T = T;
  ve->kT = V_KB * T;
  ve->mass = mass;
  const double bw = (ve->emax - ve->emin) / (n - 1);
  const unsigned long nbb = (unsigned long)(ve->emin / bw + 0.5);
  ve->nptsext = (unsigned long)n + nbb;
  ve->k = ve->d[0] / (emin * emin);
  ve->bw = bw;
  ve->ibw = 1.0 / bw;
  vsum_t s = {0.0, 0.0};
  const double onethird = 1.0 / 3.0;
  vs_add(&s, onethird * ve->d[0] * ve->emin);
  v_integrate_bins(ve, v_one, NULL, &s);
  ve->integral = vs_sum(&s);
  if (!(ve->integral > 0.0)) {
    free(ve->d);
    g_vdos_err = "VDOS integral is not positive";
    return -1;
  }
  const double sf = 1.0 / ve->integral;
  for (int i = 0; i < n; ++i) ve->d[i] *= sf;
  ve->k *= sf;
  return 0;
}
static void veval_free(veval_t* ve) { free(ve->d); ve->d = NULL; }

static double veval_eval(const veval_t* ve, double e) {
  if (e <= ve->emin) return ve->k * e * e;
  double rel = vclamp((e - ve->emin) * ve->ibw, -0.5, ve->n + 0.5);
  const int ib = (int)rel;
  if (ib >= ve->n - 1) return 0.0;
  rel = vclamp(rel - ib, 0.0, 1.0);
  return (1.0 - rel) * ve->d[ib] + rel * ve->d[ib + 1];
}

static double v_g0_f(double e, const void* c) {
  const double inv2kT = *(const double*)c;
  return 1.0 / (e * tanh(e * inv2kT));
}
static double v_xc_f(double x, const void* c) { (void)c; return safe_xcothx(x); }
static double v_x3c_f(double x, const void* c) { (void)c; return safe_x3cothx(x); }
static double v_teff_f(double e, const void* c) {
  const double inv2kT = *(const double*)c;
  return safe_xcothx(e * inv2kT);
}

static double veval_gamma0(const veval_t* ve) {
  const double twokT = 2.0 * ve->kT;
  const double inv2kT = 1.0 / twokT;
  vsum_t s = {0.0, 0.0};
  vs_add(&s, ve->k * (twokT * twokT) * romberg_lvl(v_xc_f, NULL, 0.0, ve->emin * inv2kT, 1));
  v_integrate_bins(ve, v_g0_f, &inv2kT, &s);
  return ve->emax * vs_sum(&s);
}
static double veval_teff(const veval_t* ve) {
  const double twokT = 2.0 * ve->kT;
  const double inv2kT = 1.0 / twokT;
  vsum_t s = {0.0, 0.0};
  vs_add(&s, ve->k * (twokT * twokT * twokT) * romberg_lvl(v_x3c_f, NULL, 0.0, ve->emin * inv2kT, 1));
  v_integrate_bins(ve, v_teff_f, &inv2kT, &s);
  return ve->T * vs_sum(&s);
}
static double veval_msd(const veval_t* ve, double gamma0) {
  const double convfact = 0.5 * (V_HBAR * V_HBAR) * (V_C * V_C) / V_DALTON2EVC2;
  return convfact * gamma0 / (ve->mass * ve->emax);
}
static double veval_g1sym(const veval_t* ve, double e, double gamma0) {
  if (e < 0) e = -e;
  const double twokT = 2 * ve->kT;
  const double u = e / twokT;
  if (e <= ve->emin) return (ve->k * ve->kT * ve->emax / gamma0) * safe_xdivsinhx(u);
  return veval_eval(ve, e) * ve->emax / (e * 2.0 * gamma0 * sinh(u));
}
static void veval_g1pair(const veval_t* ve, double e, double gamma0, double* neg, double* pos) {
  if (e < 200.0 * ve->kT) {
    const double g = veval_g1sym(ve, e, gamma0);
    if (!g) { *neg = 0.0; *pos = 0.0; return; }
    const double dbf = exp(e / (2 * ve->kT));
    *neg = g * dbf;
    *pos = g / dbf;
    return;
  }
  const double kkk = veval_eval(ve, e) * ve->emax / (e * gamma0);
  if (!kkk) { *neg = 0.0; *pos = 0.0; return; }
  const double edkt = e / ve->kT;
  *neg = -kkk / expm1(-edkt);
  *pos = kkk * safe_rec_expm1(+edkt);
}

/* analyseVDOS values: out = {gamma0, msd, teff, original integral} */
int orc_vdos_analyse(double emin, double emax, const double* density, int n, double T, double mass, double* out) {
  g_vdos_err = NULL;
  veval_t ve;
  if (veval_init(&ve, emin, emax, density, n, T, mass)) return -1;
  const double g0 = veval_gamma0(&ve);
  out[0] = g0;
  out[1] = veval_msd(&ve, g0);
  out[2] = veval_teff(&ve);
  out[3] = ve.integral;
  veval_free(&ve);
  return 0;
}

/* ------------------------------------------------------------------ */
/* FFT convolution (NCFastConvolve.cc:224-268): y = dt * |IFFT(FFT(a1) FFT(a2))|
   over the next power of two >= n1 + n2 - 1.  Radix-2, decimation in time,
   twiddles w_j = exp(+-2 pi i j / N) from libm. */
static void v_fft(double* re, double* im, int logn, int forward, const double* wc, const double* ws) {
  const int n = 1 << logn;
  for (int j = 1, i = 0; j < n; ++j) {
    int bit = n >> 1;
    for (; i & bit; bit >>= 1) i ^= bit;
    i ^= bit;
    if (j < i) {
      double t = re[i]; re[i] = re[j]; re[j] = t;
      t = im[i]; im[i] = im[j]; im[j] = t;
    }
  }
  for (int len = 2; len <= n; len <<= 1) {
    const int half = len >> 1;
    const int step = n / len;
    for (int s = 0; s < n; s += len) {
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
}

typedef struct { int logn; double *wc, *ws, *b; } vconv_t;
static void vconv_free(vconv_t* cv) { free(cv->wc); free(cv->ws); free(cv->b); memset(cv, 0, sizeof(*cv)); }
/* y (size n1 + n2 - 1, allocated here) */
static double* v_convolve(vconv_t* cv, const double* a1, int n1, const double* a2, int n2, double dt, int* ny) {
  const int m = n1 + n2 - 1;
  int logn = 0;
  while ((1 << logn) < m) ++logn;
  const int n = 1 << logn;
  if (cv->logn != logn || !cv->wc) {
    free(cv->wc); free(cv->ws); free(cv->b);
    cv->logn = logn;
    cv->wc = (double*)malloc(sizeof(double) * (size_t)(n / 2 + 1));
    cv->ws = (double*)malloc(sizeof(double) * (size_t)(n / 2 + 1));
    /* separate libm cos and sin calls (through volatile pointers, which no
       compiler fuses into sincos: that routine can differ from them in the
       last bit, and the truncation edges of the G_n depend on such bits) */
    double (*volatile fcos)(double) = cos;
    double (*volatile fsin)(double) = sin;
    for (int j = 0; j <= n / 2; ++j) {
      const double ang = V_2PI * (double)j / (double)n;
      cv->wc[j] = fcos(ang);
      cv->ws[j] = fsin(ang);
    }
    cv->b = (double*)malloc(sizeof(double) * (size_t)n * 4);
  }
  double *r1 = cv->b, *i1 = cv->b + n, *r2 = cv->b + 2 * n, *i2 = cv->b + 3 * n;
  memset(cv->b, 0, sizeof(double) * (size_t)n * 4);
  memcpy(r1, a1, sizeof(double) * (size_t)n1);
  memcpy(r2, a2, sizeof(double) * (size_t)n2);
  v_fft(r1, i1, logn, 1, cv->wc, cv->ws);
  v_fft(r2, i2, logn, 1, cv->wc, cv->ws);
  for (int j = 0; j < n; ++j) {
    const double a = r1[j], b = i1[j], c = r2[j], d = i2[j];
    r1[j] = a * c - b * d;
    i1[j] = a * d + b * c;
  }
  v_fft(r1, i1, logn, 0, cv->wc, cv->ws);
  double* y = (double*)malloc(sizeof(double) * (size_t)m);
  const double k = dt / n;
  for (int j = 0; j < m; ++j) y[j] = sqrt(r1[j] * r1[j] + i1[j] * i1[j]) * k;
  *ny = m;
  return y;
}

/* ------------------------------------------------------------------ */
/* VDOSGn (NCVDOSGn.cc:48-492) with the default truncation and thinning
   (minOrder 5, thinNBins 1000, truncationThreshold 1e-14) */
typedef struct { double* s; int n; double lo, bw, hi, smax; unsigned long thin; } vgn_t;
typedef struct { vgn_t* g; int norders, cap; double kT; vconv_t cv; } vgnset_t;

static void vgn_make(vgn_t* g, double* spec, int n, double lo, double bw, unsigned long thin) {
  g->s = spec;
  g->n = n;
  g->lo = lo;
  g->bw = bw;
  g->thin = thin;
  g->hi = lo + (n - 1) * bw;
  double area = 0.0;
  for (int i = 0; i < n; ++i) area += spec[i];
  area *= bw;
  const double inv = 1.0 / area;
  double mx = -INFINITY;
  for (int i = 0; i < n; ++i) { spec[i] *= inv; mx = spec[i] > mx ? spec[i] : mx; }
  g->smax = mx;
}
static double vgn_eval(const vgn_t* g, double e) {
  if (!(g->lo <= e && e <= g->hi)) return 0.0;
  const double a = (e - g->lo) * (1.0 / g->bw);
  const double fa = floor(a);
  size_t idx = (size_t)fa;
  if (idx > (size_t)(g->n - 2)) idx = (size_t)(g->n - 2);
  const double f = a - fa;
  return g->s[idx] * (1.0 - f) + f * g->s[idx + 1];
}
static void vgn_erange(const vgn_t* g, double rel, double* lo, double* hi) {
  const double thr = rel * g->smax;
  *lo = g->lo;
  *hi = g->hi;
  for (int i = 0; i < g->n; ++i)
    if (g->s[i] >= thr) { *lo = g->lo + i * g->bw; break; }
  for (int i = g->n; i > 0; --i)
    if (g->s[i - 1] >= thr) { *hi = vmin(*hi, g->lo + (i - 1) * g->bw); break; }
}

static int vgnset_init(vgnset_t* gs, const veval_t* ve) {
  memset(gs, 0, sizeof(*gs));
  gs->kT = ve->kT;
  unsigned long nbins = ve->nptsext - 1;
  const unsigned long thicken = (unsigned long)ceil(400.0 / nbins);
  nbins *= thicken;
  const unsigned long npos = nbins + 1;
  double* eg = (double*)malloc(sizeof(double) * npos);
  v_linspace(0.0, ve->emax, (unsigned)npos, eg);
  const double bw = eg[npos - 1] / nbins;
  const unsigned long nspec = npos * 2 - 1;
  double* spec = (double*)calloc(nspec, sizeof(double));
  const double g0 = veval_gamma0(ve);
  for (unsigned long i = 0; i < npos; ++i) {
    double neg, pos;
    veval_g1pair(ve, eg[i], g0, &neg, &pos);
    spec[nbins + i] = pos;
    spec[nbins - i] = neg;
  }
  free(eg);
  double lower = -ve->emax;
  long first = 0, last = (long)nspec - 1;
  while (first != last && !(spec[first] > 0.0) && !(spec[first + 1] > 0.0)) ++first;
  while (last != 0 && !(spec[last] > 0.0) && !(spec[last - 1] > 0.0)) --last;
  if (first >= last || last - first < 3) {
    free(spec);
    g_vdos_err = "Too few non-zero pts in G1 spectrum.";
    return -1;
  }
  double* s1 = spec;
  int n1 = (int)nspec;
  if (first != 0 || last != (long)nspec - 1) {
    lower += first * bw;
    n1 = (int)(last - first + 1);
    s1 = (double*)malloc(sizeof(double) * (size_t)n1);
    memcpy(s1, spec + first, sizeof(double) * (size_t)n1);
    free(spec);
  }
  gs->cap = 16;
  gs->g = (vgn_t*)calloc((size_t)gs->cap, sizeof(vgn_t));
  vgn_make(&gs->g[0], s1, n1, lower, bw, 1);
  gs->norders = 1;
  return 0;
}
static void vgnset_free(vgnset_t* gs) {
  for (int i = 0; i < gs->norders; ++i) free(gs->g[i].s);
  free(gs->g);
  vconv_free(&gs->cv);
  memset(gs, 0, sizeof(*gs));
}
static double* v_thin(const double* v, int n, unsigned long tf, int* nout) {
  const int m = (int)((n + tf - 1) / tf);
  double* o = (double*)malloc(sizeof(double) * (size_t)m);
  int j = 0;
  for (long i = 0; i < n; i += (long)tf) o[j++] = v[i];
  *nout = m;
  return o;
}
static void vgnset_grow_one(vgnset_t* gs) {
  const unsigned order = (unsigned)gs->norders + 1;
  const unsigned o2 = order / 2, o1 = order - o2;
  const vgn_t* p1 = &gs->g[o1 - 1];
  const vgn_t* p2 = &gs->g[o2 - 1];
  double dt;
  unsigned long t1 = p1->thin, t2 = p2->thin;
  const double *in1 = p1->s, *in2 = p2->s;
  int n1 = p1->n, n2 = p2->n;
  double* tmp = NULL;
  if (t1 == t2) {
    dt = p1->bw;
  } else {
    dt = vmax(p1->bw, p2->bw);
    if (t1 > t2) {
      tmp = v_thin(in2, n2, t1 / t2, &n2);
      in2 = tmp;
      t2 *= t1 / t2;
    } else {
      tmp = v_thin(in1, n1, t2 / t1, &n1);
      in1 = tmp;
      t1 *= t2 / t1;
    }
  }
  double start = p1->lo + p2->lo;
  int ny = 0;
  double* y = v_convolve(&gs->cv, in1, n1, in2, n2, dt, &ny);
  free(tmp);
  unsigned long xthin = 1;
  if (order >= 5u) {
    double mx = -INFINITY;
    for (int i = 0; i < ny; ++i) mx = y[i] > mx ? y[i] : mx;
    const double cut = 1e-14 * mx;
    int ifront = 0, iback = ny - 1;
    for (; ifront < iback; ++ifront)
      if (y[ifront] > cut) break;
    for (; iback > ifront; --iback)
      if (y[iback] > cut) break;
    if (iback > ifront) {
      const int nn = iback - ifront + 1;
      memmove(y, y + ifront, sizeof(double) * (size_t)nn);
      ny = nn;
    }
    start += ifront * dt;
    if (ny > 1000) {
      while ((unsigned long)ny > 1000ul * xthin) xthin *= 2;
      if (xthin >= 8 && order <= 10u) xthin /= 2;
      int nt = 0;
      double* yt = v_thin(y, ny, xthin, &nt);
      free(y);
      y = yt;
      ny = nt;
      dt *= xthin;
    }
  }
  if (gs->norders == gs->cap) {
    gs->cap *= 2;
    gs->g = (vgn_t*)realloc(gs->g, sizeof(vgn_t) * (size_t)gs->cap);
  }
  vgn_make(&gs->g[gs->norders], y, ny, start, dt, t1 * xthin);
  gs->norders++;
}
static void vgnset_grow(vgnset_t* gs, unsigned target) {
  while ((unsigned)gs->norders < target) vgnset_grow_one(gs);
}

/* the spectrum of G_n (on lo + i*bw, i < npts) for n = 1..order, as
   ncrystal_raw_vdos2gn (ncrystal.cc:755-790); two-call (cap < npts returns
   npts without writing) */
int orc_vdos_gn(double emin, double emax, const double* density, int n, double T, double mass, int order,
                double* lo, double* bw, double* spec, int cap) {
  g_vdos_err = NULL;
  veval_t ve;
  if (order < 1 || veval_init(&ve, emin, emax, density, n, T, mass)) return -1;
  vgnset_t gs;
  if (vgnset_init(&gs, &ve)) { veval_free(&ve); return -1; }
  vgnset_grow(&gs, (unsigned)order);
  const vgn_t* g = &gs.g[order - 1];
  const int np = g->n;
  if (cap >= np) {
    *lo = g->lo;
    *bw = g->bw;
    memcpy(spec, g->s, sizeof(double) * (size_t)np);
  }
  vgnset_free(&gs);
  veval_free(&ve);
  return np;
}

/* ------------------------------------------------------------------ */
/* createScatteringKernel helpers (NCVDOSToScatKnl.cc) */
static double v_alpha2x(double kT, double msd) {
  const double fact = (2.0 * V_NMASS_EVC2 / (V_HBAR * V_HBAR));
  return fact * kT * msd;
}
/* findRoot (NCMath.cc:291-332) on f(y) = y exp(-y) - k */
static double v_findroot_yexp(double k, double a, double b, double acc, int* ok) {
  double fa = a * exp(-a) - k, fb = b * exp(-b) - k;
  *ok = 1;
  if (!(b > a) || ((!fa) * fb < 0.0)) { *ok = 0; return 0.0; }
  acc *= 0.5;
  unsigned i = 60;
  while (--i) {
    const double dfba = fb - fa;
    double c = (a * fb - b * fa) / dfba;
    if (b - a < acc) return c;
    const double kk = 0.15 * (b - a);
    c = vmax(a + kk, vmin(b - kk, c));
    const double fc = c * exp(-c) - k;
    if (!fc) return c;
    if (fa * fc < 0) { b = c; fb = fc; } else { a = c; fa = fc; }
  }
  *ok = 0;
  return 0.0;
}
/* rangeXNexpMX (NCVDOSToScatKnl.cc:332-357) */
static int v_range_xnexpmx(unsigned n, double eps, double* lo, double* hi) {
  const double fn = (double)n;
  const double k = V_INVE * pow(eps, 1.0 / fn);
  int ok1, ok2;
  *lo = fn * v_findroot_yexp(k, 0.0, 1.0, 1e-13, &ok1);
  *hi = fn * v_findroot_yexp(k, 1.0, 700.0, 1e-13, &ok2);
  return ok1 && ok2 ? 0 : -1;
}
/* sabPointWithinAlphaPlusCurve (NCVDOSToScatKnl.cc:412-429) */
static int v_within_alphaplus(double c, double alpha, double beta) {
  const double cpb = c + beta;
  if (cpb < 0.0) return 0;
  const double t = 0.5 * (alpha - beta) - c;
  return t <= 0.0 || c * cpb >= t * t;
}
/* findExtremeSABPointWithinAlphaPlusCurve (NCVDOSToScatKnl.cc:359-410) */
static void v_extreme_point(double E, double a0, double a1, double b0, double b1, double* ea, double* eb) {
  if (b1 <= -E) { *ea = -1.0; *eb = -1.0; return; }
  const double apb1 = 2 * E + b1 + 2 * sqrt(E * (E + b1));
  if (apb1 <= a0) { *ea = -1.0; *eb = -1.0; return; }
  b0 = vmax(b0, -E);
  const double apb0 = 2 * E + b0 + 2 * sqrt(E * (E + b0));
  if (apb0 >= a1) { *ea = a1; *eb = b0; return; }
  a1 = vmin(a1, apb1);
  if (apb0 < a0) b0 = a0 - 2.0 * sqrt(E * a0);
  *ea = a1;
  *eb = b0;
}

/* setupAlphaGrid (NCVDOSToScatKnl.cc:431-548) */
typedef struct { double v; int r0; } vflag_t;
static void v_stable_sort_flags(vflag_t* a, int n) {
  /* insertion sort: stable, ordering of std::pair<double,bool> */
  for (int i = 1; i < n; ++i) {
    vflag_t x = a[i];
    int j = i - 1;
    while (j >= 0 && (x.v < a[j].v || (x.v == a[j].v && x.r0 < a[j].r0))) { a[j + 1] = a[j]; --j; }
    a[j + 1] = x;
  }
}
static void v_finalise_alpha(const double* grid, int ng, const double* r0, int n0, double* out) {
  const int n = ng + n0;
  vflag_t* m = (vflag_t*)malloc(sizeof(vflag_t) * (size_t)n);
  for (int i = 0; i < ng; ++i) { m[i].v = grid[i]; m[i].r0 = 0; }
  for (int i = 0; i < n0; ++i) { m[ng + i].v = r0[i]; m[ng + i].r0 = 1; }
  v_stable_sort_flags(m, n);
  for (int i = 1; i < n - 1; ++i)
    if (m[i].r0) m[i].v = 0.5 * (m[i - 1].v + m[i + 1].v);
  for (int i = 0; i < n; ++i) out[i] = m[i].v;
  free(m);
}
static void v_alpha_grid(double kT, double msd, double alphaMax, unsigned npts, double* out) {
  const double a2x = v_alpha2x(kT, msd);
  const double x2a = 1.0 / a2x;
  const double alphaMin = x2a * 1e-50;
  const double alphaMin2 = x2a * 1e-10;
  const double g1 = x2a * 1.0;
  const double g15 = x2a * 15.0;
  const unsigned n0 = (unsigned)(npts * 0.15 + 0.5);
  const unsigned n1 = (unsigned)(npts * 0.29 + 0.5);
  const unsigned n2 = (unsigned)(npts * 0.23 + 0.5);
  const unsigned n3 = npts - (n1 + n2 + n0);
  const unsigned n123 = n1 + n2 + n3;
  const double upscat = (n0 < 10 ? 6.0 : (n0 > 50 ? 14.0 : 10.0));
  double* r0 = (double*)malloc(sizeof(double) * n0);
  v_linspace(vmin(1e-3, alphaMax * 0.01), vmin(upscat, alphaMax * 0.99), n0, r0);
  double* g = (double*)malloc(sizeof(double) * (n123 + 8));
  unsigned ng = 0;
  if (alphaMax <= alphaMin * 100.0) {
    v_linspace(alphaMax * 0.001, alphaMax, n123, g);
    ng = n123;
  } else {
    g[ng++] = alphaMin;
    if (alphaMax <= g1 * 10.0) {
      v_linspace(alphaMin2, alphaMax, n123 - 1, g + ng);
      ng += n123 - 1;
    } else {
      v_linspace(alphaMin2, g1, n1 - 1, g + ng);
      ng += n1 - 1;
      if (alphaMax < 2.0 * g15) {
        double* ls = (double*)malloc(sizeof(double) * (n2 + n3 + 2));
        v_linspace(g1, alphaMax, n2 + n3 + 2, ls);
        for (unsigned i = 1; i + 1 < n2 + n3 + 2; ++i) g[ng++] = ls[i];
        free(ls);
      } else {
        double* ls = (double*)malloc(sizeof(double) * (n2 + 2));
        v_linspace(g1, g15, n2 + 2, ls);
        for (unsigned i = 1; i + 1 < n2 + 2; ++i) g[ng++] = ls[i];
        free(ls);
        v_geomspace(g15, alphaMax, n3, g + ng);
        ng += n3;
      }
    }
  }
  v_finalise_alpha(g, (int)ng, r0, (int)n0, out);
  free(g);
  free(r0);
}

/* reducePtsInDistribution (NCMath.cc:419-505): repeatedly drop the interior
   point of least importance; ties in the importance ordered by insertion, as
   in the reference's multimap */
typedef struct { double imp; unsigned long seq; int idx; } vheap_t;
static int vheap_less(const vheap_t* a, const vheap_t* b) {
  return a->imp < b->imp || (a->imp == b->imp && a->seq < b->seq);
}
static void vheap_push(vheap_t* h, int* n, vheap_t e) {
  int i = (*n)++;
  h[i] = e;
  while (i > 0) {
    const int p = (i - 1) / 2;
    if (!vheap_less(&h[i], &h[p])) break;
    const vheap_t t = h[i]; h[i] = h[p]; h[p] = t;
    i = p;
  }
}
static vheap_t vheap_pop(vheap_t* h, int* n) {
  const vheap_t top = h[0];
  h[0] = h[--(*n)];
  int i = 0;
  while (1) {
    const int l = 2 * i + 1, r = l + 1;
    int m = i;
    if (l < *n && vheap_less(&h[l], &h[m])) m = l;
    if (r < *n && vheap_less(&h[r], &h[m])) m = r;
    if (m == i) break;
    const vheap_t t = h[i]; h[i] = h[m]; h[m] = t;
    i = m;
  }
  return top;
}
static int v_reduce_pts(double* x, double* y, int n, int target) {
  if (target >= n) return n;
  double ymax = -INFINITY;
  for (int i = 0; i < n; ++i) ymax = y[i] > ymax ? y[i] : ymax;
  const double inv = 1.0 / ymax;
  double* lny = (double*)malloc(sizeof(double) * (size_t)n);
  int *prv = (int*)malloc(sizeof(int) * (size_t)n), *nxt = (int*)malloc(sizeof(int) * (size_t)n);
  unsigned long* cur = (unsigned long*)malloc(sizeof(unsigned long) * (size_t)n);
  char* alive = (char*)malloc((size_t)n);
  for (int i = 0; i < n; ++i) {
    lny[i] = log(vmax(1e-20, y[i] * inv));
    prv[i] = i - 1;
    nxt[i] = i + 1;
    alive[i] = 1;
  }
  int cap = 4 * n + 16, hn = 0;
  vheap_t* h = (vheap_t*)malloc(sizeof(vheap_t) * (size_t)cap);
  unsigned long seq = 0;
#define V_IMP(i) (fabs(x[prv[i]] * (y[i] - y[nxt[i]]) + x[i] * (y[nxt[i]] - y[prv[i]]) + x[nxt[i]] * (y[prv[i]] - y[i])) * \
                  fabs(x[prv[i]] * (lny[i] - lny[nxt[i]]) + x[i] * (lny[nxt[i]] - lny[prv[i]]) + x[nxt[i]] * (lny[prv[i]] - lny[i])) * \
                  fabs(x[prv[i]] * (lny[i] - lny[nxt[i]]) + x[i] * (lny[nxt[i]] - lny[prv[i]]) + x[nxt[i]] * (lny[prv[i]] - lny[i])))
  for (int i = 1; i < n - 1; ++i) {
    vheap_t e = {V_IMP(i), seq++, i};
    cur[i] = e.seq;
    vheap_push(h, &hn, e);
  }
  int remaining = n;
  while (remaining > target) {
    vheap_t e = vheap_pop(h, &hn);
    if (!alive[e.idx] || cur[e.idx] != e.seq) continue;
    const int i = e.idx, a = prv[i], b = nxt[i];
    alive[i] = 0;
    --remaining;
    nxt[a] = b;
    prv[b] = a;
    if (hn + 2 >= cap) {
      cap *= 2;
      h = (vheap_t*)realloc(h, sizeof(vheap_t) * (size_t)cap);
    }
    if (a != 0) {
      vheap_t ea = {V_IMP(a), seq++, a};
      cur[a] = ea.seq;
      vheap_push(h, &hn, ea);
    }
    if (b != n - 1) {
      vheap_t eb = {V_IMP(b), seq++, b};
      cur[b] = eb.seq;
      vheap_push(h, &hn, eb);
    }
  }
#undef V_IMP
  int j = 0;
  for (int i = 0; i < n; ++i)
    if (alive[i]) { x[j] = x[i]; y[j] = y[i]; ++j; }
  free(lny); free(prv); free(nxt); free(cur); free(alive); free(h);
  return j;
}

/* setupBetaGrid (NCVDOSToScatKnl.cc:550-785); returns the grid size */
typedef struct { double b0, b1; unsigned n; } vgap_t;
static int vgap_less(const vgap_t* s, const vgap_t* o) {
  const double a = (s->b1 - s->b0) / (s->n + 1);
  const double b = (o->b1 - o->b0) / (o->n + 1);
  if (fabs(a - b) <= (0.5 * 1e-13) * (fabs(a) + fabs(b)) + 1e-13) return s->b0 > o->b0;
  return a > b;
}
static int vdcmp(const void* a, const void* b) {
  const double x = *(const double*)a, y = *(const double*)b;
  return x < y ? -1 : (x > y ? 1 : 0);
}
static int v_beta_grid(const vgnset_t* gs, double betaMax, unsigned lux, double** out) {
  const double invkT = 1.0 / gs->kT;
  const vgn_t* g1 = &gs->g[0];
  const double G1 = fabs(g1->lo * invkT);
  const unsigned o3 = (unsigned)gs->norders < 3u ? (unsigned)gs->norders : 3u;
  double G3 = fabs(gs->g[o3 - 1].lo * invkT);
  if (G1 == G3) G3 = G1 * 1.0001;
  betaMax = vmax(betaMax, G1 * 1.01);
  G3 = vmin(G3, betaMax * 0.9999);
  const int luxm1 = (int)(lux - 1u);
  const double D = vmin(betaMax * 0.9999, vmax(G3, (luxm1 > 2 ? luxm1 : 2) * 10.0));
  const unsigned ntotal = 100u * (1u << lux);
  /* negative part of G1's grid (in beta units) */
  int ne = g1->n;
  double* ev = (double*)malloc(sizeof(double) * (size_t)ne);
  double* rs = (double*)malloc(sizeof(double) * (size_t)ne);
  v_linspace(g1->lo * invkT, g1->hi * invkT, (unsigned)ne, ev);
  memcpy(rs, g1->s, sizeof(double) * (size_t)ne);
  const double epsilon = -0.1 * g1->bw;
  while (ev[ne - 1] > epsilon) --ne;
  const unsigned n1_max = (unsigned)(ne * 1.25 + 6.5) & ~1u;
  const unsigned n0 = 1;
  double L1 = 2 * G1, L2 = 2 * (D - G1);
  const double L3 = betaMax - D;
  L1 *= 4;
  L2 *= 2;
  const double Lnorm = 1.0 / (L1 + L2 + L3);
  double f1 = vmax(0.2, L1 * Lnorm), f2 = L2 * Lnorm;
  if (f1 + f2 > 0.99) {
    const double tmp = 0.99 / (f1 + f2);
    f1 *= tmp;
    f2 *= tmp;
  }
  unsigned n1 = (unsigned)(f1 * ntotal + 0.5) / 2;
  if (n1_max < n1) n1 = n1_max;
  unsigned n2 = (unsigned)(f2 * ntotal + 0.5) / 2;
  if (n1 < 15) n1 = 15;
  if (n2 < 1) n2 = 1;
  unsigned n012;
  while (1) {
    n012 = 2 * (n1 + n2) + n0;
    if (n012 >= ntotal - 1) {
      if (n2 > n1) --n2; else --n1;
    } else {
      break;
    }
  }
  const unsigned n3 = ntotal - n012;
  double* grid = (double*)malloc(sizeof(double) * (ntotal + 16));
  unsigned ng = 0;
  {
    double* v = (double*)malloc(sizeof(double) * (n3 + 1));
    v_linspace(-betaMax, -D, n3 + 1, v);
    for (unsigned i = 0; i < n3; ++i) grid[ng++] = v[i];
    free(v);
  }
  const unsigned idx_r2 = ng;
  {
    double* v = (double*)malloc(sizeof(double) * (n2 + 1));
    v_linspace(-D, -G1, n2 + 1, v);
    for (unsigned i = 0; i < n2; ++i) grid[ng++] = v[i];
    free(v);
  }
  {
    unsigned near0 = (unsigned)(n1 * 0.2 + 0.5);
    if (near0 < 5) near0 = 5;
    unsigned nspec = n1 - near0;
    double* extra = NULL;
    int nextra = 0;
    if (nspec >= (unsigned)ne) {
      nspec = (unsigned)ne;
      near0 = n1 - nspec;
    } else {
      unsigned ngaps = 0;
      if (nspec > 30 && (unsigned)ne - nspec > 10) {
        ngaps = (unsigned)(nspec * 0.1 + 0.5);
        if (ngaps < 5) ngaps = 5;
      }
      nspec -= ngaps;
      const double g1bw = ev[1] - ev[0];
      ne = v_reduce_pts(ev, rs, ne, (int)nspec);
      if (ngaps > 0) {
        vgap_t* gaps = (vgap_t*)malloc(sizeof(vgap_t) * (size_t)ne);
        int ngp = 0;
        const double dmn = 1.5 * g1bw;
        for (int i = 0; i + 1 < ne; ++i)
          if (ev[i + 1] - ev[i] > dmn) { gaps[ngp].b0 = ev[i]; gaps[ngp].b1 = ev[i + 1]; gaps[ngp].n = 0; ++ngp; }
        while (ngaps > 0 && ngp > 0) {
          for (int i = 1; i < ngp; ++i) {  /* stable insertion sort */
            vgap_t x = gaps[i];
            int j = i - 1;
            while (j >= 0 && vgap_less(&x, &gaps[j])) { gaps[j + 1] = gaps[j]; --j; }
            gaps[j + 1] = x;
          }
          gaps[0].n += 1;
          --ngaps;
        }
        for (int i = 0; i < ngp; ++i) nextra += (int)gaps[i].n;
        extra = (double*)malloc(sizeof(double) * (size_t)(nextra + 1));
        int k = 0;
        for (int i = 0; i < ngp; ++i) {
          if (!gaps[i].n) continue;
          const double bw = (gaps[i].b1 - gaps[i].b0) / (gaps[i].n + 1.0);
          for (unsigned q = 0; q < gaps[i].n; ++q) extra[k++] = gaps[i].b0 + (q + 1) * bw;
        }
        free(gaps);
      }
      if (ngaps > 0) near0 += ngaps;
    }
    double* sp = (double*)malloc(sizeof(double) * (size_t)(ne + nextra + 1));
    memcpy(sp, ev, sizeof(double) * (size_t)ne);
    if (nextra) {
      memcpy(sp + ne, extra, sizeof(double) * (size_t)nextra);
      qsort(sp, (size_t)(ne + nextra), sizeof(double), vdcmp);
    }
    free(extra);
    for (int i = 0; i < ne + nextra; ++i) grid[ng++] = sp[i];
    free(sp);
    const double last = grid[ng - 1];
    double* v = (double*)malloc(sizeof(double) * near0);
    v_geomspace(vmin(1e-50, -0.001 * last), -last * 0.1, near0, v);
    for (unsigned i = 0; i < near0; ++i) grid[ng++] = -v[near0 - 1 - i];
    free(v);
  }
  const unsigned idx_r1back = ng - 1;
  grid[ng++] = 0.0;
  for (unsigned i = idx_r1back; i >= idx_r2; --i) {
    grid[ng++] = -grid[i];
    if (i == 0) break;
  }
  free(ev);
  free(rs);
  *out = grid;
  return (int)ng;
}

/* Stirling's series sum (NCVDOSToScatKnl.cc:38-58) */
static double v_stirling9(double x) {
  const double c1 = 1. / 12., c2 = 1. / 288., c3 = -139 / 51840., c4 = -571. / 2488320.,
               c5 = 163879. / 209018880., c6 = 5246819. / 75246796800., c7 = -534703531. / 902961561600.,
               c8 = -4483131259. / 86684309913600., c9 = 432261921612371. / 514904800886784000.;
  return 1.0 + x * (c1 + x * (c2 + x * (c3 + x * (c4 + x * (c5 + x * (c6 + x * (c7 + x * (c8 + x * c9))))))));
}
static size_t v_closest(const double* v, size_t n, double value) {
  size_t lo = 0, cnt = n;
  while (cnt > 0) {
    const size_t h = cnt / 2;
    if (v[lo + h] < value) { lo = lo + h + 1; cnt -= h + 1; } else cnt = h;
  }
  if (lo == 0) return 0;
  if (lo == n) return n - 1;
  return fabs(v[lo] - value) < fabs(v[lo - 1] - value) ? lo : lo - 1;
}
/* fillSABFromVDOS (NCVDOSToScatKnl.cc:99-297) for orders [omin, omax] added
   into sab */
static void v_fill(const vgnset_t* gs, double msd, const double* ag, int na, const double* bg, int nb,
                   unsigned omin, unsigned omax, double* sab) {
  const unsigned maxo = (unsigned)gs->norders;
  const double kT = gs->kT;
  const double a2x = v_alpha2x(kT, msd);
  double* x = (double*)malloc(sizeof(double) * (size_t)na);
  double* emh = (double*)malloc(sizeof(double) * (size_t)na);
  double* lx = (double*)malloc(sizeof(double) * (size_t)na);
  double* fxn = (double*)malloc(sizeof(double) * (size_t)na);
  double* af = (double*)malloc(sizeof(double) * (size_t)na);
  for (int i = 0; i < na; ++i) {
    x[i] = ag[i] * a2x;
    emh[i] = exp(-0.5 * x[i]);
    lx[i] = maxo >= 16 ? log(x[i]) : 0.0;
    fxn[i] = emh[i];
  }
  const size_t iz = v_closest(bg, (size_t)nb, 0.0);
  const size_t iflip = v_closest(bg, (size_t)nb, -bg[nb - 1]);
  double* eb = (double*)malloc(sizeof(double) * (iz + 1));
  for (size_t j = 0; j <= iz; ++j) eb[j] = exp(bg[j]);
  for (unsigned n = 1; n <= maxo; ++n) {
    if (n < omin) continue;
    if (n > omax) break;
    const vgn_t* g = &gs->g[n - 1];
    const double invn = 1.0 / n;
    if (n < 16) {
      for (int i = 0; i < na; ++i) fxn[i] *= x[i] * invn;
      for (int i = 0; i < na; ++i) af[i] = fxn[i] * emh[i] * kT;
    } else {
      const double gn = v_stirling9(invn);
      const double fact = kT * V_INVSQRT2PI / (sqrt(n) * gn);
      if (!fact) continue;
      const double logn = log(n);
      for (int i = 0; i < na; ++i) {
        const double ea = n * (lx[i] - logn + 1.0) - x[i];
        af[i] = fact * exp(ea);
      }
    }
    for (size_t j = 0; j <= iz; ++j) {
      const double e = bg[j] * kT;
      if (!(g->lo <= e && e <= g->hi)) continue;
      double emb = 0.0;
      size_t pj = 0;
      if (j >= iflip && bg[j] != 0.0) {
        pj = iz + (iz - j);
        emb = eb[j];
      }
      const double G = 1.0 * vgn_eval(g, e);
      if (!(G > 0.0)) continue;
      int i = 0;
      while (i != na && !(af[i] > 0.0)) ++i;
      for (; i != na; ++i) {
        if (!(af[i] > 0.0)) break;
        const double c = af[i] * G;
        sab[j * (size_t)na + (size_t)i] += c;
        if (emb) sab[pj * (size_t)na + (size_t)i] += c * emb;
      }
    }
  }
  free(x); free(emh); free(lx); free(fxn); free(af); free(eb);
}

/* createScatteringKernel (NCVDOSToScatKnl.cc:787-944) with the default
   truncation/thinning and no order weights.  Result arrays are malloc'ed;
   freed with orc_vdos_kernel_free. */
struct orc_vdos_kernel {
  double *alpha, *beta, *sab;
  int na, nb, norders;
  double suggested_emax, msd, gamma0, emax_reached;
};
void orc_vdos_kernel_free(orc_vdos_kernel* k) {
  if (!k) return;
  free(k->alpha); free(k->beta); free(k->sab); free(k);
}
orc_vdos_kernel* orc_vdos_kernel_create(double emin, double emax, const double* density, int n, double T,
                                        double mass, unsigned lux, double target_emax_req) {
  g_vdos_err = NULL;
  static char msgbuf[512];
  if (lux > 5 || !(target_emax_req >= 0.0)) { g_vdos_err = "bad vdoslux or target Emax"; return NULL; }
  const double lux2emax[6] = {0.5, 1.0, 3.0, 5.0, 8.0, 12.0};
  double targetEmax = target_emax_req > 0.0 ? target_emax_req : lux2emax[lux];
  veval_t ve;
  if (veval_init(&ve, emin, emax, density, n, T, mass)) return NULL;
  const double kT = ve.kT, invkT = 1.0 / kT;
  const double gamma0 = veval_gamma0(&ve);
  const double msd = veval_msd(&ve, gamma0);
  double tE = targetEmax * invkT;
  unsigned maxo = 4;
  vgnset_t gs;
  if (vgnset_init(&gs, &ve)) { veval_free(&ve); return NULL; }
  vgnset_grow(&gs, maxo);
  unsigned order_limit = 1000;
  if (target_emax_req > 0.0 || lux == 5) order_limit *= 10;
  if (lux == 0) order_limit /= 10;
  const double emax_lowest = target_emax_req > 0.0 ? target_emax_req : 1e-15;
  const double rel = pow(10.0, -(3.0 + 2.0 * lux));
  const double x2a = 1.0 / v_alpha2x(kT, msd);
  orc_vdos_kernel* out = NULL;
  double a0, a1, b0, b1;
#define V_RANGES(nn) do { \
    vgn_erange(&gs.g[(nn) - 1], rel, &b0, &b1); b0 *= invkT; b1 *= invkT; \
    if (v_range_xnexpmx((nn), rel, &a0, &a1)) { g_vdos_err = "root search failed"; goto done; } \
    a0 *= x2a; a1 *= x2a; } while (0)
  while (1) {
    vgnset_grow(&gs, maxo);
    V_RANGES((unsigned)gs.norders);
    if (v_within_alphaplus(tE, a0, b1)) {
      ++maxo;
    } else {
      break;
    }
    if (maxo > order_limit) {
      double red = targetEmax;
      do {
        red *= 0.99;
        if (red < emax_lowest) {
          snprintf(msgbuf, sizeof msgbuf,
                   "VDOS expansion too slow - can not reach E=%geV after %u phonon convolutions (likely causes: "
                   "either the target energy value is too high, vdoslux too low, the temperature too high, or the "
                   "VDOS is very unusual).", emax_lowest, order_limit);
          g_vdos_err = msgbuf;
          goto done;
        }
      } while (v_within_alphaplus(red * invkT, a0, b1));
      tE = red * invkT;
      targetEmax = red;
      break;
    }
  }
  vgnset_grow(&gs, maxo);
  {
    double betaMin = 0.0, alphaMax = 0.0;
    for (unsigned o = 1; o <= maxo; ++o) {
      V_RANGES(o);
      double ea, ebb;
      v_extreme_point(tE, a0, a1, b0, b1, &ea, &ebb);
      alphaMax = vmax(alphaMax, ea);
      betaMin = vmin(betaMin, ebb);
    }
    if (!(betaMin < 0.0 && alphaMax > 0.0)) { g_vdos_err = "no kinematically accessible region"; goto done; }
    const double ub = -betaMin * 1.01, ua = alphaMax * 1.01;
    double* bg = NULL;
    const int nb = v_beta_grid(&gs, ub, lux, &bg);
    const int na = nb / 2;
    double* ag = (double*)malloc(sizeof(double) * (size_t)na);
    v_alpha_grid(kT, msd, ua, (unsigned)na, ag);
    double* sab = (double*)calloc((size_t)na * (size_t)nb, sizeof(double));
    const unsigned nord = (unsigned)gs.norders;
    const unsigned njobs = nord <= 16 ? 1 : nord / 16;
    if (njobs == 1) {
      v_fill(&gs, msd, ag, na, bg, nb, 1, 0xFFFFFFFFu, sab);
    } else {
      const unsigned per = nord / njobs;
      double* part = (double*)malloc(sizeof(double) * (size_t)na * (size_t)nb);
      unsigned next = 1;
      for (unsigned j = 0; j < njobs; ++j) {
        const unsigned lo = next;
        next += per;
        const unsigned hi = (next - 1) < nord ? next - 1 : nord;
        double* tgt = j == 0 ? sab : part;
        if (j) memset(part, 0, sizeof(double) * (size_t)na * (size_t)nb);
        v_fill(&gs, msd, ag, na, bg, nb, lo, hi, tgt);
        if (j)
          for (size_t q = 0; q < (size_t)na * (size_t)nb; ++q) sab[q] += part[q];
      }
      free(part);
    }
    out = (orc_vdos_kernel*)calloc(1, sizeof(*out));
    out->alpha = ag;
    out->beta = bg;
    out->sab = sab;
    out->na = na;
    out->nb = nb;
    out->norders = (int)nord;
    out->suggested_emax = targetEmax;
    out->msd = msd;
    out->gamma0 = gamma0;
    out->emax_reached = targetEmax;
  }
#undef V_RANGES
done:
  vgnset_free(&gs);
  veval_free(&ve);
  return out;
}
int orc_vdos_kernel_dims(const orc_vdos_kernel* k, int* na, int* nb, int* norders, double* info3) {
  *na = k->na;
  *nb = k->nb;
  *norders = k->norders;
  info3[0] = k->suggested_emax;
  info3[1] = k->msd;
  info3[2] = k->gamma0;
  return 0;
}
void orc_vdos_kernel_copy(const orc_vdos_kernel* k, double* alpha, double* beta, double* sab) {
  memcpy(alpha, k->alpha, sizeof(double) * (size_t)k->na);
  memcpy(beta, k->beta, sizeof(double) * (size_t)k->nb);
  memcpy(sab, k->sab, sizeof(double) * (size_t)k->na * (size_t)k->nb);
}
