// |F(hkl)|^2 and d-spacing of every (h,k,l) in the enumeration domain of a
// cell (no space group), or of a list of
// symmetry representatives (space group), one lane per (h,k,l).
//
// Follows ncrystal_core/src/extd_utils/NCFillHKL.cc:299-383 in operation
// order (so that d-spacings are bit-identical and |F|^2 agrees to the last
// few ulp of exp): the reciprocal lattice
// (ncrystal_core/src/utils/NCLatticeUtils.cc:95-149) times (h,k,l), Debye-Waller factors
// exp(-0.5 ksq msd) with the per-species early-abort thresholds, phases with
// the polynomial sincos of ncrystal_core/include/NCrystal/internal/utils/NCMath.hh:318-326
// and ncrystal_core/src/utils/NCMath.cc:96-125, Neumaier sums per species and
// over species.
#include "nc_kernels.hh"
#include "nc_math.hh"

namespace ncamd {

NCD void sincos_2pix(double xx, double& s, double& c) {
  double A = (xx - floor(xx)) * k2Pi;
  A -= kPi;
  const double Aabs = fabs(A);
  sincos_mpi2pi2(dmin(Aabs, kPi - Aabs), c, s);
  c = copysign(c, Aabs - kPiHalf);
  s = copysign(s, -A);
}

constexpr int kMaxHKLSpecies = 32;

__global__ void __launch_bounds__(256) k_hkl_eval(HKLBuild p, int64_t first, int64_t count,
                                                  double* __restrict__ fsq_out, double* __restrict__ dsp_out,
                                                  unsigned char* __restrict__ valid) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= count) return;
  const int64_t q = first + t;
  const int64_t nl = 2 * (int64_t)p.max_l + 1;
  const int64_t nk_full = 2 * (int64_t)p.max_k + 1;
  const int64_t n_h0 = ((int64_t)p.max_k + 1) * nl;
  int h, k, l;
  if (q < n_h0) {
    h = 0;
    k = (int)(q / nl);
    l = (int)(q % nl) - p.max_l;
  } else {
    const int64_t qq = q - n_h0;
    const int64_t per_h = nk_full * nl;
    h = 1 + (int)(qq / per_h);
    const int64_t r = qq % per_h;
    k = (int)(r / nl) - p.max_k;
    l = (int)(r % nl) - p.max_l;
  }
  unsigned char ok = 0;
  double fsq = 0.0, dsp = 0.0;
  const double dh = (double)h, dk = (double)k, dl = (double)l;
  double wx, wy, wz;
  {
    // RotMatrix::operator* order without contraction (NCRotMatrix.hh)
#pragma clang fp contract(off)
    wx = p.rm[0] * dh + p.rm[1] * dk + p.rm[2] * dl;
    wy = p.rm[3] * dh + p.rm[4] * dk + p.rm[5] * dl;
    wz = p.rm[6] * dh + p.rm[7] * dk + p.rm[8] * dl;
  }
  const double ksq = wx * wx + wy * wy + wz * wz;
  do {
    if (!in_interval(p.ksq_lo, p.ksq_hi, ksq)) break;
    const double kk2 = 0.5 * ksq;
    double fac[kMaxHKLSpecies];
    double upper = 0.0;
    for (int i = 0; i < p.nspecies; ++i) {
      const double w = kk2 * p.msd[i];
      if (w > p.thr[i]) { fac[i] = 0.0; continue; }
      const double f = p.csl[i] * exp(-w);
      fac[i] = f;
      upper += (double)(p.pos_off[i + 1] - p.pos_off[i]) * fabs(f);
    }
    if (upper * upper * 2.0 < p.fsquarecut) break;
    if (h == 0 && k == 0 && l <= 0) break;
    StableSum re, im;
    for (int i = 0; i < p.nspecies; ++i) {
      const double f = fac[i];
      if (!f) continue;
      StableSum cs, sn;
      for (int j = p.pos_off[i]; j < p.pos_off[i + 1]; ++j) {
        const double ph = dh * p.pos[3 * j] + dk * p.pos[3 * j + 1] + dl * p.pos[3 * j + 2];
        double sv, cv;
        sincos_2pix(ph, sv, cv);
        cs.add(cv);
        sn.add(sv);
      }
      re.add(cs.sum() * f);
      im.add(sn.sum() * f);
    }
    const double r = re.sum(), ii = im.sum();
    fsq = r * r + ii * ii;
    if (fsq < p.fsquarecut) break;
    const double kval = sqrt(ksq);
    const double invk = 1.0 / kval;
    dsp = k2Pi * invk;
    if (!in_interval(p.dcut_lo, p.dcut_hi, dsp)) break;
    ok = 1;
  } while (false);
  fsq_out[t] = fsq;
  dsp_out[t] = dsp;
  valid[t] = ok;
}

// |F|^2 and d-spacing of listed (h,k,l) representatives of the space-group
// path, ncrystal_core/src/extd_utils/NCFillHKL.cc:609-681 in operation
// order: wave vector by the reciprocal-lattice matrix
// (ncrystal_core/include/NCrystal/internal/utils/NCRotMatrix.hh:103-110),
// the same Debye-Waller and phase sums as above, d = 2 pi / sqrt(ksq).
__global__ void __launch_bounds__(256) k_hkl_eval_keys(HKLBuild p, HKLMatrix rm, const int* __restrict__ hkl,
                                                       int64_t count, double* __restrict__ fsq_out,
                                                       double* __restrict__ dsp_out,
                                                       unsigned char* __restrict__ valid) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= count) return;
  const double dh = (double)hkl[3 * t], dk = (double)hkl[3 * t + 1], dl = (double)hkl[3 * t + 2];
  const double* m = rm.m;
  const double wx = m[0] * dh + m[1] * dk + m[2] * dl;
  const double wy = m[3] * dh + m[4] * dk + m[5] * dl;
  const double wz = m[6] * dh + m[7] * dk + m[8] * dl;
  const double ksq = wx * wx + wy * wy + wz * wz;
  unsigned char ok = 0;
  double fsq = 0.0, dsp = 0.0;
  do {
    if (!in_interval(p.ksq_lo, p.ksq_hi, ksq)) break;
    const double kk2 = 0.5 * ksq;
    double fac[kMaxHKLSpecies];
    double upper = 0.0;
    for (int i = 0; i < p.nspecies; ++i) {
      const double w = kk2 * p.msd[i];
      if (w > p.thr[i]) { fac[i] = 0.0; continue; }
      const double f = p.csl[i] * exp(-w);
      fac[i] = f;
      upper += (double)(p.pos_off[i + 1] - p.pos_off[i]) * fabs(f);
    }
    if (upper * upper * 2.0 < p.fsquarecut) break;
    StableSum re, im;
    for (int i = 0; i < p.nspecies; ++i) {
      const double f = fac[i];
      if (!f) continue;
      StableSum cs, sn;
      for (int j = p.pos_off[i]; j < p.pos_off[i + 1]; ++j) {
        const double ph = dh * p.pos[3 * j] + dk * p.pos[3 * j + 1] + dl * p.pos[3 * j + 2];
        double sv, cv;
        sincos_2pix(ph, sv, cv);
        cs.add(cv);
        sn.add(sv);
      }
      re.add(cs.sum() * f);
      im.add(sn.sum() * f);
    }
    const double r = re.sum(), ii = im.sum();
    fsq = r * r + ii * ii;
    if (fsq < p.fsquarecut) break;
    dsp = k2Pi / sqrt(ksq);
    if (!in_interval(p.dcut_lo, p.dcut_hi, dsp)) break;
    ok = 1;
  } while (false);
  fsq_out[t] = fsq;
  dsp_out[t] = dsp;
  valid[t] = ok;
}

void launch_hkl_eval_keys(const HKLBuild& p, const HKLMatrix& rm, const int* hkl, int64_t count, double* fsq,
                          double* dsp, unsigned char* valid, hipStream_t st) {
  if (count <= 0) return;
  hipLaunchKernelGGL(k_hkl_eval_keys, dim3((unsigned)((count + 255) / 256)), dim3(256), 0, st, p, rm, hkl, count,
                     fsq, dsp, valid);
}

void launch_hkl_eval(const HKLBuild& p, int64_t first, int64_t count, double* fsq, double* dsp,
                     unsigned char* valid, hipStream_t st) {
  if (count <= 0) return;
  hipLaunchKernelGGL(k_hkl_eval, dim3((unsigned)((count + 255) / 256)), dim3(256), 0, st, p, first, count, fsq,
                     dsp, valid);
}

}  // namespace ncamd
