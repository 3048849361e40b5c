// Layered-crystal Bragg scattering by the reference models of
// ncrystal_core/src/extd_utils/NCLCRefModels.cc: a mosaic single crystal
// (the SCBragg kernels) averaged over rotations of the crystal around the
// layer axis.  A rotation of the crystal by phi is a rotation of the neutron
// by -phi in the crystal's frame, so each neutron is expanded into N
// rotated directions (one lane per neutron and rotation), the SCBragg
// kernels evaluate all of them in one launch, and small per-neutron kernels
// average (cross sections) or pick one rotation by its cross section and
// rotate the scattered direction back (sampling).
#include "nc_kernels.hh"
#include "nc_dev.hh"
#include "nc_scb.hh"

namespace ncamd {
namespace {

// Rodrigues' rotation of the unit vector v by (c, s) = (cos phi, sin phi)
// around the unit axis a, in the operation order of
// ncrystal_core/include/NCrystal/internal/utils/NCRotMatrix.hh:172-184:
// v c + (a x v) s + a ((a . v)(1 - c))
NCD void lc_rotate(const LCView& L, double vx, double vy, double vz, double c, double s, double& ox, double& oy,
                   double& oz) {
  const double cx = L.ay * vz - L.az * vy, cy = L.az * vx - L.ax * vz, cz = L.ax * vy - L.ay * vx;
  const double f = (L.ax * vx + L.ay * vy + L.az * vz) * (1.0 - c);
  ox = vx * c + cx * s + L.ax * f;
  oy = vy * c + cy * s + L.ay * f;
  oz = vz * c + cz * s + L.az * f;
}

// lane t = i * N + j: neutron i rotated by rotation j, from the fixed table
// cs (equally spaced angles) or, with cs null, by a uniform random angle
// (Philox draw t of (seed, stream)), whose (cos, sin) goes to ocs
__global__ void __launch_bounds__(256) k_lc_expand(const double* __restrict__ ekin, const double* __restrict__ dirs,
                                                   int64_t n, int N, LCView L, const double2* __restrict__ cs,
                                                   uint64_t seed, uint32_t stream, double* __restrict__ oE,
                                                   double* __restrict__ oD, double2* __restrict__ ocs) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n * (int64_t)N) return;
  const int64_t i = t / N;
  const int j = (int)(t - i * N);
  double c, s;
  if (cs) {
    const double2 v = cs[j];
    c = v.x;
    s = v.y;
  } else {
    Rng r;
    r.init(seed, stream, (uint64_t)t);
    sincos(k2Pi * r.gen() - kPi, &s, &c);
    ocs[t] = make_double2(c, s);
  }
  double ux = dirs[3 * i], uy = dirs[3 * i + 1], uz = dirs[3 * i + 2];
  const double inv = 1.0 / sqrt(ux * ux + uy * uy + uz * uz);
  ux *= inv; uy *= inv; uz *= inv;
  double rx, ry, rz;
  lc_rotate(L, ux, uy, uz, c, s, rx, ry, rz);
  oE[t] = ekin[i];
  oD[3 * t] = rx;
  oD[3 * t + 1] = ry;
  oD[3 * t + 2] = rz;
}

// mean of each neutron's N cross sections, compensated (Neumaier) sum as the
// reference's StableSum
__global__ void __launch_bounds__(256) k_lc_mean(const double* __restrict__ xs, int64_t n, int N,
                                                 double* __restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double sum = 0.0, comp = 0.0;
  for (int j = 0; j < N; ++j) {
    const double x = xs[i * N + j];
    const double t = sum + x;
    comp += fabs(sum) >= fabs(x) ? (sum - t) + x : (x - t) + sum;
    sum = t;
  }
  out[i] = (sum + comp) / (double)N;
}

// one rotation per neutron with probability proportional to its cross
// section (running sums and the first one above u x total, as
// pickRandIdxByWeight); neutrons whose rotations all have zero cross section
// get sel = -1.  The chosen rotated direction goes to sD, the energy to sE.
__global__ void __launch_bounds__(256) k_lc_pick(const double* __restrict__ xs, const double* __restrict__ rE,
                                                 const double* __restrict__ rD, int64_t n, int N, uint64_t seed,
                                                 uint32_t stream, double* __restrict__ sE, double* __restrict__ sD,
                                                 int* __restrict__ sel) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double tot = 0.0;
  for (int j = 0; j < N; ++j) tot += xs[i * N + j];
  int pick = -1;
  if (tot > 0.0) {
    Rng r;
    r.init(seed, stream, (uint64_t)i);
    const double u = r.gen() * tot;
    double run = 0.0;
    pick = N - 1;
    for (int j = 0; j < N; ++j) {
      run += xs[i * N + j];
      if (u < run && xs[i * N + j] > 0.0) { pick = j; break; }
    }
    while (pick > 0 && !(xs[i * N + pick] > 0.0)) --pick;  // u == tot: the last rotation that contributes
  }
  const int64_t t = i * N + (pick < 0 ? 0 : pick);
  sel[i] = pick;
  sE[i] = rE[t];
  sD[3 * i] = rD[3 * t];
  sD[3 * i + 1] = rD[3 * t + 1];
  sD[3 * i + 2] = rD[3 * t + 2];
}

// scattered directions rotated back by the chosen rotation (reverse sense);
// neutrons without a rotation leave unchanged
__global__ void __launch_bounds__(256) k_lc_unrotate(const double* __restrict__ scD, const int* __restrict__ sel,
                                                     const double2* __restrict__ cs, const double* __restrict__ ekin,
                                                     const double* __restrict__ dirs, int64_t n, int N, LCView L,
                                                     double* __restrict__ oE, double* __restrict__ oD) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int j = sel[i];
  oE[i] = ekin[i];
  if (j < 0) {
    oD[3 * i] = dirs[3 * i];
    oD[3 * i + 1] = dirs[3 * i + 1];
    oD[3 * i + 2] = dirs[3 * i + 2];
    return;
  }
  const double2 v = cs[i * N + j];
  double rx, ry, rz;
  lc_rotate(L, scD[3 * i], scD[3 * i + 1], scD[3 * i + 2], v.x, -v.y, rx, ry, rz);
  oD[3 * i] = rx;
  oD[3 * i + 1] = ry;
  oD[3 * i + 2] = rz;
}

unsigned lc_blocks(int64_t n) { return (unsigned)((n + 255) / 256); }

// ---- lcmode=0: ncrystal_core/src/extd_utils/NCLCUtils.cc

// values on the 2^-40 grid of the model's cache keys (:45-52)
NCD double lc_discr(double x) {
  return (double)(uint64_t)(x * 1099511627776.0 + 0.5) * (1.0 / 1099511627776.0);
}

// GaussMos::calcRawCrossSectionValue after InteractionPars::set
// (NCGaussMos.hh:204-214, NCGaussMos.cc:252-280): Q times the
// Gaussian-on-sphere circle integral, Q = wl^3 / (sin cos theta_Bragg) x
// fsq / 2
NCD double lc0_raw(const SCBView& v, double wl, double inv2d, double fsq, double cosg) {
  const double w = gm_round(wl), i2 = gm_round(inv2d);
  const double sinpt = w * i2;
  const double cospt = sqrt(1 - sinpt * sinpt);
  const double tmp2 = cospt * sinpt;
  const double Qprime = tmp2 > 0 ? (w * w * w) / tmp2 : ((sinpt > 0.5 && fsq != 0.0) ? -2.0 : 0.0);
  if (!(Qprime > 0.0)) return Qprime != 0.0 ? __builtin_huge_val() : 0.0;
  const double Q = Qprime * (fsq * 0.5);
  const double cc = dclamp(cosg, -1.0, 1.0);
  const double sc = sqrt(1.0 - cc * cc);
  double val;
  if (!gos_circle_fast_ok(v, cc, sc, sinpt, cospt, val)) val = gos_circle_slow(v, cc, sc, sinpt, cospt);
  return Q * val;
}

// levels 5 and up of the Romberg integration below (rare at the model's
// accuracy)
NCD double lc0_romberg_tail(const SCBView& v, double wl, const LC0Set& ps, double k1, double k0, double acc, double a,
                            double h, double r40, double r41, double r42, double r43, double r44) {
  double prev[16], row[16];
  prev[0] = r40; prev[1] = r41; prev[2] = r42; prev[3] = r43; prev[4] = r44;
  int nj = 8;
  for (int i = 5; i < 16; ++i) {
    const double hh = h;
    h *= 0.5;
    nj *= 2;
    double c = 0.0;
    for (int j = 0; j < nj; ++j) c += lc0_raw(v, wl, ps.inv2d, ps.fsq, k1 * cos(a + h + hh * (double)j) + k0);
    row[0] = h * c + 0.5 * prev[0];
    double nk = 1.0;
    for (int j = 0; j < i; ++j) {
      nk *= 4.0;
      row[j + 1] = (nk * row[j] - prev[j]) / (nk - 1.0);
    }
    if (fabs(row[i] - prev[i - 1]) <= acc * fabs(row[i])) return row[i];
    for (int j = 0; j <= i; ++j) prev[j] = row[j];
  }
  return prev[15];
}

// Romberg integration of NCRomberg.cc:58-140 (17 points, halvings up to
// level 15, accept |R_k - R_(k-1)| <= acc |R_k|) of the cross section over
// crystallite rotations phi, cos(gamma) = k1 cos(phi) + k0
NCD double lc0_integrate(const SCBView& v, double wl, const LC0Set& ps, double k1, double k0, double acc, double a,
                         double b) {
  double h = b - a;
  const double delta = h * 0.0625;
  // the 17 points summed into the groups of the first five levels (ends,
  // midpoint, quarters, eighths, sixteenths), in one loop
  double s_end = 0.0, s_mid = 0.0, s_q = 0.0, s_e = 0.0, s_s = 0.0;
  for (int i = 0; i < 17; ++i) {
    const double f = lc0_raw(v, wl, ps.inv2d, ps.fsq, k1 * cos(a + delta * (double)i) + k0);
    if (i == 0 || i == 16) s_end += f;
    else if (i == 8) s_mid += f;
    else if (i == 4 || i == 12) s_q += f;
    else if ((i & 1) == 0) s_e += f;
    else s_s += f;
  }
  h *= 0.5;
  const double r00 = s_end * h;
  const double r10 = h * s_mid + 0.5 * r00;
  const double r11 = (4.0 * r10 - r00) / 3.0;
  h *= 0.5;
  const double r20 = h * s_q + 0.5 * r10;
  const double r21 = (4.0 * r20 - r10) / 3.0;
  const double r22 = (16.0 * r21 - r11) / 15.0;
  h *= 0.5;
  const double r30 = h * s_e + 0.5 * r20;
  const double r31 = (4.0 * r30 - r20) / 3.0;
  const double r32 = (16.0 * r31 - r21) / 15.0;
  const double r33 = (64.0 * r32 - r22) / 63.0;
  h *= 0.5;
  const double r40 = h * s_s + 0.5 * r30;
  const double r41 = (4.0 * r40 - r30) / 3.0;
  const double r42 = (16.0 * r41 - r31) / 15.0;
  const double r43 = (64.0 * r42 - r32) / 63.0;
  const double r44 = (256.0 * r43 - r33) / 255.0;
  if (fabs(r44 - r33) <= acc * fabs(r44)) return r44;
  return lc0_romberg_tail(v, wl, ps, k1, k0, acc, a, h, r40, r41, r42, r43, r44);
}

// a rotation range (ROI) of a plane set: kind 0 normal along the axis, 1
// neutron along the axis, 2 the range [r0, r1] of phi
struct LC0Roi {
  int set, kind;
  double sg, r0, r1;
};

// the ROIs of one neutron (LCROIFinder::findROIs, NCLCUtils.cc:184-303) and
// their cross sections (:396-439) in order; returns the sum, and with
// target >= 0 stops at the first ROI whose running sum exceeds it (pick)
NCD double lc0_walk(const LC0View& L, const SCBView& v, double wl, double c3, double s3, double target, LC0Roi& pick) {
  double tot = 0.0;
  for (int k = 0; k < L.nsets; ++k) {
    const LC0Set ps = L.sets[k];
    if (wl > ps.twodsp) break;
    const double c2 = ps.inv2d * wl;
    const double c23 = c3 * c2;
    const double s2 = sqrt(1.0 - c2 * c2);
    const double s23 = s3 * s2;
    const double lo = c23 - s23, hi = c23 + s23;
    if (dmax(lo, ps.c1plus) > dmin(hi, ps.c1minus)) continue;
    const bool anti = dmax(lo, -ps.c1minus) <= dmin(hi, -ps.c1plus);
    for (int side = 0; side < (anti ? 2 : 1); ++side) {
      const double sg = side ? -1.0 : 1.0;
      LC0Roi roi{k, 2, sg, 0.0, 0.0};
      double x;
      if (ps.s1 == 0.0) {
        roi.kind = 0;
        x = lc0_raw(v, wl, ps.inv2d, ps.fsq, sg * c3);
      } else if (fabs(s3) < 1e-10) {
        roi.kind = 1;
        x = lc0_raw(v, wl, ps.inv2d, ps.fsq, (ps.c1 * c3) * sg);
      } else {
        const double c2low = c2 * L.cta - s2 * L.sta;
        const double c2high = s2 < L.sta ? 1.0 : c2 * L.cta + s2 * L.sta;
        const double b = 1.0 / (ps.s1 * s3);
        const double mab = -(ps.c1 * c3) * b;
        const double p1 = dclamp(mab + sg * (c2low * b), -1.0, 1.0);
        const double p2 = dclamp(mab + sg * (c2high * b), -1.0, 1.0);
        if (!(fabs(p1 - p2) > 1e-10)) continue;
        roi.r0 = acos(dmax(p1, p2));
        roi.r1 = acos(dmin(p1, p2));
        x = lc0_integrate(v, wl, ps, ps.s1 * s3 * sg, ps.c1 * sg * c3, L.acc, roi.r0, roi.r1) * (1.0 / kPi);
      }
      tot += x;
      if (target >= 0.0 && tot > target) {
        pick = roi;
        return tot;
      }
      pick = roi;  // the last ROI, should rounding leave the target unreached
    }
  }
  return tot;
}

// discretised wavelength and |cos| to the layer axis (:332-344)
NCD void lc0_neutron(const LC0View& L, double E, double& ux, double& uy, double& uz, double& wl, double& c3raw,
                     double& c3, double& s3) {
  const double inv = 1.0 / sqrt(ux * ux + uy * uy + uz * uz);
  ux *= inv; uy *= inv; uz *= inv;
  wl = lc_discr(sqrt(kWl2Ekin / E));
  c3raw = L.ax * ux + L.ay * uy + L.az * uz;
  c3 = dmin(lc_discr(fabs(c3raw)), 1.0);
  s3 = sqrt(fabs(1.0 - c3 * c3));
}

__global__ void __launch_bounds__(256) k_lc0_xs(LC0View L, SCBView v, const double* __restrict__ ekin,
                                                const double* __restrict__ dirs, int64_t n, double* __restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double E = ekin[i];
  double tot = 0.0;
  if (E > 0.0) {
    double ux = dirs[3 * i], uy = dirs[3 * i + 1], uz = dirs[3 * i + 2], wl, c3raw, c3, s3;
    lc0_neutron(L, E, ux, uy, uz, wl, c3raw, c3, s3);
    LC0Roi roi;
    tot = lc0_walk(L, v, wl, c3, s3, -1.0, roi);
  }
  out[i] = L.xsfact * tot;
}

// scattering (LCHelper::genScatter, NCLCUtils.cc:518-690): a ROI by its
// cross section; in it a rotation phi by rejection under the reference's
// envelope (8 bins over the range, each the larger end value times 1.7 plus
// 2% of the largest, at most 1000 tries), uniform for a neutron along the
// axis; a random sign of sin(phi); the Gaussian-mosaic scattering of the
// reversed neutron (-s3, 0, -c3) off that normal in the frame of the axis
// (LCStdFrame::genScat, :755-772), taken to the lab frame by the neutron and
// the axis (rotateToFrame) and reversed
__global__ void __launch_bounds__(256) k_lc0_sample(LC0View L, SCBView v, const double* __restrict__ ekin,
                                                    const double* __restrict__ dirs, int64_t n, uint64_t seed,
                                                    uint32_t stream, double* __restrict__ oE, double* __restrict__ oD) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double E = ekin[i];
  double ox = dirs[3 * i], oy = dirs[3 * i + 1], oz = dirs[3 * i + 2];
  oE[i] = E;
  if (E > 0.0) {
    double ux = ox, uy = oy, uz = oz, wl, c3raw, c3, s3;
    lc0_neutron(L, E, ux, uy, uz, wl, c3raw, c3, s3);
    LC0Roi roi;
    const double tot = lc0_walk(L, v, wl, c3, s3, -1.0, roi);
    if (tot > 0.0) {
      Rng r;
      r.init(seed, stream, (uint64_t)i);
      lc0_walk(L, v, wl, c3, s3, r.gen() * tot, roi);
      const LC0Set ps = L.sets[roi.set];
      const V3 indir_std{-s3, 0.0, -c3};
      V3 normal{0.0, 0.0, roi.sg};
      if (roi.kind != 0) {
        double phi, cosphi;
        if (roi.kind == 1) {
          phi = r.gen() * kPi;
          cosphi = cos(phi);
        } else {
          const double k1 = ps.s1 * s3 * roi.sg, k0 = ps.c1 * roi.sg * c3;
          const double len = roi.r1 - roi.r0;
          double tmp[9];
          double mx = 0.0;
          for (int j = 0; j < 9; ++j) {
            tmp[j] = lc0_raw(v, wl, ps.inv2d, ps.fsq, k1 * cos(roi.r0 + j * (len * 0.125)) + k0);
            mx = dmax(mx, tmp[j]);
          }
          float data[8];
          float sum = 0.0f;
          for (int j = 0; j < 8; ++j) data[j] = (sum += (float)(dmax(tmp[j], tmp[j + 1]) * 1.7 + 0.02 * mx));
          phi = roi.r0;
          cosphi = cos(phi);
          for (int tries = 0; tries < 1000; ++tries) {
            const float t = data[7] * (float)r.gen();
            int b = 0;
            while (b < 7 && data[b] < t) ++b;
            const double ov = b ? (double)data[b] - (double)data[b - 1] : (double)data[0];
            phi = roi.r0 + ((b + r.gen()) * 0.125) * len;
            cosphi = cos(phi);
            const double xsphi = lc0_raw(v, wl, ps.inv2d, ps.fsq, k1 * cosphi + k0);
            if (xsphi > ov * r.gen()) break;
          }
        }
        const double sgn = r.gen() < 0.5 ? 1.0 : -1.0;
        const double sinphi = sgn * sqrt(dmax(0.0, 1.0 - cosphi * cosphi));
        const double ns1 = roi.sg * ps.s1;
        normal = V3{ns1 * cosphi, ns1 * sinphi, roi.sg * ps.c1};
      }
      const V3 w = gm_genscat(v, r, normal, gm_round(ps.inv2d), wl, indir_std);
      const double isg = c3raw >= 0.0 ? 1.0 : -1.0;
      const V3 a{ux, uy, uz}, b{L.ax * isg, L.ay * isg, L.az * isg};
      const V3 o = rotate_to_frame(s3, c3, a, b, w, r);
      ox = -o.x;
      oy = -o.y;
      oz = -o.z;
    }
  }
  oD[3 * i] = ox;
  oD[3 * i + 1] = oy;
  oD[3 * i + 2] = oz;
}

}  // namespace

void launch_lc0_xs(const LC0View& L, const SCBView& v, const double* ekin, const double* dirs, int64_t n, double* out,
                   hipStream_t st) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_lc0_xs, dim3(lc_blocks(n)), dim3(256), 0, st, L, v, ekin, dirs, n, out);
}
void launch_lc0_sample(const LC0View& L, const SCBView& v, const double* ekin, const double* dirs, int64_t n,
                       uint64_t seed, uint32_t stream, double* oE, double* oD, hipStream_t st) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_lc0_sample, dim3(lc_blocks(n)), dim3(256), 0, st, L, v, ekin, dirs, n, seed, stream, oE, oD);
}

void launch_lc_expand(const double* ekin, const double* dirs, int64_t n, int N, const LCView& L, const double2* cs,
                      uint64_t seed, uint32_t stream, double* oE, double* oD, double2* ocs, hipStream_t st) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_lc_expand, dim3(lc_blocks(n * N)), dim3(256), 0, st, ekin, dirs, n, N, L, cs, seed, stream, oE,
                     oD, ocs);
}
void launch_lc_mean(const double* xs, int64_t n, int N, double* out, hipStream_t st) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_lc_mean, dim3(lc_blocks(n)), dim3(256), 0, st, xs, n, N, out);
}
void launch_lc_pick(const double* xs, const double* rE, const double* rD, int64_t n, int N, uint64_t seed,
                    uint32_t stream, double* sE, double* sD, int* sel, hipStream_t st) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_lc_pick, dim3(lc_blocks(n)), dim3(256), 0, st, xs, rE, rD, n, N, seed, stream, sE, sD, sel);
}
void launch_lc_unrotate(const double* scD, const int* sel, const double2* cs, const double* ekin, const double* dirs,
                        int64_t n, int N, const LCView& L, double* oE, double* oD, hipStream_t st) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_lc_unrotate, dim3(lc_blocks(n)), dim3(256), 0, st, scD, sel, cs, ekin, dirs, n, N, L, oE, oD);
}

}  // namespace ncamd
