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

}  // namespace

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
