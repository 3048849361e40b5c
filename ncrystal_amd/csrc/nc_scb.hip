// Per-neutron kernels of the mosaic single-crystal Bragg process: oriented
// cross sections and scatter sampling, one lane per neutron.
//
// Each kernel pair is a main kernel without the Romberg circle integral and
// a queue kernel that redoes, with it, the neutrons whose scan met a
// contribution outside the spline's validity (backscattering within the
// mosaic spread).  Cross sections involve no random numbers and sampling
// draws from the neutron's own Philox stream from its first draw, so the
// split does not change any result.
//
// Reference: ncrystal_core/src/scbragg/NCSCBragg.cc:233-323 (crossSection,
// sampleScatter), ncrystal_core/src/cinterface/ncrystal.cc:1248-1282 (batched
// layout).
#include "nc_kernels.hh"

namespace ncamd {

namespace {

template <bool AllowSlow>
NCD void scb_xs_one(const SCBView& v, bool valid, int64_t i, const double* __restrict__ ekin,
                    const double* __restrict__ dirs, double* __restrict__ out, int64_t* __restrict__ queue,
                    unsigned long long* __restrict__ qcount) {
  ScbLane L;
  L.wl = 1.0;
  L.u = V3{0.0, 0.0, 1.0};
  L.nf = 0;
  bool can = false;
  if (valid) can = scb_lane_setup(v, ekin[i], dirs[3 * i], dirs[3 * i + 1], dirs[3 * i + 2], L);
  scb_walk<false, AllowSlow>(v, L, 0, 0.0);
  if (!valid) return;
  if (!AllowSlow && L.slow) {
    const unsigned long long slot = atomicAdd(qcount, 1ull);
    queue[slot] = i;
    return;
  }
  out[i] = (can && L.count) ? L.total : 0.0;
}

template <bool AllowSlow>
NCD void scb_sample_one(const SCBView& v, bool valid, int64_t i, const double* __restrict__ ekin,
                        const double* __restrict__ dirs, uint64_t seed, uint32_t stream, double* __restrict__ oE,
                        double* __restrict__ oD, int64_t* __restrict__ queue,
                        unsigned long long* __restrict__ qcount) {
  double E = 0.0;
  V3 din{0.0, 0.0, 1.0};
  ScbLane L;
  L.wl = 1.0;
  L.u = din;
  L.nf = 0;
  bool can = false;
  if (valid) {
    E = ekin[i];
    din = V3{dirs[3 * i], dirs[3 * i + 1], dirs[3 * i + 2]};
    can = scb_lane_setup(v, E, din.x, din.y, din.z, L);
  }
  // pass 1: number of contributions and their cumulative total
  scb_walk<false, AllowSlow>(v, L, 0, 0.0);
  bool queued = false;
  if (valid && !AllowSlow && L.slow) {
    const unsigned long long slot = atomicAdd(qcount, 1ull);
    queue[slot] = i;
    queued = true;
  }
  const bool scatter = valid && can && !queued && L.count > 0 && L.total > 0.0;
  Rng rng;
  rng.init(seed, stream, (uint64_t)i);
  double target = 0.0;
  if (scatter && L.count > 1) target = L.total * rng.gen();
  // pass 2: the contribution whose cumulative value passes the target
  ScbLane S = L;
  if (!scatter) S.nf = 0;
  scb_walk<true, AllowSlow>(v, S, L.count, target);
  if (!valid || queued) return;
  V3 o = din;
  if (scatter && S.found) o = gm_genscat(v, rng, S.sel_normal, S.sel_inv2d_r, L.wl, L.u);
  oE[i] = E;
  oD[3 * i] = o.x;
  oD[3 * i + 1] = o.y;
  oD[3 * i + 2] = o.z;
}

}  // namespace

// Grid-stride loops step over whole blocks so that every lane of a
// wavefront reaches the wave-uniform scan together.
__global__ void __launch_bounds__(256) k_scb_xs_main(SCBView v, const double* __restrict__ ekin,
                                                     const double* __restrict__ dirs, int64_t n,
                                                     double* __restrict__ out, int64_t* __restrict__ queue,
                                                     unsigned long long* __restrict__ qcount) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t base = (int64_t)blockIdx.x * blockDim.x; base < n; base += stride) {
    const int64_t i = base + threadIdx.x;
    scb_xs_one<false>(v, i < n, i, ekin, dirs, out, queue, qcount);
  }
}
__global__ void __launch_bounds__(256) k_scb_xs_slow(SCBView v, const double* __restrict__ ekin,
                                                     const double* __restrict__ dirs, double* __restrict__ out,
                                                     const int64_t* __restrict__ queue,
                                                     const unsigned long long* __restrict__ qcount) {
  const int64_t nq = (int64_t)*qcount;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t base = (int64_t)blockIdx.x * blockDim.x; base < nq; base += stride) {
    const int64_t q = base + threadIdx.x;
    const bool valid = q < nq;
    const int64_t i = valid ? queue[q] : 0;
    scb_xs_one<true>(v, valid, i, ekin, dirs, out, nullptr, nullptr);
  }
}
__global__ void __launch_bounds__(256) k_scb_sample_main(SCBView v, const double* __restrict__ ekin,
                                                         const double* __restrict__ dirs, int64_t n, uint64_t seed,
                                                         uint32_t stream, double* __restrict__ oE,
                                                         double* __restrict__ oD, int64_t* __restrict__ queue,
                                                         unsigned long long* __restrict__ qcount) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t base = (int64_t)blockIdx.x * blockDim.x; base < n; base += stride) {
    const int64_t i = base + threadIdx.x;
    scb_sample_one<false>(v, i < n, i, ekin, dirs, seed, stream, oE, oD, queue, qcount);
  }
}
__global__ void __launch_bounds__(256) k_scb_sample_slow(SCBView v, const double* __restrict__ ekin,
                                                         const double* __restrict__ dirs, uint64_t seed,
                                                         uint32_t stream, double* __restrict__ oE,
                                                         double* __restrict__ oD, const int64_t* __restrict__ queue,
                                                         const unsigned long long* __restrict__ qcount) {
  const int64_t nq = (int64_t)*qcount;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t base = (int64_t)blockIdx.x * blockDim.x; base < nq; base += stride) {
    const int64_t q = base + threadIdx.x;
    const bool valid = q < nq;
    const int64_t i = valid ? queue[q] : 0;
    scb_sample_one<true>(v, valid, i, ekin, dirs, seed, stream, oE, oD, nullptr, nullptr);
  }
}

__global__ void k_zero_u64_scb(unsigned long long* p) { *p = 0ull; }

static inline unsigned scb_grid(int64_t n) {
  int64_t g = (n + 255) / 256;
  const int64_t gmax = 256 * 64;
  if (g > gmax) g = gmax;
  if (g < 1) g = 1;
  return (unsigned)g;
}

void launch_scb_xs(const SCBView& v, const double* ekin, const double* dirs, int64_t n, double* out,
                   int64_t* queue, unsigned long long* qcount, hipStream_t st) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_zero_u64_scb, dim3(1), dim3(1), 0, st, qcount);
  hipLaunchKernelGGL(k_scb_xs_main, dim3(scb_grid(n)), dim3(256), 0, st, v, ekin, dirs, n, out, queue, qcount);
  hipLaunchKernelGGL(k_scb_xs_slow, dim3(256), dim3(256), 0, st, v, ekin, dirs, out, queue, qcount);
}

void launch_scb_sample(const SCBView& v, const double* ekin, const double* dirs, int64_t n, uint64_t seed,
                       uint32_t stream, double* oE, double* oD, int64_t* queue, unsigned long long* qcount,
                       hipStream_t st) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_zero_u64_scb, dim3(1), dim3(1), 0, st, qcount);
  hipLaunchKernelGGL(k_scb_sample_main, dim3(scb_grid(n)), dim3(256), 0, st, v, ekin, dirs, n, seed, stream, oE, oD,
                     queue, qcount);
  hipLaunchKernelGGL(k_scb_sample_slow, dim3(256), dim3(256), 0, st, v, ekin, dirs, seed, stream, oE, oD, queue,
                     qcount);
}

}  // namespace ncamd
