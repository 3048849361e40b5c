// MiniMC on the device: a neutron beam through a sphere of material, with
// the forced-collision scheme of NCrystal's standard engine
// (ncrystal_core/src/minimc/NCMMC_SimEngine.cc:170-449): in each step every
// live neutron splits into the part transmitted to the surface without
// scattering (weighted by the scattering transmission probability and the
// absorption attenuation, and tallied) and the part that scatters at a
// distance drawn from the exponential truncated to the path inside the
// volume (weighted by 1 - transmission and the absorption up to that point,
// with Russian roulette of low-weight multiply scattered neutrons).  The
// scattered neutrons of a step are appended to the next step's arrays
// (structure of arrays, wave-aggregated appends) so that the library's
// cross-section and scattering kernels run on them unchanged.
#include "nc_kernels.hh"

namespace ncamd {

namespace {

constexpr double kMmcRadToDeg = 57.295779513082320876798154814105170332405472466564;

// Philox streams of the engine: (seed, kind + step, source index)
constexpr uint32_t kMmcStreamSrc = 0x6d6d6300u;
constexpr uint32_t kMmcStreamStep = 0x6d6d6400u;

// distance along u from p to the surface of a sphere of radius^2 r2, p inside
NCD double sphere_exit(double r2, double x, double y, double z, double ux, double uy, double uz) {
  const double b = x * ux + y * uy + z * uz;
  const double c = x * x + y * y + z * z - r2;
  const double disc = b * b - c;
  const double d = -b + sqrt(disc > 0.0 ? disc : 0.0);
  return d > 0.0 ? d : 0.0;
}

// distance along u from p to where it enters the sphere: 0 inside (or on the
// surface moving inwards), -1 when it misses (NCMMC_Sphere.hh distToVolumeEntry)
NCD double sphere_entry(double r2, double x, double y, double z, double ux, double uy, double uz) {
  const double b = x * ux + y * uy + z * uz;
  const double c = x * x + y * y + z * z - r2;
  if (c <= 0.0) return (c < b ? c : b) < 0.0 ? 0.0 : -1.0;
  const double disc = b * b - c;
  if (disc < 0.0) return -1.0;
  const double t = -b - sqrt(disc);
  return t >= 0.0 ? t : -1.0;
}

// scattering-history category of the tally breakdown (NCMMC_StdTallies.cc:
// init_dethistidvect): no scattering, one elastic, one inelastic, several
// all elastic, several otherwise
NCD int mmc_category(int nscat, int ninel) {
  if (nscat > 1) return ninel ? 4 : 3;
  if (nscat == 1) return ninel ? 2 : 1;
  return 0;
}

// wavelength [Aa] of energy [eV] (ekin2wl) and |k|^2 [1/Aa^2] per eV
constexpr double kMmcWl2Ekin = 0.081804209605330899;
constexpr double kMmcEkin2Ksq = 39.478417604357434475337963999505 / kMmcWl2Ekin;

// the tallied quantity of an outgoing neutron (NCMMC_StdTallies.cc:328-520,
// 590-670): angle or cosine to the beam, scatterings (-1: missed the
// volume), weight, energy, wavelength, energy transfer, momentum transfer
NCD double mmc_value(const MmcTally& t, double ux, double uy, double uz, double w, int nscat, double E) {
  const double mu0 = ux * t.u0x + uy * t.u0y + uz * t.u0z;
  const double mu = mu0 < -1.0 ? -1.0 : (mu0 > 1.0 ? 1.0 : mu0);
  switch (t.kind) {
    case kMmcTallyTheta: return acos(mu) * kMmcRadToDeg;
    case kMmcTallyMu: return mu;
    case kMmcTallyNscat:
    case kMmcTallyNscatUW: return (double)nscat;
    case kMmcTallyE: return E;
    case kMmcTallyL: return sqrt(kMmcWl2Ekin / (E > 4.9406564584124654e-324 ? E : 4.9406564584124654e-324));
    case kMmcTallyDE: return t.e0 - E;
    case kMmcTallyQ: {
      const double q2 = (E + t.e0 - 2.0 * mu * sqrt(E * t.e0)) * kMmcEkin2Ksq;
      return q2 > 0.0 ? sqrt(q2) : 0.0;
    }
    default: return w;
  }
}
// the w and nscat_uw tallies count records, the others sum weights
NCD double mmc_weight(const MmcTally& t, double w) {
  return (t.kind == kMmcTallyW || t.kind == kMmcTallyNscatUW) ? 1.0 : w;
}

// per-workgroup histograms in LDS: [1 + 5 categories][sum w, sum w^2][underflow, bins, overflow]
struct MmcLdsHist {
  double h[kMmcCats * 2 * (kMmcMaxBins + 2)];
};

// slot of a value among (underflow, bins, overflow) as the reference's
// histograms bin it (NCHists.hh valueToBin): the upper edge itself goes to
// the last bin
NCD int mmc_slot(const MmcTally& t, double v) {
  if (!(v >= t.xmin)) return 0;  // underflow (and NaN)
  if (v >= t.xmax) return v == t.xmax ? t.nbins : t.nbins + 1;
  const double f = (v - t.xmin) * t.inv_bw;
  const int b = f < (double)t.nbins ? (int)f : t.nbins;
  return 1 + b;
}

// weighted fills skip weights that are not positive (NCHists.hh fill)
NCD void mmc_fill(MmcLdsHist& L, const MmcTally& t, double v, double w, int cat) {
  if (!(w > 0.0)) return;
  const int s = mmc_slot(t, v);
  const int stride = t.nbins + 2;
  atomicAdd(&L.h[s], w);
  atomicAdd(&L.h[stride + s], w * w);
  if (t.breakdown) {
    const int o = (1 + cat) * 2 * stride;
    atomicAdd(&L.h[o + s], w);
    atomicAdd(&L.h[o + stride + s], w * w);
  }
}

NCD void mmc_hist_clear(MmcLdsHist& L, const MmcTally& t) {
  const int m = kMmcCats * 2 * (t.nbins + 2);
  for (int i = (int)threadIdx.x; i < m; i += (int)blockDim.x) L.h[i] = 0.0;
  __syncthreads();
}

// each workgroup adds its histograms to its own slice of the partial sums
// (launches on one stream, at most kMmcBlocks workgroups: no two writers of
// a slice at a time); k_mmc_reduce sums the slices at the end
NCD void mmc_hist_flush(MmcLdsHist& L, const MmcTally& t) {
  __syncthreads();
  const int m = (t.breakdown ? kMmcCats : 1) * 2 * (t.nbins + 2);
  double* part = t.part + (size_t)blockIdx.x * kMmcCats * 2 * (t.nbins + 2);
  for (int i = (int)threadIdx.x; i < m; i += (int)blockDim.x)
    if (L.h[i] != 0.0) part[i] += L.h[i];
}

// source neutrons id0 .. id0 + n - 1: position and direction, moved to the
// sphere's surface; those missing it are tallied unless ignored
constexpr int kMmcThreads = 512;
constexpr int kMmcWaves = kMmcThreads / 64;

__global__ void __launch_bounds__(kMmcThreads) k_mmc_source(MmcSource src, double r2, MmcTally t, int ignoremiss, uint64_t seed,
                                                    int64_t id0, int64_t n, MmcBasket out,
                                                    unsigned long long* __restrict__ count) {
  __shared__ MmcLdsHist L;
  __shared__ unsigned long long s_app[kMmcWaves + 1];
  mmc_hist_clear(L, t);
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i0 = (int64_t)blockIdx.x * blockDim.x; i0 < n; i0 += stride) {
    const int64_t i = i0 + threadIdx.x;
    bool take = false;
    double x = src.x, y = src.y, z = src.z, d = 0.0;
    if (i < n) {
      if (src.kind == kMmcSrcCircular && src.r > 0.0) {
        Rng rng;
        rng.init(seed, kMmcStreamSrc, (uint64_t)(id0 + i));
        const double rad = src.r * sqrt(rng.gen());
        const double phi = 6.283185307179586476925286766559 * rng.gen();
        const double cp = cos(phi), sp = sin(phi);
        x += rad * (cp * src.ax + sp * src.bx);
        y += rad * (cp * src.ay + sp * src.by);
        z += rad * (cp * src.az + sp * src.bz);
      }
      d = sphere_entry(r2, x, y, z, src.ux, src.uy, src.uz);
      if (d < 0.0) {
        if (!ignoremiss) mmc_fill(L, t, mmc_value(t, src.ux, src.uy, src.uz, 1.0, -1, src.ekin), 1.0, 0);
      } else {
        take = true;
      }
    }
    const unsigned long long pos = block_append<kMmcWaves>(count, take, s_app);
    if (take) {
      out.x[pos] = x + d * src.ux;
      out.y[pos] = y + d * src.uy;
      out.z[pos] = z + d * src.uz;
      out.d[3 * pos] = src.ux;
      out.d[3 * pos + 1] = src.uy;
      out.d[3 * pos + 2] = src.uz;
      out.E[pos] = src.ekin;
      out.w[pos] = 1.0;
      out.id[pos] = id0 + i;
      out.nscat[pos] = 0;
      out.ninel[pos] = 0;
    }
  }
  mmc_hist_flush(L, t);
}

// one step of the forced-collision engine for the n neutrons of `in`
__global__ void __launch_bounds__(kMmcThreads) k_mmc_step(MmcBasket in, int64_t n, const double* __restrict__ xs_s,
                                                  const double* __restrict__ xs_a, double macro, double r2,
                                                  MmcTally t, MmcEngine e, uint64_t seed, uint32_t step,
                                                  MmcBasket out, unsigned long long* __restrict__ count) {
  __shared__ MmcLdsHist L;
  __shared__ unsigned long long s_app[kMmcWaves + 1];
  mmc_hist_clear(L, t);
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i0 = (int64_t)blockIdx.x * blockDim.x; i0 < n; i0 += stride) {
    const int64_t i = i0 + threadIdx.x;
    bool take = false;
    double x = 0.0, y = 0.0, z = 0.0, ux = 0.0, uy = 0.0, uz = 0.0, E = 0.0, ws = 0.0, dsc = 0.0;
    int nscat = 0, ninel = 0;
    int64_t id = 0;
    if (i < n) {
      x = in.x[i]; y = in.y[i]; z = in.z[i];
      ux = in.d[3 * i]; uy = in.d[3 * i + 1]; uz = in.d[3 * i + 2];
      E = in.E[i];
      const double w = in.w[i];
      id = in.id[i];
      nscat = in.nscat[i];
      ninel = in.ninel[i];
      const double D = sphere_exit(r2, x, y, z, ux, uy, uz);
      const bool limited = e.nscatlimit >= 0 && nscat >= e.nscatlimit;
      const double ss = limited ? 0.0 : macro * xs_s[i];
      const double sa = xs_a ? macro * xs_a[i] : 0.0;
      const double ptr = ss > 0.0 ? exp(-ss * D) : 1.0;
      // the part reaching the surface unscattered
      const double wt = w * (sa > 0.0 ? exp(-sa * D) : 1.0) * ptr;
      mmc_fill(L, t, mmc_value(t, ux, uy, uz, wt, nscat, E), mmc_weight(t, wt), mmc_category(nscat, ninel));
      // the part scattering inside
      if (w > 0.0 && ss > 0.0) {
        Rng rng;
        rng.init(seed, kMmcStreamStep + step, (uint64_t)id);
        uint32_t r0, r1, r2_, r3;
        rng.gen4(r0, r1, r2_, r3);
        const double u_roul = Rng::to01(((uint64_t)r1 << 32) | r0);
        const double u_dist = Rng::to01(((uint64_t)r3 << 32) | r2_);
        double rf = 1.0;
        bool alive = true;
        if (nscat >= e.roul_nscat && w < e.roul_wthr) {
          if (u_roul > e.roul_psurv) alive = false;
          else rf = 1.0 / e.roul_psurv;
        }
        if (alive) {
          // exponential with rate ss truncated to [0, D]
          dsc = -log1p(u_dist * expm1(-ss * D)) / ss;
          dsc = dsc < 0.0 ? 0.0 : (dsc > D ? D : dsc);
          ws = w * rf * (sa > 0.0 ? exp(-sa * dsc) : 1.0) * (1.0 - ptr);
          take = ws > 0.0;
        }
      }
    }
    const unsigned long long pos = block_append<kMmcWaves>(count, take, s_app);
    if (take) {
      out.x[pos] = x + dsc * ux;
      out.y[pos] = y + dsc * uy;
      out.z[pos] = z + dsc * uz;
      out.d[3 * pos] = ux;
      out.d[3 * pos + 1] = uy;
      out.d[3 * pos + 2] = uz;
      out.E[pos] = E;
      out.w[pos] = ws;
      out.id[pos] = id;
      out.nscat[pos] = nscat;
      out.ninel[pos] = ninel;
    }
  }
  mmc_hist_flush(L, t);
}

// after the scattering call: new energy and direction, scattering counts
// (inelastic when the energy changed)
__global__ void k_mmc_post(MmcBasket b, int64_t n, const double* __restrict__ oE, const double* __restrict__ oD) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double E2 = oE[i];
  const bool elastic = E2 == b.E[i];
  b.nscat[i] += 1;
  if (!elastic) b.ninel[i] += 1;
  b.E[i] = E2;
  b.d[3 * i] = oD[3 * i];
  b.d[3 * i + 1] = oD[3 * i + 1];
  b.d[3 * i + 2] = oD[3 * i + 2];
}

// one workgroup per 64 consecutive histogram entries: wave w sums slices
// w, w + 16, ... (the lanes read consecutive entries of a slice), then the
// 16 wave sums are added in a fixed order, so the result does not depend on
// scheduling
constexpr int kMmcRedWaves = 16;
__global__ void __launch_bounds__(64 * kMmcRedWaves) k_mmc_reduce(const double* __restrict__ part, int m, int nblocks,
                                                                  double* __restrict__ hist) {
  __shared__ double s_sum[kMmcRedWaves][64];
  const int lane = (int)(threadIdx.x & 63), wv = (int)(threadIdx.x >> 6);
  const int i = (int)blockIdx.x * 64 + lane;
  double s = 0.0;
  if (i < m) {
#pragma unroll 8
    for (int b = wv; b < nblocks; b += kMmcRedWaves) s += part[(size_t)b * m + i];
  }
  s_sum[wv][lane] = s;
  __syncthreads();
  if (wv == 0 && i < m) {
    double tot = 0.0;
    for (int w = 0; w < kMmcRedWaves; ++w) tot += s_sum[w][lane];
    hist[i] = tot;
  }
}

unsigned mmc_grid(int64_t n) {
  const int64_t g = (n + kMmcThreads - 1) / kMmcThreads;
  return (unsigned)(g < 1 ? 1 : (g > kMmcBlocks ? kMmcBlocks : g));
}

}  // namespace

void launch_mmc_source(const MmcSource& src, double r2, const MmcTally& t, int ignoremiss, uint64_t seed, int64_t id0,
                       int64_t n, const MmcBasket& out, unsigned long long* count, hipStream_t st) {
  hipLaunchKernelGGL(k_mmc_source, dim3(mmc_grid(n)), dim3(kMmcThreads), 0, st, src, r2, t, ignoremiss, seed, id0, n,
                     out, count);
}
void launch_mmc_step(const MmcBasket& in, int64_t n, const double* xs_s, const double* xs_a, double macro, double r2,
                     const MmcTally& t, const MmcEngine& e, uint64_t seed, uint32_t step, const MmcBasket& out,
                     unsigned long long* count, hipStream_t st) {
  hipLaunchKernelGGL(k_mmc_step, dim3(mmc_grid(n)), dim3(kMmcThreads), 0, st, in, n, xs_s, xs_a, macro, r2, t, e, seed,
                     step, out, count);
}
void launch_mmc_reduce(const MmcTally& t, hipStream_t st) {
  const int m = kMmcCats * 2 * (t.nbins + 2);
  hipLaunchKernelGGL(k_mmc_reduce, dim3((unsigned)((m + 63) / 64)), dim3(64 * kMmcRedWaves), 0, st, t.part, m,
                     kMmcBlocks, t.hist);
}
void launch_mmc_post(const MmcBasket& b, int64_t n, const double* oE, const double* oD, hipStream_t st) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_mmc_post, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, b, n, oE, oD);
}

}  // namespace ncamd
