// Per-neutron kernels: cross sections and scatter sampling of a process
// composition (PowderBragg, incoherent elastic, S(alpha,beta)) with one lane
// per neutron and a Philox stream per neutron.
//
// Composition semantics: ncrystal_core/src/interfaces/NCProcImpl.cc:166-204
// (cumulative component xs), :353-362 (xs), :379-389 (component pick +
// sampling); batched layout: ncrystal_core/src/cinterface/ncrystal.cc:1107-1246.
#include "nc_kernels.hh"

namespace ncamd {

// Variant kCompPBOnly: a single PowderBragg component (the same sum as
// comp_xs: 0 + scale * xs); kCompFG: FreeGas / absorption components
template <bool FP32PB, int Variant>
__global__ void __launch_bounds__(256) k_xs_iso(CompView c, const double* __restrict__ ekin,
                                                int64_t n, int64_t repeat, double* __restrict__ out) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    const double E = ekin[i];
    double xs;
    if constexpr (Variant == kCompPBOnly) {
      xs = comp_outside(c, E) ? 0.0 : 0.0 + c.scale[0] * (pb_active(c.pb, E) ? pb_xs_sel<FP32PB>(c.pb, E) : 0.0);
    } else {
      xs = comp_xs<FP32PB, Variant == kCompFG>(c, E, nullptr);  // kCompFG: any composition with gas components
    }
    for (int64_t r = 0; r < repeat; ++r) out[r * n + i] = xs;
  }
}

template <int Variant>
__global__ void __launch_bounds__(256) k_sample_iso(CompView c, const double* __restrict__ ekin,
                                                    int64_t n, int64_t total, uint64_t seed,
                                                    uint32_t stream, double* __restrict__ oE,
                                                    double* __restrict__ oMu,
                                                    unsigned long long* __restrict__ nerr) {
  const int64_t gstride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < total; k += gstride) {
    const int64_t i = (total == n) ? k : (k % n);
    Rng rng;
    rng.init(seed, stream, (uint64_t)k);
    double Ef, mu;
    const int rc = comp_sample_iso<Variant>(c, rng, ekin[i], Ef, mu);
    if (rc) { Ef = -1.0; mu = -999.0; atomicAdd(nerr, 1ull); }
    oE[k] = Ef;
    oMu[k] = mu;
  }
}

template <int Variant>
__global__ void __launch_bounds__(256) k_sample_dirs(CompView c, const double* __restrict__ ekin,
                                                     const double* __restrict__ dirs, int64_t n,
                                                     uint64_t seed, uint32_t stream,
                                                     double* __restrict__ oE, double* __restrict__ oD,
                                                     unsigned long long* __restrict__ nerr) {
  const int64_t gstride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gstride) {
    Rng rng;
    rng.init(seed, stream, (uint64_t)k);
    double Ef, mu;
    const int rc = comp_sample_iso<Variant>(c, rng, ekin[k], Ef, mu);
    double ox = 0.0, oy = 0.0, oz = 0.0;
    if (rc) {
      Ef = -1.0;
      atomicAdd(nerr, 1ull);
    } else {
      rand_dir_given_mu(rng, mu, dirs[3 * k], dirs[3 * k + 1], dirs[3 * k + 2], ox, oy, oz);
    }
    oE[k] = Ef;
    oD[3 * k] = ox; oD[3 * k + 1] = oy; oD[3 * k + 2] = oz;
  }
}

// FreeGas (kKindFG) and 1/v absorption (kKindAbs) processes: cross section
// over the repeat layout, and FreeGas sampling (isotropic, or with a new
// direction from mu), one Philox stream per output slot.
__global__ void __launch_bounds__(256) k_simple_xs(int kind, FGView g, double abs_c, const double* __restrict__ ekin,
                                                   int64_t n, int64_t total, double* __restrict__ out) {
  const int64_t gstride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < total; k += gstride) {
    const double E = ekin[(total == n) ? k : (k % n)];
    out[k] = kind == kKindFG ? fg_xs(g, E) : absoov_xs(abs_c, E);
  }
}
template <bool Dirs>
__global__ void __launch_bounds__(256) k_fg_sample(FGView g, const double* __restrict__ ekin,
                                                   const double* __restrict__ dirs, int64_t n, int64_t total,
                                                   uint64_t seed, uint32_t stream, double* __restrict__ oE,
                                                   double* __restrict__ oX, unsigned long long* __restrict__ nerr) {
  const int64_t gstride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < total; k += gstride) {
    const int64_t i = (total == n) ? k : (k % n);
    Rng rng;
    rng.init(seed, stream, (uint64_t)k);
    double Ef, mu;
    const int rc = fg_sample_demu(g, rng, ekin[i], Ef, mu);
    if (rc) atomicAdd(nerr, 1ull);
    if constexpr (Dirs) {
      double ox = 0.0, oy = 0.0, oz = 0.0;
      if (rc) Ef = -1.0;
      else rand_dir_given_mu(rng, mu, dirs[3 * k], dirs[3 * k + 1], dirs[3 * k + 2], ox, oy, oz);
      oE[k] = Ef;
      oX[3 * k] = ox; oX[3 * k + 1] = oy; oX[3 * k + 2] = oz;
    } else {
      if (rc) { Ef = -1.0; mu = -999.0; }
      oE[k] = Ef;
      oX[k] = mu;
    }
  }
}

// Compositions with an S(alpha,beta) part, in stages so that no kernel
// carries every sampler: k_comp_pick samples PowderBragg / ElInc picks in
// place and queues the neutrons that pick S(alpha,beta); k_comp_sab<false>
// samples those below the table's Emax and queues the rest for
// k_comp_sab<true> (free-gas extender).  The S(alpha,beta) stages restart
// the neutron's Philox stream and replay the pick's draw (`skip`), so every
// neutron gets what one kernel doing all of it would give.
// outputs of one neutron: (E', mu), or (E', direction) for the oriented call
// (random azimuth about the incoming direction, NCProcImpl.cc:29-37)
template <bool Dirs>
NCD void comp_store(int64_t k, int rc, double Ef, double mu, Rng& rng, const double* __restrict__ dirs,
                    double* __restrict__ oE, double* __restrict__ o2, unsigned long long* __restrict__ nerr) {
  if constexpr (Dirs) {
    double ox = 0.0, oy = 0.0, oz = 0.0;
    if (rc) { Ef = -1.0; atomicAdd(nerr, 1ull); }
    else rand_dir_given_mu(rng, mu, dirs[3 * k], dirs[3 * k + 1], dirs[3 * k + 2], ox, oy, oz);
    oE[k] = Ef;
    o2[3 * k] = ox; o2[3 * k + 1] = oy; o2[3 * k + 2] = oz;
  } else {
    if (rc) { Ef = -1.0; mu = -999.0; atomicAdd(nerr, 1ull); }
    oE[k] = Ef;
    o2[k] = mu;
  }
}

template <bool Dirs, bool Gas>
__global__ void __launch_bounds__(256) k_comp_pick(CompView c, const double* __restrict__ ekin,
                                                   const double* __restrict__ dirs, int64_t n, int64_t total,
                                                   uint64_t seed, uint32_t stream, double* __restrict__ oE,
                                                   double* __restrict__ o2, unsigned long long* __restrict__ nerr,
                                                   int64_t* __restrict__ sq, unsigned long long* __restrict__ sqc,
                                                   const double* __restrict__ ext, int64_t* __restrict__ cq,
                                                   unsigned long long* __restrict__ cqc) {
  __shared__ unsigned long long s_app[5];
  const int64_t gstride = (int64_t)gridDim.x * blockDim.x;
  // workgroup-uniform loop: the queue appends synchronise the workgroup
  for (int64_t k0 = (int64_t)blockIdx.x * blockDim.x; k0 < total; k0 += gstride) {
    const int64_t k = k0 + threadIdx.x;
    const bool valid = k < total;
    Rng rng;
    double Ef = 0.0, mu = 0.0;
    int slot = 0, rc = kCompToStage + kCompToSCB + 1;  // neither queue
    if (valid) {
      const int64_t i = (total == n) ? k : (k % n);
      rng.init(seed, stream, (uint64_t)k);
      rc = comp_pick_sample<Gas>(c, rng, ekin[i], Ef, mu, slot, ext ? ext[k] : 0.0);
    }
    // queue appends, one atomic per workgroup and queue: one queue per
    // staged component, and the SCBragg queue
    const bool staged = valid && rc == kCompToStage;
    for (int s = 0; s < c.nstaged; ++s) {
      const bool want = staged && slot == s;
      const unsigned long long pos = block_append<4>(&sqc[s], want, s_app);
      if (want) sq[(int64_t)s * total + (int64_t)pos] = k;
    }
    const bool to_scb = valid && rc == kCompToSCB;
    if (cqc) {
      const unsigned long long pos = block_append<4>(cqc, to_scb, s_app);
      if (to_scb) cq[pos] = k;
    }
    if (valid && !staged && !to_scb) comp_store<Dirs>(k, rc, Ef, mu, rng, dirs, oE, o2, nerr);
  }
}

template <bool HighE, bool Dirs>
__global__ void __launch_bounds__(256) k_comp_sab(SABView s, const double* __restrict__ ekin,
                                                  const double* __restrict__ dirs, int64_t n, int64_t total,
                                                  uint64_t seed, uint32_t stream, int skip, double* __restrict__ oE,
                                                  double* __restrict__ o2, unsigned long long* __restrict__ nerr,
                                                  const int64_t* __restrict__ inq,
                                                  const unsigned long long* __restrict__ inqc,
                                                  int64_t* __restrict__ hq, unsigned long long* __restrict__ hqc) {
  const int64_t gstride = (int64_t)gridDim.x * blockDim.x;
  const int64_t nq = (int64_t)*inqc;
  const double emax = s.egrid[s.ne - 1];
  for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < nq; q += gstride) {
    const int64_t k = inq[q];
    const double E = ekin[(total == n) ? k : (k % n)];
    if constexpr (!HighE) {
      if (!(E < emax)) {
        hq[atomicAdd(hqc, 1ull)] = k;
        continue;
      }
    }
    Rng rng;
    rng.init(seed, stream, (uint64_t)k);
    if (skip) rng.gen();
    double Ef, mu;
    const int rc = sab_sample_iso<HighE>(s, rng, E, Ef, mu);
    comp_store<Dirs>(k, rc, Ef, mu, rng, dirs, oE, o2, nerr);
  }
}

// stage kernel of a FreeGas component beside S(alpha,beta) components: the
// queued neutrons restart their stream, replay the pick's draw and sample
// the free-gas collision
template <bool Dirs>
__global__ void __launch_bounds__(256) k_comp_fg(FGView g, const double* __restrict__ ekin,
                                                 const double* __restrict__ dirs, int64_t n, int64_t total,
                                                 uint64_t seed, uint32_t stream, int skip, double* __restrict__ oE,
                                                 double* __restrict__ o2, unsigned long long* __restrict__ nerr,
                                                 const int64_t* __restrict__ inq,
                                                 const unsigned long long* __restrict__ inqc) {
  const int64_t gstride = (int64_t)gridDim.x * blockDim.x;
  const int64_t nq = (int64_t)*inqc;
  for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < nq; q += gstride) {
    const int64_t k = inq[q];
    Rng rng;
    rng.init(seed, stream, (uint64_t)k);
    if (skip) rng.gen();
    double Ef, mu;
    const int rc = fg_sample_demu(g, rng, ekin[(total == n) ? k : (k % n)], Ef, mu);
    comp_store<Dirs>(k, rc, Ef, mu, rng, dirs, oE, o2, nerr);
  }
}

__global__ void __launch_bounds__(256) k_sab_ab(SABView s, const double* __restrict__ ekin, int64_t n,
                                                uint64_t seed, uint32_t stream, double* __restrict__ oa,
                                                double* __restrict__ ob,
                                                unsigned long long* __restrict__ nerr) {
  const int64_t gstride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gstride) {
    Rng rng;
    rng.init(seed, stream, (uint64_t)k);
    double a, b;
    const int rc = sab_sample_ab(s, ekin[k], rng, a, b);
    if (rc) { a = -1.0; b = 0.0; atomicAdd(nerr, 1ull); }
    oa[k] = a;
    ob[k] = b;
  }
}

// S(alpha,beta)-only materials: the free-gas extender (E >= Emax) is rare and
// register-hungry, so the main kernel leaves those neutrons in a queue that a
// second kernel drains.  Both draw from the neutron's own Philox stream from
// its first draw, so the split does not change any result.
__global__ void __launch_bounds__(256) k_sab_sample_iso_main(SABView s, const double* __restrict__ ekin,
                                                             int64_t n, int64_t total, uint64_t seed,
                                                             uint32_t stream, double* __restrict__ oE,
                                                             double* __restrict__ oMu,
                                                             unsigned long long* __restrict__ nerr,
                                                             int64_t* __restrict__ queue,
                                                             unsigned long long* __restrict__ qcount,
                                                             int domain_check) {
  const int64_t gstride = (int64_t)gridDim.x * blockDim.x;
  const double emax = s.egrid[s.ne - 1];
  for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < total; k += gstride) {
    const int64_t i = (total == n) ? k : (k % n);
    const double E = ekin[i];
    double Ef, mu;
    if (domain_check && !(E >= 0.0 && E <= __builtin_huge_val())) {
      // a composition outside its energy domain does not scatter
      Ef = E; mu = 1.0;
    } else if (!(E < emax)) {
      // the free-gas kernel, and the sampler's own handling of non-finite E
      const unsigned long long slot = atomicAdd(qcount, 1ull);
      queue[slot] = k;
      continue;
    } else {
      Rng rng;
      rng.init(seed, stream, (uint64_t)k);
      const int rc = sab_sample_iso<false>(s, rng, E, Ef, mu);
      if (rc) { Ef = -1.0; mu = -999.0; atomicAdd(nerr, 1ull); }
    }
    oE[k] = Ef;
    oMu[k] = mu;
  }
}

__global__ void __launch_bounds__(256) k_sab_sample_iso_highe(SABView s, const double* __restrict__ ekin,
                                                              int64_t n, int64_t total, uint64_t seed,
                                                              uint32_t stream, double* __restrict__ oE,
                                                              double* __restrict__ oMu,
                                                              unsigned long long* __restrict__ nerr,
                                                              const int64_t* __restrict__ queue,
                                                              const unsigned long long* __restrict__ qcount) {
  const int64_t gstride = (int64_t)gridDim.x * blockDim.x;
  const int64_t nq = (int64_t)*qcount;
  for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < nq; q += gstride) {
    const int64_t k = queue[q];
    const int64_t i = (total == n) ? k : (k % n);
    Rng rng;
    rng.init(seed, stream, (uint64_t)k);
    double Ef, mu;
    const int rc = sab_sample_iso<true>(s, rng, ekin[i], Ef, mu);
    if (rc) { Ef = -1.0; mu = -999.0; atomicAdd(nerr, 1ull); }
    oE[k] = Ef;
    oMu[k] = mu;
  }
}

// search-acceleration tables of an S(alpha,beta) process (built once)
__global__ void k_sab_cdf_guides(SABView s, unsigned short* __restrict__ guide) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t per = kCdfGuide + 1;
  if (t >= (int64_t)s.ne * per) return;
  const int ie = (int)(t / per), g = (int)(t % per);
  const int4 hdr = s.shdr[ie];
  unsigned short r = 0;
  if (hdr.x == 1) {
    const double* cdf = s.scdf + (size_t)ie * (s.nb + 1);
    r = (unsigned short)lower_bound_d(cdf, 0, hdr.z, (double)g / (double)kCdfGuide);
  }
  guide[t] = r;
}
__global__ void k_sab_row_guides(SABView s, unsigned short* __restrict__ guide, double* __restrict__ scale) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t per = kRowGuide + 1;
  if (t >= (int64_t)s.nb * per) return;
  const int ib = (int)(t / per), g = (int)(t % per);
  const double* row = s.cumul + (size_t)ib * s.na;
  const double rmax = row[s.na - 1];
  const double rs = rmax > 0.0 ? (double)kRowGuide / rmax : 0.0;
  if (g == 0) scale[ib] = rs;
  // the last bucket is closed by na itself so that values at or above the
  // row maximum always have their answer inside the bracket
  const int r = (g == kRowGuide) ? s.na : upper_bound_d(row, 0, s.na, (double)g / rs);
  guide[t] = rs > 0.0 ? (unsigned short)r : (unsigned short)0;
}

static inline unsigned grid_for(int64_t n, int bs) {
  int64_t g = (n + bs - 1) / bs;
  // enough waves for the 256 CUs, bounded so a grid-stride loop covers the rest
  const int64_t gmax = 256 * 64;
  if (g > gmax) g = gmax;
  if (g < 1) g = 1;
  return (unsigned)g;
}

__global__ void k_zero_u64(unsigned long long* p) { *p = 0ull; }

__global__ void k_zero_u64_n(unsigned long long* p, int n) {
  if ((int)threadIdx.x < n) p[threadIdx.x] = 0ull;
}

// the stage launches of a composition with an S(alpha,beta) part, after the
// pick kernel filled one queue per staged component (queues + s * total,
// counter qcount[s]); the high-energy queue after them (counter
// qcount[kMaxStaged]) is reused by each S(alpha,beta) component in turn
template <bool Dirs>
void launch_stages(const CompView& c, const double* ekin, const double* dirs, int64_t n, int64_t total,
                   uint64_t seed, uint32_t stream, double* oE, double* o2, unsigned long long* nerr,
                   int64_t* queues, unsigned long long* qcount, hipStream_t st) {
  const int skip = c.n > 1 ? 1 : 0;
  int64_t* hq = queues + (int64_t)c.nstaged * total;
  unsigned long long* hqc = qcount + kMaxStaged;
  const dim3 g(grid_for(total, 256)), b(256);
  for (int sl = 0; sl < c.nstaged; ++sl) {
    const int comp = c.stage_comp[sl];
    int64_t* q = queues + (int64_t)sl * total;
    if (c.kind[comp] == kKindSAB) {
      hipLaunchKernelGGL(k_zero_u64, dim3(1), dim3(1), 0, st, hqc);
      hipLaunchKernelGGL((k_comp_sab<false, Dirs>), g, b, 0, st, c.sab[sl], ekin, dirs, n, total, seed, stream, skip,
                         oE, o2, nerr, q, qcount + sl, hq, hqc);
      hipLaunchKernelGGL((k_comp_sab<true, Dirs>), g, b, 0, st, c.sab[sl], ekin, dirs, n, total, seed, stream, skip,
                         oE, o2, nerr, hq, hqc, nullptr, nullptr);
    } else {
      hipLaunchKernelGGL(k_comp_fg<Dirs>, g, b, 0, st, c.fg[comp], ekin, dirs, n, total, seed, stream, skip, oE, o2,
                         nerr, q, qcount + sl);
    }
  }
}


// synthetic neutron energies, log-uniform, from a dedicated Philox stream
__global__ void __launch_bounds__(256) k_fill_loguniform(double* __restrict__ out, int64_t n, double lemin,
                                                         double lemax, uint64_t seed) {
  const int64_t gstride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gstride) {
    Rng rng;
    rng.init(seed, 0xFFFFFFFFu, (uint64_t)k);
    const double u = rng.gen();
    out[k] = exp(lemin + u * (lemax - lemin));
  }
}


void launch_simple_xs(int kind, const FGView& g, double abs_c, const double* ekin, int64_t n, int64_t repeat,
                      double* out, hipStream_t st) {
  const int64_t total = n * repeat;
  if (total <= 0) return;
  hipLaunchKernelGGL(k_simple_xs, dim3(grid_for(total, 256)), dim3(256), 0, st, kind, g, abs_c, ekin, n, total, out);
}
void launch_fg_sample(const FGView& g, const double* ekin, const double* dirs, int64_t n, int64_t repeat,
                      uint64_t seed, uint32_t stream, double* oE, double* oX, unsigned long long* nerr,
                      hipStream_t st) {
  const int64_t total = n * repeat;
  if (total <= 0) return;
  if (dirs)
    hipLaunchKernelGGL(k_fg_sample<true>, dim3(grid_for(total, 256)), dim3(256), 0, st, g, ekin, dirs, n, total, seed,
                       stream, oE, oX, nerr);
  else
    hipLaunchKernelGGL(k_fg_sample<false>, dim3(grid_for(total, 256)), dim3(256), 0, st, g, ekin, dirs, n, total,
                       seed, stream, oE, oX, nerr);
}


void launch_xs_iso(const CompView& c, const double* ekin, int64_t n, int64_t repeat, double* out,
                   bool fp32, hipStream_t st) {
  if (n <= 0 || repeat <= 0) return;
  const dim3 g(grid_for(n, 256)), b(256);
  const int v = comp_variant(c);
  if (v == kCompPBOnly) {
    if (fp32) hipLaunchKernelGGL((k_xs_iso<true, kCompPBOnly>), g, b, 0, st, c, ekin, n, repeat, out);
    else hipLaunchKernelGGL((k_xs_iso<false, kCompPBOnly>), g, b, 0, st, c, ekin, n, repeat, out);
  } else if (v == kCompFG || comp_has_gas(c)) {
    if (fp32) hipLaunchKernelGGL((k_xs_iso<true, kCompFG>), g, b, 0, st, c, ekin, n, repeat, out);
    else hipLaunchKernelGGL((k_xs_iso<false, kCompFG>), g, b, 0, st, c, ekin, n, repeat, out);
  } else {
    if (fp32) hipLaunchKernelGGL((k_xs_iso<true, kCompNoSAB>), g, b, 0, st, c, ekin, n, repeat, out);
    else hipLaunchKernelGGL((k_xs_iso<false, kCompNoSAB>), g, b, 0, st, c, ekin, n, repeat, out);
  }
}
void launch_sample_iso(const CompView& c, const double* ekin, int64_t n, int64_t repeat, uint64_t seed,
                       uint32_t stream, double* oE, double* oMu, unsigned long long* nerr, int64_t* queues,
                       unsigned long long* qcount, hipStream_t st) {
  const int64_t total = n * repeat;
  if (total <= 0) return;
  switch (comp_variant(c)) {
    case kCompPBOnly:
      hipLaunchKernelGGL(k_sample_iso<kCompPBOnly>, dim3(grid_for(total, 256)), dim3(256), 0, st, c, ekin, n, total,
                         seed, stream, oE, oMu, nerr);
      break;
    case kCompNoSAB:
      hipLaunchKernelGGL(k_sample_iso<kCompNoSAB>, dim3(grid_for(total, 256)), dim3(256), 0, st, c, ekin, n, total,
                         seed, stream, oE, oMu, nerr);
      break;
    case kCompFG:
      hipLaunchKernelGGL(k_sample_iso<kCompFG>, dim3(grid_for(total, 256)), dim3(256), 0, st, c, ekin, n, total,
                         seed, stream, oE, oMu, nerr);
      break;
    default: {
      hipLaunchKernelGGL(k_zero_u64_n, dim3(1), dim3(64), 0, st, qcount, kMaxStaged + 1);
      const dim3 g(grid_for(total, 256)), b(256);
      if (comp_has_gas(c))
        hipLaunchKernelGGL((k_comp_pick<false, true>), g, b, 0, st, c, ekin, (const double*)nullptr, n, total, seed,
                           stream, oE, oMu, nerr, queues, qcount, (const double*)nullptr, (int64_t*)nullptr,
                           (unsigned long long*)nullptr);
      else
        hipLaunchKernelGGL((k_comp_pick<false, false>), g, b, 0, st, c, ekin, (const double*)nullptr, n, total,
                           seed, stream, oE, oMu, nerr, queues, qcount, (const double*)nullptr, (int64_t*)nullptr,
                           (unsigned long long*)nullptr);
      launch_stages<false>(c, ekin, nullptr, n, total, seed, stream, oE, oMu, nerr, queues, qcount, st);
    }
  }
}
__global__ void __launch_bounds__(256) k_xs_ext(CompView c, const double* __restrict__ ekin,
                                                const double* __restrict__ ext, int64_t n, double* __restrict__ out) {
  const int64_t gstride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gstride)
    out[k] = comp_xs<false, true>(c, ekin[k], nullptr, ext[k]);  // FreeGas components may sit beside SCBragg
}
void launch_xs_ext(const CompView& c, const double* ekin, const double* ext, int64_t n, double* out,
                   hipStream_t st) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_xs_ext, dim3(grid_for(n, 256)), dim3(256), 0, st, c, ekin, ext, n, out);
}

void launch_sample_dirs(const CompView& c, const double* ekin, const double* dirs, int64_t n, uint64_t seed,
                        uint32_t stream, double* oE, double* oD, unsigned long long* nerr, int64_t* queues,
                        unsigned long long* qcount, hipStream_t st, const double* ext, int64_t* cq,
                        unsigned long long* cqc) {
  if (cqc) hipLaunchKernelGGL(k_zero_u64, dim3(1), dim3(1), 0, st, cqc);
  if (n <= 0) return;
  switch (comp_variant(c)) {
    case kCompPBOnly:
      hipLaunchKernelGGL(k_sample_dirs<kCompPBOnly>, dim3(grid_for(n, 256)), dim3(256), 0, st, c, ekin, dirs, n,
                         seed, stream, oE, oD, nerr);
      break;
    case kCompNoSAB:
      hipLaunchKernelGGL(k_sample_dirs<kCompNoSAB>, dim3(grid_for(n, 256)), dim3(256), 0, st, c, ekin, dirs, n,
                         seed, stream, oE, oD, nerr);
      break;
    case kCompFG:
      hipLaunchKernelGGL(k_sample_dirs<kCompFG>, dim3(grid_for(n, 256)), dim3(256), 0, st, c, ekin, dirs, n,
                         seed, stream, oE, oD, nerr);
      break;
    default: {
      hipLaunchKernelGGL(k_zero_u64_n, dim3(1), dim3(64), 0, st, qcount, kMaxStaged + 1);
      const dim3 g(grid_for(n, 256)), b(256);
      if (comp_has_gas(c))
        hipLaunchKernelGGL((k_comp_pick<true, true>), g, b, 0, st, c, ekin, dirs, n, n, seed, stream, oE, oD, nerr,
                           queues, qcount, ext, cq, cqc);
      else
        hipLaunchKernelGGL((k_comp_pick<true, false>), g, b, 0, st, c, ekin, dirs, n, n, seed, stream, oE, oD, nerr,
                           queues, qcount, ext, cq, cqc);
      launch_stages<true>(c, ekin, dirs, n, n, seed, stream, oE, oD, nerr, queues, qcount, st);
    }
  }
}
void launch_sab_ab(const SABView& s, const double* ekin, int64_t n, uint64_t seed, uint32_t stream, double* oa,
                   double* ob, unsigned long long* nerr, hipStream_t st) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_sab_ab, dim3(grid_for(n, 256)), dim3(256), 0, st, s, ekin, n, seed, stream, oa, ob, nerr);
}

void launch_sab_sample_iso(const SABView& s, const double* ekin, int64_t n, int64_t repeat, uint64_t seed,
                           uint32_t stream, double* oE, double* oMu, unsigned long long* nerr, int64_t* queue,
                           unsigned long long* qcount, int domain_check, hipStream_t st) {
  const int64_t total = n * repeat;
  if (total <= 0) return;
  // kernel instead of a memset node, so that callers may capture this in a graph
  hipLaunchKernelGGL(k_zero_u64, dim3(1), dim3(1), 0, st, qcount);
  hipLaunchKernelGGL(k_sab_sample_iso_main, dim3(grid_for(total, 256)), dim3(256), 0, st, s, ekin, n, total, seed,
                     stream, oE, oMu, nerr, queue, qcount, domain_check);
  hipLaunchKernelGGL(k_sab_sample_iso_highe, dim3(grid_for(total, 256)), dim3(256), 0, st, s, ekin, n, total, seed,
                     stream, oE, oMu, nerr, queue, qcount);
}
void launch_zero_u64(unsigned long long* p, hipStream_t st) {
  hipLaunchKernelGGL(k_zero_u64, dim3(1), dim3(1), 0, st, p);
}
void launch_sab_queue(const SABView& s, const double* ekin, int64_t n, int64_t total, uint64_t seed, uint32_t stream,
                      double* oE, double* oMu, unsigned long long* nerr, const int64_t* queue,
                      const unsigned long long* qcount, hipStream_t st) {
  hipLaunchKernelGGL(k_sab_sample_iso_highe, dim3(grid_for(total, 256)), dim3(256), 0, st, s, ekin, n, total, seed,
                     stream, oE, oMu, nerr, queue, qcount);
}
void launch_sab_guides(const SABView& s, unsigned short* cdf_guide, unsigned short* row_guide, double* row_scale,
                       hipStream_t st) {
  const int64_t n1 = (int64_t)s.ne * (kCdfGuide + 1);
  hipLaunchKernelGGL(k_sab_cdf_guides, dim3((unsigned)((n1 + 255) / 256)), dim3(256), 0, st, s, cdf_guide);
  const int64_t n2 = (int64_t)s.nb * (kRowGuide + 1);
  hipLaunchKernelGGL(k_sab_row_guides, dim3((unsigned)((n2 + 255) / 256)), dim3(256), 0, st, s, row_guide,
                     row_scale);
}
void launch_fill_loguniform(double* out, int64_t n, double lemin, double lemax, uint64_t seed, hipStream_t st) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_fill_loguniform, dim3(grid_for(n, 256)), dim3(256), 0, st, out, n, lemin, lemax, seed);
}

}  // namespace ncamd
