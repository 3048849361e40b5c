// Per-neutron kernels of the mosaic single-crystal Bragg process: oriented
// cross sections and scatter sampling, one lane per neutron.
//
// Scheduling: the cost of a neutron is the number of demi-normals of the
// families with 2d > lambda, which varies as lambda^-3 across a batch.  A
// counting sort by that number (k_scb_keys / k_scb_offsets / k_scb_scatter)
// gives every wavefront neutrons of nearly equal cost; the kernels then take
// neutron perm[k] in lane k.  Results are written at the neutron's own index
// and every neutron draws from its own Philox stream, so the order changes
// no result.
//
// Each kernel pair is a main kernel without the Romberg circle integral and
// a queue kernel that redoes, with it, the neutrons whose scan met a
// contribution outside the spline's validity (backscattering within the
// mosaic spread).  Cross sections involve no random numbers and sampling
// draws from the neutron's own Philox stream from its first draw, so the
// split does not change any result either.
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
                    unsigned long long* __restrict__ qcount, uint16_t* __restrict__ list) {
  ScbLane L;
  L.wl = 1.0;
  L.u = V3{0.0, 0.0, 1.0};
  L.jlim = 0;
  bool can = false;
  if (valid) can = scb_lane_setup(v, ekin[i], dirs[3 * i], dirs[3 * i + 1], dirs[3 * i + 2], L);
  int nlist;
  bool complete;
  scb_walk<false, AllowSlow>(v, L, 0, 0.0, list, nlist, complete);
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
                        unsigned long long* __restrict__ qcount, uint16_t* __restrict__ list,
                        const int64_t* __restrict__ idmap, int skip) {
  double E = 0.0;
  V3 din{0.0, 0.0, 1.0};
  ScbLane L;
  L.wl = 1.0;
  L.u = din;
  L.jlim = 0;
  bool can = false;
  if (valid) {
    E = ekin[i];
    din = V3{dirs[3 * i], dirs[3 * i + 1], dirs[3 * i + 2]};
    can = scb_lane_setup(v, E, din.x, din.y, din.z, L);
  }
  // pass 1: number of contributions and their cumulative total
  int nlist;
  bool complete;
  scb_walk<false, AllowSlow>(v, L, 0, 0.0, list, nlist, complete);
  bool queued = false;
  if (valid && !AllowSlow && L.slow) {
    const unsigned long long slot = atomicAdd(qcount, 1ull);
    queue[slot] = i;
    queued = true;
  }
  const bool scatter = valid && can && !queued && L.count > 0 && L.total > 0.0;
  const int64_t o = idmap && valid ? idmap[i] : i;  // output slot and Philox index
  Rng rng;
  rng.init(seed, stream, (uint64_t)o);
  for (int q = 0; q < skip; ++q) (void)rng.gen();
  double target = 0.0;
  if (scatter && L.count > 1) target = L.total * rng.gen();
  // pass 2: the contribution whose cumulative value passes the target, from
  // the candidate list of pass 1 when it held them all (wave-uniform), else
  // by a new scan
  ScbLane S = L;
  if (!scatter) S.jlim = 0;
  if (complete) {
    scb_select_list<AllowSlow>(v, S, L.count, target, list, nlist);
  } else {
    int nl2;
    bool c2;
    scb_walk<true, AllowSlow>(v, S, L.count, target, list, nl2, c2);
  }
  if (!valid || queued) return;
  V3 d = din;
  if (scatter && S.found) d = gm_genscat(v, rng, S.sel_normal, S.sel_inv2d_r, L.wl, L.u);
  oE[o] = E;
  oD[3 * o] = d.x;
  oD[3 * o + 1] = d.y;
  oD[3 * o + 2] = d.z;
}

}  // namespace

// ---- counting sort of the batch by scan length
__global__ void __launch_bounds__(256) k_scb_keys(SCBView v, const double* __restrict__ ekin,
                                                  const double* __restrict__ dirs, int64_t n,
                                                  unsigned* __restrict__ keys, unsigned* __restrict__ hist) {
  __shared__ unsigned lh[kScbBuckets];
  for (int b = threadIdx.x; b < kScbBuckets; b += blockDim.x) lh[b] = 0u;
  __syncthreads();
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    ScbLane L;
    L.jlim = 0;
    scb_lane_setup(v, ekin[i], dirs[3 * i], dirs[3 * i + 1], dirs[3 * i + 2], L);
    const unsigned b = (unsigned)(((int64_t)L.jlim * kScbBuckets) / ((int64_t)v.nnormals + 1));
    keys[i] = b;
    atomicAdd(&lh[b], 1u);
  }
  __syncthreads();
  for (int b = threadIdx.x; b < kScbBuckets; b += blockDim.x)
    if (lh[b]) atomicAdd(&hist[b], lh[b]);
}

// exclusive prefix sum of the bucket counts (one block); cursors start at
// the bucket offsets, the histogram is cleared for the next call
__global__ void __launch_bounds__(kScbBuckets) k_scb_offsets(unsigned* __restrict__ hist,
                                                              unsigned long long* __restrict__ cursor) {
  __shared__ unsigned long long s[kScbBuckets];
  const int t = threadIdx.x;
  s[t] = hist[t];
  __syncthreads();
  for (int o = 1; o < kScbBuckets; o <<= 1) {
    const unsigned long long x = t >= o ? s[t - o] : 0ull;
    __syncthreads();
    s[t] += x;
    __syncthreads();
  }
  cursor[t] = s[t] - hist[t];
  hist[t] = 0u;
}

// each block takes kScbScatterItems x 256 consecutive neutrons, ranks them
// inside the block with LDS atomics and reserves one range per bucket with
// a single global atomic, so that global atomics do not pile up on the few
// buckets a batch usually fills
constexpr int kScbScatterItems = 16;
__global__ void __launch_bounds__(256) k_scb_scatter(const unsigned* __restrict__ keys, int64_t n,
                                                     unsigned long long* __restrict__ cursor,
                                                     int64_t* __restrict__ perm) {
  __shared__ unsigned cnt[kScbBuckets];
  __shared__ unsigned long long base[kScbBuckets];
  const int64_t chunk = (int64_t)blockDim.x * kScbScatterItems;
  for (int64_t c0 = (int64_t)blockIdx.x * chunk; c0 < n; c0 += (int64_t)gridDim.x * chunk) {
    for (int b = threadIdx.x; b < kScbBuckets; b += blockDim.x) cnt[b] = 0u;
    __syncthreads();
    unsigned key[kScbScatterItems], rank[kScbScatterItems];
#pragma unroll
    for (int j = 0; j < kScbScatterItems; ++j) {
      const int64_t i = c0 + (int64_t)j * blockDim.x + threadIdx.x;
      key[j] = i < n ? keys[i] : 0u;
      rank[j] = i < n ? atomicAdd(&cnt[key[j]], 1u) : 0u;
    }
    __syncthreads();
    for (int b = threadIdx.x; b < kScbBuckets; b += blockDim.x)
      base[b] = cnt[b] ? atomicAdd(&cursor[b], (unsigned long long)cnt[b]) : 0ull;
    __syncthreads();
#pragma unroll
    for (int j = 0; j < kScbScatterItems; ++j) {
      const int64_t i = c0 + (int64_t)j * blockDim.x + threadIdx.x;
      if (i < n) perm[base[key[j]] + rank[j]] = i;
    }
    __syncthreads();
  }
}

__global__ void k_scb_clear(unsigned* __restrict__ hist, unsigned long long* __restrict__ qcount) {
  for (int b = threadIdx.x; b < kScbBuckets; b += blockDim.x) hist[b] = 0u;
  if (threadIdx.x == 0) *qcount = 0ull;
}

// ---- main and queue kernels.  Grid-stride loops step over whole blocks so
// that every lane of a wavefront reaches the wave-uniform scan together.
__global__ void __launch_bounds__(256) k_scb_xs_main(SCBView v, const double* __restrict__ ekin,
                                                     const double* __restrict__ dirs, int64_t n,
                                                     const int64_t* __restrict__ perm, double* __restrict__ out,
                                                     int64_t* __restrict__ queue,
                                                     unsigned long long* __restrict__ qcount) {
  __shared__ uint16_t lists[4 * kScbListWave];
  uint16_t* list = lists + (threadIdx.x >> 6) * kScbListWave;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t base = (int64_t)blockIdx.x * blockDim.x; base < n; base += stride) {
    const int64_t k = base + threadIdx.x;
    const bool valid = k < n;
    scb_xs_one<false>(v, valid, valid ? perm[k] : 0, ekin, dirs, out, queue, qcount, list);
  }
}
__global__ void __launch_bounds__(256) k_scb_xs_slow(SCBView v, const double* __restrict__ ekin,
                                                     const double* __restrict__ dirs, double* __restrict__ out,
                                                     const int64_t* __restrict__ queue,
                                                     const unsigned long long* __restrict__ qcount) {
  __shared__ uint16_t lists[4 * kScbListWave];
  uint16_t* list = lists + (threadIdx.x >> 6) * kScbListWave;
  const int64_t nq = (int64_t)*qcount;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t base = (int64_t)blockIdx.x * blockDim.x; base < nq; base += stride) {
    const int64_t q = base + threadIdx.x;
    const bool valid = q < nq;
    scb_xs_one<true>(v, valid, valid ? queue[q] : 0, ekin, dirs, out, nullptr, nullptr, list);
  }
}
// Sampling in three kernels over the sorted batch (fewer registers each
// than one kernel doing it all): the first counts the contributions and
// their total for sorted position k (cnt[k], tot[k]; 0 for a neutron that
// does not scatter, whose output it writes, or that goes to the Romberg
// queue); the second draws the target and repeats the scan until the
// selected contribution (sel[k]: normal of the reflecting side, rounded
// 1/(2d); 0 if none); the third generates the scattered direction, with the
// neutron's Philox stream continued after the target draw.
__global__ void __launch_bounds__(256) k_scb_sample_count(SCBView v, const double* __restrict__ ekin,
                                                          const double* __restrict__ dirs, int64_t n,
                                                          const int64_t* __restrict__ perm, double* __restrict__ oE,
                                                          double* __restrict__ oD, int* __restrict__ cnt,
                                                          double* __restrict__ tot, int64_t* __restrict__ queue,
                                                          unsigned long long* __restrict__ qcount,
                                                          const int64_t* __restrict__ idmap) {
  __shared__ uint16_t lists[4 * kScbListWave];
  uint16_t* list = lists + (threadIdx.x >> 6) * kScbListWave;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t base = (int64_t)blockIdx.x * blockDim.x; base < n; base += stride) {
    const int64_t k = base + threadIdx.x;
    const bool valid = k < n;
    const int64_t i = valid ? perm[k] : 0;
    double E = 0.0;
    V3 din{0.0, 0.0, 1.0};
    ScbLane L;
    L.wl = 1.0;
    L.u = din;
    L.jlim = 0;
    bool can = false;
    if (valid) {
      E = ekin[i];
      din = V3{dirs[3 * i], dirs[3 * i + 1], dirs[3 * i + 2]};
      can = scb_lane_setup(v, E, din.x, din.y, din.z, L);
    }
    int nlist;
    bool complete;
    scb_walk<false, false>(v, L, 0, 0.0, list, nlist, complete);
    if (!valid) continue;
    if (L.slow) {
      const unsigned long long slot = atomicAdd(qcount, 1ull);
      queue[slot] = i;
      cnt[k] = 0;
      continue;
    }
    const bool scatter = can && L.count > 0 && L.total > 0.0;
    cnt[k] = scatter ? L.count : 0;
    tot[k] = L.total;
    if (!scatter) {
      const int64_t o = idmap ? idmap[i] : i;
      oE[o] = E;
      oD[3 * o] = din.x;
      oD[3 * o + 1] = din.y;
      oD[3 * o + 2] = din.z;
    }
  }
}
__global__ void __launch_bounds__(256) k_scb_sample_pick(SCBView v, const double* __restrict__ ekin,
                                                         const double* __restrict__ dirs, int64_t n,
                                                         const int64_t* __restrict__ perm, uint64_t seed,
                                                         uint32_t stream, const int* __restrict__ cnt,
                                                         const double* __restrict__ tot, double4* __restrict__ sel,
                                                         const int64_t* __restrict__ idmap, int skip) {
  __shared__ uint16_t lists[4 * kScbListWave];
  uint16_t* list = lists + (threadIdx.x >> 6) * kScbListWave;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t base = (int64_t)blockIdx.x * blockDim.x; base < n; base += stride) {
    const int64_t k = base + threadIdx.x;
    const int c = k < n ? cnt[k] : 0;
    const int64_t i = c > 0 ? perm[k] : 0;
    double E = 0.0;
    V3 din{0.0, 0.0, 1.0};
    ScbLane L;
    L.wl = 1.0;
    L.u = din;
    L.jlim = 0;
    if (c > 0) {
      E = ekin[i];
      din = V3{dirs[3 * i], dirs[3 * i + 1], dirs[3 * i + 2]};
      scb_lane_setup(v, E, din.x, din.y, din.z, L);
    }
    Rng rng;
    rng.init(seed, stream, (uint64_t)(idmap && c > 0 ? idmap[i] : i));
    for (int q = 0; q < skip; ++q) (void)rng.gen();
    double target = 0.0;
    if (c > 1) target = tot[k] * rng.gen();
    int nl;
    bool complete;
    scb_walk<true, false>(v, L, c, target, list, nl, complete);
    if (c <= 0) continue;
    sel[k] = L.found ? make_double4(L.sel_normal.x, L.sel_normal.y, L.sel_normal.z, L.sel_inv2d_r)
                     : make_double4(0.0, 0.0, 0.0, 0.0);
  }
}
__global__ void __launch_bounds__(256) k_scb_sample_scatter(SCBView v, const double* __restrict__ ekin,
                                                            const double* __restrict__ dirs, int64_t n,
                                                            const int64_t* __restrict__ perm, uint64_t seed,
                                                            uint32_t stream, double* __restrict__ oE,
                                                            double* __restrict__ oD, const int* __restrict__ cnt,
                                                            const double4* __restrict__ sel,
                                                            const int64_t* __restrict__ idmap, int skip) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += stride) {
    const int c = cnt[k];
    if (c <= 0) continue;
    const int64_t i = perm[k];
    const double E = ekin[i];
    const V3 din{dirs[3 * i], dirs[3 * i + 1], dirs[3 * i + 2]};
    ScbLane L;
    scb_lane_setup(v, E, din.x, din.y, din.z, L);
    const int64_t oi = idmap ? idmap[i] : i;
    Rng rng;
    rng.init(seed, stream, (uint64_t)oi);
    for (int q = 0; q < skip; ++q) (void)rng.gen();
    if (c > 1) (void)rng.gen();  // the target draw of the selection
    const double4 s4 = sel[k];
    V3 o = din;
    if (s4.w != 0.0) o = gm_genscat(v, rng, V3{s4.x, s4.y, s4.z}, s4.w, L.wl, L.u);
    oE[oi] = E;
    oD[3 * oi] = o.x;
    oD[3 * oi + 1] = o.y;
    oD[3 * oi + 2] = o.z;
  }
}
__global__ void __launch_bounds__(256) k_scb_sample_slow(SCBView v, const double* __restrict__ ekin,
                                                         const double* __restrict__ dirs, uint64_t seed,
                                                         uint32_t stream, double* __restrict__ oE,
                                                         double* __restrict__ oD, const int64_t* __restrict__ queue,
                                                         const unsigned long long* __restrict__ qcount,
                                                         const int64_t* __restrict__ idmap, int skip) {
  __shared__ uint16_t lists[4 * kScbListWave];
  uint16_t* list = lists + (threadIdx.x >> 6) * kScbListWave;
  const int64_t nq = (int64_t)*qcount;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t base = (int64_t)blockIdx.x * blockDim.x; base < nq; base += stride) {
    const int64_t q = base + threadIdx.x;
    const bool valid = q < nq;
    scb_sample_one<true>(v, valid, valid ? queue[q] : 0, ekin, dirs, seed, stream, oE, oD, nullptr, nullptr, list,
                         idmap, skip);
  }
}

static inline unsigned scb_grid(int64_t n) {
  int64_t g = (n + 255) / 256;
  const int64_t gmax = 256 * 64;
  if (g > gmax) g = gmax;
  if (g < 1) g = 1;
  return (unsigned)g;
}

// sorts the batch by scan length into w.perm and clears the queue counter
static void scb_order(const SCBView& v, const double* ekin, const double* dirs, int64_t n, const ScbWork& w,
                      hipStream_t st) {
  hipLaunchKernelGGL(k_scb_clear, dim3(1), dim3(256), 0, st, w.hist, w.qcount);
  hipLaunchKernelGGL(k_scb_keys, dim3(scb_grid(n)), dim3(256), 0, st, v, ekin, dirs, n, w.keys, w.hist);
  hipLaunchKernelGGL(k_scb_offsets, dim3(1), dim3(kScbBuckets), 0, st, w.hist, w.cursor);
  hipLaunchKernelGGL(k_scb_scatter, dim3(scb_grid((n + kScbScatterItems - 1) / kScbScatterItems)), dim3(256), 0, st,
                     w.keys, n, w.cursor, w.perm);
}

void launch_scb_xs(const SCBView& v, const double* ekin, const double* dirs, int64_t n, double* out,
                   const ScbWork& w, hipStream_t st) {
  if (n <= 0) return;
  scb_order(v, ekin, dirs, n, w, st);
  hipLaunchKernelGGL(k_scb_xs_main, dim3(scb_grid(n)), dim3(256), 0, st, v, ekin, dirs, n, w.perm, out, w.queue,
                     w.qcount);
  hipLaunchKernelGGL(k_scb_xs_slow, dim3(256), dim3(256), 0, st, v, ekin, dirs, out, w.queue, w.qcount);
}

// gathers the neutrons listed in q (count *qc, at most nmax) into compact
// arrays, with their original indices in idmap
__global__ void __launch_bounds__(256) k_scb_gather(const int64_t* __restrict__ q,
                                                    const unsigned long long* __restrict__ qc,
                                                    const double* __restrict__ ekin, const double* __restrict__ dirs,
                                                    double* __restrict__ cE, double* __restrict__ cD,
                                                    int64_t* __restrict__ idmap) {
  const int64_t m = (int64_t)*qc;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < m; t += stride) {
    const int64_t k = q[t];
    cE[t] = ekin[k];
    cD[3 * t] = dirs[3 * k];
    cD[3 * t + 1] = dirs[3 * k + 1];
    cD[3 * t + 2] = dirs[3 * k + 2];
    idmap[t] = k;
  }
}
void launch_scb_gather(const int64_t* q, const unsigned long long* qc, int64_t nmax, const double* ekin,
                       const double* dirs, double* cE, double* cD, int64_t* idmap, hipStream_t st) {
  if (nmax <= 0) return;
  hipLaunchKernelGGL(k_scb_gather, dim3(scb_grid(nmax)), dim3(256), 0, st, q, qc, ekin, dirs, cE, cD, idmap);
}

void launch_scb_sample(const SCBView& v, const double* ekin, const double* dirs, int64_t n, uint64_t seed,
                       uint32_t stream, double* oE, double* oD, const ScbWork& w, hipStream_t st,
                       const int64_t* idmap, int skip) {
  if (n <= 0) return;
  scb_order(v, ekin, dirs, n, w, st);
  hipLaunchKernelGGL(k_scb_sample_count, dim3(scb_grid(n)), dim3(256), 0, st, v, ekin, dirs, n, w.perm, oE, oD,
                     w.cnt, w.tot, w.queue, w.qcount, idmap);
  hipLaunchKernelGGL(k_scb_sample_pick, dim3(scb_grid(n)), dim3(256), 0, st, v, ekin, dirs, n, w.perm, seed, stream,
                     w.cnt, w.tot, w.sel, idmap, skip);
  hipLaunchKernelGGL(k_scb_sample_scatter, dim3(scb_grid(n)), dim3(256), 0, st, v, ekin, dirs, n, w.perm, seed,
                     stream, oE, oD, w.cnt, w.sel, idmap, skip);
  hipLaunchKernelGGL(k_scb_sample_slow, dim3(256), dim3(256), 0, st, v, ekin, dirs, seed, stream, oE, oD, w.queue,
                     w.qcount, idmap, skip);
}

}  // namespace ncamd
