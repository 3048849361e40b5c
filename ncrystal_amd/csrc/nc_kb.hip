// Box-envelope sampler of S(alpha,beta): the kernel that bounds t over every
// box, and the sampling kernels (device code in nc_kb.hh).
#include "nc_kernels.hh"

namespace ncamd {

// S along table row ib at alpha inside cell ia (log-linear in alpha, linear
// where a corner value is zero), as the sampler's cell evaluation computes it
NCD double kb_row(const SABView& s, int ib, int ia, double a) {
  const double2* r = s.sl + (size_t)ib * s.na + ia;
  const double2 c0 = r[0], c1 = r[1];
  const double a0 = s.alpha[ia], a1 = s.alpha[ia + 1];
  return kb_loglin(a0, c0.x, a1, c1.x, a, c0.y, c1.y, 1.0 / (a1 - a0));
}

// lower index of the cell containing x (clamped into the grid)
NCD int kb_cell_clamped(const double* g, int n, double x) {
  const int up = upper_bound_d(g, 0, n, x);
  return up - 1 < 0 ? 0 : (up - 1 > n - 2 ? n - 2 : up - 1);
}

// Bounds of S over the rectangle [alo,ahi] x [blo,bhi] inside the table grid.
// On a piece (rectangle and one cell) S = (1 - tau) s0(alpha) + tau s1(alpha)
// with the two rows s0, s1 log-linear (or linear) in alpha: each row is
// convex and monotone on the cell, and S is linear in tau.  So S is convex in
// alpha and linear in beta, and its maximum over the piece is at one of the
// piece's corners; its minimum is at least min over the tau end points of
// (1 - tau) min(s0) + tau min(s1), with each row's minimum at an alpha end
// point of the piece.
NCD void kb_sbounds(const SABView& s, double alo, double ahi, double blo, double bhi, double& smax, double& smin) {
  const int ia0 = kb_cell_clamped(s.alpha, s.na, alo), ia1 = kb_cell_clamped(s.alpha, s.na, ahi);
  const int ib0 = kb_cell_clamped(s.beta, s.nb, blo), ib1 = kb_cell_clamped(s.beta, s.nb, bhi);
  double mx = 0.0, mn = __builtin_huge_val();
  for (int ib = ib0; ib <= ib1; ++ib) {
    const double b0 = s.beta[ib], b1 = s.beta[ib + 1];
    const double tA = (dmax(blo, b0) - b0) / (b1 - b0), tB = (dmin(bhi, b1) - b0) / (b1 - b0);
    for (int ia = ia0; ia <= ia1; ++ia) {
      const double aA = dmax(alo, s.alpha[ia]), aB = dmin(ahi, s.alpha[ia + 1]);
      const double r0A = kb_row(s, ib, ia, aA), r0B = kb_row(s, ib, ia, aB);
      const double r1A = kb_row(s, ib + 1, ia, aA), r1B = kb_row(s, ib + 1, ia, aB);
      mx = dmax(mx, dmax(dmax(r0A + (r1A - r0A) * tA, r0A + (r1A - r0A) * tB),
                         dmax(r0B + (r1B - r0B) * tA, r0B + (r1B - r0B) * tB)));
      const double m0 = dmin(r0A, r0B), m1 = dmin(r1A, r1B);
      mn = dmin(mn, dmin(m0 + (m1 - m0) * tA, m0 + (m1 - m0) * tB));
    }
  }
  smax = mx;
  smin = dmax(mn, 0.0);
}

// Bounds of S over the (alpha, beta) image of the sub-box [w0,w1] x [k0,k1]
// of an interval's unit square, through the image's bounding rectangle
// (widened by margins for the rounding of the sampler's own arithmetic).
// Points of the rectangle outside the table have S = 0.
NCD void kb_subbox_bounds(const SABView& s, double eps, double seps, double wmax, double w0, double w1, double k0,
                          double k1, double& smax, double& smin) {
  const double r0 = w0 * wmax, r1 = w1 * wmax;
  double blo = r0 * r0 - eps, bhi = r1 * r1 - eps;
  // alpha(r, kappa) = r^2 + eps - 2 sqrt(eps) r (1 - 2 kappa): linear in
  // kappa, convex in r (minimum at r = sqrt(eps)(1 - 2 kappa))
  double alo = __builtin_huge_val(), ahi = -__builtin_huge_val();
  for (int c = 0; c < 2; ++c) {
    const double kap = c ? k1 : k0;
    const double f = 2.0 * seps * (1.0 - 2.0 * kap);
    const double a_r0 = r0 * r0 + eps - f * r0, a_r1 = r1 * r1 + eps - f * r1;
    ahi = dmax(ahi, dmax(a_r0, a_r1));
    double amin = dmin(a_r0, a_r1);
    const double rs = 0.5 * f;
    if (rs > r0 && rs < r1) amin = dmin(amin, eps - 0.25 * f * f);
    alo = dmin(alo, amin);
  }
  const double ma = 1e-9 * (1.0 + fabs(alo) + fabs(ahi)) + 1e-9 * eps;
  const double mb = 1e-9 * (1.0 + fabs(blo) + fabs(bhi)) + 1e-9 * eps;
  alo -= ma; ahi += ma; blo -= mb; bhi += mb;
  const double A0 = s.alpha[0], A1 = s.alpha[s.na - 1], B0 = s.beta[0], B1 = s.beta[s.nb - 1];
  if (ahi < A0 || alo > A1 || bhi < B0 || blo > B1) { smax = 0.0; smin = 0.0; return; }
  const bool inside = alo >= A0 && ahi <= A1 && blo >= B0 && bhi <= B1;
  kb_sbounds(s, dmax(alo, A0), dmin(ahi, A1), dmax(blo, B0), dmin(bhi, B1), smax, smin);
  if (!inside) smin = 0.0;
}

// Bounds of t = S * J over each box (boxes of every interval, possibly of
// several candidate grids: interval = (index / nbox) % nn).  The box is split
// into kKBSub x kKBSub sub-boxes; J = jfac w^2 is bounded by its values at
// the sub-box's w edges.  Also records the table cell of the lower corner of
// the box's (alpha, beta) range as the start of the sampler's cell searches.
constexpr int kKBSub = 4;
__global__ void k_kb_bounds(SABView s, const double4* __restrict__ node_par, int nn, int nbox, int64_t total,
                            KBBox* __restrict__ box) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= total) return;
  const int n = (int)((t / nbox) % nn);
  const double4 par = node_par[n];
  const double eps = par.x, seps = par.y, wmax = par.z, jfac = par.w;
  KBBox b = box[t];
  const double w0 = b.w0, w1 = b.w1, k0 = b.k0, k1 = b.k1;
  double env = 0.0, tlow = __builtin_huge_val();
  for (int sw = 0; sw < kKBSub; ++sw) {
    const double sw0 = w0 + (w1 - w0) * ((double)sw / kKBSub);
    const double sw1 = (sw + 1 == kKBSub) ? w1 : w0 + (w1 - w0) * ((double)(sw + 1) / kKBSub);
    for (int sk = 0; sk < kKBSub; ++sk) {
      const double sk0 = k0 + (k1 - k0) * ((double)sk / kKBSub);
      const double sk1 = (sk + 1 == kKBSub) ? k1 : k0 + (k1 - k0) * ((double)(sk + 1) / kKBSub);
      double smax, smin;
      kb_subbox_bounds(s, eps, seps, wmax, sw0, sw1, sk0, sk1, smax, smin);
      env = dmax(env, smax * (jfac * sw1 * sw1));
      tlow = dmin(tlow, smin * (jfac * sw0 * sw0));
    }
  }
  b.env = __double2float_ru(env * (1.0 + 1e-9));
  b.tlow = __double2float_rd(tlow * (1.0 - 1e-9));
  // cells covered by the box's (alpha, beta) range (with the margins of the
  // sampler's rounding); the envelope sampler requires grids of at most
  // 65536 points
  double alo = __builtin_huge_val(), ahi = -__builtin_huge_val();
  const double r0 = w0 * wmax, r1 = w1 * wmax;
  for (int c = 0; c < 2; ++c) {
    const double f = 2.0 * seps * (1.0 - 2.0 * (c ? k1 : k0));
    const double a_r0 = r0 * r0 + eps - f * r0, a_r1 = r1 * r1 + eps - f * r1;
    ahi = dmax(ahi, dmax(a_r0, a_r1));
    alo = dmin(alo, dmin(a_r0, a_r1));
    const double rs = 0.5 * f;
    if (rs > r0 && rs < r1) alo = dmin(alo, eps - 0.25 * f * f);
  }
  double blo = r0 * r0 - eps, bhi = r1 * r1 - eps;
  const double ma = 1e-9 * (1.0 + fabs(alo) + fabs(ahi)) + 1e-9 * eps;
  const double mb = 1e-9 * (1.0 + fabs(blo) + fabs(bhi)) + 1e-9 * eps;
  alo -= ma; ahi += ma; blo -= mb; bhi += mb;
  const uint32_t ia0 = (uint32_t)kb_cell_clamped(s.alpha, s.na, alo);
  const uint32_t ia1 = (uint32_t)kb_cell_clamped(s.alpha, s.na, ahi);
  const uint32_t ib0 = (uint32_t)kb_cell_clamped(s.beta, s.nb, blo);
  const uint32_t ib1 = (uint32_t)kb_cell_clamped(s.beta, s.nb, bhi);
  b.arange = ia0 | (ia1 << 16);
  b.brange = ib0 | (ib1 << 16);
  box[t] = b;
}

// Sampling: one lane per neutron at a time, one proposal per lane per loop
// step, so that lanes whose neutron needs more proposals do not hold up the
// others.  A wave owns chunks of kKBChunk consecutive neutrons (chunk c of
// the wave w: c = w, w + nwaves, ...); lanes that finish a neutron take the
// next ones of the chunk in lane order, so that a step's loads and stores
// fall on neighbouring entries.  Neutrons outside the sampler's energy range
// or without an accepted proposal after maxtries go to the queue of the
// plain-sampler kernel, which restarts their Philox stream from draw 0.
// Mode 0: outputs (E', mu); mode 1: raw (alpha, beta).
constexpr int64_t kKBChunk = 1024;
template <int Mode>
__global__ void __launch_bounds__(256) k_kb_sample(SABView s, KBView k, const double* __restrict__ ekin, int64_t n,
                                                   int64_t total, uint64_t seed, uint32_t stream,
                                                   double* __restrict__ o1, double* __restrict__ o2,
                                                   unsigned long long* __restrict__ nerr,
                                                   int64_t* __restrict__ queue,
                                                   unsigned long long* __restrict__ qcount, int maxtries) {
  const int lane = (int)(threadIdx.x & 63);
  const uint64_t below = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
  const int64_t wave = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const int64_t nwaves = ((int64_t)gridDim.x * blockDim.x) >> 6;
  // wave-uniform chunk state
  int64_t chunk = wave;
  int64_t next = chunk * kKBChunk;
  int64_t cend = next + kKBChunk < total ? next + kKBChunk : total;
  // lane state: the neutron, its energy (and E/kT), interval and the draw
  // counter of its Philox stream (key and stream are the kernel's)
  bool active = false;
  int64_t kk = 0;
  int tries = 0, iv = 0;
  uint32_t draw = 0;
  double E = 0.0, ediv = 0.0;
  while (true) {
    // idle lanes take the next neutrons of the chunk, moving on to the
    // wave's next chunk when this one is used up
    uint64_t idle = __ballot(!active);
    while (idle && next < total) {
      const int64_t take = (int64_t)__popcll(idle) < cend - next ? (int64_t)__popcll(idle) : cend - next;
      if (!active) {
        const int64_t rank = (int64_t)__popcll(idle & below);
        if (rank < take) {
          kk = next + rank;
          E = ekin[(total == n) ? kk : (kk % n)];
          iv = kb_interval(k, E);
          if (iv < 0) {
            queue[atomicAdd(qcount, 1ull)] = kk;
          } else {
            ediv = E * k.inv_kT;
            draw = 0;
            tries = 0;
            active = true;
          }
        }
      }
      next += take;
      if (next >= cend) {
        chunk += nwaves;
        next = chunk * kKBChunk;
        if (next > total) next = total;
        cend = next + kKBChunk < total ? next + kKBChunk : total;
      }
      idle = __ballot(!active);
    }
    if (!__ballot(active)) break;
    if (active) {
      Rng rng;
      rng.init(seed, stream, (uint64_t)kk);
      rng.c0 = draw;
      uint32_t r0, r1, r2, r3;
      rng.gen4(r0, r1, r2, r3);
      ++draw;
      double a, b;
      if (kb_propose(s, k, iv, ediv, r0, r1, r2, r3, a, b)) {
        if constexpr (Mode == 0) {
          double Ef, mu;
          if (sab_ab_to_emu(s, rng, E, a, b, Ef, mu)) { Ef = -1.0; mu = -999.0; atomicAdd(nerr, 1ull); }
          o1[kk] = Ef;
          o2[kk] = mu;
        } else {
          o1[kk] = a;
          o2[kk] = b;
        }
        active = false;
      } else if (++tries >= maxtries) {
        queue[atomicAdd(qcount, 1ull)] = kk;
        active = false;
      }
    }
  }
}

// raw (alpha, beta) of the plain sampler for the neutrons the envelope
// sampler handed back (from the first draw of their stream)
__global__ void k_kb_queue_ab(SABView s, const double* __restrict__ ekin, uint64_t seed, uint32_t stream,
                              double* __restrict__ oa, double* __restrict__ ob, const int64_t* __restrict__ queue,
                              const unsigned long long* __restrict__ qcount) {
  const int64_t nq = (int64_t)*qcount;
  for (int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; j < nq; j += (int64_t)gridDim.x * blockDim.x) {
    const int64_t kk = queue[j];
    Rng rng;
    rng.init(seed, stream, (uint64_t)kk);
    double a = -1.0, b = 0.0;
    if (sab_sample_ab<true>(s, ekin[kk], rng, a, b)) { a = -1.0; b = 0.0; }
    oa[kk] = a;
    ob[kk] = b;
  }
}

static inline unsigned kb_grid(int64_t n) {
  int64_t g = (n + 255) / 256;
  const int64_t gmax = 256 * 64;
  if (g > gmax) g = gmax;
  if (g < 1) g = 1;
  return (unsigned)g;
}

void launch_kb_bounds(const SABView& s, const double4* node_par, int nn, int nbox, int64_t total, KBBox* box,
                      hipStream_t st) {
  if (total <= 0) return;
  hipLaunchKernelGGL(k_kb_bounds, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, st, s, node_par, nn, nbox,
                     total, box);
}

// persistent grid: as many blocks as fit on the device at once (the chunks
// are handed out statically), fewer for small batches
template <int Mode>
static unsigned kb_sample_grid(int64_t total) {
  static int per_cu = 0, cus = 0;
  if (!per_cu) {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus < 1)
      cus = 256;
    int nb = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_kb_sample<Mode>, 256, 0) != hipSuccess || nb < 1) nb = 2;
    per_cu = nb;
  }
  const int64_t chunks = (total + kKBChunk - 1) / kKBChunk;
  int64_t g = (chunks + 3) / 4;  // 4 waves per block, one chunk per wave at least
  const int64_t gmax = (int64_t)per_cu * cus;
  if (g > gmax) g = gmax;
  if (g < 1) g = 1;
  return (unsigned)g;
}

void launch_kb_sample_iso(const SABView& s, const KBView& k, const double* ekin, int64_t n, int64_t repeat,
                          uint64_t seed, uint32_t stream, double* oE, double* oMu, unsigned long long* nerr,
                          int64_t* queue, unsigned long long* qcount, int maxtries, hipStream_t st) {
  const int64_t total = n * repeat;
  if (total <= 0) return;
  launch_zero_u64(qcount, st);
  hipLaunchKernelGGL(k_kb_sample<0>, dim3(kb_sample_grid<0>(total)), dim3(256), 0, st, s, k, ekin, n, total, seed,
                     stream, oE, oMu, nerr, queue, qcount, maxtries);
  launch_sab_queue(s, ekin, n, total, seed, stream, oE, oMu, nerr, queue, qcount, st);
}

void launch_kb_sample_ab(const SABView& s, const KBView& k, const double* ekin, int64_t n, uint64_t seed,
                         uint32_t stream, double* oa, double* ob, int64_t* queue, unsigned long long* qcount,
                         int maxtries, hipStream_t st) {
  if (n <= 0) return;
  launch_zero_u64(qcount, st);
  hipLaunchKernelGGL(k_kb_sample<1>, dim3(kb_sample_grid<1>(n)), dim3(256), 0, st, s, k, ekin, n, n, seed, stream, oa,
                     ob, (unsigned long long*)nullptr, queue, qcount, maxtries);
  hipLaunchKernelGGL(k_kb_queue_ab, dim3(kb_grid(n)), dim3(256), 0, st, s, ekin, seed, stream, oa, ob, queue,
                     qcount);
}

}  // namespace ncamd
