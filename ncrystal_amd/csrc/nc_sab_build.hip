// S(alpha,beta) table construction on the GPU, FP64 throughout.
//
//  k_sab_derive  : log S and the cumulative per-beta-row alpha integrals
//                  (ncrystal_core/src/sab/NCSABIntegrator.cc:105-140), one lane
//                  per beta row.
//  k_sab_analyse : one workgroup per energy point (NCSABIntegrator.cc:352-559):
//                  lanes split the beta grid (kinematic alpha range by binary
//                  search -- equivalent to the incremental walk of
//                  ncrystal_core/src/sab/NCSABUtils.cc:460-536 -- and the tailed
//                  breakdown of NCSABUtils.cc:538-633); lane 0 then runs the
//                  ordered trapezoid sum over beta (first-bin 4/3 rule) and the
//                  beta-distribution CDF (ncrystal_core/src/utils/NCPointwiseDist.cc:32-68)
//                  so that the summation order is the reference's.
#include "nc_kernels.hh"

namespace ncamd {

__global__ void k_sab_derive(SABBuildBufs b) {
  const int ib = blockIdx.x * blockDim.x + threadIdx.x;
  if (ib >= b.nb) return;
  const int na = b.na;
  const size_t row = (size_t)ib * na;
  double prev_s = b.sab[row];
  double prev_l = prev_s > 0.0 ? log(prev_s) : -__builtin_huge_val();
  b.sl[row] = make_double2(prev_s, prev_l);
  b.cumul[row] = 0.0;
  double cum = 0.0;
  for (int ai = 0; ai < na - 1; ++ai) {
    const double s1 = b.sab[row + ai + 1];
    const double l1 = s1 > 0.0 ? log(s1) : -__builtin_huge_val();
    const double integ = integrate_alpha_interval_fast(b.alpha[ai], prev_s, b.alpha[ai + 1], s1, prev_l, l1);
    cum += integ;
    b.sl[row + ai + 1] = make_double2(s1, l1);
    b.cumul[row + ai + 1] = cum;
    prev_s = s1;
    prev_l = l1;
  }
}

struct TB {
  double xs_front, xs_middle, xs_back;
  int imiddle_low, imiddle_upp;
  double f_alpha, f_sval, f_logsval, b_alpha, b_sval, b_logsval;
  bool narrow;
};

NCD void set_tail(const double* ag, const double2* sl, int aidx, double alpha, double& a, double& s, double& ls) {
  a = alpha;
  const double2 q0 = sl[aidx], q1 = sl[aidx + 1];
  s = interp_loglin_fast(ag[aidx], q0.x, ag[aidx + 1], q1.x, alpha, q0.y, q1.y);
  ls = log(dmax(s, kDblMin));
}

NCD TB tailed_breakdown(const double* ag, int na, const double2* sl, const double* cumul, double alow,
                        double aupp, int ilo, int iup) {
  TB tb{};
  alow = dclamp(alow, ag[0], ag[na - 1]);
  aupp = dclamp(aupp, ag[0], ag[na - 1]);
  if (ilo == iup || alow == aupp) return tb;
  if (ilo + 1 == iup) {
    tb.narrow = true;
    set_tail(ag, sl, ilo, alow, tb.f_alpha, tb.f_sval, tb.f_logsval);
    set_tail(ag, sl, ilo, aupp, tb.b_alpha, tb.b_sval, tb.b_logsval);
    tb.xs_front = integrate_alpha_interval_fast(tb.f_alpha, tb.f_sval, tb.b_alpha, tb.b_sval, tb.f_logsval,
                                                tb.b_logsval);
    return tb;
  }
  tb.imiddle_low = ilo;
  tb.imiddle_upp = iup;
  if (alow >= ag[ilo]) {
    set_tail(ag, sl, ilo, alow, tb.f_alpha, tb.f_sval, tb.f_logsval);
    const double2 q = sl[ilo + 1];
    tb.xs_front = integrate_alpha_interval_fast(tb.f_alpha, tb.f_sval, ag[ilo + 1], q.x, tb.f_logsval, q.y);
    ++tb.imiddle_low;
  }
  if (aupp <= ag[iup]) {
    set_tail(ag, sl, iup - 1, aupp, tb.b_alpha, tb.b_sval, tb.b_logsval);
    const double2 q = sl[iup - 1];
    tb.xs_back = integrate_alpha_interval_fast(ag[iup - 1], q.x, tb.b_alpha, tb.b_sval, q.y, tb.b_logsval);
    --tb.imiddle_upp;
  }
  tb.xs_middle = (tb.imiddle_upp > tb.imiddle_low ? cumul[tb.imiddle_upp] - cumul[tb.imiddle_low] : 0.0);
  return tb;
}

template <bool DoSampler>
__global__ void __launch_bounds__(256) k_sab_analyse(SABBuildBufs b, const double* __restrict__ energies,
                                                     double* __restrict__ xs_out, int4* __restrict__ shdr,
                                                     double* __restrict__ sfirst, double* __restrict__ scdf,
                                                     double2* __restrict__ sxy, AInfo* __restrict__ sinfo) {
  extern __shared__ double lds_xsb[];  // [nb]
  __shared__ int s_ibl, s_jskip;
  __shared__ double s_bll;
  __shared__ int s_starts;
  const int ie = blockIdx.x;
  const int tid = threadIdx.x;
  const int nb = b.nb, na = b.na;
  const double* __restrict__ bg = b.beta;
  const double* __restrict__ ag = b.alpha;
  const double ekin = energies[ie];
  const double ediv = ekin / b.kT;
  const double af = ag[0], ab = ag[na - 1];

  if (tid == 0) { s_ibl = nb; s_jskip = -1; }
  __syncthreads();
  // first beta index with a kinematically accessible alpha range
  for (int ib = tid; ib < nb; ib += blockDim.x) {
    const double beta = bg[ib];
    double alow = -1.0, aupp = -2.0;
    if (beta > -ediv) { const Lims l = alpha_limits(ediv, beta); alow = l.lo; aupp = l.hi; }
    const bool inacc = (ab <= alow || af >= aupp || aupp < alow);
    if (!inacc) atomicMin(&s_ibl, ib);
  }
  __syncthreads();
  const int ibl = s_ibl;
  if (tid == 0) {
    double bll = -ediv;
    int starts = 1;
    if (bll < bg[0]) {
      const double c1 = bg[0] - (bg[1] - bg[0]) * 1e-6;
      const double c2 = bg[0] - fabs(bg[0]) * 1e-13;
      const double c3 = nextafter(bg[0], bll);
      bll = dmin(dmin(c1, c2), c3);
      starts = 0;
    }
    if (ibl < nb && ibl > 0 && bll < bg[ibl - 1]) { bll = bg[ibl - 1]; starts = 0; }
    s_bll = bll;
    s_starts = starts;
  }
  __syncthreads();
  if (ibl >= nb) {
    if (tid == 0) {
      xs_out[ie] = 0.0;
      if (DoSampler) { shdr[ie] = make_int4(0, nb, 0, 0); sfirst[ie] = 1.0; }
    }
    return;
  }
  const double bll = s_bll;
  const int nrel = nb - ibl;
  for (int j = tid; j < nrel; j += blockDim.x)
    if (bg[ibl + j] == bll) s_jskip = j;
  __syncthreads();
  const int jskip = s_jskip;

  for (int j = tid; j < nrel; j += blockDim.x) {
    if (j == jskip) { lds_xsb[j] = 0.0; continue; }
    const int ib = ibl + j;
    const double beta = bg[ib];
    // accessible alpha range (NCSABUtils.cc:460-536)
    double alow0 = -1.0, aupp0 = -2.0;
    if (beta > -ediv) { const Lims l = alpha_limits(ediv, beta); alow0 = l.lo; aupp0 = l.hi; }
    int ilo, iup;
    if (ab <= alow0 || af >= aupp0 || aupp0 < alow0) {
      ilo = na; iup = na;
    } else {
      ilo = upper_bound_d(ag, 0, na, alow0) - 1;
      if (ilo < 0) ilo = 0;
      if (ilo > na - 1) ilo = na - 1;
      iup = lower_bound_d(ag, 0, na, aupp0);
      if (iup > na - 1) iup = na - 1;
    }
    const Lims lims = alpha_limits(ediv, beta);
    const double alow = lims.lo, aupp = lims.hi;
    double xsb = 0.0;
    TB tb{};
    if (iup > ilo && aupp > alow) {
      const size_t off = (size_t)na * ib;
      tb = tailed_breakdown(ag, na, b.sl + off, b.cumul + off, alow, aupp, ilo, iup);
      xsb = tb.xs_front + tb.xs_back + tb.xs_middle;
    }
    lds_xsb[j] = xsb;
    if (DoSampler) {
      AInfo inf{};
      if (xsb > 0.0) {
        inf.f_alpha = tb.f_alpha; inf.f_sval = tb.f_sval; inf.f_logsval = tb.f_logsval;
        inf.b_alpha = tb.b_alpha; inf.b_sval = tb.b_sval; inf.b_logsval = tb.b_logsval;
        if (tb.narrow) {
          inf.prob_front = 1.0;
        } else {
          inf.prob_front = tb.xs_front / xsb;
          inf.prob_notback = 1.0 - tb.xs_back / xsb;
          inf.f_idx = tb.imiddle_low;
          inf.b_idx = tb.imiddle_upp;
        }
      } else {
        inf.prob_front = 2.0;
        inf.f_alpha = alow;
        inf.b_alpha = aupp;
      }
      const int dst = j - ((jskip >= 0 && j > jskip) ? 1 : 0);
      sinfo[(size_t)ie * nb + dst] = inf;
    }
  }
  __syncthreads();
  if (tid != 0) return;

  // ordered part: trapezoid over beta with the first-bin 4/3 rule
  double* __restrict__ cdf = DoSampler ? scdf + (size_t)ie * (nb + 1) : nullptr;
  double2* __restrict__ xy = DoSampler ? sxy + (size_t)ie * (nb + 1) : nullptr;
  double prev_b = bll, prev_xs = 0.0;
  int nv = 0;
  if (DoSampler) { xy[0] = make_double2(prev_b, 0.0); nv = 1; }
  bool next_kin = s_starts != 0;
  StableSum tot;
  for (int j = 0; j < nrel; ++j) {
    if (j == jskip) continue;
    const double beta = bg[ibl + j];
    const double xsb = lds_xsb[j];
    if (next_kin) {
      next_kin = false;
      const double db_real = (beta - prev_b);
      prev_b -= db_real * (1.0 / 3.0);
      if (DoSampler) xy[0].x = prev_b;
    }
    tot.add(0.5 * (beta - prev_b) * (xsb + prev_xs));
    prev_b = beta;
    prev_xs = xsb;
    if (DoSampler) { xy[nv] = make_double2(prev_b, prev_xs); ++nv; }
  }
  double xs_total = tot.sum() * b.sigma_bound / (4 * ediv);
  if (!(xs_total >= 0.0)) xs_total = 0.0;
  xs_out[ie] = xs_total;
  if (!DoSampler) return;
  sfirst[ie] = s_starts ? bll : 1.0;
  if (xs_total == 0.0) { shdr[ie] = make_int4(0, ibl, 0, 0); return; }
  // PointwiseDist construction
  StableSum area;
  cdf[0] = 0.0;
  for (int i = 1; i < nv; ++i) {
    const double2 p0 = xy[i - 1], p1 = xy[i];
    area.add((p1.x - p0.x) * 0.5 * (p1.y + p0.y));
    cdf[i] = area.sum();
  }
  const double tv = area.sum();
  int kind = 1;
  if (!(tv > 0.0)) kind = -1;
  const double nf = 1.0 / tv;
  for (int i = 0; i < nv; ++i) { cdf[i] *= nf; xy[i].y *= nf; }
  cdf[nv - 1] = 1.0;
  shdr[ie] = make_int4(kind, ibl, nv, nrel - (jskip >= 0 ? 1 : 0));
}

__global__ void k_fg_xs(double sigma_free, double ca, const double* e, int n, double* out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = sigma_free * freegas_xs_shape(ca * e[i]);
}

void launch_sab_derive(const SABBuildBufs& b, hipStream_t st) {
  hipLaunchKernelGGL(k_sab_derive, dim3((b.nb + 63) / 64), dim3(64), 0, st, b);
}
void launch_sab_xs_points(const SABBuildBufs& b, const double* energies, int n, double* xs_out, hipStream_t st) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_sab_analyse<false>, dim3(n), dim3(256), sizeof(double) * b.nb, st, b, energies, xs_out,
                     nullptr, nullptr, nullptr, nullptr, nullptr);
}
void launch_sab_analyse(const SABBuildBufs& b, const double* egrid, int ne, double* xs_out, int4* shdr,
                        double* sfirst, double* scdf, double2* sxy, AInfo* sinfo, hipStream_t st) {
  hipLaunchKernelGGL(k_sab_analyse<true>, dim3(ne), dim3(256), sizeof(double) * b.nb, st, b, egrid, xs_out, shdr,
                     sfirst, scdf, sxy, sinfo);
}
void launch_fg_xs(double sigma_free, double ca, const double* e, int n, double* out, hipStream_t st) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_fg_xs, dim3((n + 255) / 256), dim3(256), 0, st, sigma_free, ca, e, n, out);
}

}  // namespace ncamd
