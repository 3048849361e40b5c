// Host layer and C ABI of libncrystal_amd (declared in include/ncrystal_amd.h).
//
// Process handles own their device tables.  Material compilation (hkl
// families, PowderBragg cumulative tables, S(alpha,beta) integration) runs
// once per handle; every per-neutron evaluation runs in the HIP kernels of
// nc_kernels.hip.  There is no CPU evaluation path: without a usable GPU every
// entry point raises an error.
//
// Error convention and batched-call semantics follow
// ncrystal_core/src/cinterface/ncrystal.cc:275-323 (setError/haltonerror) and
// :1107-1282 (_many functions, poisoned outputs on error).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <unordered_set>
#include <stdexcept>
#include <string>
#include <vector>
#include <thread>
#include <atomic>

#include "../../include/ncrystal_amd.h"
#include "../../include/ncrystal_dropin.h"
#include "nc_kernels.hh"
#include "nc_math.hh"
#include "nc_vdos.hh"

namespace ncamd {
namespace {

// ---------------------------------------------------------------- errors ---
int g_quiet = 0, g_halt = 1, g_waserror = 0;
char g_errmsg[512];
char g_errtype[64];
void (*g_errhandler)(char*, char*) = nullptr;  // called with (type, message)

void set_error(const char* msg, const char* etype) noexcept {
  strncpy(g_errmsg, msg, sizeof(g_errmsg) - 1);
  strncpy(g_errtype, etype, sizeof(g_errtype) - 1);
  g_errmsg[sizeof(g_errmsg) - 1] = '\0';
  g_errtype[sizeof(g_errtype) - 1] = '\0';
  if (g_errhandler) g_errhandler(g_errtype, g_errmsg);
  g_waserror = 1;
  if (!g_quiet) fprintf(stderr, "ncrystal_amd ERROR [%s]: %s\n", g_errtype, g_errmsg);
  if (g_halt) {
    fprintf(stderr, "ncrystal_amd terminating due to ERROR\n");
    exit(1);
  }
}

struct Error : std::runtime_error {
  std::string type;
  Error(const std::string& t, const std::string& m) : std::runtime_error(m), type(t) {}
};
[[noreturn]] void raise(const char* type, const std::string& msg) { throw Error(type, msg); }

#define NCAMD_HIP(call)                                                                       \
  do {                                                                                        \
    hipError_t e_ = (call);                                                                   \
    if (e_ != hipSuccess) raise("HIPError", std::string(#call) + ": " + hipGetErrorString(e_)); \
  } while (0)

#define NCAMD_TRY try {
#define NCAMD_CATCH                                             \
  }                                                             \
  catch (const Error& e) { set_error(e.what(), e.type.c_str()); } \
  catch (const std::exception& e) { set_error(e.what(), "std::exception"); }

void ensure_device() {
  static int ok = -1;
  static std::mutex mtx;
  std::lock_guard<std::mutex> lk(mtx);
  if (ok == 1) return;
  int n = 0;
  hipError_t e = hipGetDeviceCount(&n);
  if (e != hipSuccess || n < 1) {
    ok = 0;
    raise("NoDevice", "ncrystal_amd requires a gfx950 GPU; none is available (no CPU fallback exists)");
  }
  ok = 1;
}

// -------------------------------------------------------- device buffers ---
struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
  DevBuf() = default;
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  ~DevBuf() { if (p) (void)hipFree(p); }
  void alloc(size_t b) {
    if (p) { (void)hipFree(p); p = nullptr; }
    bytes = b;
    if (b) NCAMD_HIP(hipMalloc(&p, b));
  }
  void ensure(size_t b) { if (b > bytes) alloc(b); }
  template <class T> T* as() const { return static_cast<T*>(p); }
  template <class T> void upload(const T* h, size_t n) {
    alloc(n * sizeof(T));
    if (n) NCAMD_HIP(hipMemcpy(p, h, n * sizeof(T), hipMemcpyHostToDevice));
  }
};

// -------------------------------------------------------------- numerics ---
// Romberg + Debye MSD (ncrystal_core/src/utils/NCRomberg.cc:58-140,
// ncrystal_core/src/phys_utils/NCDebyeMSD.cc:27-86); init-time scalar.
template <class F, class A>
double romberg(const F& f, const A& acc, double a, double b) {
  double h = (b - a);
  double fv[17];
  const double delta = h * 0.0625;
  for (int i = 0; i < 17; ++i) fv[i] = f(a + delta * (double)i);
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
  if (acc(4, R33, R44)) return R44;
  double c5 = 0.0;
  {
    const double off = a + h * 0.5;
    for (double i = 0; i < 16.0; ++i) c5 += f(off + h * i);
  }
  h *= 0.5;
  const double R50 = h * c5 + 0.5 * R40;
  const double R51 = (4. / 3.) * R50 + (-1. / 3.) * R40;
  const double R52 = (16. / 15.) * R51 + (-1. / 15.) * R41;
  const double R53 = (64. / 63.) * R52 + (-1. / 63.) * R42;
  const double R54 = (256. / 255.) * R53 + (-1. / 255.) * R43;
  const double R55 = (1024. / 1023.) * R54 + (-1. / 1023.) * R44;
  if (acc(5, R44, R55)) return R55;
  constexpr unsigned maxlevel = 16;
  double cache1[maxlevel], cache2[maxlevel];
  double *row_prev = cache1, *row = cache2;
  row_prev[0] = R50; row_prev[1] = R51; row_prev[2] = R52;
  row_prev[3] = R53; row_prev[4] = R54; row_prev[5] = R55;
  unsigned nj = 16;
  for (unsigned i = 6; i < maxlevel; ++i) {
    const double hh = h;
    h *= 0.5;
    nj *= 2;
    double c = 0.0;
    {
      const double off = a + h, nn = nj;
      for (double j = 0; j < nn; ++j) c += f(off + hh * j);
    }
    row[0] = h * c + 0.5 * row_prev[0];
    double n_k = 1.;
    for (unsigned j = 0; j < i; ++j) {
      n_k *= 4.0;
      row[j + 1] = (n_k * row[j] - row_prev[j]) / (n_k - 1.0);
    }
    if (acc(i, row_prev[i - 1], row[i])) return row[i];
    std::swap(row_prev, row);
  }
  raise("CalcError", "Romberg integration did not converge");
}

double debye_msd(double dt, double t, double am) {
  if (!(dt > 0.0 && dt < 1e5) || !(t >= 0.0) || !(am >= 1.007 && am < 500))
    raise("BadInput", "debye_msd: parameters out of range");
  constexpr double hbar = 4.135667662e-15 * 0.159154943091895335768883763372514362034459646;
  constexpr double c = 299792458e10;
  const double kk = 3.0 * hbar * hbar * c * c / (931494095.17 * kBoltzmann);
  const double scale = kk / (am * dt);
  const double x = t / dt;
  if (x < 1e-50) return scale * 0.25;
  auto f = [](double xx) {
    if (fabs(xx) < 1e-4) {
      const double x2 = xx * xx;
      return 1 - xx * 0.5 + x2 * 0.08333333333333333333333333333333 - x2 * x2 * 0.00138888888888888888888888888888888889;
    }
    if (xx <= 700) return xx / expm1(xx);
    return (xx > 760 ? 0.0 : exp(-xx + log(-xx / expm1(-xx))));
  };
  auto acc = [](unsigned lvl, double prev, double est) { return lvl > 11 || (lvl > 7 && fabs(est - prev) < 1e-10); };
  return scale * (0.25 + x * x * romberg(f, acc, 0.0, 1.0 / x));
}

// ------------------------------------------------------------- processes ---
constexpr uint32_t kMagic = 0x6e63616dU;

struct Proc {
  uint32_t magic = kMagic;
  int refcount = 1;
  int kind = kKindNone;  // kKind* of a single process, or kKindNone for a composition
  uint64_t seed = 0;
  uint64_t callctr = 0;
  // owned device memory
  std::vector<std::unique_ptr<DevBuf>> bufs;
  DevBuf d_err;
  // staging for the host-pointer API
  DevBuf st_in, st_in2, st_out1, st_out2;
  // high-energy queue of the S(alpha,beta)-only sampling path (also the
  // Romberg queue of SCBragg)
  DevBuf q_idx, q_count;
  // SCBragg batch ordering
  DevBuf scb_keys, scb_perm, scb_hist, scb_cursor, scb_cnt, scb_tot, scb_sel;
  // component data
  PBView pb{};
  ElIncView ei{};
  SABView sab{};
  FGView fg{};         // FreeGas process
  double abs_c = 0.0;  // 1/v absorption: xs = abs_c / sqrt(E)
  // oriented composition (one SCBragg component): its cross sections per
  // neutron, the neutrons that pick it, their compacted inputs
  bool oriented_comp = false;
  DevBuf oc_ext, oc_q, oc_qc, oc_cE, oc_cD, oc_idmap;
  double pb_v0n = 0.0;
  std::vector<double> pb_e2d_h, pb_fdm_h;
  std::vector<float> pb_fdm32_h;
  bool has_fp32 = false;
  // S(alpha,beta) host copies for read-back
  int sab_na = 0, sab_nb = 0;
  std::vector<double> sab_alpha_h, sab_beta_h;
  // box-envelope sampler (optional sampling mode of an S(alpha,beta)
  // process), on a uniform grid or on a grid placed by the learned model
  bool use_kb = false;
  int kb_maxtries = 64;
  KBView kb{};
  std::vector<std::unique_ptr<DevBuf>> kb_bufs;
  std::vector<double> kb_node_h;
  std::vector<KBBox> kb_box_h;
  // per interval: envelope mass of the grid in use, of the uniform grid,
  // blend weight of the uniform grid (1: uniform), squeeze mass
  std::vector<double> kb_stats_h;
  std::vector<double> sab_egrid_h, sab_xs_h;
  // single-crystal Bragg (host copies of the device tables for read-back)
  SCBView scb{};
  std::vector<double> scb_inv2d_h, scb_xsfact_h, scb_normals_h, scb_lt1_h, scb_lt2_h;
  std::vector<int64_t> scb_off_h;
  double scb_gos[15] = {};
  // layered crystal by a reference model: the SCBragg process is comps[0];
  // lc_mode > 0 averages cross sections over lc_nprime equally spaced
  // rotations and samples over lc_n random ones, lc_mode < 0 uses lc_n
  // random rotations for both
  int lc_mode = 0, lc_n = 0, lc_nprime = 0;
  LCView lc{};
  LC0View lc0{};  // lcmode=0: plane sets (device copy in lc0_sets)
  DevBuf lc0_sets;
  DevBuf lc_cs, lc_E, lc_D, lc_xs, lc_csr, lc_sE, lc_sD, lc_sel, lc_oE, lc_oD;
  // composition
  std::vector<Proc*> comps;
  std::vector<double> scales;
  // a clone reads the tables of its source, which it keeps alive; the source
  // counts its live clones (its tables may not be rebuilt meanwhile)
  Proc* src = nullptr;
  int nclones = 0;

  DevBuf* newbuf() {
    bufs.emplace_back(new DevBuf());
    return bufs.back().get();
  }
  ~Proc() {
    magic = 0;
    for (auto* c : comps) {
      if (c && --c->refcount == 0) delete c;
    }
    if (src) {
      --src->nclones;
      if (--src->refcount == 0) delete src;
    }
  }
};

Proc* extract(ncamd_process_t h) {
  Proc* p = static_cast<Proc*>(h.internal);
  if (!p || p->magic != kMagic) raise("LogicError", "invalid ncamd_process_t handle");
  return p;
}

ncamd_process_t wrap(Proc* p) {
  ncamd_process_t h;
  h.internal = p;
  return h;
}

void init_errbuf(Proc* p) {
  p->d_err.alloc(sizeof(unsigned long long));
  NCAMD_HIP(hipMemset(p->d_err.p, 0, sizeof(unsigned long long)));
}

CompView make_view(Proc* p) {
  CompView v{};
  std::vector<std::pair<Proc*, double>> list;
  if (p->kind == kKindNone) {
    for (size_t i = 0; i < p->comps.size(); ++i) list.emplace_back(p->comps[i], p->scales[i]);
  } else {
    list.emplace_back(p, 1.0);
  }
  v.n = (int)list.size();
  double lo = __builtin_huge_val(), hi = -__builtin_huge_val();
  for (int i = 0; i < v.n; ++i) {
    Proc* c = list[i].first;
    v.kind[i] = c->kind;
    v.scale[i] = list[i].second;
    double dlo = 0.0, dhi = __builtin_huge_val();
    switch (c->kind) {
      case kKindPB:
        v.pb = c->pb;
        if (c->pb.n > 0) dlo = c->pb.threshold; else { dlo = __builtin_huge_val(); dhi = __builtin_huge_val(); }
        break;
      case kKindElInc: v.ei = c->ei; break;
      case kKindSAB: break;  // staged below
      case kKindLC: raise("BadInput", "oriented process: use the calls with neutron directions");
      case kKindSCB:
        if (p->kind != kKindNone) raise("BadInput", "oriented process: use the calls with neutron directions");
        v.has_scb = 1;
        v.scb_threshold = c->scb.threshold;
        dlo = c->scb.threshold;
        break;
      case kKindFG:
      case kKindAbs:
        // a single FreeGas or absorption process has its own kernels
        if (p->kind != kKindNone) raise("LogicError", "FreeGas and absorption processes are evaluated on their own");
        if (c->kind == kKindFG) v.fg[i] = c->fg;
        else v.absc[i] = c->abs_c;
        break;
      default: raise("LogicError", "bad component kind");
    }
    if (dlo < dhi) { lo = std::min(lo, dlo); hi = std::max(hi, dhi); }
  }
  v.use_domain = p->kind == kKindNone ? 1 : 0;
  v.domain_lo = lo;
  v.domain_hi = hi;
  // staged components: every S(alpha,beta) component, and FreeGas components
  // beside S(alpha,beta) or SCBragg ones (gas mixtures without them sample
  // free gas in place)
  bool staged_kernels = false;
  for (int i = 0; i < v.n; ++i) staged_kernels = staged_kernels || v.kind[i] == kKindSAB || v.kind[i] == kKindSCB;
  v.nstaged = 0;
  for (int i = 0; i < v.n; ++i) {
    v.stage_of[i] = -1;
    if (v.kind[i] == kKindSAB || (staged_kernels && v.kind[i] == kKindFG)) {
      if (v.nstaged >= kMaxStaged)
        raise("BadInput", "at most 4 S(alpha,beta) components (and FreeGas components beside them) per composition");
      const int sl = v.nstaged++;
      v.stage_comp[sl] = i;
      v.stage_of[i] = sl;
      if (v.kind[i] == kKindSAB) v.sab[sl] = list[i].first->sab;
    }
  }
  return v;
}

// ------------------------------------------------------------ PowderBragg ---
// init: ncrystal_core/src/powderbragg/NCPowderBragg.cc:68-110 (+ the exact-d
// merge of :122-145)
// FP32 cumulative table of the PowderBragg contributions: a compensated
// prefix sum in float-float (hi, lo) pairs, whose hi + lo carries the FP64
// sum to about 2^-44 relative, so each output is the FP64 cumulative rounded
// to float (to within one ulp).  One workgroup of kCumulBlock lanes: each
// lane sums a contiguous chunk, the chunk totals are scanned with wavefront
// shuffles and across the waves through LDS, then each lane walks its chunk
// again from its exclusive prefix.
struct F2 {
  float hi, lo;
};
// float-float addition: Knuth's TwoSum of the leading words, the trailing
// words added to its error term, renormalised
__device__ inline F2 f2_add(F2 a, F2 b) {
  const float s = a.hi + b.hi;
  const float bb = s - a.hi;
  const float e = (a.hi - (s - bb)) + (b.hi - bb);
  const float t = e + (a.lo + b.lo);
  const float hi = s + t;
  return F2{hi, t - (hi - s)};
}
__device__ inline F2 f2_of(double x) {
  const float hi = (float)x;
  return F2{hi, (float)(x - (double)hi)};
}
constexpr int kCumulBlock = 1024;
__global__ void __launch_bounds__(kCumulBlock) k_fp32_cumul_scan(const double* __restrict__ contrib, int n,
                                                                 float* __restrict__ out) {
  __shared__ F2 s_wave[kCumulBlock / 64];
  const int tid = (int)threadIdx.x, lane = tid & 63, wid = tid >> 6;
  const int chunk = (n + kCumulBlock - 1) / kCumulBlock;
  const int b = tid * chunk, e = min(n, b + chunk);
  F2 tot{0.0f, 0.0f};
  for (int i = b; i < e; ++i) tot = f2_add(tot, f2_of(contrib[i]));
  // inclusive scan of the chunk totals within the wavefront
  F2 inc = tot;
  for (int d = 1; d < 64; d <<= 1) {
    const F2 up{__shfl_up(inc.hi, d, 64), __shfl_up(inc.lo, d, 64)};
    if (lane >= d) inc = f2_add(up, inc);
  }
  if (lane == 63) s_wave[wid] = inc;
  __syncthreads();
  F2 pre{0.0f, 0.0f};
  for (int w = 0; w < wid; ++w) pre = f2_add(pre, s_wave[w]);
  const F2 left{__shfl_up(inc.hi, 1, 64), __shfl_up(inc.lo, 1, 64)};
  if (lane > 0) pre = f2_add(pre, left);
  for (int i = b; i < e; ++i) {
    pre = f2_add(pre, f2_of(contrib[i]));
    out[i] = pre.hi + pre.lo;
  }
}

// search guides of a PowderBragg table: IEEE-bit buckets of the energy
// axis for the threshold search and equal-value buckets of the cumulative
// sum for the plane pick.  Each bucket stores the search answer at its lower
// edge, so a bracket [bucket, next bucket] (widened by one bucket for the
// rounding of target*scale) always contains the plain binary-search answer.
void pb_guides(Proc* p, const std::vector<double>& e2d, const std::vector<double>& fdm) {
  const int n = (int)e2d.size();
  auto bits = [](double x) { unsigned long long u; std::memcpy(&u, &x, 8); return u; };
  auto fromb = [](unsigned long long u) { double x; std::memcpy(&x, &u, 8); return x; };
  // about 4 buckets per plane over the [e2d[0], e2d[n-1]] range, at most 2^20
  int shift = 52, m = 0;
  long long base = 0, nbkt = 1;
  for (m = 20; m >= 0; --m) {
    shift = 52 - m;
    base = (long long)(bits(e2d.front()) >> shift);
    nbkt = (long long)(bits(e2d.back()) >> shift) - base + 1;
    if (nbkt <= std::max<long long>(64, 4LL * n) && nbkt <= (1 << 20)) break;
  }
  if (m < 0) { m = 0; shift = 52; base = (long long)(bits(e2d.front()) >> shift);
               nbkt = (long long)(bits(e2d.back()) >> shift) - base + 1; }
  std::vector<int> eb((size_t)nbkt + 1);
  for (long long b = 0; b <= nbkt; ++b) {
    const double edge = fromb((unsigned long long)(base + b) << shift);
    eb[(size_t)b] = (int)(std::upper_bound(e2d.begin(), e2d.end(), edge) - e2d.begin());
  }
  eb[(size_t)nbkt] = n;
  const int cn = (int)std::min<long long>(1 << 20, std::max<long long>(64, 4LL * n));
  const double scale = (double)cn / fdm.back();
  std::vector<int> cg((size_t)cn + 1);
  for (int g = 0; g < cn; ++g)
    cg[(size_t)g] = (int)(std::lower_bound(fdm.begin(), fdm.end(), (double)g / scale) - fdm.begin());
  cg[(size_t)cn] = n;
  DevBuf* b1 = p->newbuf(); b1->upload(eb.data(), eb.size());
  DevBuf* b2 = p->newbuf(); b2->upload(cg.data(), cg.size());
  p->pb.e_bucket = b1->as<int>();
  p->pb.e_base = base;
  p->pb.e_shift = shift;
  p->pb.e_nbkt = (int)nbkt;
  p->pb.c_guide = b2->as<int>();
  p->pb.c_n = cn;
  p->pb.c_scale = scale;
}

void pb_upload(Proc* p, const std::vector<double>& e2d, const std::vector<double>& fdm,
               const std::vector<double>& contrib, bool fp32);

Proc* create_pb(double v0n, const double* dsp, const double* fsqm, int64_t n, bool fp32) {
  ensure_device();
  if (!(v0n > 0)) raise("BadInput", "v0_times_natoms is not a positive number.");
  std::vector<std::pair<double, double>> data;
  data.reserve((size_t)n);
  for (int64_t i = 0; i < n; ++i) {
    if (fsqm[i] < 0) raise("CalcError", "Inconsistent data implies negative |F|^2*multiplicity.");
    if (data.empty() || data.back().first != dsp[i]) data.emplace_back(dsp[i], fsqm[i]);
    else data.back().second += fsqm[i];
  }
  double xsectfact = 0.5 / v0n;
  xsectfact *= kWl2Ekin / (1.0 * 1.0);
  std::sort(data.begin(), data.end(), std::greater<std::pair<double, double>>());
  std::unique_ptr<Proc> p(new Proc());
  p->kind = kKindPB;
  p->pb_v0n = v0n;
  std::vector<double> e2d, fdm, contrib;
  StableSum s;
  double prev = -__builtin_huge_val();
  for (auto& it : data) {
    if (!(it.first > 0.0)) raise("CalcError", "Inconsistent plane data implies non-positive (or NaN) d_spacing.");
    const double c = it.first * it.second * xsectfact;
    if (fabs(prev - it.first) < 1e-11) {
      s.add(c);
      fdm.back() = s.sum();
      contrib.back() += c;
    } else {
      prev = it.first;
      s.add(c);
      fdm.push_back(s.sum());
      contrib.push_back(c);
      const double wl = 2.0 * it.first;
      e2d.push_back(kWl2Ekin / (wl * wl));
    }
  }
  if (fdm.empty() || fdm.back() <= 0.0) { fdm.clear(); e2d.clear(); contrib.clear(); }
  pb_upload(p.get(), e2d, fdm, contrib, fp32);
  return p.release();
}

// Device tables of a PowderBragg process from its host tables: 2dE ascending,
// cumulative fdm, and the per-plane contributions the FP32 cumulative is
// summed from.
void pb_upload(Proc* p, const std::vector<double>& e2d, const std::vector<double>& fdm,
               const std::vector<double>& contrib, bool fp32) {
  p->pb.n = (int)fdm.size();
  p->pb.threshold = fdm.empty() ? __builtin_huge_val() : e2d.front();
  DevBuf* be = p->newbuf();
  be->upload(e2d.data(), e2d.size());
  DevBuf* bf = p->newbuf();
  bf->upload(fdm.data(), fdm.size());
  p->pb.e2d = be->as<double>();
  p->pb.fdm = bf->as<double>();
  p->pb.fdm32 = nullptr;
  if (!fdm.empty()) pb_guides(p, e2d, fdm);
  p->pb_e2d_h = e2d;
  p->pb_fdm_h = fdm;
  if (fp32 && !fdm.empty()) {
    DevBuf* bc = p->newbuf();
    bc->upload(contrib.data(), contrib.size());
    DevBuf* b32 = p->newbuf();
    b32->alloc(sizeof(float) * contrib.size());
    hipLaunchKernelGGL(k_fp32_cumul_scan, dim3(1), dim3(kCumulBlock), 0, 0, bc->as<double>(), (int)contrib.size(),
                       b32->as<float>());
    NCAMD_HIP(hipGetLastError());
    p->pb_fdm32_h.resize(contrib.size());
    NCAMD_HIP(hipMemcpy(p->pb_fdm32_h.data(), b32->p, sizeof(float) * contrib.size(), hipMemcpyDeviceToHost));
    p->pb.fdm32 = b32->as<float>();
    p->has_fp32 = true;
  }
  init_errbuf(p);
}

// ------------------------------------------------------------------ ElInc ---
Proc* create_elinc(int n, const double* msd, const double* bixs, const double* scale) {
  ensure_device();
  if (n < 1 || n > kMaxElInc) raise("BadInput", "ElIncScatter supports 1..16 elements");
  std::unique_ptr<Proc> p(new Proc());
  p->kind = kKindElInc;
  p->ei.n = n;
  for (int i = 0; i < n; ++i) {
    if (!(msd[i] >= 0.0 && msd[i] < 1e6) || !(bixs[i] >= 0.0 && bixs[i] < 1e6) || !(scale[i] >= 0.0 && scale[i] <= 1e6))
      raise("BadInput", "ElIncScatter element data out of range");
    p->ei.msd[i] = msd[i];
    p->ei.bixs[i] = bixs[i] * scale[i];
  }
  init_errbuf(p.get());
  return p.release();
}

// ------------------------------------------------------- merged processes ---
// A composition absorbs the scales of two processes of one kind into a single
// process where the kind allows it (ProcImpl::Process::createMerged).

// PowderBragg: ncrystal_core/src/powderbragg/NCPowderBragg.cc:218-331.  The
// 2dE lists are walked in ascending order; points whose d-spacings differ by
// less than 1e-11 Aa become one point at the mean 2dE; the cumulative fdm is
// rebuilt with a stable sum of the scaled per-point differences.
Proc* merge_pb(const Proc* a, double sa, const Proc* b, double sb) {
  const auto& a_e = a->pb_e2d_h;
  const auto& b_e = b->pb_e2d_h;
  std::vector<double> e2d, fdm, contrib;
  const bool fp32 = a->has_fp32 && b->has_fp32;
  auto diff = [](const std::vector<double>& c, size_t i) { return i ? c[i] - c[i - 1] : c[0]; };
  auto dsp = [](double e) { return 0.5 * (e ? std::sqrt(kWl2Ekin / e) : __builtin_huge_val()); };
  if (a_e.empty() || b_e.empty()) {
    const Proc* o = a_e.empty() ? b : a;
    const double so = a_e.empty() ? sb : sa;
    e2d = o->pb_e2d_h;
    for (size_t i = 0; i < e2d.size(); ++i) {
      fdm.push_back(o->pb_fdm_h[i] * so);
      contrib.push_back(so * diff(o->pb_fdm_h, i));
    }
  } else {
    StableSum s;
    size_t i = 0, j = 0;
    auto take = [&](const Proc* o, double so, size_t k) {
      const double c = so * diff(o->pb_fdm_h, k);
      e2d.push_back(o->pb_e2d_h[k]);
      s.add(c);
      fdm.push_back(s.sum());
      contrib.push_back(c);
    };
    while (i < a_e.size() && j < b_e.size()) {
      if (std::fabs(dsp(a_e[i]) - dsp(b_e[j])) < 1e-11) {
        const double ca = sa * diff(a->pb_fdm_h, i), cb = sb * diff(b->pb_fdm_h, j);
        e2d.push_back(0.5 * (a_e[i] + b_e[j]));
        s.add(ca);
        s.add(cb);
        fdm.push_back(s.sum());
        contrib.push_back(ca + cb);
        ++i, ++j;
      } else if (a_e[i] < b_e[j]) {
        take(a, sa, i++);
      } else {
        take(b, sb, j++);
      }
    }
    while (i < a_e.size()) take(a, sa, i++);
    while (j < b_e.size()) take(b, sb, j++);
  }
  std::unique_ptr<Proc> p(new Proc());
  p->kind = kKindPB;
  pb_upload(p.get(), e2d, fdm, contrib, fp32);
  return p.release();
}

// ElIncScatter: ncrystal_core/src/phys_utils/NCElIncXS.cc:204-233 and
// NCElIncScatter.cc:211-223.  The (msd, scaled strength) pairs of both lists,
// zero strengths dropped, are sorted, and neighbours whose msd agree within
// |a-b| <= 0.5e-15 (|a|+|b|) + 1e-6 (NCMath.hh:538-546) are summed.
Proc* merge_elinc(const Proc* a, double sa, const Proc* b, double sb) {
  std::vector<std::pair<double, double>> t;
  for (int i = 0; i < a->ei.n; ++i)
    if (a->ei.bixs[i] * sa != 0.0) t.emplace_back(a->ei.msd[i], a->ei.bixs[i] * sa);
  for (int i = 0; i < b->ei.n; ++i)
    if (b->ei.bixs[i] * sb != 0.0) t.emplace_back(b->ei.msd[i], b->ei.bixs[i] * sb);
  std::sort(t.begin(), t.end());
  std::vector<std::pair<double, double>> r;
  for (const auto& e : t) {
    const double x = r.empty() ? 0.0 : r.back().first;
    const bool same = !r.empty() && std::fabs(x - e.first) <= 0.5e-15 * (std::fabs(x) + std::fabs(e.first)) + 1e-6;
    if (same) r.back().second += e.second;
    else r.push_back(e);
  }
  if (r.empty()) raise("BadInput", "merged ElIncScatter has no element with a nonzero cross section");
  if ((int)r.size() > kMaxElInc) raise("BadInput", "merged ElIncScatter exceeds 16 elements");
  std::unique_ptr<Proc> p(new Proc());
  p->kind = kKindElInc;
  p->ei.n = (int)r.size();
  for (size_t i = 0; i < r.size(); ++i) {
    p->ei.msd[i] = r[i].first;
    p->ei.bixs[i] = r[i].second;
  }
  init_errbuf(p.get());
  return p.release();
}

// 1/v absorption: ncrystal_core/src/absoov/NCAbsOOV.cc:60-70.
Proc* merge_abs(const Proc* a, double sa, const Proc* b, double sb) {
  std::unique_ptr<Proc> p(new Proc());
  p->kind = kKindAbs;
  p->abs_c = a->abs_c * sa + b->abs_c * sb;
  init_errbuf(p.get());
  return p.release();
}

// The merged process of a and b (scales sa, sb), or null when the kinds do
// not merge: different kinds, and the kinds without a createMerged
// (S(alpha,beta) of distinct tables, SCBragg, FreeGas).
Proc* try_merge(const Proc* a, double sa, const Proc* b, double sb) {
  if (a->kind != b->kind) return nullptr;
  switch (a->kind) {
    case kKindPB: return merge_pb(a, sa, b, sb);
    case kKindElInc: return merge_elinc(a, sa, b, sb);
    case kKindAbs: return merge_abs(a, sa, b, sb);
    default: return nullptr;
  }
}

// ------------------------------------------------------------------- SAB ---
// egrid: ncrystal_core/src/sab/NCSABIntegrator.cc:147-281,
// geomspace/logspace: ncrystal_core/src/utils/NCMath.cc:43-66
std::vector<double> geomspace(double start, double stop, unsigned num) {
  std::vector<double> v(num);
  double ls = log10(start);
  const double lst = log10(stop);
  const double interval = (lst - ls) / (num - 1);
  for (unsigned i = 0; i < num; ++i) { v[i] = pow(10.0, ls); ls += interval; }
  v[num - 1] = pow(10.0, lst);
  v[0] = start;
  v[num - 1] = stop;
  return v;
}

// transformKernelToStdFormat (ncrystal_core/src/sab/NCSABUtils.cc:134-325)
// with the checks of validateScatKnlData (NCScatKnlData.cc:26-80): trim
// all-zero edges (detail_trimZeroEdgesFromKernel, NCSABUtils.cc:30-123), expand
// a symmetric table to negative beta (expandBetaAndSABToAllBetas,
// NCSABUtils.cc:331-380), unscale S = S_scaled exp(-beta/2) (in log space
// where exp(-beta/2) overflows), and thicken beta grids of fewer than
// min_nbeta rows unless the grid is marked optimised.  Kernel types:
// 0 S(alpha,beta), 1 scaled S, 2 scaled symmetric S (beta >= 0).
struct SabStd {
  std::vector<double> alpha, beta, sab;
  int64_t ntrimmed = 0;
};
SabStd sab_std_format(int type, bool beta_optimised, int na, int nb, const double* a, const double* b,
                      const double* s, int min_nbeta) {
  if (type < 0 || type > 2) raise("BadInput", "kernel type must be 0 (S), 1 (scaled S) or 2 (scaled symmetric S)");
  for (int pass = 0; pass < 2; ++pass) {
    const int n = pass ? nb : na;
    const double* g = pass ? b : a;
    const char* lbl = pass ? "beta" : "alpha";
    if (n < 5 || n > 65534)
      raise("BadInput", std::string("invalid ") + lbl + " grid (must have 5 to 65534 entries)");
    for (int i = 0; i < n; ++i)
      if (!std::isfinite(g[i]) || (i && !(g[i] > g[i - 1])))
        raise("BadInput", std::string("invalid ") + lbl + " grid (must consist of sorted, unique, regular numbers)");
  }
  if (!(a[0] > 0.0)) raise("BadInput", "non-positive entries in alpha grid");
  if (type == 2 ? b[0] != 0.0 : !(b[0] < 0.0))
    raise("BadInput", type == 2 ? "a symmetric table's beta grid must start with 0.0"
                                : "beta grid must start with a negative entry");
  for (size_t i = 0; i < (size_t)na * nb; ++i)
    if (!(s[i] >= 0.0) || std::isinf(s[i])) raise("BadInput", "S values must be finite and non-negative");
  SabStd t;
  // trim all-zero edges: upper alpha columns, upper (beta > 0) and lower (beta < 0) rows
  auto row_zero = [&](int j) {
    for (int i = 0; i < na; ++i)
      if (s[(size_t)j * na + i] != 0.0) return false;
    return true;
  };
  auto col_zero = [&](int i) {
    for (int j = 0; j < nb; ++j)
      if (s[(size_t)j * na + i] != 0.0) return false;
    return true;
  };
  int ta = 0, tbu = 0, tbl = 0;
  for (int i = na - 1; i >= 0 && a[i] > 0.0 && col_zero(i); --i) ++ta;
  for (int j = nb - 1; j >= 0 && b[j] > 0.0 && row_zero(j); --j) ++tbu;
  for (int j = 0; j < nb && b[j] < 0.0 && row_zero(j); ++j) ++tbl;
  if (ta >= na) ta = tbu = tbl = 0;
  t.ntrimmed = ta + tbu + tbl;
  const int na2 = na - ta;
  t.alpha.assign(a, a + na2);
  t.beta.assign(b + tbl, b + (nb - tbu));
  for (int j = tbl; j < nb - tbu; ++j) t.sab.insert(t.sab.end(), s + (size_t)j * na, s + (size_t)j * na + na2);
  const size_t nal = t.alpha.size();
  if (type == 2) {
    // S(alpha, -beta) = S(alpha, beta); the zero row appears once
    const size_t nh = t.beta.size(), npos = nh - 1;
    std::vector<double> fb, fs;
    fb.reserve(2 * npos + 1);
    for (size_t k = nh; k-- > 0;) fb.push_back(-t.beta[k]);
    fb.back() = 0.0;
    for (size_t k = 1; k < nh; ++k) fb.push_back(t.beta[k]);
    fs.reserve((2 * npos + 1) * nal);
    for (size_t k = 0; k < npos; ++k)
      fs.insert(fs.end(), t.sab.begin() + (npos - k) * nal, t.sab.begin() + (npos - k + 1) * nal);
    fs.insert(fs.end(), t.sab.begin(), t.sab.end());
    t.beta.swap(fb);
    t.sab.swap(fs);
    type = 1;
  }
  if (type == 1) {
    for (size_t j = 0; j < t.beta.size(); ++j) {
      const double ea = -0.5 * t.beta[j];
      double* row = &t.sab[j * nal];
      if (ea < 700.0) {
        const double f = std::exp(ea);
        for (size_t i = 0; i < nal; ++i) row[i] *= f;
      } else {
        for (size_t i = 0; i < nal; ++i) {
          if (row[i] == 0.0) continue;
          const double c = ea + std::log(row[i]);
          if (!(c < 700.0)) raise("BadInput", "unscaling S(alpha,beta) overflows (exp of " + std::to_string(c) + ")");
          row[i] = std::exp(c);
        }
      }
    }
  }
  const size_t nbt = t.beta.size();
  if (!beta_optimised && nbt < (size_t)min_nbeta) {
    const size_t nextra = (size_t)min_nbeta / nbt;
    std::vector<double> fb, fs;
    fb.reserve(1 + (nbt - 1) * (1 + nextra));
    fs.reserve(fb.capacity() * nal);
    for (size_t j = 0; j + 1 < nbt; ++j) {
      const double b0 = t.beta[j], b1 = t.beta[j + 1];
      const double* r0 = &t.sab[j * nal];
      const double* r1 = &t.sab[(j + 1) * nal];
      fb.push_back(b0);
      fs.insert(fs.end(), r0, r0 + nal);
      const double db = (b1 - b0) / (double)(nextra + 1);
      for (size_t e = 0; e < nextra; ++e) {
        const double bb = b0 + (double)(e + 1) * db;
        fb.push_back(bb);
        for (size_t i = 0; i < nal; ++i) {
          const double k = (bb - b0) / (b1 - b0);
          fs.push_back(r0[i] * (1.0 - k) + k * r1[i]);
        }
      }
    }
    fs.insert(fs.end(), t.sab.end() - nal, t.sab.end());
    fb.push_back(t.beta.back());
    t.beta.swap(fb);
    t.sab.swap(fs);
  }
  return t;
}

Proc* create_sab(int na, int nb, const double* alpha, const double* beta, const double* sab, double T,
                 double mass, double sigma_bound, double suggested_emax, int n_egrid, const double* egrid_in) {
  ensure_device();
  if (na < 2 || nb < 2) raise("BadInput", "S(alpha,beta) grids need at least 2 points");
  if (na >= 65535) raise("BadInput", "alpha grid too large");
  for (int i = 1; i < na; ++i) if (!(alpha[i] > alpha[i - 1])) raise("BadInput", "alpha grid not increasing");
  for (int i = 1; i < nb; ++i) if (!(beta[i] > beta[i - 1])) raise("BadInput", "beta grid not increasing");
  for (size_t i = 0; i < (size_t)na * nb; ++i)
    if (!(sab[i] >= 0.0) || std::isinf(sab[i])) raise("BadInput", "S(alpha,beta) must be finite and non-negative");
  if (!(T > 0.0) || !(mass > 0.0) || !(sigma_bound > 0.0)) raise("BadInput", "invalid T, mass or sigma");
  std::unique_ptr<Proc> p(new Proc());
  p->kind = kKindSAB;
  p->sab_na = na;
  p->sab_nb = nb;
  SABView& s = p->sab;
  s.na = na;
  s.nb = nb;
  s.T = T;
  s.kT = kBoltzmann * T;
  s.mass = mass;
  {
    const double r = mass / (kNeutronMassAmu + mass);
    s.sigma_free = (r * r) * sigma_bound;
    const double A = (1.0 / kNeutronMassAmu) * mass;
    s.fg_A = A;
    s.fg_ca = A / s.kT;
  }
  s.egrid_margin = 1.05;
  p->sab_alpha_h.assign(alpha, alpha + na);
  p->sab_beta_h.assign(beta, beta + nb);
  DevBuf* ba = p->newbuf(); ba->upload(alpha, (size_t)na);
  DevBuf* bb = p->newbuf(); bb->upload(beta, (size_t)nb);
  DevBuf* bs = p->newbuf(); bs->upload(sab, (size_t)na * nb);
  DevBuf* bsl = p->newbuf(); bsl->alloc(sizeof(double2) * (size_t)na * nb);
  DevBuf* bcu = p->newbuf(); bcu->alloc(sizeof(double) * (size_t)na * nb);
  SABBuildBufs B{};
  B.na = na; B.nb = nb;
  B.alpha = ba->as<double>(); B.beta = bb->as<double>(); B.sab = bs->as<double>();
  B.sl = bsl->as<double2>(); B.cumul = bcu->as<double>();
  B.kT = s.kT; B.sigma_bound = sigma_bound;
  launch_sab_derive(B, 0);
  NCAMD_HIP(hipGetLastError());

  DevBuf tmpE, tmpX, tmpF;
  auto xs_points = [&](const std::vector<double>& es, std::vector<double>& xs, std::vector<double>* fg) {
    tmpE.upload(es.data(), es.size());
    tmpX.alloc(sizeof(double) * es.size());
    launch_sab_xs_points(B, tmpE.as<double>(), (int)es.size(), tmpX.as<double>(), 0);
    NCAMD_HIP(hipGetLastError());
    xs.resize(es.size());
    NCAMD_HIP(hipMemcpy(xs.data(), tmpX.p, sizeof(double) * es.size(), hipMemcpyDeviceToHost));
    if (fg) {
      tmpF.alloc(sizeof(double) * es.size());
      launch_fg_xs(s.sigma_free, s.fg_ca, tmpE.as<double>(), (int)es.size(), tmpF.as<double>(), 0);
      NCAMD_HIP(hipGetLastError());
      fg->resize(es.size());
      NCAMD_HIP(hipMemcpy(fg->data(), tmpF.p, sizeof(double) * es.size(), hipMemcpyDeviceToHost));
    }
  };

  std::vector<double> egrid;
  if (n_egrid > 3) {
    egrid.assign(egrid_in, egrid_in + n_egrid);
  } else {
    if (n_egrid != 0 && n_egrid != 3)
      raise("BadInput", "SABIntegrator invalid energy grid. It must either be a complete array, empty, or {emin, emax, npts}");
    double emin = 0, emax = 0;
    unsigned npts = 0;
    if (n_egrid == 3) {
      emin = egrid_in[0]; emax = egrid_in[1];
      npts = (unsigned)egrid_in[2];
      if ((double)npts != egrid_in[2]) raise("BadInput", "egrid npts must be integral");
    }
    if (!(emin >= 0.0) || !(emax >= 0.0) || !(emax == 0.0 || emin == 0.0 || emax > emin))
      raise("BadInput", "SABIntegrator invalid energy grid values");
    if (npts == 0) npts = 300;
    if (emax == 0.0 && suggested_emax > 0.0) emax = suggested_emax;
    if (emax == 0.0) {
      // determineEMax: candidates eul*0.95^k down to eul*1e-4
      const double bmin = beta[0], amax = alpha[na - 1];
      const double eul = s.kT * (bmin - amax) * (bmin - amax) / (4 * amax);
      std::vector<double> cand;
      for (double e = eul, elow = eul * 1e-4; e > elow; e *= 0.95) cand.push_back(e);
      std::vector<double> xs, fg;
      xs_points(cand, xs, &fg);
      double prev = __builtin_huge_val();
      emax = 0.0;
      for (size_t k = 0; k < cand.size(); ++k) {
        const double dist = fabs(xs[k] - fg[k]);
        if (dist > prev) { emax = 0.95 * cand[k]; break; }
        prev = dist;
      }
      if (!(emax > 0.0)) emax = 0.5 * eul;
    }
    if (emin == 0.0) {
      // determineEMin: candidates e_upp*0.9 halved down to 1e-30*e_upp
      const double e_upp = std::min(emax * 0.01, 0.01 * s.kT);
      std::vector<double> cand;
      double e = e_upp * 0.9;
      cand.push_back(e);
      while (true) {
        const double en = 0.5 * e;
        if (en <= 1e-30 * e_upp) break;
        cand.push_back(en);
        e = en;
      }
      std::vector<double> xs;
      xs_points(cand, xs, nullptr);
      double e_old = cand[0], f_old = sqrt(e_old) * xs[0];
      emin = -1.0;
      for (size_t k = 1; k < cand.size(); ++k) {
        const double f_new = sqrt(cand[k]) * xs[k];
        if (f_new == 0.0) { emin = 0.001 * e_upp; break; }
        if (fabs(f_old / f_new - 1.0) < 1e-3) { emin = e_old; break; }
        e_old = cand[k];
        f_old = f_new;
      }
      if (emin < 0.0) emin = std::min(e_old, e_upp * 0.01);
    } else if (!(emax > emin)) {
      raise("BadInput", "energy grid does not have emax>emin");
    }
    egrid = geomspace(emin, emax, npts);
  }
  if (egrid.size() < 10) raise("BadInput", "SABIntegrator invalid energy grid - must have at least 10 points.");
  for (size_t i = 1; i < egrid.size(); ++i)
    if (!(egrid[i] > egrid[i - 1])) raise("BadInput", "SABIntegrator invalid energy grid - must be sorted");
  if (!(egrid.front() > 0.0)) raise("BadInput", "SABIntegrator invalid energy grid - must be positive");

  const int ne = (int)egrid.size();
  s.ne = ne;
  DevBuf* beg = p->newbuf(); beg->upload(egrid.data(), egrid.size());
  DevBuf* bxs = p->newbuf(); bxs->alloc(sizeof(double) * ne);
  DevBuf* bhdr = p->newbuf(); bhdr->alloc(sizeof(int4) * ne);
  DevBuf* bfst = p->newbuf(); bfst->alloc(sizeof(double) * ne);
  DevBuf* bcdf = p->newbuf(); bcdf->alloc(sizeof(double) * (size_t)ne * (nb + 1));
  DevBuf* bxy = p->newbuf(); bxy->alloc(sizeof(double2) * (size_t)ne * (nb + 1));
  DevBuf* binf = p->newbuf(); binf->alloc(sizeof(AInfo) * (size_t)ne * nb);
  NCAMD_HIP(hipMemset(bcdf->p, 0, bcdf->bytes));
  NCAMD_HIP(hipMemset(bxy->p, 0, bxy->bytes));
  NCAMD_HIP(hipMemset(binf->p, 0, binf->bytes));
  launch_sab_analyse(B, beg->as<double>(), ne, bxs->as<double>(), bhdr->as<int4>(), bfst->as<double>(),
                     bcdf->as<double>(), bxy->as<double2>(), binf->as<AInfo>(), 0);
  NCAMD_HIP(hipGetLastError());
  std::vector<double> xs(ne);
  NCAMD_HIP(hipMemcpy(xs.data(), bxs->p, sizeof(double) * ne, hipMemcpyDeviceToHost));
  std::vector<int4> hdr(ne);
  NCAMD_HIP(hipMemcpy(hdr.data(), bhdr->p, sizeof(int4) * ne, hipMemcpyDeviceToHost));
  for (int i = 0; i < ne; ++i)
    if (hdr[i].x < 0) raise("CalcError", "S(alpha,beta) beta distribution has no area");
  // high-E constants (NCSABXSProvider.cc:35-51, NCSABSampler.cc:40-57)
  {
    std::vector<double> em{egrid.back()}, dummy, fg;
    tmpE.upload(em.data(), 1);
    tmpF.alloc(sizeof(double));
    launch_fg_xs(s.sigma_free, s.fg_ca, tmpE.as<double>(), 1, tmpF.as<double>(), 0);
    NCAMD_HIP(hipGetLastError());
    double extxs = 0.0;
    NCAMD_HIP(hipMemcpy(&extxs, tmpF.p, sizeof(double), hipMemcpyDeviceToHost));
    const double emax = egrid.back();
    s.k_ext = (xs.back() - extxs) * emax;
    s.k1 = xs.back() * emax;
    s.k2 = extxs * emax;
  }
  // egrid IEEE-bit bucket table: buckets narrower than half the smallest
  // relative egrid spacing so a bracket holds at most ~2 grid points
  {
    double rmin = 1e300;
    for (size_t i = 1; i < egrid.size(); ++i) rmin = std::min(rmin, egrid[i] / egrid[i - 1] - 1.0);
    int m = (int)std::ceil(std::log2(2.0 / std::max(rmin, 1e-12)));
    m = std::max(0, std::min(m, 20));
    auto bits = [](double x) { unsigned long long u; std::memcpy(&u, &x, 8); return u; };
    auto fromb = [](unsigned long long u) { double x; std::memcpy(&x, &u, 8); return x; };
    int shift;
    long long base, nbkt;
    while (true) {
      shift = 52 - m;
      base = (long long)(bits(egrid.front()) >> shift);
      nbkt = (long long)(bits(egrid.back()) >> shift) - base + 1;
      if (nbkt <= (1 << 20) || m == 0) break;
      --m;
    }
    std::vector<int> tab((size_t)nbkt + 1);
    for (long long b = 0; b <= nbkt; ++b) {
      const double edge = fromb((unsigned long long)(base + b) << shift);
      tab[(size_t)b] = (int)(std::upper_bound(egrid.begin(), egrid.end(), edge) - egrid.begin());
    }
    DevBuf* btab = p->newbuf();
    btab->upload(tab.data(), tab.size());
    s.eg_bucket = btab->as<int>();
    s.eg_nbkt = (int)nbkt;
    s.eg_shift = shift;
    s.eg_base = base;
  }
  s.alpha = ba->as<double>();
  s.beta = bb->as<double>();
  s.sl = bsl->as<double2>();
  s.cumul = bcu->as<double>();
  s.egrid = beg->as<double>();
  s.xs = bxs->as<double>();
  s.shdr = bhdr->as<int4>();
  s.sfirst = bfst->as<double>();
  s.scdf = bcdf->as<double>();
  s.sxy = bxy->as<double2>();
  s.sinfo = binf->as<AInfo>();
  {
    DevBuf* bcg = p->newbuf(); bcg->alloc(sizeof(unsigned short) * (size_t)ne * (kCdfGuide + 1));
    DevBuf* brg = p->newbuf(); brg->alloc(sizeof(unsigned short) * (size_t)nb * (kRowGuide + 1));
    DevBuf* brs = p->newbuf(); brs->alloc(sizeof(double) * (size_t)nb);
    if (nb + 1 > 65535) raise("BadInput", "beta grid too large (at most 65534 points)");
    launch_sab_guides(s, bcg->as<unsigned short>(), brg->as<unsigned short>(), brs->as<double>(), 0);
    NCAMD_HIP(hipGetLastError());
    NCAMD_HIP(hipDeviceSynchronize());
    s.cdf_guide = bcg->as<unsigned short>();
    s.row_guide = brg->as<unsigned short>();
    s.row_scale = brs->as<double>();
  }
  p->sab_egrid_h = egrid;
  p->sab_xs_h = xs;
  init_errbuf(p.get());
  return p.release();
}

// ------------------------------------------------------------------- hkl ---
// ncrystal_core/src/extd_utils/NCFillHKL.cc:34-49 (keygen), :230-450 (loop and
// family grouping); ncrystal_core/src/utils/NCLatticeUtils.cc:177-216
// (estimateHKLRange); sort ncrystal_core/src/infobld/NCInfoBuilder.cc:98-117,452;
// HKL ordering ncrystal_core/include/NCrystal/core/NCTypes.hh:978-981.
unsigned famkey(double fsq, double dsp) {
  const int ef = (int)std::ceil(std::log10(fsq));
  const double mf = fsq * std::pow(10.0, (double)(-ef));
  const int ed = (int)std::ceil(std::log10(dsp));
  const double md = dsp * std::pow(10.0, (double)(-ed));
  const unsigned part = (unsigned)(mf * 1000 + 0.5) * 4000000u + (unsigned)(md * 1000 + 0.5) * 4000u;
  return (unsigned)((double)part + std::max((double)(3000 + ef * 30 + ed), 0.0));
}
int floorhkl(double x) {
  const double xc = std::floor(x * (1.0 + 1e-12));
  if (xc <= 1.0) return 1;
  if (xc >= 2147483647.0) return 2147483647;
  return (int)(xc + 0.5);
}
struct HKL3 {
  int h, k, l;
};
bool hkl_less(const HKL3& a, const HKL3& o) {
  return (a.h != o.h ? o.h < a.h : (a.k != o.k ? o.k < a.k : o.l < a.l));
}
struct Fam {
  HKL3 rep;
  int mult;
  double fsq, dsp;
  std::vector<HKL3> members;
};

// species with a non-zero coherent scattering length (NCFillHKL.cc:175-186)
// and the per-species early-abort thresholds of the Debye-Waller factors
// (:203-209), uploaded for the |F|^2 kernels; false if there are none
struct HKLSpecies {
  DevBuf pos, csl, msd, thr, off;
};
bool hkl_species(HKLBuild& P, HKLSpecies& S, int nspecies, const int* natoms, const double* positions,
                 const double* csl, const double* msd, double dcutoff, double fsquarecut) {
  if (fsquarecut >= 0.0) fsquarecut = std::max(fsquarecut, 1.0e-300);
  P.fsquarecut = fsquarecut;
  std::vector<double> vpos, vcsl, vmsd, vthr;
  std::vector<int> off{0};
  const double* pp = positions;
  for (int i = 0; i < nspecies; ++i) {
    if (fabs(csl[i]) > 0.0) {
      vpos.insert(vpos.end(), pp, pp + 3 * natoms[i]);
      vcsl.push_back(csl[i]);
      vmsd.push_back(msd[i]);
      vthr.push_back((fsquarecut < 0.01 && fsquarecut > 1.0e-300) ? std::log(fabs(csl[i]) / fsquarecut)
                                                                  : __builtin_huge_val());
      off.push_back(off.back() + natoms[i]);
    }
    pp += 3 * natoms[i];
  }
  const int ns = (int)vcsl.size();
  if (ns > 32) raise("BadInput", "at most 32 species with non-zero coherent scattering length");
  if (ns == 0) return false;
  // ksq and d intervals (NCFillHKL.cc:220-224)
  const double DMIN = 2.2250738585072014e-308, DMAX = 1.7976931348623157e308;
  P.ksq_lo = std::min(std::max((39.4784176043574344753379639995046045412547976 * (1.0 - 1e-14)) /
                                   (__builtin_huge_val() * __builtin_huge_val()), DMIN), DMAX);
  P.ksq_hi = std::min(std::max((39.4784176043574344753379639995046045412547976 * (1.0 + 1e-14)) /
                                   (dcutoff * dcutoff), DMIN), DMAX);
  P.dcut_lo = std::min(std::max(dcutoff, DMIN), DMAX);
  P.dcut_hi = DMAX;
  S.pos.upload(vpos.data(), vpos.size());
  S.csl.upload(vcsl.data(), vcsl.size());
  S.msd.upload(vmsd.data(), vmsd.size());
  S.thr.upload(vthr.data(), vthr.size());
  S.off.upload(off.data(), off.size());
  P.nspecies = ns;
  P.pos = S.pos.as<double>(); P.csl = S.csl.as<double>(); P.msd = S.msd.as<double>();
  P.thr = S.thr.as<double>(); P.pos_off = S.off.as<int>();
  return true;
}

// InfoBuilder's family order (NCInfoBuilder.cc:98-117, std::stable_sort at
// :452) and the output arrays; members (first = representative) if asked
int64_t hkl_output(std::vector<Fam>& fams, int64_t maxout, int* out_hkl, int* out_mult, double* out_dsp,
                   double* out_fsq, int64_t* nrequired, int64_t maxmem, int64_t* out_mem_off, int* out_mem_hkl,
                   int64_t* nmem_required) {
  auto before = [](const Fam& rh, const Fam& lh) {
    if (fabs(lh.dsp - rh.dsp) > 1e-6) return lh.dsp < rh.dsp;
    const double lfm = lh.fsq * lh.mult, rfm = rh.fsq * rh.mult;
    if (fabs(lfm - rfm) > 1e-6) return lfm < rfm;
    if (lh.mult != rh.mult) return lh.mult < rh.mult;
    return hkl_less(lh.rep, rh.rep);
  };
  std::stable_sort(fams.begin(), fams.end(), before);
  const int64_t nf = (int64_t)fams.size();
  if (nrequired) *nrequired = nf;
  if (nf > maxout) return -1;
  int64_t nmem = 0;
  for (const auto& f : fams) nmem += (int64_t)f.members.size();
  if (nmem_required) *nmem_required = nmem;
  if (out_mem_off && nmem > maxmem) return -1;
  for (int64_t i = 0; i < nf; ++i) {
    out_hkl[3 * i] = fams[i].rep.h; out_hkl[3 * i + 1] = fams[i].rep.k; out_hkl[3 * i + 2] = fams[i].rep.l;
    out_mult[i] = fams[i].mult; out_dsp[i] = fams[i].dsp; out_fsq[i] = fams[i].fsq;
  }
  if (out_mem_off) {
    int64_t m = 0;
    for (int64_t i = 0; i < nf; ++i) {
      out_mem_off[i] = m;
      for (const auto& x : fams[i].members) {
        out_mem_hkl[3 * m] = x.h; out_mem_hkl[3 * m + 1] = x.k; out_mem_hkl[3 * m + 2] = x.l;
        ++m;
      }
    }
    out_mem_off[nf] = m;
  }
  return nf;
}

bool reci_lattice(double a, double b, double c, double alpha_deg, double beta_deg, double gamma_deg, double* m);
void hkl_bounds(double a, double b, double c, double alpha_deg, double beta_deg, double gamma_deg, double dcutoff,
                int* mx);

// enumeration without symmetry (NCFillHKL.cc:230-450) for any cell: the
// reciprocal lattice of NCLatticeUtils.cc:95-149 times (h,k,l), loop bounds
// from the lattice rows
int64_t hkl_nosg(double a, double b, double c, double alpha, double beta, double gamma, int nspecies,
                 const int* natoms, const double* positions, const double* csl, const double* msd, double dcutoff,
                 double fsquarecut, double merge_tol, int64_t maxout, int* out_hkl, int* out_mult, double* out_dsp,
                 double* out_fsq, int64_t* nrequired, int64_t maxmem = 0, int64_t* out_mem_off = nullptr,
                 int* out_mem_hkl = nullptr, int64_t* nmem_required = nullptr) {
  ensure_device();
  if (!(dcutoff > 0.0)) raise("BadInput", "dcutoff must be positive");
  if (!(a > 0 && b > 0 && c > 0)) raise("BadInput", "lattice parameters must be positive");
  HKLBuild P{};
  if (!reci_lattice(a, b, c, alpha, beta, gamma, P.rm))
    raise("BadInput", "invalid cell: angles must lie in (0, 180) degrees and admit a cell");
  int mx[3];
  hkl_bounds(a, b, c, alpha, beta, gamma, dcutoff, mx);
  P.max_h = mx[0]; P.max_k = mx[1]; P.max_l = mx[2];
  HKLSpecies S;
  if (nrequired) *nrequired = 0;
  if (!hkl_species(P, S, nspecies, natoms, positions, csl, msd, dcutoff, fsquarecut)) return 0;
  const int64_t nl = 2 * (int64_t)P.max_l + 1, nkf = 2 * (int64_t)P.max_k + 1;
  const int64_t total = ((int64_t)P.max_k + 1) * nl + (int64_t)P.max_h * nkf * nl;
  const int64_t chunk = std::min<int64_t>(total, (int64_t)1 << 23);
  DevBuf dF, dD, dV;
  dF.alloc(sizeof(double) * chunk);
  dD.alloc(sizeof(double) * chunk);
  dV.alloc(chunk);
  std::vector<double> hF(chunk), hD(chunk);
  std::vector<unsigned char> hV(chunk);
  std::vector<Fam> fams;
  std::multimap<unsigned, size_t> fmap;
  const int64_t n_h0 = ((int64_t)P.max_k + 1) * nl;
  for (int64_t first = 0; first < total; first += chunk) {
    const int64_t cnt = std::min(chunk, total - first);
    launch_hkl_eval(P, first, cnt, dF.as<double>(), dD.as<double>(), dV.as<unsigned char>(), 0);
    NCAMD_HIP(hipGetLastError());
    NCAMD_HIP(hipMemcpy(hV.data(), dV.p, cnt, hipMemcpyDeviceToHost));
    NCAMD_HIP(hipMemcpy(hF.data(), dF.p, sizeof(double) * cnt, hipMemcpyDeviceToHost));
    NCAMD_HIP(hipMemcpy(hD.data(), dD.p, sizeof(double) * cnt, hipMemcpyDeviceToHost));
    for (int64_t t = 0; t < cnt; ++t) {
      if (!hV[t]) continue;
      const int64_t q = first + t;
      HKL3 cur;
      if (q < n_h0) {
        cur.h = 0; cur.k = (int)(q / nl); cur.l = (int)(q % nl) - P.max_l;
      } else {
        const int64_t qq = q - n_h0, per_h = nkf * nl;
        cur.h = 1 + (int)(qq / per_h);
        const int64_t r = qq % per_h;
        cur.k = (int)(r / nl) - P.max_k;
        cur.l = (int)(r % nl) - P.max_l;
      }
      const double fsq = hF[t], dsp = hD[t];
      const unsigned key = famkey(fsq, dsp);
      auto itLB = fmap.lower_bound(key);
      bool isnew = true;
      for (auto it = itLB; it != fmap.end() && it->first == key; ++it) {
        Fam& f = fams[it->second];
        if (fabs(fsq - f.fsq) < merge_tol * (fsq + f.fsq) && fabs(dsp - f.dsp) < merge_tol * (dsp + f.dsp)) {
          f.mult += 2;
          f.members.push_back(cur);
          isnew = false;
          break;
        }
      }
      if (isnew) {
        if (fams.size() > 1000000) return -2;
        Fam f;
        f.rep = cur; f.mult = 2; f.fsq = fsq; f.dsp = dsp;
        f.members.push_back(cur);
        fmap.insert(itLB, std::make_pair(key, fams.size()));
        fams.push_back(std::move(f));
      }
    }
  }
  for (auto& f : fams) {
    std::sort(f.members.begin(), f.members.end(), hkl_less);
    f.rep = f.members.front();
  }
  return hkl_output(fams, maxout, out_hkl, out_mult, out_dsp, out_fsq, nrequired, maxmem, out_mem_off,
                    out_mem_hkl, nmem_required);
}

// ---- space group: reflections grouped by symmetry equivalence
// (ncrystal_core/src/extd_utils/NCFillHKL.cc:545-709).
// Equivalent reflections by Laue class, the rule sets of
// ncrystal_core/src/phys_utils/NCEqRefl.cc:25-150: each entry is stored as
// the smaller (HKL ordering) of the pair (x, -x)
// (ncrystal_core/include/NCrystal/internal/phys_utils/NCEqRefl.hh:125-132).
HKL3 eq_norm(int h, int k, int l) {
  const HKL3 x{h, k, l}, xm{-h, -k, -l};
  return hkl_less(xm, x) ? xm : x;
}
int eqrefl_raw(int sg, int h, int k, int l, HKL3* o) {
  int n = 0;
  auto add = [&](int x, int y, int z) { o[n++] = eq_norm(x, y, z); };
  if (sg < 3) {  // triclinic
    add(h, k, l);
  } else if (sg < 16) {  // monoclinic
    add(h, k, l); add(h, -k, l);
  } else if (sg < 75) {  // orthorhombic
    add(h, k, l); add(h, -k, -l); add(h, -k, l); add(h, k, -l);
  } else if (sg < 89) {  // tetragonal 4/m
    add(h, k, l); add(k, -h, -l); add(h, k, -l); add(k, -h, l);
  } else if (sg < 143) {  // tetragonal 4/mmm
    add(h, k, l); add(k, h, l); add(k, -h, -l); add(h, k, -l);
    add(k, -h, l); add(h, -k, -l); add(k, h, -l); add(h, -k, l);
  } else if (sg < 149) {  // trigonal -3
    add(h, k, l); add(h + k, -h, -l); add(k, -h - k, l);
  } else if (sg < 168) {  // trigonal -3m
    add(h, k, l); add(h + k, -h, -l); add(k, -h - k, l); add(k, h, -l); add(h + k, -k, l); add(h, -h - k, -l);
  } else if (sg < 177) {  // hexagonal 6/m
    add(h, k, l); add(h, k, -l); add(k, -h - k, -l); add(k, -h - k, l); add(h + k, -h, l); add(h + k, -h, -l);
  } else if (sg < 195) {  // hexagonal 6/mmm
    add(h, k, l); add(k, -h - k, -l); add(h + k, -h, -l); add(h, k, -l); add(k, -h - k, l);
    add(h + k, -h, l); add(k, h, l); add(h + k, -k, l); add(h, -h - k, l);
    add(k, h, -l); add(h + k, -k, -l); add(h, -h - k, -l);
  } else if (sg < 207) {  // cubic m-3
    add(h, k, l); add(h, -k, -l); add(h, -k, l); add(h, k, -l);
    add(k, l, h); add(k, -l, -h); add(k, -l, h); add(k, l, -h);
    add(l, h, k); add(l, -h, -k); add(l, -h, k); add(l, h, -k);
  } else {  // cubic m-3m
    add(h, k, l); add(h, -k, -l); add(h, -k, l); add(h, k, -l);
    add(k, l, h); add(k, -l, -h); add(k, -l, h); add(k, l, -h);
    add(l, h, k); add(l, -h, -k); add(l, -h, k); add(l, h, -k);
    add(k, h, l); add(k, -h, -l); add(k, -h, l); add(k, h, -l);
    add(l, k, h); add(l, -k, -h); add(l, -k, h); add(l, k, -h);
    add(h, l, k); add(h, -l, -k); add(h, -l, k); add(h, l, -k);
  }
  return n;
}

// lattice matrix (row-major; angles in radians) of a cell outside the
// closed forms, as ncrystal_core/src/utils/NCLatticeUtils.cc:29-90
// (getLatticeRot) builds it: rows (a, 0, 0), (b cos g, b sin g, q),
// (c cos b, q, r), where q = c (cos a - cos b cos g) / sin g stands in both
// off-diagonal places (1,2) and (2,1) and r = c sqrt(sin^2 b - (q/c)^2); the
// reference's d-spacings and cell volumes of triclinic and rhombohedral
// cells follow from this matrix, so it is kept as it is.  Cosines and sines
// of 90 and 120 degrees are exact.
void lattice_general(double a, double b, double c, double al, double be, double ga, double* L) {
  constexpr double kSin120 = 0.86602540378443864676372317075293618347140262690519;
  const double tol = 1e-14;
  const bool g90 = fabs(ga - kPiHalf) < tol, g120 = fabs(ga - kPi * (2 / 3.)) < tol;
  const double cg = g90 ? 0.0 : (g120 ? -0.5 : std::cos(ga));
  const double sg = g90 ? 1.0 : (g120 ? kSin120 : std::sin(ga));
  const double ca = fabs(al - kPiHalf) < tol ? 0.0 : std::cos(al);
  const bool b90 = fabs(be - kPiHalf) < tol;
  const double cb = b90 ? 0.0 : std::cos(be), sb = b90 ? 1.0 : std::sin(be);
  const double t = (ca - cb * cg) / sg, q = c * t;
  const double r = q == 0.0 ? (cb == 0.0 ? c : c * sb) : c * std::sqrt(sb * sb - t * t);
  const double rows[9] = {a, 0.0, 0.0, b * cg, b * sg, q, c * cb, q, r};
  for (int i = 0; i < 9; ++i) L[i] = rows[i];
}

// hkl loop bounds: |h_i| = |row_i . G| / (2 pi) <= |row_i| / dcutoff with the
// rows of the lattice matrix (lengths a, b, c for the closed-form cells); a
// range covering the reference's estimate (NCLatticeUtils.cc:177-242, the
// same rows mapped over points of a sphere, times 1.05), so the same
// reflections are met in the same order
void hkl_bounds(double a, double b, double c, double alpha_deg, double beta_deg, double gamma_deg, double dcutoff,
                int* mx) {
  constexpr double kDeg = 0.0174532925199432957692369076848861271344287189;
  double L[9];
  lattice_general(a, b, c, alpha_deg * kDeg, beta_deg * kDeg, gamma_deg * kDeg, L);
  for (int i = 0; i < 3; ++i)
    mx[i] = floorhkl(std::sqrt(L[3 * i] * L[3 * i] + L[3 * i + 1] * L[3 * i + 1] + L[3 * i + 2] * L[3 * i + 2]) /
                     dcutoff);
}

// reciprocal lattice (row-major) of a cell given in degrees: the closed
// forms of ncrystal_core/src/utils/NCLatticeUtils.cc:105-144 (all angles 90;
// alpha = beta = 90 and gamma = 120; alpha = gamma = 90), else 2 pi times the
// inverse of the lattice matrix (NCLatticeUtils.cc:145-148; an adjugate
// inverse, equal to the reference's Gauss-Jordan one to rounding); false
// for angles outside (0, 180) degrees or that admit no cell
bool reci_lattice(double a, double b, double c, double alpha_deg, double beta_deg, double gamma_deg, double* m) {
  constexpr double kDeg = 0.0174532925199432957692369076848861271344287189;
  const double al = alpha_deg * kDeg, be = beta_deg * kDeg, ga = gamma_deg * kDeg;
  const double tol = 1e-14;
  const bool a90 = fabs(al - kPiHalf) < tol, b90 = fabs(be - kPiHalf) < tol, g90 = fabs(ga - kPiHalf) < tol;
  const bool g120 = fabs(ga - kPi * (2 / 3.)) < tol;
  for (int i = 0; i < 9; ++i) m[i] = 0.0;
  if (a90 && b90 && g90) {
    m[0] = k2Pi / a; m[4] = k2Pi / b; m[8] = k2Pi / c;
  } else if (a90 && b90 && g120) {
    constexpr double k2PiDivSqrt3 = 3.62759872846843570118815651528431146456813249618553;
    constexpr double k4PiDivSqrt3 = 7.25519745693687140237631303056862292913626499237106;
    m[0] = k2Pi / a; m[3] = k2PiDivSqrt3 / a; m[4] = k4PiDivSqrt3 / b; m[8] = k2Pi / c;
  } else if (a90 && g90) {
    const double sb = std::sin(be), tb = std::tan(be);
    const double cotb = 1.0 / tb;
    m[0] = k2Pi / a; m[4] = k2Pi / b; m[6] = -cotb * k2Pi / a; m[8] = k2Pi / (c * sb);
  } else {
    if (!(al > 0.0 && al < kPi && be > 0.0 && be < kPi && ga > 0.0 && ga < kPi)) return false;
    double L[9];
    lattice_general(a, b, c, al, be, ga, L);
    if (!(L[8] > 0.0)) return false;
    const double adj[9] = {L[4] * L[8] - L[5] * L[7], L[2] * L[7] - L[1] * L[8], L[1] * L[5] - L[2] * L[4],
                           L[5] * L[6] - L[3] * L[8], L[0] * L[8] - L[2] * L[6], L[2] * L[3] - L[0] * L[5],
                           L[3] * L[7] - L[4] * L[6], L[1] * L[6] - L[0] * L[7], L[0] * L[4] - L[1] * L[3]};
    const double f = k2Pi / (L[0] * adj[0] + L[1] * adj[3] + L[2] * adj[6]);
    for (int i = 0; i < 9; ++i) m[i] = adj[i] * f;
  }
  return true;
}

int64_t hkl_spacegroup(int sg, double a, double b, double c, double alpha, double beta, double gamma, int nspecies,
                       const int* natoms, const double* positions, const double* csl, const double* msd,
                       double dcutoff, double fsquarecut, int64_t maxout, int* out_hkl, int* out_mult,
                       double* out_dsp, double* out_fsq, int64_t* nrequired, int64_t maxmem,
                       int64_t* out_mem_off, int* out_mem_hkl, int64_t* nmem_required) {
  ensure_device();
  if (sg < 1 || sg > 230) raise("BadInput", "Space group number is not in the range 1 to 230");
  if (!(dcutoff > 0.0)) raise("BadInput", "dcutoff must be positive");
  if (!(a > 0 && b > 0 && c > 0)) raise("BadInput", "lattice parameters must be positive");
  HKLMatrix rm;
  if (!reci_lattice(a, b, c, alpha, beta, gamma, rm.m))
    raise("BadInput", "invalid cell: angles must lie in (0, 180) degrees and admit a cell");
  int mx[3];
  hkl_bounds(a, b, c, alpha, beta, gamma, dcutoff, mx);
  const int max_h = mx[0], max_k = mx[1], max_l = mx[2];
  HKLBuild P{};
  HKLSpecies S;
  if (nrequired) *nrequired = 0;
  if (!hkl_species(P, S, nspecies, natoms, positions, csl, msd, dcutoff, fsquarecut)) return 0;
  // first occurrence of each symmetry representative in the loop order of
  // the reference (h >= 0, then k, then l), (0,0,0) excluded; the ksq
  // preselection is applied here so that only candidates go to the GPU
  std::unordered_set<uint64_t> seen;
  auto pack = [](const HKL3& x) {
    return ((uint64_t)(uint32_t)(x.h + (1 << 20)) << 42) | ((uint64_t)(uint32_t)(x.k + (1 << 20)) << 21) |
           (uint64_t)(uint32_t)(x.l + (1 << 20));
  };
  auto rep_of = [&](int h, int k, int l) {
    HKL3 o[24];
    const int n = eqrefl_raw(sg, h, k, l, o);
    int lo = 0;
    for (int i = 1; i < n; ++i)
      if (hkl_less(o[i], o[lo])) lo = i;
    return o[lo];
  };
  seen.insert(pack(rep_of(0, 0, 0)));
  std::vector<int> keys;
  for (int h = 0; h <= max_h; ++h)
    for (int k = (h ? -max_k : 0); k <= max_k; ++k)
      for (int l = -max_l; l <= max_l; ++l) {
        const HKL3 key = rep_of(h, k, l);
        if (!seen.insert(pack(key)).second) continue;
        const double dh = (double)key.h, dk = (double)key.k, dl = (double)key.l;
        const double wx = rm.m[0] * dh + rm.m[1] * dk + rm.m[2] * dl;
        const double wy = rm.m[3] * dh + rm.m[4] * dk + rm.m[5] * dl;
        const double wz = rm.m[6] * dh + rm.m[7] * dk + rm.m[8] * dl;
        const double ksq = wx * wx + wy * wy + wz * wz;
        if (!(P.ksq_lo <= ksq && ksq <= P.ksq_hi)) continue;
        keys.push_back(key.h); keys.push_back(key.k); keys.push_back(key.l);
      }
  const int64_t nk = (int64_t)keys.size() / 3;
  std::vector<Fam> fams;
  if (nk > 0) {
    DevBuf dK, dF, dD, dV;
    dK.upload(keys.data(), keys.size());
    dF.alloc(sizeof(double) * nk);
    dD.alloc(sizeof(double) * nk);
    dV.alloc(nk);
    launch_hkl_eval_keys(P, rm, dK.as<int>(), nk, dF.as<double>(), dD.as<double>(), dV.as<unsigned char>(), 0);
    NCAMD_HIP(hipGetLastError());
    std::vector<double> hF(nk), hD(nk);
    std::vector<unsigned char> hV(nk);
    NCAMD_HIP(hipMemcpy(hV.data(), dV.p, nk, hipMemcpyDeviceToHost));
    NCAMD_HIP(hipMemcpy(hF.data(), dF.p, sizeof(double) * nk, hipMemcpyDeviceToHost));
    NCAMD_HIP(hipMemcpy(hD.data(), dD.p, sizeof(double) * nk, hipMemcpyDeviceToHost));
    for (int64_t t = 0; t < nk; ++t) {
      if (!hV[t]) continue;
      if (fams.size() > 1000000) return -2;
      HKL3 o[24];
      const int n = eqrefl_raw(sg, keys[3 * t], keys[3 * t + 1], keys[3 * t + 2], o);
      std::sort(o, o + n, hkl_less);
      Fam f;
      for (int i = 0; i < n; ++i)
        if (f.members.empty() || !(o[i].h == f.members.back().h && o[i].k == f.members.back().k &&
                                   o[i].l == f.members.back().l))
          f.members.push_back(o[i]);
      f.rep = f.members.front();
      f.mult = 2 * (int)f.members.size();
      f.fsq = hF[t];
      f.dsp = hD[t];
      fams.push_back(std::move(f));
    }
  }
  return hkl_output(fams, maxout, out_hkl, out_mult, out_dsp, out_fsq, nrequired, maxmem, out_mem_off,
                    out_mem_hkl, nmem_required);
}

// ------------------------------------------------- single-crystal Bragg ---
// Construction of an SCBragg process (init, host FP64, same operation order
// as the reference so that every table is bit-identical):
//  - families: ncrystal_core/src/scbragg/NCSCBragg.cc:125-231 (planes keyed
//    by (d, |F|^2) rounded to 2^-30, std::greater order, original values kept
//    when unambiguous, demi-normals rotated to the lab frame)
//  - demi-normals from explicit hkl: ncrystal_core/src/extd_utils/NCPlaneProvider.cc:111-175
//  - GaussMos / truncation: ncrystal_core/src/phys_utils/NCGaussMos.cc:38-75,
//    ncrystal_core/src/phys_utils/NCGaussOnSphere.cc:510-547
//  - GaussOnSphere::set (normalisation, fast-path bounds, spline tables):
//    ncrystal_core/src/phys_utils/NCGaussOnSphere.cc:181-378
//  - cubic spline with end-point derivatives: ncrystal_core/src/utils/NCSpline.cc:31-108
//  - estimateSingleSidedDerivative / findRoot: ncrystal_core/src/utils/NCMath.cc:271-334
constexpr double kFWHM2Sigma = 0.424660900144009521360751417051444809857570547;
constexpr double kArcSecH = 0.00000484813681109535993589914102357947975956353302;

struct HostSpline {
  double a = 0, b = 0, invdelta = 0;
  std::vector<double> y, y2;
};

void spline_fit(HostSpline& t, const std::vector<double>& y, double dleft, double dright) {
  const int n = (int)y.size(), nm1 = n - 1;
  std::vector<double> u(nm1), y2(n);
  y2[0] = -0.5;
  u[0] = 3.0 * ((y[1] - y[0]) - dleft);
  for (int i = 1; i < nm1; ++i) {
    const double p = 1.0 / (2.0 + 0.5 * y2[i - 1]);
    y2[i] = -0.5 * p;
    const double uu = y[i + 1] - 2.0 * y[i] + y[i - 1];
    u[i] = (3.0 * uu - 0.5 * u[i - 1]) * p;
  }
  const double qn = 0.5;
  const double un = 3.0 * (dright - (y[nm1] - y[n - 2]));
  y2[nm1] = (un - qn * u[n - 2]) / (qn * y2[n - 2] + 1.0);
  for (int k = nm1; k; --k) {
    y2[k - 1] *= y2[k];
    y2[k - 1] += u[k - 1];
  }
  t.y = y;
  t.y2 = std::move(y2);
}

template <class F>
void spline_table(HostSpline& t, const F& f, double a, double b, double fpa, double fpb, unsigned npts) {
  t.a = a;
  t.b = b;
  std::vector<double> fv(npts);
  const double delta = (b - a) / (npts - 1);
  for (unsigned i = 0; i + 1 < npts; ++i) fv[i] = f(a + i * delta);
  fv[npts - 1] = f(b);
  spline_fit(t, fv, fpa * delta, fpb * delta);
  t.invdelta = 1.0 / delta;
}

template <class F>
double deriv_single_sided4(const F& f, double x, double h) {
  return -(2 / 3.) * (31.5 * f(x) - 64. * f(x + 0.125 * h) + 40. * f(x + 0.25 * h) - 8. * f(x + 0.5 * h) +
                      0.5 * f(x + h)) / h;
}

template <class F>
double find_root(const F& f, double a, double b, double acc) {
  double fa = f(a), fb = f(b);
  acc *= 0.5;
  unsigned i = 60;
  while (--i) {
    double c = (a * fb - b * fa) / (fb - fa);
    if (b - a < acc) return c;
    const double k = 0.15 * (b - a);
    c = dmax(a + k, dmin(b - k, c));
    const double fc = f(c);
    if (!fc) return c;
    if (fa * fc < 0) { b = c; fb = fc; } else { a = c; fa = fc; }
  }
  raise("CalcError", "root search did not converge");
}

struct GosHost {
  double sigma, trunc, prec, cta, sta, expfact, norm, k1, k2, numacc;
  unsigned nlt;
  HostSpline evalcosx, sofcosd;
};

void gos_setup(GosHost& g, double sigma, double trunc, double prec) {
  prec = prec <= 1 ? dclamp(prec, 1e-7, 0.1) : dclamp(prec, 1.0, 10000.0);
  g.trunc = trunc; g.sigma = sigma; g.prec = prec;
  g.cta = std::cos(trunc);
  g.sta = std::sin(trunc);
  g.expfact = -0.5 / (sigma * sigma);
  {
    const double c = -0.5 / (sigma * sigma);
    auto f = [c](double x) { return dmax(0.0, std::sin(x) * std::exp(c * x * x)); };
    auto acc = [](unsigned lvl, double prev, double est) { return lvl > 12 || fabs(prev - est) < 1e-12 * fabs(est); };
    const double integral = k2Pi * romberg(f, acc, 0.0, dmin(trunc, 20.0 * sigma));
    g.norm = 1.0 / integral;
  }
  unsigned nlt;
  if (prec < 1.0) {
    const double tp = dmin(0.999999, prec * 0.5);
    auto fce = [tp](double x) { return (1.0 - (1.0 - 0.5 * x * x) / cos_mpi2pi2(x)) - tp; };
    const double amax = find_root(fce, 0.0, 0.999999 * kPiHalf, 1e-13);
    if (g.trunc > amax || amax <= 1e-10) {
      g.k1 = __builtin_huge_val(); g.k2 = 0.0;
    } else {
      g.k1 = cos_mpi2pi2(amax); g.k2 = g.cta - 1e-5;
    }
    g.numacc = dclamp(prec * 0.1, 1e-7, 1e-4);
    const unsigned v = (unsigned)(10.0 / std::sqrt(prec) + 0.5);
    nlt = std::max(40u, std::min(v, 10000u));
  } else {
    const unsigned v = (unsigned)(prec + 0.5);
    nlt = std::max(v, 20u);
    g.numacc = 1e-4;
    g.k1 = 0.0; g.k2 = __builtin_huge_val();
  }
  if (g.sigma < 10 * kArcSecH) {
    g.numacc = dmax(g.numacc, 1e-6);
    nlt = std::min(nlt, 1000u);
    if (g.sigma < kArcSecH) {
      nlt = std::min(nlt, 500u);
      g.numacc = dmax(g.numacc, 1e-5);
      if (g.sigma < 0.1 * kArcSecH) { nlt = std::min(nlt, 200u); g.numacc = dmax(g.numacc, 1e-4); }
    }
  }
  g.nlt = nlt;
  {
    const double ltmin = cos_mpi2pi2(dmin(kPiHalf, trunc));
    const double ltinvdelta = (nlt - 1.0) / (1.0 - ltmin);
    const double norm = g.norm, ef = g.expfact;
    auto f = [norm, ef](double cosx) {
      const double x = std::acos(dmin(1.0, dmax(-1.0, cosx)));
      return norm * std::exp(ef * x * x);
    };
    double h = 0.01 / ltinvdelta;
    h = dmax(h, dmin(1e-4, (1.0 - ltmin) * 0.1));
    const double fpb = deriv_single_sided4(f, 1.0, -h);
    const double fpa = deriv_single_sided4(f, ltmin, h);
    spline_table(g.evalcosx, f, ltmin, 1.0, fpa, fpb, nlt);
  }
  {
    const double cta = cos_mpi2pi2(trunc);
    const double kk = g.norm * 2.506628274631000502415765284811045253006986740609938316629923576 * sigma;
    const double ef = g.expfact, tasq = trunc * trunc;
    auto f = [kk, ef, tasq](double cosd) {
      const double d = std::acos(dmin(1.0, dmax(-1.0, cosd)));
      const double d2 = d * d;
      return kk * std::exp(ef * d2) * std::erf(std::sqrt(dmax(0.0, -ef * (tasq - d2))));
    };
    double h = 0.01 * (1.0 - cta) / nlt;
    h = dmax(h, dmin(1e-4, (1.0 - cta) * 0.1));
    const double fpb = deriv_single_sided4(f, 1.0, -h);
    const double fpa = f(cta + (1.0 - cta) / nlt) / ((1.0 - cta) / nlt);
    spline_table(g.sofcosd, f, cta, 1.0, fpa, fpb, nlt);
  }
}

Proc* create_scb(double v0n, int64_t nfam, const double* dsp, const double* fsq, const int64_t* mem_off,
                 const int* mem_hkl, const double* L, const double* R, double mos_fwhm, double prec,
                 double ntrunc) {
  ensure_device();
  if (!(v0n > 0)) raise("BadInput", "v0_times_natoms is not a positive number.");
  if (nfam < 1) raise("BadInput", "SCBragg needs at least one reflection family");
  if (!(mos_fwhm > 0.0) || !(prec >= 0.0) || !(ntrunc >= 0.0)) raise("BadInput", "invalid mosaicity parameters");
  if (mem_off[0] != 0) raise("BadInput", "member offsets must start at 0");
  const int64_t nplanes = mem_off[nfam];
  if (nplanes < 1 || nplanes >= (int64_t)1 << 31) raise("BadInput", "number of demi-normals out of range");
  constexpr double two30 = 1073741824.0;
  // planes in provider order, keyed by rounded (d, |F|^2)
  struct Key { uint64_t kd, kf; int64_t src; };
  std::vector<Key> keys((size_t)nplanes);
  std::vector<double> pn(3 * (size_t)nplanes);
  std::map<uint64_t, double> orig_d, orig_f;
  double maxd = 0.0;
  int64_t q = 0;
  for (int64_t f = 0; f < nfam; ++f) {
    if (mem_off[f + 1] < mem_off[f]) raise("BadInput", "member offsets must be non-decreasing");
    if (!(dsp[f] > 0.0 && fsq[f] > 0.0 && dsp[f] < 1e7 && fsq[f] < 1e7))
      raise("BadInput", "d-spacing and |F|^2 must be in (0, 1e7)");
    for (int64_t m = mem_off[f]; m < mem_off[f + 1]; ++m, ++q) {
      const double h = mem_hkl[3 * m], k = mem_hkl[3 * m + 1], l = mem_hkl[3 * m + 2];
      double vx = L[0] * h + L[1] * k + L[2] * l;
      double vy = L[3] * h + L[4] * k + L[5] * l;
      double vz = L[6] * h + L[7] * k + L[8] * l;
      const double m2 = vx * vx + vy * vy + vz * vz;
      constexpr double eps = 2.220446049250313080847263336181640625e-16;
      if (!(m2 >= 1.0 - 2.0 * eps && m2 <= 1.0 + 2.0 * eps)) {
        if (!(m2 > 0.0)) raise("BadInput", "hkl (0,0,0) in a family member list");
        const double ff = 1.0 / std::sqrt(m2);
        vx *= ff; vy *= ff; vz *= ff;
      }
      pn[3 * q] = vx; pn[3 * q + 1] = vy; pn[3 * q + 2] = vz;
      if (dsp[f] > maxd) maxd = dsp[f];
      const uint64_t kd = (uint64_t)(dsp[f] * two30 + 0.5), kf = (uint64_t)(fsq[f] * two30 + 0.5);
      keys[q] = Key{kd, kf, q};
      auto itd = orig_d.find(kd);
      if (itd == orig_d.end()) orig_d[kd] = dsp[f];
      else if (fabs(dsp[f] - itd->second) > 1e-12) itd->second = -1;
      auto itf = orig_f.find(kf);
      if (itf == orig_f.end()) orig_f[kf] = fsq[f];
      else if (fabs(fsq[f] - itf->second) > 1e-12) itf->second = -1;
    }
  }
  // std::greater on (kd, kf); planes of one key stay in provider order
  std::stable_sort(keys.begin(), keys.end(), [](const Key& x, const Key& y) {
    return x.kd != y.kd ? x.kd > y.kd : x.kf > y.kf;
  });
  std::vector<double> inv2d, inv2d_r, xsfact, xsf, normals(3 * (size_t)nplanes);
  std::vector<int64_t> off;
  for (int64_t i = 0; i < nplanes; ++i) {
    if (i == 0 || keys[i].kd != keys[i - 1].kd || keys[i].kf != keys[i - 1].kf) {
      const double dd = orig_d[keys[i].kd], fv = orig_f[keys[i].kf];
      const double dspv = dd > 0 ? dd : keys[i].kd / two30;
      const double fsqv = fv > 0 ? fv : keys[i].kf / two30;
      xsfact.push_back(fsqv / v0n);
      inv2d.push_back(0.5 / dspv);
      off.push_back(i);
    }
    const double* n = &pn[3 * keys[i].src];
    normals[3 * i] = R[0] * n[0] + R[1] * n[1] + R[2] * n[2];
    normals[3 * i + 1] = R[3] * n[0] + R[4] * n[1] + R[5] * n[2];
    normals[3 * i + 2] = R[6] * n[0] + R[7] * n[1] + R[8] * n[2];
  }
  off.push_back(nplanes);
  const int nf = (int)inv2d.size();
  for (int f = 0; f < nf; ++f) {
    inv2d_r.push_back(std::floor(dmax(inv2d[f], 1e-15) * 1e15 + 0.5) * 1e-15);
    xsf.push_back(xsfact[f] * 0.5);
  }
  // GaussMos truncation and GaussOnSphere tables
  double nt = ntrunc;
  if (nt == 0) {
    if (prec == 0) nt = 8.0;
    else if (prec >= 1.0) nt = 3.0;
    else nt = dclamp(1.1 * std::sqrt(-2.0 * std::log(dmax(1e-300, prec))), 3.0, 8.0);
  }
  const double sigma = mos_fwhm * kFWHM2Sigma;
  const double trunc = nt * sigma;
  if (!(trunc < kPiHalf))
    raise("BadInput", "Mosaicity too large, truncation angle (sigma*Ntrunc) must be less than pi/2");
  GosHost g;
  gos_setup(g, sigma, trunc, prec);

  std::unique_ptr<Proc> p(new Proc());
  p->kind = kKindSCB;
  std::vector<int> off32(off.begin(), off.end());
  // per demi-normal scan records (normal, rounded 1/(2d) of its family) and family index
  std::vector<double> nrec(4 * (size_t)(nplanes + kScbPad), 0.0);
  std::vector<float> nrec32(4 * (size_t)(nplanes + kScbPad), 0.0f);
  // (family cross-section factor, family index) per demi-normal
  std::vector<double> nfx(2 * (size_t)(nplanes + kScbPad), 0.0);
  for (int f = 0; f < nf; ++f)
    for (int64_t j = off[f]; j < off[f + 1]; ++j) {
      nrec[4 * j] = normals[3 * j];
      nrec[4 * j + 1] = normals[3 * j + 1];
      nrec[4 * j + 2] = normals[3 * j + 2];
      nrec[4 * j + 3] = inv2d_r[f];
      nfx[2 * j] = xsf[f];
      nfx[2 * j + 1] = (double)f;
    }
  for (size_t j = 0; j < nrec.size(); ++j) nrec32[j] = (float)nrec[j];
  for (int64_t j = nplanes; j < nplanes + kScbPad; ++j) {  // padding records never pass the pre-filter
    nrec[4 * j + 3] = 3.0e38;
    nrec32[4 * j + 3] = 3.0e38f;
  }
  std::vector<double> lt1, lt2;
  for (size_t i = 0; i < g.evalcosx.y.size(); ++i) { lt1.push_back(g.evalcosx.y[i]); lt1.push_back(g.evalcosx.y2[i]); }
  for (size_t i = 0; i < g.sofcosd.y.size(); ++i) { lt2.push_back(g.sofcosd.y[i]); lt2.push_back(g.sofcosd.y2[i]); }
  DevBuf* b_inv2d = p->newbuf(); b_inv2d->upload(inv2d.data(), inv2d.size());
  DevBuf* b_xsf = p->newbuf(); b_xsf->upload(xsf.data(), xsf.size());
  DevBuf* b_off = p->newbuf(); b_off->upload(off32.data(), off32.size());
  DevBuf* b_nrec = p->newbuf(); b_nrec->upload(nrec.data(), nrec.size());
  DevBuf* b_nrec32 = p->newbuf(); b_nrec32->upload(nrec32.data(), nrec32.size());
  DevBuf* b_nfx = p->newbuf(); b_nfx->upload(nfx.data(), nfx.size());
  DevBuf* b_lt1 = p->newbuf(); b_lt1->upload(lt1.data(), lt1.size());
  DevBuf* b_lt2 = p->newbuf(); b_lt2->upload(lt2.data(), lt2.size());
  SCBView& v = p->scb;
  v.nfam = nf;
  v.nnormals = (int)nplanes;
  v.inv2d = b_inv2d->as<double>();
  v.xsf = b_xsf->as<double>();
  v.off = b_off->as<int>();
  v.nrec = b_nrec->as<double4>();
  v.nrec32 = b_nrec32->as<float4>();
  v.nfx = b_nfx->as<double2>();
  // |sin(phi) - sin(theta)| <= |phi - theta| < truncation angle for every
  // contributing normal; the margins cover the rounding of cos(truncation)
  // (~1e-14) and of the FP32 records and products of the pre-filter (< 3e-6)
  v.band32 = (float)(trunc * (1.0 + 1e-6) + 1e-5);
  {
    const double w = maxd * 2.0, w2 = w * w;
    v.threshold = w2 ? kWl2Ekin / w2 : __builtin_huge_val();
  }
  v.cta = g.cta; v.k1 = g.k1; v.k2 = g.k2; v.numacc = g.numacc;
  v.lt_evalcosx = SplineView{g.evalcosx.a, g.evalcosx.invdelta, (int)g.evalcosx.y.size(), b_lt1->as<double2>()};
  v.lt_sofcosd = SplineView{g.sofcosd.a, g.sofcosd.invdelta, (int)g.sofcosd.y.size(), b_lt2->as<double2>()};
  p->scb_inv2d_h = inv2d;
  p->scb_xsfact_h = xsfact;
  p->scb_off_h = off;
  p->scb_normals_h = normals;
  p->scb_lt1_h = lt1;
  p->scb_lt2_h = lt2;
  const double gd[15] = {g.sigma, g.trunc, g.prec, g.cta, g.sta, g.expfact, g.norm, g.k1, g.k2, g.numacc,
                         (double)g.nlt, g.evalcosx.a, g.evalcosx.invdelta, g.sofcosd.a, g.sofcosd.invdelta};
  std::copy(gd, gd + 15, p->scb_gos);
  init_errbuf(p.get());
  return p.release();
}

// ------------------------------------------------------- material batches ---
// Many materials in one persistent launch (nc_batch.hip).  The batch holds
// references to its materials, their device views, the tile list and the
// tuned launch parameters; each call uploads the materials' (seed, call
// counter) pairs and replays a captured graph of the reset + persistent
// kernels.
constexpr uint32_t kBatchMagic = 0x6e636262U;

struct Batch {
  uint32_t magic = kBatchMagic;
  std::vector<Proc*> mats;
  std::vector<int64_t> counts, offsets;
  std::vector<double> cost;  // estimated relative cost per neutron
  int64_t total = 0;
  DevBuf d_mats, d_rngs, d_tiles, d_next, d_err;
  DevBuf d_sab, d_tcount, d_highe, d_hcount;  // staged sampling (BatchWork)
  std::vector<BatchRng> h_rngs;
  int tile = 4096, nblocks = 1024, ordering = 1;
  int ntiles = 0;
  bool tiles_valid = false;
  // one captured graph per mode (0 cross sections, 1 sampling), each kept
  // with the pointers and tuning it was captured for, so that callers that
  // alternate the two calls replay both graphs
  struct GraphSlot {
    hipGraphExec_t exec = nullptr;
    int tile = 0, nblocks = 0, ordering = -1;
    const void *in = nullptr, *o1 = nullptr, *o2 = nullptr;
    long long captures = 0, replays = 0;
  };
  GraphSlot graphs[2];
  hipStream_t cap_stream = nullptr;
  void drop_graphs() {
    for (GraphSlot& g : graphs)
      if (g.exec) { (void)hipGraphExecDestroy(g.exec); g.exec = nullptr; }
  }
  ~Batch() {
    magic = 0;
    drop_graphs();
    if (cap_stream) (void)hipStreamDestroy(cap_stream);
    for (auto* m : mats)
      if (m && --m->refcount == 0) delete m;
  }
};

Batch* extract_batch(ncamd_batch_t h) {
  Batch* b = static_cast<Batch*>(h.internal);
  if (!b || b->magic != kBatchMagic) raise("LogicError", "invalid ncamd_batch_t handle");
  return b;
}

// per-neutron cost estimate of a material (relative units): sampling an
// S(alpha,beta) component dominates, plane searches grow with log2(planes)
double batch_cost(const CompView& v) {
  double c = 1.0;
  for (int i = 0; i < v.n; ++i) {
    if (v.kind[i] == kKindSAB) c += 8.0;
    if (v.kind[i] == kKindPB) c += 0.1 * std::log2(2.0 + v.pb.n);
    if (v.kind[i] == kKindElInc) c += 1.0;
  }
  return c;
}

void batch_build_tiles(Batch* b) {
  std::vector<BatchTile> tiles;
  std::vector<int> order(b->mats.size());
  for (size_t i = 0; i < order.size(); ++i) order[i] = (int)i;
  if (b->ordering == 1)  // most expensive materials first (longest-processing-time order)
    std::stable_sort(order.begin(), order.end(), [b](int x, int y) {
      return b->cost[x] * (double)b->counts[x] > b->cost[y] * (double)b->counts[y];
    });
  for (int m : order)
    for (int64_t k = 0; k < b->counts[m]; k += b->tile)
      tiles.push_back(BatchTile{m, k, std::min<int64_t>(b->counts[m], k + b->tile)});
  // the high-energy queue packs (tile << 40 | neutron)
  if (tiles.size() >= ((size_t)1 << 23)) raise("BadInput", "too many tiles (use a larger tile size)");
  // captured graphs refer to the old tile list
  b->drop_graphs();
  b->ntiles = (int)tiles.size();
  b->d_tiles.upload(tiles.data(), tiles.size());
  b->d_tcount.ensure(sizeof(int) * std::max<size_t>(tiles.size(), 1));
  b->tiles_valid = true;
}

void batch_upload_rngs(Batch* b, bool advance, hipStream_t st) {
  for (size_t i = 0; i < b->mats.size(); ++i) {
    b->h_rngs[i].seed = b->mats[i]->seed;
    b->h_rngs[i].stream = advance ? b->mats[i]->callctr++ : b->mats[i]->callctr;
  }
  NCAMD_HIP(hipMemcpyAsync(b->d_rngs.p, b->h_rngs.data(), sizeof(BatchRng) * b->h_rngs.size(),
                           hipMemcpyHostToDevice, st));
}

// launch through the mode's captured graph (re-captured when its pointers or
// the tuning change)
void batch_run(Batch* b, int mode, const double* in, double* o1, double* o2, hipStream_t st) {
  if (!b->tiles_valid) batch_build_tiles(b);
  if (!b->cap_stream) NCAMD_HIP(hipStreamCreateWithFlags(&b->cap_stream, hipStreamNonBlocking));
  Batch::GraphSlot& gs = b->graphs[mode ? 1 : 0];
  const bool same = gs.exec && gs.in == in && gs.o1 == o1 && gs.o2 == o2 && gs.tile == b->tile &&
                    gs.nblocks == b->nblocks && gs.ordering == b->ordering;
  if (!same) {
    if (gs.exec) { NCAMD_HIP(hipGraphExecDestroy(gs.exec)); gs.exec = nullptr; }
    hipGraph_t g = nullptr;
    NCAMD_HIP(hipStreamBeginCapture(b->cap_stream, hipStreamCaptureModeThreadLocal));
    BatchWork w;
    w.next = b->d_next.as<unsigned>();
    w.sab = b->d_sab.as<int64_t>();
    w.tcount = b->d_tcount.as<int>();
    w.highe = b->d_highe.as<int64_t>();
    w.hcount = b->d_hcount.as<unsigned long long>();
    launch_batch(b->d_mats.as<BatchMat>(), b->d_rngs.as<BatchRng>(), b->d_tiles.as<BatchTile>(), b->ntiles,
                 b->nblocks, mode, in, o1, o2, w, b->d_err.as<unsigned long long>(), b->cap_stream);
    const hipError_t e1 = hipGetLastError();
    const hipError_t e2 = hipStreamEndCapture(b->cap_stream, &g);
    NCAMD_HIP(e1);
    NCAMD_HIP(e2);
    const hipError_t e3 = hipGraphInstantiate(&gs.exec, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    NCAMD_HIP(e3);
    gs.in = in; gs.o1 = o1; gs.o2 = o2;
    gs.tile = b->tile; gs.nblocks = b->nblocks; gs.ordering = b->ordering;
    ++gs.captures;
  } else {
    ++gs.replays;
  }
  NCAMD_HIP(hipGraphLaunch(gs.exec, st));
}

// ------------------------------------------- S(alpha,beta) envelope sampler ---
// Tables of the box-envelope sampler (device side in nc_kb.hh): interval
// edges log-spaced in [emin, emax), per interval the coordinate parameters,
// the grid of boxes with the bounds of t over each (GPU) and the alias table
// over box masses (host, Walker/Vose), plus exact guides for the alpha/beta
// cell searches.
DevBuf* kb_newbuf(Proc* p) {
  p->kb_bufs.emplace_back(new DevBuf());
  return p->kb_bufs.back().get();
}

double std_normal_cdf(double z) { return 0.5 * std::erfc(-z * 0.7071067811865476); }

// Edges of a grid of nb cells on [0,1] at the quantiles of
// H(x) = lam x + (1 - lam) G(x), with G the distribution function `cdf`
// restricted to [0,1] (G = x where it has no mass there).  Edges are floats
// shared by neighbouring boxes, strictly increasing, with 0 and 1 at the ends.
template <class F>
void kb_edges(int nb, double lam, int ng, F cdf, float* out) {
  out[0] = 0.0f;
  out[nb] = 1.0f;
  const double c0 = lam < 1.0 ? cdf(0.0) : 0.0, c1 = lam < 1.0 ? cdf(1.0) : 1.0;
  const double mass = c1 - c0;
  if (lam >= 1.0 || !(mass > 1e-300) || !std::isfinite(mass)) {
    for (int e = 1; e < nb; ++e) out[e] = (float)((double)e / nb);
    return;
  }
  std::vector<double> H((size_t)ng + 1);
  for (int i = 0; i <= ng; ++i) {
    const double x = (double)i / ng;
    const double g = std::min(1.0, std::max(0.0, (cdf(x) - c0) / mass));
    H[(size_t)i] = lam * x + (1.0 - lam) * g;
  }
  int j = 0;
  for (int e = 1; e < nb; ++e) {
    const double target = (double)e / nb;
    while (j < ng - 1 && H[(size_t)j + 1] < target) ++j;
    const double dh = H[(size_t)j + 1] - H[(size_t)j];
    const double f = dh > 0.0 ? std::min(1.0, std::max(0.0, (target - H[(size_t)j]) / dh)) : 0.5;
    out[e] = (float)((j + f) / ng);
  }
  for (int e = 1; e < nb; ++e) out[e] = std::max(out[e], std::nextafter(out[e - 1], 2.0f));
  for (int e = nb - 1; e >= 1; --e) out[e] = std::min(out[e], std::nextafter(out[e + 1], -1.0f));
}

// One interval's mixture (from the model's 48 outputs): weights, means,
// standard deviations and correlations of 8 bivariate Gaussians in (w, kappa).
struct KBMix {
  double pi[kMlpComp], mw[kMlpComp], mk[kMlpComp], sw[kMlpComp], sk[kMlpComp], r[kMlpComp];
  explicit KBMix(const float* y) {
    double lmax = -1e300, se = 0.0;
    for (int j = 0; j < kMlpComp; ++j) lmax = std::max(lmax, (double)y[6 * j]);
    for (int j = 0; j < kMlpComp; ++j) se += std::exp((double)y[6 * j] - lmax);
    for (int j = 0; j < kMlpComp; ++j) {
      pi[j] = std::exp((double)y[6 * j] - lmax) / se;
      mw[j] = y[6 * j + 1];
      mk[j] = y[6 * j + 2];
      sw[j] = std::exp(std::min(3.0, std::max(-9.0, (double)y[6 * j + 3])));
      sk[j] = std::exp(std::min(3.0, std::max(-9.0, (double)y[6 * j + 4])));
      r[j] = 0.95 * std::tanh((double)y[6 * j + 5]);
    }
  }
};

// grid of one interval: w-edges from the mixture's w-marginal, kappa-edges of
// each strip from the mixture's kappa distribution given w in the strip
// (components weighted by their mass in the strip, conditioned at its centre)
void kb_grid_boxes(const KBMix* mix, double lam, int bw, int bk, KBBox* out) {
  std::vector<float> we((size_t)bw + 1), ke((size_t)bk + 1);
  kb_edges(bw, mix ? lam : 1.0, 2048,
           [mix](double x) {
             double c = 0.0;
             for (int j = 0; j < kMlpComp; ++j) c += mix->pi[j] * std_normal_cdf((x - mix->mw[j]) / mix->sw[j]);
             return c;
           },
           we.data());
  for (int iw = 0; iw < bw; ++iw) {
    const double wa = we[(size_t)iw], wb = we[(size_t)iw + 1], wc = 0.5 * (wa + wb);
    double rho[kMlpComp], mu[kMlpComp], sd[kMlpComp];
    if (mix)
      for (int j = 0; j < kMlpComp; ++j) {
        rho[j] = mix->pi[j] * (std_normal_cdf((wb - mix->mw[j]) / mix->sw[j]) -
                               std_normal_cdf((wa - mix->mw[j]) / mix->sw[j]));
        mu[j] = mix->mk[j] + mix->r[j] * mix->sk[j] * (wc - mix->mw[j]) / mix->sw[j];
        sd[j] = mix->sk[j] * std::sqrt(1.0 - mix->r[j] * mix->r[j]);
      }
    kb_edges(bk, mix ? lam : 1.0, 256,
             [&](double x) {
               double c = 0.0;
               for (int j = 0; j < kMlpComp; ++j) c += rho[j] * std_normal_cdf((x - mu[j]) / sd[j]);
               return c;
             },
             ke.data());
    for (int ik = 0; ik < bk; ++ik) {
      KBBox& b = out[(size_t)iw * bk + ik];
      b.w0 = we[(size_t)iw]; b.w1 = we[(size_t)iw + 1];
      b.k0 = ke[(size_t)ik]; b.k1 = ke[(size_t)ik + 1];
      b.env = 0.0f; b.tlow = 0.0f; b.arange = 0; b.brange = 0;
    }
  }
}

// Builds the sampler; with `mix` (one mixture per interval, at the
// interval's upper energy) the grid of each interval is the candidate with
// the smallest envelope mass among the uniform grid and the learned grids
// blended with it, so it is never worse than the uniform grid.
void build_kb(Proc* p, int nn, double emin, double emax, int bw, int bk, int maxtries,
              const std::vector<KBMix>* mix) {
  if (p->kind != kKindSAB) raise("BadInput", "not an S(alpha,beta) process");
  const SABView& s = p->sab;
  // only inside the table's energy grid does the reference sample the table
  // density itself; above it mixes in the free-gas extender, below it
  // applies its low-energy rule, and there the plain sampler takes over
  if (emax <= 0.0 || emax > p->sab_egrid_h.back()) emax = p->sab_egrid_h.back();
  if (emin <= 0.0 || emin < p->sab_egrid_h.front()) emin = p->sab_egrid_h.front();
  if (!(emin > 0.0 && emax > emin)) raise("BadInput", "envelope sampler: need 0 < emin < emax");
  if (nn < 1 || nn > 1 << 16 || bw < 1 || bk < 1 || (int64_t)bw * bk > 1 << 16 || maxtries < 1)
    raise("BadInput", "envelope sampler: invalid grid sizes");
  if (!(p->sab_alpha_h.front() > 0.0)) raise("BadInput", "envelope sampler: alpha grid must be positive");
  if (p->sab_alpha_h.size() > 65536 || p->sab_beta_h.size() > 65536)
    raise("BadInput", "envelope sampler: alpha and beta grids of at most 65536 points");
  const double lams_learned[] = {1.0, 0.6, 0.45, 0.35, 0.28, 0.2, 0.1};
  const int ncand = mix ? 7 : 1;
  if ((int64_t)ncand * nn * bw * bk > (int64_t)1 << 26) raise("BadInput", "envelope sampler: too many boxes");
  if (mix && (int)mix->size() != nn) raise("LogicError", "one mixture per interval expected");
  p->use_kb = false;
  p->kb_bufs.clear();
  KBView k{};
  k.nn = nn; k.bw = bw; k.bk = bk; k.nbox = bw * bk; k.squeeze = 1;
  const double lr = std::log(emax / emin) / nn;
  k.elo = emin; k.ehi = emax; k.lr_inv = 1.0 / lr;
  k.inv_elo = 1.0 / emin; k.inv_kT = 1.0 / s.kT;
  std::vector<double> node((size_t)nn + 1);
  for (int i = 0; i <= nn; ++i) node[(size_t)i] = emin * std::exp(lr * i);
  node[0] = emin; node[(size_t)nn] = emax;
  const double bmax = p->sab_beta_h.back();
  std::vector<double> par(4 * (size_t)nn);
  for (int i = 0; i < nn; ++i) {
    const double eps = node[(size_t)i + 1] / s.kT;
    const double seps = std::sqrt(eps), wmax = std::sqrt(bmax + eps);
    par[4 * i] = eps; par[4 * i + 1] = seps; par[4 * i + 2] = wmax; par[4 * i + 3] = 8.0 * seps * wmax * wmax * wmax;
  }
  DevBuf* bn = kb_newbuf(p); bn->upload(node.data(), node.size());
  DevBuf* bp = kb_newbuf(p); bp->upload(par.data(), par.size());
  k.node_e = bn->as<double>();
  k.node_par = bp->as<double4>();
  // candidate grids (candidate-major, then interval, then box) and their
  // bounds on the GPU
  const size_t nper = (size_t)nn * k.nbox;
  std::vector<KBBox> cand((size_t)ncand * nper);
  {
    const int njobs = ncand * nn;
    const int nth = mix ? (int)std::max(1u, std::min(16u, std::thread::hardware_concurrency())) : 1;
    std::atomic<int> next{0};
    auto work = [&]() {
      for (int jb = next++; jb < njobs; jb = next++) {
        const int c = jb / nn, i = jb % nn;
        kb_grid_boxes(mix ? &(*mix)[(size_t)i] : nullptr, lams_learned[c], bw, bk,
                      cand.data() + (size_t)c * nper + (size_t)i * k.nbox);
      }
    };
    std::vector<std::thread> th;
    for (int t = 1; t < nth; ++t) th.emplace_back(work);
    work();
    for (auto& t : th) t.join();
  }
  {
    DevBuf dcand;
    dcand.upload(cand.data(), cand.size());
    launch_kb_bounds(s, k.node_par, nn, k.nbox, (int64_t)cand.size(), dcand.as<KBBox>(), 0);
    NCAMD_HIP(hipGetLastError());
    NCAMD_HIP(hipMemcpy(cand.data(), dcand.p, sizeof(KBBox) * cand.size(), hipMemcpyDeviceToHost));
  }
  auto area = [](const KBBox& b) { return ((double)b.w1 - (double)b.w0) * ((double)b.k1 - (double)b.k0); };
  std::vector<KBBox> boxes(nper);
  std::vector<double> stats(4 * (size_t)nn);
  for (int i = 0; i < nn; ++i) {
    int best = 0;
    double best_mass = 0.0, uni_mass = 0.0;
    for (int c = 0; c < ncand; ++c) {
      const KBBox* bx = cand.data() + (size_t)c * nper + (size_t)i * k.nbox;
      double m = 0.0;
      for (int j = 0; j < k.nbox; ++j) m += bx[j].env * area(bx[j]);
      if (c == 0) uni_mass = m;
      if (c == 0 || m < best_mass) { best = c; best_mass = m; }
    }
    const KBBox* bx = cand.data() + (size_t)best * nper + (size_t)i * k.nbox;
    double sq = 0.0;
    for (int j = 0; j < k.nbox; ++j) sq += bx[j].tlow * area(bx[j]);
    std::copy(bx, bx + k.nbox, boxes.begin() + (ptrdiff_t)((size_t)i * k.nbox));
    stats[4 * (size_t)i] = best_mass;
    stats[4 * (size_t)i + 1] = uni_mass;
    stats[4 * (size_t)i + 2] = lams_learned[best];
    stats[4 * (size_t)i + 3] = sq;
  }
  cand.clear();
  cand.shrink_to_fit();
  // alias tables (Walker/Vose) per interval over the box masses env x area,
  // stored with the boxes in one record per slot
  std::vector<KBSlot> alias(nper);
  std::vector<double> q((size_t)k.nbox);
  std::vector<int> small, large;
  for (int i = 0; i < nn; ++i) {
    const KBBox* bx = boxes.data() + (size_t)i * k.nbox;
    KBSlot* al = alias.data() + (size_t)i * k.nbox;
    double tot = 0.0;
    for (int j = 0; j < k.nbox; ++j) tot += bx[j].env * area(bx[j]);
    for (int j = 0; j < k.nbox; ++j) {
      al[j] = KBSlot{};
      al[j].prob = 1.0; al[j].inv_p = 1.0; al[j].inv_q = 0.0;
      al[j].box = bx[j];
      al[j].alias = j;
    }
    if (!(tot > 0.0) || !std::isfinite(tot)) {
      // nothing to sample (a zero bound would accept every proposal, u * 0
      // <= 0): NaN bounds fail every comparison, so every proposal is
      // rejected and the neutrons go to the plain sampler
      for (int j = 0; j < k.nbox; ++j) al[j].box.env = al[j].box.tlow = __builtin_nanf("");
      continue;
    }
    small.clear(); large.clear();
    for (int j = 0; j < k.nbox; ++j) {
      q[(size_t)j] = bx[j].env * area(bx[j]) * k.nbox / tot;
      (q[(size_t)j] < 1.0 ? small : large).push_back(j);
    }
    while (!small.empty() && !large.empty()) {
      const int sj = small.back(); small.pop_back();
      const int lj = large.back(); large.pop_back();
      al[sj].prob = q[(size_t)sj];
      al[sj].alias = lj;
      q[(size_t)lj] = (q[(size_t)lj] + q[(size_t)sj]) - 1.0;
      (q[(size_t)lj] < 1.0 ? small : large).push_back(lj);
    }
    for (int j = 0; j < k.nbox; ++j) {
      al[j].inv_p = al[j].prob > 0.0 ? 1.0 / al[j].prob : 0.0;
      al[j].inv_q = al[j].prob < 1.0 ? 1.0 / (1.0 - al[j].prob) : 0.0;
    }
  }
  DevBuf* ba = kb_newbuf(p); ba->upload(alias.data(), alias.size());
  k.slot = ba->as<KBSlot>();
  p->kb = k;
  p->kb_maxtries = maxtries;
  p->kb_node_h = node;
  p->kb_box_h = std::move(boxes);
  p->kb_stats_h = std::move(stats);
  p->use_kb = true;
}

// --------------------------------------------------------- MLP proposal ---
// Host side of the MLP (kernels in nc_mlp.hip): parameters, Adam state and
// activation buffers live on the device.  Training draws the samples of a
// run of steps with the plain S(alpha,beta) sampler in one launch, then runs
// one fused kernel per step (features, forward, mixture NLL, backward, Adam).
constexpr uint32_t kMlpMagic = 0x6e636d6cU;

struct Mlp {
  uint32_t magic = kMlpMagic;
  DevBuf w, g, m, v, loss;
  int step = 0;
  double lelo = 0.0, linv = 1.0;  // feature scaling: x = (log E - lelo) * linv
  int cap = 0;
  DevBuf X, T, Z1, H1, Z2, H2, Z3, H3, Y;
  DevBuf E;
  DevBuf counter;               // finished workgroups of a fused step
  DevBuf tE, ta, tb, tloss;     // training samples of a run of steps, loss per step
  size_t tcap = 0, lcap = 0;
  ~Mlp() { magic = 0; }
  void ensure(int n) {
    if (n <= cap) return;
    cap = n;
    const size_t f = sizeof(float);
    X.alloc(f * n * kMlpIn); T.alloc(f * n * 2);
    Z1.alloc(f * n * kMlpHidden); H1.alloc(f * n * kMlpHidden);
    Z2.alloc(f * n * kMlpHidden); H2.alloc(f * n * kMlpHidden);
    Z3.alloc(f * n * kMlpHidden); H3.alloc(f * n * kMlpHidden);
    Y.alloc(f * n * kMlpOut);
    E.alloc(sizeof(double) * n);
  }
  void ensure_train(size_t nsamp, size_t steps) {
    if (nsamp > tcap) {
      tE.alloc(sizeof(double) * nsamp); ta.alloc(sizeof(double) * nsamp); tb.alloc(sizeof(double) * nsamp);
      tcap = nsamp;
    }
    if (steps > lcap) {
      tloss.alloc(sizeof(float) * steps);
      lcap = steps;
    }
  }
  float* P(int off) { return w.as<float>() + off; }
  float* G(int off) { return g.as<float>() + off; }
};

Mlp* extract_mlp(ncamd_mlp_t h) {
  Mlp* m = static_cast<Mlp*>(h.internal);
  if (!m || m->magic != kMlpMagic) raise("LogicError", "invalid ncamd_mlp_t handle");
  return m;
}

// forward pass of n rows of features X (in m->X) into m->Y
void mlp_forward(Mlp* m, int n, hipStream_t st) {
  const int H = kMlpHidden;
  launch_gemm_bf16(m->X.as<float>(), kMlpIn, false, m->P(kOffW1), H, false, m->Z1.as<float>(), H, m->P(kOffB1), n, H,
                   kMlpIn, false, st);
  launch_silu(m->Z1.as<float>(), m->H1.as<float>(), (int64_t)n * H, st);
  launch_gemm_bf16(m->H1.as<float>(), H, false, m->P(kOffW2), H, false, m->Z2.as<float>(), H, m->P(kOffB2), n, H, H,
                   false, st);
  launch_silu(m->Z2.as<float>(), m->H2.as<float>(), (int64_t)n * H, st);
  launch_gemm_bf16(m->H2.as<float>(), H, false, m->P(kOffW3), H, false, m->Z3.as<float>(), H, m->P(kOffB3), n, H, H,
                   false, st);
  launch_silu(m->Z3.as<float>(), m->H3.as<float>(), (int64_t)n * H, st);
  launch_gemm_bf16(m->H3.as<float>(), H, false, m->P(kOffW4), kMlpOut, false, m->Y.as<float>(), kMlpOut,
                   m->P(kOffB4), n, kMlpOut, H, false, st);
}

Mlp* create_mlp(uint64_t seed, double emin, double emax) {
  ensure_device();
  if (!(emin > 0.0 && emax > emin)) raise("BadInput", "MLP: need 0 < emin < emax");
  std::unique_ptr<Mlp> m(new Mlp());
  m->lelo = std::log(emin);
  m->linv = 1.0 / (std::log(emax) - std::log(emin));
  std::vector<float> w((size_t)kMlpParams, 0.f);
  uint64_t s = seed * 0x9E3779B97F4A7C15ull + 0x1234567ull;
  auto unif = [&s]() {  // splitmix64 -> [0,1)
    s += 0x9E3779B97F4A7C15ull;
    uint64_t z = s;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    return (double)(z >> 11) * 0x1.0p-53;
  };
  auto init = [&](int off, int fin, int fout, double gain) {
    const double lim = gain * std::sqrt(6.0 / (fin + fout));
    for (int i = 0; i < fin * fout; ++i) w[(size_t)off + i] = (float)((2.0 * unif() - 1.0) * lim);
  };
  init(kOffW1, kMlpIn, kMlpHidden, 1.0);
  init(kOffW2, kMlpHidden, kMlpHidden, 1.0);
  init(kOffW3, kMlpHidden, kMlpHidden, 1.0);
  init(kOffW4, kMlpHidden, kMlpOut, 0.1);
  for (int c = 0; c < kMlpComp; ++c) {  // broad initial components spread along w
    float* b4 = w.data() + kOffB4 + 6 * c;
    b4[0] = 0.f; b4[1] = (c + 0.5f) / kMlpComp; b4[2] = 0.5f; b4[3] = std::log(0.2f); b4[4] = std::log(0.3f); b4[5] = 0.f;
  }
  m->w.upload(w.data(), w.size());
  m->g.alloc(sizeof(float) * kMlpParams);
  std::vector<float> z((size_t)kMlpParams, 0.f);
  m->m.upload(z.data(), z.size());
  m->v.upload(z.data(), z.size());
  m->loss.alloc(sizeof(float));
  m->counter.alloc(sizeof(unsigned));
  NCAMD_HIP(hipMemset(m->counter.p, 0, sizeof(unsigned)));
  NCAMD_HIP(hipMemset(m->g.p, 0, sizeof(float) * kMlpParams));
  return m.release();
}


// ----------------------------------------------- learned proposal sampler ---
// The envelope sampler on grids placed by the MLP: its mixture at each
// interval's upper energy gives the quantiles at which the box edges go.
void build_learned(Proc* p, Mlp* m, int nn, double emin, double emax, int bw, int bk, int maxtries) {
  if (p->kind != kKindSAB) raise("BadInput", "not an S(alpha,beta) process");
  // the same energy range as the envelope sampler: inside the table's grid
  if (emax <= 0.0 || emax > p->sab_egrid_h.back()) emax = p->sab_egrid_h.back();
  if (emin <= 0.0 || emin < p->sab_egrid_h.front()) emin = p->sab_egrid_h.front();
  if (!(emin > 0.0 && emax > emin)) raise("BadInput", "learned sampler: need 0 < emin < emax");
  if (nn < 1 || nn > 1 << 16) raise("BadInput", "learned sampler: invalid number of intervals");
  const double lr = std::log(emax / emin) / nn;
  std::vector<double> eu((size_t)nn);
  for (int i = 0; i < nn; ++i) eu[(size_t)i] = (i + 1 == nn) ? emax : emin * std::exp(lr * (i + 1));
  m->ensure(nn);
  NCAMD_HIP(hipMemcpy(m->E.p, eu.data(), sizeof(double) * nn, hipMemcpyHostToDevice));
  launch_mlp_features(m->E.as<double>(), nn, m->lelo, m->linv, m->X.as<float>(), 0);
  mlp_forward(m, nn, 0);
  NCAMD_HIP(hipGetLastError());
  std::vector<float> y((size_t)nn * kMlpOut);
  NCAMD_HIP(hipMemcpy(y.data(), m->Y.p, sizeof(float) * y.size(), hipMemcpyDeviceToHost));
  std::vector<KBMix> mix;
  mix.reserve((size_t)nn);
  for (int i = 0; i < nn; ++i) mix.emplace_back(y.data() + (size_t)i * kMlpOut);
  build_kb(p, nn, emin, emax, bw, bk, maxtries, &mix);
}

// ---------------------------------------------------- batched evaluation ---
hipStream_t as_stream(void* s) { return static_cast<hipStream_t>(s); }

uint32_t next_stream(Proc* p) { return (uint32_t)(p->callctr++); }

unsigned long long read_reset_err(Proc* p, hipStream_t st) {
  unsigned long long n = 0;
  NCAMD_HIP(hipMemcpyAsync(&n, p->d_err.p, sizeof(n), hipMemcpyDeviceToHost, st));
  NCAMD_HIP(hipStreamSynchronize(st));
  if (n) NCAMD_HIP(hipMemsetAsync(p->d_err.p, 0, sizeof(n), st));
  return n;
}

// isotropic sampling dispatch: S(alpha,beta)-only handles take the split
// main/high-energy kernels, everything else the composition kernel
// two queues of up to `total` neutrons (and their counters) for the staged
// composition kernels
// queues of a staged composition: one per staged component and the
// high-energy queue, total neutrons each; kMaxStaged + 1 counters
void comp_queues(Proc* p, int64_t total, hipStream_t st) {
  const int nst = make_view(p).nstaged;
  const size_t need = (size_t)(nst + 1) * sizeof(int64_t) * (size_t)std::max<int64_t>(total, 1);
  if (p->q_idx.bytes < need) {
    NCAMD_HIP(hipStreamSynchronize(st));  // the old queue may still be in use on this stream
    p->q_idx.alloc(need);
  }
  if (p->q_count.bytes < (kMaxStaged + 1) * sizeof(unsigned long long))
    p->q_count.alloc((kMaxStaged + 1) * sizeof(unsigned long long));
}

void check_fp32(Proc* p) {
  CompView v = make_view(p);
  for (int i = 0; i < v.n; ++i)
    if (v.kind[i] == kKindPB && !v.pb.fdm32)
      raise("BadInput", "process was created without FP32 PowderBragg tables");
}

ScbWork scb_work(Proc* p, int64_t n, hipStream_t st);
Proc* comp_scb(Proc* p) {
  for (Proc* c : p->comps)
    if (c->kind == kKindSCB) return c;
  raise("LogicError", "oriented composition without an SCBragg component");
  return nullptr;
}
void oc_buffers(Proc* p, int64_t n, hipStream_t st) {
  const size_t m = (size_t)std::max<int64_t>(n, 1);
  if (p->oc_ext.bytes < sizeof(double) * m) {
    NCAMD_HIP(hipStreamSynchronize(st));  // the old buffers may still be in use on this stream
    p->oc_ext.ensure(sizeof(double) * m);
    p->oc_q.ensure(sizeof(int64_t) * m);
    p->oc_cE.ensure(sizeof(double) * m);
    p->oc_cD.ensure(sizeof(double) * 3 * m);
    p->oc_idmap.ensure(sizeof(int64_t) * m);
  }
  if (!p->oc_qc.p) p->oc_qc.alloc(sizeof(unsigned long long));
}
// oriented composition (NCProcImpl.cc:340-389 with the SCBragg component's
// oriented calls): the SCBragg cross sections of every neutron first, then
// the composition kernels with them in the SCBragg slot
void oriented_xs(Proc* p, const double* ekin, const double* dirs, int64_t n, double* out, hipStream_t st) {
  if (n <= 0) return;
  Proc* c = comp_scb(p);
  oc_buffers(p, n, st);
  const ScbWork w = scb_work(c, n, st);
  launch_scb_xs(c->scb, ekin, dirs, n, p->oc_ext.as<double>(), w, st);
  launch_xs_ext(make_view(p), ekin, p->oc_ext.as<double>(), n, out, st);
}
// the pick and the isotropic components' sampling as for any composition;
// the neutrons that pick the SCBragg component are gathered and sampled by
// the SCBragg kernels on their own Philox streams after the pick's draw
// (the host waits for their number)
void oriented_sample(Proc* p, const double* ekin, const double* dirs, int64_t n, double* oE, double* oD,
                     hipStream_t st) {
  const uint32_t stream = next_stream(p);
  if (n <= 0) return;
  Proc* c = comp_scb(p);
  oc_buffers(p, n, st);
  const ScbWork w = scb_work(c, n, st);
  launch_scb_xs(c->scb, ekin, dirs, n, p->oc_ext.as<double>(), w, st);
  const CompView v = make_view(p);
  comp_queues(p, n, st);
  launch_sample_dirs(v, ekin, dirs, n, p->seed, stream, oE, oD, p->d_err.as<unsigned long long>(),
                     p->q_idx.as<int64_t>(), p->q_count.as<unsigned long long>(), st, p->oc_ext.as<double>(),
                     p->oc_q.as<int64_t>(), p->oc_qc.as<unsigned long long>());
  unsigned long long m = 0;
  NCAMD_HIP(hipMemcpyAsync(&m, p->oc_qc.p, sizeof(m), hipMemcpyDeviceToHost, st));
  NCAMD_HIP(hipStreamSynchronize(st));
  if (!m) return;
  launch_scb_gather(p->oc_q.as<int64_t>(), p->oc_qc.as<unsigned long long>(), n, ekin, dirs, p->oc_cE.as<double>(),
                    p->oc_cD.as<double>(), p->oc_idmap.as<int64_t>(), st);
  launch_scb_sample(c->scb, p->oc_cE.as<double>(), p->oc_cD.as<double>(), (int64_t)m, p->seed, stream, oE, oD, w, st,
                    p->oc_idmap.as<int64_t>(), v.n > 1 ? 1 : 0);
}

// cross sections of any isotropic process or composition (the direction
// does not enter)
void xs_dispatch(Proc* p, const double* ekin, int64_t n, int64_t repeat, double* out, bool fp32, hipStream_t st) {
  if (p->oriented_comp) raise("BadInput", "oriented process: use the calls with neutron directions");
  if (p->kind == kKindFG || p->kind == kKindAbs) {
    if (fp32) raise("BadInput", "process was created without FP32 PowderBragg tables");
    launch_simple_xs(p->kind, p->fg, p->abs_c, ekin, n, repeat, out, st);
    return;
  }
  if (fp32) check_fp32(p);
  launch_xs_iso(make_view(p), ekin, n, repeat, out, fp32, st);
}

// an absorption process, or a composition of absorption processes
bool is_absorption(const Proc* p) {
  return p->kind == kKindAbs || (p->kind == kKindNone && !p->comps.empty() && p->comps[0]->kind == kKindAbs);
}

// the sampler mode of an S(alpha,beta) process rebuilds tables that clones
// read, so it is fixed while clones exist (and on clones)
void check_tables_mutable(const Proc* p) {
  if (p->src) raise("BadInput", "a clone cannot change its sampler (change the source before cloning)");
  if (p->nclones) raise("BadInput", "a process with live clones cannot change its sampler");
}

// Process::domain of the reference: PowderBragg and SCBragg from their
// threshold (NCPowderBragg.cc:147-150, NCSCBragg.cc:290-293; a PowderBragg
// without planes is null: (inf, inf)), the other processes everywhere
// (NCProcImpl.hh:446); a composition spans the union of its components'
// non-null domains, (0, 0) when it has none (NCProcImpl.hh:370,
// NCProcImpl.cc:285-296)
void proc_domain(const Proc* p, double& lo, double& hi) {
  const double inf = __builtin_huge_val();
  switch (p->kind) {
    case kKindPB:
      lo = p->pb.n > 0 ? p->pb.threshold : inf;
      hi = inf;
      return;
    case kKindSCB: lo = p->scb.threshold; hi = inf; return;
    case kKindLC: proc_domain(p->comps[0], lo, hi); return;
    case kKindNone: {
      lo = 0.0; hi = 0.0;
      for (const Proc* c : p->comps) {
        double a, b;
        proc_domain(c, a, b);
        if (a >= b) continue;
        if (lo >= hi) { lo = a; hi = b; }
        else { lo = std::min(lo, a); hi = std::max(hi, b); }
      }
      return;
    }
    default: lo = 0.0; hi = inf; return;
  }
}

// ProcImpl::Process::name of the reference's classes
const char* proc_name(const Proc* p) {
  switch (p->kind) {
    case kKindPB: return "PowderBragg";
    case kKindElInc: return "ElIncScatter";
    case kKindSAB: return "SABScatter";
    case kKindSCB: return "SCBragg";
    case kKindLC: return "LCBragg";
    case kKindFG: return "FreeGas";
    case kKindAbs: return "AbsOOV";
    case kKindNone: return "ProcComposition";
    default: return "unknown";
  }
}

// a process reading the tables of `s` with its own stream and staging memory
uint64_t clone_seed(uint64_t seed, uint64_t salt);
Proc* clone_proc(Proc* s, uint64_t seed) {
  ensure_device();
  Proc* root = s->src ? s->src : s;  // clones of clones read the root's tables
  std::unique_ptr<Proc> c(new Proc());
  c->kind = s->kind;
  c->seed = seed;
  c->pb = s->pb; c->ei = s->ei; c->sab = s->sab; c->fg = s->fg; c->abs_c = s->abs_c;
  c->oriented_comp = s->oriented_comp;
  c->pb_v0n = s->pb_v0n; c->pb_e2d_h = s->pb_e2d_h; c->pb_fdm_h = s->pb_fdm_h; c->pb_fdm32_h = s->pb_fdm32_h;
  c->has_fp32 = s->has_fp32;
  c->sab_na = s->sab_na; c->sab_nb = s->sab_nb; c->sab_alpha_h = s->sab_alpha_h; c->sab_beta_h = s->sab_beta_h;
  c->use_kb = s->use_kb; c->kb_maxtries = s->kb_maxtries; c->kb = s->kb;
  c->kb_node_h = s->kb_node_h; c->kb_box_h = s->kb_box_h; c->kb_stats_h = s->kb_stats_h;
  c->sab_egrid_h = s->sab_egrid_h; c->sab_xs_h = s->sab_xs_h;
  c->scb = s->scb;
  c->scb_inv2d_h = s->scb_inv2d_h; c->scb_xsfact_h = s->scb_xsfact_h; c->scb_normals_h = s->scb_normals_h;
  c->scb_lt1_h = s->scb_lt1_h; c->scb_lt2_h = s->scb_lt2_h; c->scb_off_h = s->scb_off_h;
  std::copy(s->scb_gos, s->scb_gos + 15, c->scb_gos);
  c->lc_mode = s->lc_mode; c->lc_n = s->lc_n; c->lc_nprime = s->lc_nprime; c->lc = s->lc;
  c->lc0 = s->lc0;  // plane sets read from the source's buffer (the clone keeps the source alive)
  c->comps = s->comps;
  c->scales = s->scales;
  init_errbuf(c.get());
  for (Proc* x : c->comps) x->refcount++;
  if (s->kind == kKindLC) {  // the crystal's work buffers are per process: the clone gets its own crystal
    Proc* own = clone_proc(s->comps[0], clone_seed(seed, 1));
    --c->comps[0]->refcount;
    c->comps[0] = own;
    c->lc_cs.alloc(s->lc_cs.bytes);
    NCAMD_HIP(hipMemcpy(c->lc_cs.p, s->lc_cs.p, s->lc_cs.bytes, hipMemcpyDeviceToDevice));
  }
  c->src = root;
  root->refcount++;
  root->nclones++;
  return c.release();
}

// seed of a clone: splitmix64 of (source seed, salt)
uint64_t clone_seed(uint64_t seed, uint64_t salt) {
  uint64_t z = seed ^ (0x9E3779B97F4A7C15ull * (salt + 1));
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

void sample_dirs_dispatch(Proc* p, const double* ekin, const double* dirs, int64_t n, double* oE, double* oD,
                          hipStream_t st) {
  if (p->oriented_comp) {
    oriented_sample(p, ekin, dirs, n, oE, oD, st);
    return;
  }
  if (is_absorption(p)) raise("BadInput", "an absorption process does not scatter");
  if (p->kind == kKindFG) {
    launch_fg_sample(p->fg, ekin, dirs, n, 1, p->seed, next_stream(p), oE, oD, p->d_err.as<unsigned long long>(),
                     st);
    return;
  }
  const CompView v = make_view(p);
  comp_queues(p, n, st);
  launch_sample_dirs(v, ekin, dirs, n, p->seed, next_stream(p), oE, oD, p->d_err.as<unsigned long long>(),
                     p->q_idx.as<int64_t>(), p->q_count.as<unsigned long long>(), st);
}

// ext_q / ext_qc: caller-provided queue (2 x n x repeat entries) and two
// counters; then no device memory is allocated here
void sample_iso_dispatch(Proc* p, const double* ekin, int64_t n, int64_t repeat, double* oE, double* oMu,
                         hipStream_t st, int64_t* ext_q = nullptr, unsigned long long* ext_qc = nullptr) {
  if (p->oriented_comp) raise("BadInput", "oriented process: use the calls with neutron directions");
  if (is_absorption(p)) raise("BadInput", "an absorption process does not scatter");
  if (p->kind == kKindFG) {
    launch_fg_sample(p->fg, ekin, nullptr, n, repeat, p->seed, next_stream(p), oE, oMu,
                     p->d_err.as<unsigned long long>(), st);
    return;
  }
  const uint32_t stream = next_stream(p);
  Proc* single = p->kind == kKindSAB ? p : nullptr;
  if (!single && p->kind == kKindNone && p->comps.size() == 1 && p->comps[0]->kind == kKindSAB &&
      p->scales[0] == 1.0)
    single = p->comps[0];
  if (single) {
    if (!ext_q) {
      const size_t need = sizeof(int64_t) * (size_t)std::max<int64_t>(n * repeat, 1);
      if (p->q_idx.bytes < need) {
        NCAMD_HIP(hipStreamSynchronize(st));  // the old queue may still be in use on this stream
        p->q_idx.alloc(need);
      }
      if (!p->q_count.p) p->q_count.alloc(2 * sizeof(unsigned long long));
    }
    int64_t* q = ext_q ? ext_q : p->q_idx.as<int64_t>();
    unsigned long long* qc = ext_q ? ext_qc : p->q_count.as<unsigned long long>();
    if (single->use_kb)
      launch_kb_sample_iso(single->sab, single->kb, ekin, n, repeat, p->seed, stream, oE, oMu,
                           p->d_err.as<unsigned long long>(), q, qc, single->kb_maxtries, st);
    else
      launch_sab_sample_iso(single->sab, ekin, n, repeat, p->seed, stream, oE, oMu,
                            p->d_err.as<unsigned long long>(), q, qc, p != single ? 1 : 0, st);
  } else {
    if (!ext_q) comp_queues(p, n * repeat, st);
    launch_sample_iso(make_view(p), ekin, n, repeat, p->seed, stream, oE, oMu, p->d_err.as<unsigned long long>(),
                      ext_q ? ext_q : p->q_idx.as<int64_t>(),
                      ext_q ? ext_qc : p->q_count.as<unsigned long long>(), st);
  }
}

// ------------------------------------------ allocation-free sampling call ---
// Fixed scratch for ncamd_dev_samplescatterisotropic_many_noalloc: two
// counters (in the first 256 bytes) and the sampling queues (two int64
// entries per neutron) for at most kNoAllocMaxNeutrons neutrons per call,
// 32 MiB + 256 B, held by every GPU the library runs on until the process
// ends.  Host code uses the address from hipGetSymbolAddress, never the
// array's name.
constexpr int64_t kNoAllocMaxNeutrons = int64_t(1) << 21;
constexpr size_t kNoAllocCounterBytes = 256;
constexpr size_t kNoAllocBytes = kNoAllocCounterBytes + 2 * sizeof(int64_t) * (size_t)kNoAllocMaxNeutrons;
static __device__ __align__(256) unsigned char scratch[kNoAllocBytes];

// the device address of `scratch` on the current device, looked up once per
// device (with that device current)
unsigned char* noalloc_scratch() {
  int dev = 0;
  NCAMD_HIP(hipGetDevice(&dev));
  static std::mutex mtx;
  static std::map<int, unsigned char*> ptrs;
  std::lock_guard<std::mutex> lk(mtx);
  auto it = ptrs.find(dev);
  if (it != ptrs.end()) return it->second;
  void* ptr = nullptr;
  const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !ptr)
    raise("HIPError", std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
  ptrs[dev] = static_cast<unsigned char*>(ptr);
  return ptrs[dev];
}

// work buffers of the SCBragg kernels (batch ordering, Romberg queue), sized
// for n neutrons
ScbWork scb_work(Proc* p, int64_t n, hipStream_t st) {
  const size_t m = (size_t)std::max<int64_t>(n, 1);
  if (p->q_idx.bytes < sizeof(int64_t) * m || p->scb_keys.bytes < sizeof(unsigned) * m ||
      p->scb_sel.bytes < sizeof(double4) * m) {
    NCAMD_HIP(hipStreamSynchronize(st));  // the old buffers may still be in use on this stream
    p->q_idx.ensure(sizeof(int64_t) * m);
    p->scb_perm.ensure(sizeof(int64_t) * m);
    p->scb_keys.ensure(sizeof(unsigned) * m);
    p->scb_cnt.ensure(sizeof(int) * m);
    p->scb_tot.ensure(sizeof(double) * m);
    p->scb_sel.ensure(sizeof(double4) * m);
  }
  if (!p->q_count.p) p->q_count.alloc(2 * sizeof(unsigned long long));
  if (!p->scb_hist.p) {
    p->scb_hist.alloc(sizeof(unsigned) * kScbBuckets);
    NCAMD_HIP(hipMemset(p->scb_hist.p, 0, sizeof(unsigned) * kScbBuckets));
    p->scb_cursor.alloc(sizeof(unsigned long long) * kScbBuckets);
  }
  ScbWork w;
  w.keys = p->scb_keys.as<unsigned>();
  w.perm = p->scb_perm.as<int64_t>();
  w.hist = p->scb_hist.as<unsigned>();
  w.cursor = p->scb_cursor.as<unsigned long long>();
  w.queue = p->q_idx.as<int64_t>();
  w.qcount = p->q_count.as<unsigned long long>();
  w.cnt = p->scb_cnt.as<int>();
  w.tot = p->scb_tot.as<double>();
  w.sel = p->scb_sel.as<double4>();
  return w;
}

// ---- layered crystal by the reference models (NCLCRefModels.cc:25-171)
constexpr int64_t kLcLanes = int64_t(1) << 23;  // rotated neutrons per pass

// work arrays for `lanes` rotated neutrons and m neutrons
void lc_buffers(Proc* p, int64_t lanes, int64_t m, hipStream_t st) {
  const size_t L = (size_t)std::max<int64_t>(lanes, 1), M = (size_t)std::max<int64_t>(m, 1);
  if (p->lc_E.bytes < sizeof(double) * L || p->lc_sE.bytes < sizeof(double) * M) {
    NCAMD_HIP(hipStreamSynchronize(st));  // the old buffers may still be in use on this stream
    p->lc_E.ensure(sizeof(double) * L);
    p->lc_D.ensure(3 * sizeof(double) * L);
    p->lc_xs.ensure(sizeof(double) * L);
    p->lc_csr.ensure(sizeof(double2) * L);
    p->lc_sE.ensure(sizeof(double) * M);
    p->lc_sD.ensure(3 * sizeof(double) * M);
    p->lc_sel.ensure(sizeof(int) * M);
    p->lc_oE.ensure(sizeof(double) * M);
    p->lc_oD.ensure(3 * sizeof(double) * M);
  }
}

// cross sections (LCBraggRef::crossSection: mean over equally spaced
// rotations phi_i = 2 pi i / N - pi, N the smallest prime >= nsample;
// LCBraggRndmRot::crossSection: mean over nsample random rotations, new ones
// on every call)
void lc_xs(Proc* p, const double* E, const double* D, int64_t n, double* out, hipStream_t st) {
  Proc* sc = p->comps[0];
  if (p->lc_mode == 0) {
    launch_lc0_xs(p->lc0, sc->scb, E, D, n, out, st);
    return;
  }
  const bool fixed = p->lc_mode > 0;
  const int N = fixed ? p->lc_nprime : p->lc_n;
  const int64_t chunk = std::max<int64_t>(1, kLcLanes / N);
  for (int64_t i0 = 0; i0 < n; i0 += chunk) {
    const int64_t m = std::min(chunk, n - i0);
    lc_buffers(p, m * N, m, st);
    launch_lc_expand(E + i0, D + 3 * i0, m, N, p->lc, fixed ? p->lc_cs.as<double2>() : nullptr, p->seed,
                     fixed ? 0u : next_stream(p), p->lc_E.as<double>(), p->lc_D.as<double>(), p->lc_csr.as<double2>(),
                     st);
    const ScbWork w = scb_work(sc, m * N, st);
    launch_scb_xs(sc->scb, p->lc_E.as<double>(), p->lc_D.as<double>(), m * N, p->lc_xs.as<double>(), w, st);
    launch_lc_mean(p->lc_xs.as<double>(), m, N, out + i0, st);
  }
}

// scattering (LCBraggRef / LCBraggRndmRot::sampleScatter): nsample random
// rotations, one picked with probability proportional to its cross section,
// the crystal's scattering of the rotated neutron, the outgoing direction
// rotated back; neutrons without cross section at every rotation leave
// unchanged.  The random-rotation model draws new rotations here instead of
// reusing those of a preceding cross-section call (the same distribution).
void lc_sample(Proc* p, const double* E, const double* D, int64_t n, double* oE, double* oD, hipStream_t st) {
  Proc* sc = p->comps[0];
  if (p->lc_mode == 0) {
    launch_lc0_sample(p->lc0, sc->scb, E, D, n, p->seed, next_stream(p), oE, oD, st);
    return;
  }
  const int N = p->lc_n;
  const int64_t chunk = std::max<int64_t>(1, kLcLanes / N);
  for (int64_t i0 = 0; i0 < n; i0 += chunk) {
    const int64_t m = std::min(chunk, n - i0);
    lc_buffers(p, m * N, m, st);
    launch_lc_expand(E + i0, D + 3 * i0, m, N, p->lc, nullptr, p->seed, next_stream(p), p->lc_E.as<double>(),
                     p->lc_D.as<double>(), p->lc_csr.as<double2>(), st);
    const ScbWork wx = scb_work(sc, m * N, st);
    launch_scb_xs(sc->scb, p->lc_E.as<double>(), p->lc_D.as<double>(), m * N, p->lc_xs.as<double>(), wx, st);
    launch_lc_pick(p->lc_xs.as<double>(), p->lc_E.as<double>(), p->lc_D.as<double>(), m, N, p->seed, next_stream(p),
                   p->lc_sE.as<double>(), p->lc_sD.as<double>(), p->lc_sel.as<int>(), st);
    const ScbWork ws = scb_work(sc, m, st);
    launch_scb_sample(sc->scb, p->lc_sE.as<double>(), p->lc_sD.as<double>(), m, p->seed, next_stream(p),
                      p->lc_oE.as<double>(), p->lc_oD.as<double>(), ws, st);
    launch_lc_unrotate(p->lc_oD.as<double>(), p->lc_sel.as<int>(), p->lc_csr.as<double2>(), E + i0, D + 3 * i0, m, N,
                       p->lc, oE + i0, oD + 3 * i0, st);
  }
}


}  // namespace
}  // namespace ncamd

using namespace ncamd;

extern "C" {

const char* ncamd_version_str(void) { return "0.1.0 (NCrystal 4.4.2 physics)"; }
int ncamd_device_count(void) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) return 0;
  return n;
}
int ncamd_error(void) { return g_waserror; }
const char* ncamd_lasterror(void) { return g_waserror ? g_errmsg : ""; }
const char* ncamd_lasterrortype(void) { return g_waserror ? g_errtype : ""; }
void ncamd_clearerror(void) { g_waserror = 0; g_errmsg[0] = '\0'; g_errtype[0] = '\0'; }
int ncamd_sethaltonerror(int h) { const int o = g_halt; g_halt = h; return o; }
int ncamd_setquietonerror(int q) { const int o = g_quiet; g_quiet = q; return o; }
void ncamd_seterrhandler(void (*handler)(char*, char*)) { g_errhandler = handler; }

void ncamd_ref(ncamd_process_t h) {
  NCAMD_TRY extract(h)->refcount++; NCAMD_CATCH
}
void ncamd_unref(ncamd_process_t* h) {
  NCAMD_TRY
  Proc* p = extract(*h);
  if (--p->refcount == 0) delete p;
  h->internal = nullptr;
  NCAMD_CATCH
}
int ncamd_refcount(ncamd_process_t h) {
  NCAMD_TRY return extract(h)->refcount; NCAMD_CATCH
  return -1;
}
int ncamd_valid(ncamd_process_t h) {
  Proc* p = static_cast<Proc*>(h.internal);
  return (p && p->magic == kMagic) ? 1 : 0;
}
int ncamd_isoriented(ncamd_process_t h) {
  NCAMD_TRY
  const Proc* p = extract(h);
  return (p->kind == kKindSCB || p->kind == kKindLC || p->oriented_comp) ? 1 : 0;
  NCAMD_CATCH
  return -1;
}

double ncamd_debye_msd(double dt, double t, double am) {
  NCAMD_TRY return debye_msd(dt, t, am); NCAMD_CATCH
  return -1.0;
}

int64_t ncamd_hkl_orthogonal(double a, double b, double c, int nspecies, const int* natoms, const double* positions,
                             const double* csl, const double* msd, double dcutoff, double fsquarecut,
                             double merge_tolerance, int64_t maxout, int* out_hkl, int* out_mult, double* out_dsp,
                             double* out_fsq, int64_t* nrequired) {
  NCAMD_TRY
  return hkl_nosg(a, b, c, 90.0, 90.0, 90.0, nspecies, natoms, positions, csl, msd, dcutoff, fsquarecut, merge_tolerance, maxout,
                        out_hkl, out_mult, out_dsp, out_fsq, nrequired);
  NCAMD_CATCH
  return -3;
}
int64_t ncamd_hkl_orthogonal_members(double a, double b, double c, int nspecies, const int* natoms,
                                     const double* positions, const double* csl, const double* msd, double dcutoff,
                                     double fsquarecut, double merge_tolerance, int64_t maxout, int* out_hkl,
                                     int* out_mult, double* out_dsp, double* out_fsq, int64_t* nrequired,
                                     int64_t maxmem, int64_t* out_mem_off, int* out_mem_hkl,
                                     int64_t* nmem_required) {
  NCAMD_TRY
  if (!out_mem_off || !out_mem_hkl) raise("BadInput", "member output arrays are required");
  return hkl_nosg(a, b, c, 90.0, 90.0, 90.0, nspecies, natoms, positions, csl, msd, dcutoff, fsquarecut,
                  merge_tolerance, maxout, out_hkl, out_mult, out_dsp, out_fsq, nrequired, maxmem, out_mem_off,
                  out_mem_hkl, nmem_required);
  NCAMD_CATCH
  return -3;
}

int64_t ncamd_hkl_cell_members(double a, double b, double c, double alpha, double beta, double gamma, int nspecies,
                               const int* natoms, const double* positions, const double* csl, const double* msd,
                               double dcutoff, double fsquarecut, double merge_tolerance, int64_t maxout,
                               int* out_hkl, int* out_mult, double* out_dsp, double* out_fsq, int64_t* nrequired,
                               int64_t maxmem, int64_t* out_mem_off, int* out_mem_hkl, int64_t* nmem_required) {
  NCAMD_TRY
  if (!out_mem_off || !out_mem_hkl) raise("BadInput", "member output arrays are required");
  return hkl_nosg(a, b, c, alpha, beta, gamma, nspecies, natoms, positions, csl, msd, dcutoff, fsquarecut,
                  merge_tolerance, maxout, out_hkl, out_mult, out_dsp, out_fsq, nrequired, maxmem, out_mem_off,
                  out_mem_hkl, nmem_required);
  NCAMD_CATCH
  return -3;
}

int64_t ncamd_hkl_spacegroup_members(int spacegroup, double a, double b, double c, double alpha, double beta,
                                     double gamma, int nspecies, const int* natoms, const double* positions,
                                     const double* csl, const double* msd, double dcutoff, double fsquarecut,
                                     int64_t maxout, int* out_hkl, int* out_mult, double* out_dsp, double* out_fsq,
                                     int64_t* nrequired, int64_t maxmem, int64_t* out_mem_off, int* out_mem_hkl,
                                     int64_t* nmem_required) {
  NCAMD_TRY
  if (!out_mem_off != !out_mem_hkl) raise("BadInput", "give both member output arrays or neither");
  return hkl_spacegroup(spacegroup, a, b, c, alpha, beta, gamma, nspecies, natoms, positions, csl, msd, dcutoff,
                        fsquarecut, maxout, out_hkl, out_mult, out_dsp, out_fsq, nrequired, maxmem, out_mem_off,
                        out_mem_hkl, nmem_required);
  NCAMD_CATCH
  return -3;
}

int ncamd_reci_lattice(double a, double b, double c, double alpha, double beta, double gamma, double* out9) {
  NCAMD_TRY
  if (!(a > 0 && b > 0 && c > 0)) raise("BadInput", "lattice parameters must be positive");
  if (!reci_lattice(a, b, c, alpha, beta, gamma, out9))
    raise("BadInput", "invalid cell: angles must lie in (0, 180) degrees and admit a cell");
  return 0;
  NCAMD_CATCH
  return -1;
}

// LCBragg by a reference model (NCLCBragg.cc:62-73, NCLCRefModels.cc:25-171):
// mode > 0 LCBraggRef with nsample = mode, mode < 0 LCBraggRndmRot with
// nsample = -mode; the layer axis in the lab frame
ncamd_process_t ncamd_create_lcbragg_ref(ncamd_process_t scbragg, const double* lcaxis_lab, int mode) {
  NCAMD_TRY
  ensure_device();
  Proc* sc = extract(scbragg);
  if (sc->kind != kKindSCB) raise("BadInput", "LCBragg: the crystal must be an SCBragg process");
  if (mode == 0) raise("BadInput", "LCBragg: mode 0 (the layered-crystal helper) is not available; use a reference model (mode != 0)");
  if (mode > 100000 || mode < -100000) raise("BadInput", "LCBragg: at most 100000 rotations per neutron");
  const double ax = lcaxis_lab[0], ay = lcaxis_lab[1], az = lcaxis_lab[2];
  const double norm = std::sqrt(ax * ax + ay * ay + az * az);
  if (!(norm > 0.0) || !std::isfinite(norm)) raise("BadInput", "LCBragg: invalid layer axis");
  std::unique_ptr<Proc> p(new Proc());
  p->kind = kKindLC;
  p->lc_mode = mode;
  p->lc_n = mode > 0 ? mode : -mode;
  int np = p->lc_n;
  auto prime = [](int k) {
    if (k < 2) return false;
    for (int d = 2; (int64_t)d * d <= k; ++d)
      if (k % d == 0) return false;
    return true;
  };
  while (!prime(np)) ++np;
  p->lc_nprime = np;
  p->lc = LCView{ax / norm, ay / norm, az / norm};
  if (mode > 0) {  // rotations 2 pi i / N - pi of the cross sections
    std::vector<double> cs(2 * (size_t)np);
    const double dphi = k2Pi / np;
    for (int i = 0; i < np; ++i) {
      const double phi = i * dphi - kPi;
      cs[2 * i] = std::cos(phi);
      cs[2 * i + 1] = std::sin(phi);
    }
    p->lc_cs.upload(cs.data(), cs.size());
  }
  sc->refcount++;
  p->comps.push_back(sc);
  p->scales.push_back(1.0);
  init_errbuf(p.get());
  return wrap(p.release());
  NCAMD_CATCH
  return wrap(nullptr);
}

// LCBragg (NCLCBragg.cc:36-73) from the inputs of an SCBragg of the crystal
// and the layer axis in the crystal frame: mode 0 the layered-crystal model
// (LCHelper, NCLCUtils.cc:58-131 plane sets), else a reference model
ncamd_process_t ncamd_create_lcbragg(double v0n, int64_t nfam, const double* dsp, const double* fsq,
                                     const int64_t* mem_off, const int* mem_hkl, const double* reci_lattice,
                                     const double* cry2lab, const double* lcaxis_crystal, double mos_fwhm,
                                     double prec, double ntrunc, int mode) {
  NCAMD_TRY
  ensure_device();
  double ac[3] = {lcaxis_crystal[0], lcaxis_crystal[1], lcaxis_crystal[2]};
  const double an = std::sqrt(ac[0] * ac[0] + ac[1] * ac[1] + ac[2] * ac[2]);
  if (!(an > 0.0) || !std::isfinite(an)) raise("BadInput", "LCBragg: invalid layer axis");
  for (double& x : ac) x /= an;
  double al[3];
  for (int r = 0; r < 3; ++r) al[r] = cry2lab[3 * r] * ac[0] + cry2lab[3 * r + 1] * ac[1] + cry2lab[3 * r + 2] * ac[2];
  std::unique_ptr<Proc> sc(create_scb(v0n, nfam, dsp, fsq, mem_off, mem_hkl, reci_lattice, cry2lab, mos_fwhm, prec,
                                      ntrunc));
  if (mode != 0) {
    ncamd_process_t h = ncamd_create_lcbragg_ref(wrap(sc.get()), al, mode);
    if (!h.internal) return h;  // error state already set
    --sc->refcount;  // the LCBragg process holds the crystal
    sc.release();
    return h;
  }
  // plane sets: crystal-frame demi-normals by the angle of the one in the
  // layer axis' hemisphere, merged on the 2^-40 grid of (d, angle), sorted
  // by decreasing (d, angle); values from the keys
  constexpr double kDiscr = 1099511627776.0;
  std::map<std::pair<uint64_t, uint64_t>, double, std::greater<std::pair<uint64_t, uint64_t>>> sets;
  for (int64_t f = 0; f < nfam; ++f) {
    if (!(dsp[f] > 0.0 && dsp[f] < 1e7) || !(fsq[f] > 0.0)) raise("BadInput", "LCBragg: invalid family");
    for (int64_t j = mem_off[f]; j < mem_off[f + 1]; ++j) {
      const double h = mem_hkl[3 * j], k = mem_hkl[3 * j + 1], l = mem_hkl[3 * j + 2];
      double nv[3];
      for (int r = 0; r < 3; ++r)
        nv[r] = reci_lattice[3 * r] * h + reci_lattice[3 * r + 1] * k + reci_lattice[3 * r + 2] * l;
      const double nn = std::sqrt(nv[0] * nv[0] + nv[1] * nv[1] + nv[2] * nv[2]);
      const double ca = std::fabs((ac[0] * nv[0] + ac[1] * nv[1] + ac[2] * nv[2]) / nn);
      const double alpha = ca > 0.9999999 ? 0.0 : std::acos(ca);
      const auto key = std::make_pair((uint64_t)(dsp[f] * kDiscr + 0.5), (uint64_t)(alpha * kDiscr + 0.5));
      sets[key] += fsq[f];
    }
  }
  const double trunc = sc->scb_gos[1];
  std::vector<LC0Set> hs;
  for (const auto& e : sets) {
    const double d = e.first.first * (1.0 / kDiscr);
    const double alpha = std::min(e.first.second * (1.0 / kDiscr), kPiHalf);
    LC0Set ps{};
    ps.twodsp = 2.0 * d;
    ps.inv2d = 0.5 / d;
    ps.c1 = alpha != 0.0 ? std::cos(alpha) : 1.0;
    ps.s1 = alpha != 0.0 ? std::sin(alpha) : 0.0;
    ps.c1minus = alpha > trunc ? std::cos(alpha - trunc) : 1.0;
    ps.c1plus = std::cos(alpha + trunc);
    ps.fsq = e.second;
    hs.push_back(ps);
  }
  std::unique_ptr<Proc> p(new Proc());
  p->kind = kKindLC;
  p->lc_mode = 0;
  p->lc = LCView{al[0], al[1], al[2]};
  const double aln = std::sqrt(al[0] * al[0] + al[1] * al[1] + al[2] * al[2]);
  p->lc0_sets.alloc(sizeof(LC0Set) * std::max<size_t>(hs.size(), 1));
  if (!hs.empty()) NCAMD_HIP(hipMemcpy(p->lc0_sets.p, hs.data(), sizeof(LC0Set) * hs.size(), hipMemcpyHostToDevice));
  p->lc0.sets = p->lc0_sets.as<LC0Set>();
  p->lc0.nsets = (int)hs.size();
  p->lc0.ax = al[0] / aln; p->lc0.ay = al[1] / aln; p->lc0.az = al[2] / aln;
  p->lc0.cta = sc->scb_gos[3];
  p->lc0.sta = sc->scb_gos[4];
  p->lc0.acc = dclamp(sc->scb_gos[2], 1e-7, 1e-2);
  p->lc0.xsfact = 1.0 / v0n;
  p->comps.push_back(sc.release());
  p->scales.push_back(1.0);
  init_errbuf(p.get());
  return wrap(p.release());
  NCAMD_CATCH
  return wrap(nullptr);
}

ncamd_process_t ncamd_create_scbragg(double v0n, int64_t nfam, const double* dsp, const double* fsq,
                                     const int64_t* mem_off, const int* mem_hkl, const double* reci_lattice,
                                     const double* cry2lab, double mos_fwhm, double prec, double ntrunc) {
  NCAMD_TRY
  return wrap(create_scb(v0n, nfam, dsp, fsq, mem_off, mem_hkl, reci_lattice, cry2lab, mos_fwhm, prec, ntrunc));
  NCAMD_CATCH
  return wrap(nullptr);
}
int ncamd_scbragg_info(ncamd_process_t h, int64_t* nfam, int64_t* nnormals, double* threshold, double* gos) {
  NCAMD_TRY
  Proc* p = extract(h);
  if (p->kind != kKindSCB) raise("BadInput", "not an SCBragg process");
  *nfam = p->scb.nfam;
  *nnormals = p->scb_off_h.back();
  *threshold = p->scb.threshold;
  std::copy(p->scb_gos, p->scb_gos + 15, gos);
  return 0;
  NCAMD_CATCH
  return -1;
}
int ncamd_scbragg_tables(ncamd_process_t h, double* inv2d, double* xsfact, int64_t* off, double* normals,
                         double* lt_evalcosx, double* lt_sofcosd) {
  NCAMD_TRY
  Proc* p = extract(h);
  if (p->kind != kKindSCB) raise("BadInput", "not an SCBragg process");
  std::copy(p->scb_inv2d_h.begin(), p->scb_inv2d_h.end(), inv2d);
  std::copy(p->scb_xsfact_h.begin(), p->scb_xsfact_h.end(), xsfact);
  std::copy(p->scb_off_h.begin(), p->scb_off_h.end(), off);
  std::copy(p->scb_normals_h.begin(), p->scb_normals_h.end(), normals);
  std::copy(p->scb_lt1_h.begin(), p->scb_lt1_h.end(), lt_evalcosx);
  std::copy(p->scb_lt2_h.begin(), p->scb_lt2_h.end(), lt_sofcosd);
  return 0;
  NCAMD_CATCH
  return -1;
}

ncamd_process_t ncamd_create_powderbragg(double v0n, const double* dsp, const double* fsqm, int64_t n, int fp32) {
  NCAMD_TRY return wrap(create_pb(v0n, dsp, fsqm, n, fp32 != 0)); NCAMD_CATCH
  return wrap(nullptr);
}
ncamd_process_t ncamd_create_elincscatter(int n, const double* msd, const double* bixs, const double* scale) {
  NCAMD_TRY return wrap(create_elinc(n, msd, bixs, scale)); NCAMD_CATCH
  return wrap(nullptr);
}
ncamd_process_t ncamd_create_freegas(double temperature, double mass_amu, double sigma_free) {
  NCAMD_TRY
  ensure_device();
  if (!(temperature > 0.0) || !std::isfinite(temperature)) raise("BadInput", "temperature must be > 0");
  if (!(mass_amu > 0.0) || !std::isfinite(mass_amu)) raise("BadInput", "target mass must be > 0");
  if (!(sigma_free > 0.0) || !std::isfinite(sigma_free)) raise("BadInput", "free cross section must be > 0");
  std::unique_ptr<Proc> p(new Proc());
  p->kind = kKindFG;
  const double kT = kBoltzmann * temperature;
  p->fg.kT = kT;
  p->fg.A = (1.0 / kNeutronMassAmu) * mass_amu;
  p->fg.sigma_free = sigma_free;
  p->fg.ca = (mass_amu / kNeutronMassAmu) / kT;  // NCFreeGasUtils.cc:45-53: a^2 = ca E
  init_errbuf(p.get());
  return wrap(p.release());
  NCAMD_CATCH
  return wrap(nullptr);
}

ncamd_process_t ncamd_create_absoov(double sigma_abs) {
  NCAMD_TRY
  ensure_device();
  if (!(sigma_abs >= 0.0) || !std::isfinite(sigma_abs)) raise("BadInput", "absorption cross section must be >= 0");
  std::unique_ptr<Proc> p(new Proc());
  p->kind = kKindAbs;
  // const_ekin_2200m_s (ncrystal_core/include/NCrystal/internal/utils/NCMath.hh:45-46)
  const double mevc2 = kNeutronMassAmu * kDalton2eVc2 / (kSpeedOfLightAas * kSpeedOfLightAas);
  const double ekin2200 = 0.5 * mevc2 * 2200.0e10 * 2200.0e10;
  p->abs_c = std::sqrt(ekin2200) * sigma_abs;
  init_errbuf(p.get());
  return wrap(p.release());
  NCAMD_CATCH
  return wrap(nullptr);
}

ncamd_process_t ncamd_create_sabscatter(int na, int nb, const double* alpha, const double* beta, const double* sab,
                                        double T, double mass, double sigma_bound, double suggested_emax,
                                        int n_egrid, const double* egrid) {
  NCAMD_TRY
  return wrap(create_sab(na, nb, alpha, beta, sab, T, mass, sigma_bound, suggested_emax, n_egrid, egrid));
  NCAMD_CATCH
  return wrap(nullptr);
}
// ProcComposition::addComponent (ncrystal_core/src/interfaces/NCProcImpl.cc:
// 263-331): zero scales and null processes are dropped, a nested composition
// adds its components with multiplied scales, a process already present adds
// to its scale, a process that merges with a present one replaces it by the
// merged process at scale 1, anything else is appended.
void comp_add(Proc* p, Proc* c, double scale) {
  if (!(scale >= 0.0) || std::isinf(scale)) raise("BadInput", "component scales must be finite and >= 0");
  if (c == p) raise("BadInput", "a composition cannot contain itself");
  if (scale == 0.0 || (c->kind == kKindPB && c->pb.n == 0)) return;
  if (c->kind == kKindNone) {
    for (size_t i = 0; i < c->comps.size(); ++i) comp_add(p, c->comps[i], c->scales[i] * scale);
    return;
  }
  for (size_t i = 0; i < p->comps.size(); ++i) {
    Proc* e = p->comps[i];
    if (e == c) {
      p->scales[i] += scale;
      return;
    }
    if (Proc* m = try_merge(e, p->scales[i], c, scale)) {
      if (--e->refcount == 0) delete e;
      p->comps[i] = m;
      p->scales[i] = 1.0;
      return;
    }
  }
  c->refcount++;
  p->comps.push_back(c);
  p->scales.push_back(scale);
}

ncamd_process_t ncamd_create_composition(int n, const ncamd_process_t* procs, const double* scales) {
  NCAMD_TRY
  ensure_device();
  if (n < 1) raise("BadInput", "composition needs at least one component");
  std::unique_ptr<Proc> p(new Proc());
  p->kind = kKindNone;
  for (int i = 0; i < n; ++i) comp_add(p.get(), extract(procs[i]), scales[i]);
  if (p->comps.empty()) raise("BadInput", "composition has no non-null component");
  if (p->comps.size() > (size_t)kMaxComp) raise("BadInput", "composition has more than 8 components after merging");
  int seen[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  for (Proc* c : p->comps) {
    if (c->kind == kKindLC) raise("BadInput", "an LCBragg process cannot be a component of a composition");
    if (c->kind == kKindSCB) p->oriented_comp = true;  // oriented composition: calls with directions only
    if (c->kind == kKindSCB && seen[c->kind]) raise("BadInput", "at most one SCBragg process per composition");
    seen[c->kind]++;
  }
  // one process type per composition (NCProcImpl.cc:269-272)
  if (seen[kKindAbs] && seen[kKindAbs] != (int)p->comps.size())
    raise("BadInput", "a composition holds either scattering or absorption processes");
  if (!seen[kKindAbs]) (void)make_view(p.get());  // checks the staged-component limit
  init_errbuf(p.get());
  return wrap(p.release());
  NCAMD_CATCH
  return wrap(nullptr);
}

int ncamd_composition_component(ncamd_process_t h, int i, ncamd_process_t* component, double* scale) {
  NCAMD_TRY
  Proc* p = extract(h);
  if (p->kind != kKindNone) raise("BadInput", "not a composition");
  const int n = (int)p->comps.size();
  if (i >= 0 && i < n) {
    if (!component || !scale) raise("BadInput", "null output pointer");
    p->comps[i]->refcount++;
    *component = wrap(p->comps[i]);
    *scale = p->scales[i];
  }
  return n;
  NCAMD_CATCH
  return -1;
}

int64_t ncamd_powderbragg_tables(ncamd_process_t h, double* e2d, double* fdm, float* fdm32, int64_t maxn) {
  NCAMD_TRY
  Proc* p = extract(h);
  if (p->kind != kKindPB) raise("BadInput", "not a PowderBragg process");
  const int64_t n = (int64_t)p->pb_e2d_h.size();
  if (n > maxn) return n;
  std::copy(p->pb_e2d_h.begin(), p->pb_e2d_h.end(), e2d);
  std::copy(p->pb_fdm_h.begin(), p->pb_fdm_h.end(), fdm);
  if (fdm32 && p->has_fp32) std::copy(p->pb_fdm32_h.begin(), p->pb_fdm32_h.end(), fdm32);
  return n;
  NCAMD_CATCH
  return -1;
}
int ncamd_sab_egrid(ncamd_process_t h, double* egrid, double* xs, int maxn) {
  NCAMD_TRY
  Proc* p = extract(h);
  if (p->kind != kKindSAB) raise("BadInput", "not an S(alpha,beta) process");
  const int n = (int)p->sab_egrid_h.size();
  if (n > maxn) return n;
  std::copy(p->sab_egrid_h.begin(), p->sab_egrid_h.end(), egrid);
  std::copy(p->sab_xs_h.begin(), p->sab_xs_h.end(), xs);
  return n;
  NCAMD_CATCH
  return -1;
}
int ncamd_sab_derived(ncamd_process_t h, double* logsab, double* cumul) {
  NCAMD_TRY
  Proc* p = extract(h);
  if (p->kind != kKindSAB) raise("BadInput", "not an S(alpha,beta) process");
  const size_t N = (size_t)p->sab_na * p->sab_nb;
  std::vector<double2> sl(N);
  NCAMD_HIP(hipMemcpy(sl.data(), p->sab.sl, sizeof(double2) * N, hipMemcpyDeviceToHost));
  for (size_t i = 0; i < N; ++i) logsab[i] = sl[i].y;
  NCAMD_HIP(hipMemcpy(cumul, p->sab.cumul, sizeof(double) * N, hipMemcpyDeviceToHost));
  return 0;
  NCAMD_CATCH
  return -1;
}
int ncamd_sab_sampler(ncamd_process_t h, int ie, int* kind, int* ibl, int* nvals, double* firstbin, double* bx,
                      double* by, double* bcdf, double* info_dbl, int* info_idx) {
  NCAMD_TRY
  Proc* p = extract(h);
  if (p->kind != kKindSAB) raise("BadInput", "not an S(alpha,beta) process");
  const SABView& s = p->sab;
  if (ie < 0 || ie >= s.ne) raise("BadInput", "energy grid index out of range");
  int4 hdr;
  NCAMD_HIP(hipMemcpy(&hdr, s.shdr + ie, sizeof(int4), hipMemcpyDeviceToHost));
  NCAMD_HIP(hipMemcpy(firstbin, s.sfirst + ie, sizeof(double), hipMemcpyDeviceToHost));
  *kind = hdr.x; *ibl = hdr.y; *nvals = hdr.z;
  if (hdr.x != 1 || !bx) return 0;
  const int nv = hdr.z;
  std::vector<double2> xy(nv);
  NCAMD_HIP(hipMemcpy(xy.data(), s.sxy + (size_t)ie * (s.nb + 1), sizeof(double2) * nv, hipMemcpyDeviceToHost));
  NCAMD_HIP(hipMemcpy(bcdf, s.scdf + (size_t)ie * (s.nb + 1), sizeof(double) * nv, hipMemcpyDeviceToHost));
  for (int i = 0; i < nv; ++i) { bx[i] = xy[i].x; by[i] = xy[i].y; }
  const int ni = hdr.w;
  std::vector<AInfo> inf(ni);
  NCAMD_HIP(hipMemcpy(inf.data(), s.sinfo + (size_t)ie * s.nb, sizeof(AInfo) * ni, hipMemcpyDeviceToHost));
  for (int i = 0; i < ni; ++i) {
    const AInfo& f = inf[i];
    double* d = info_dbl + 8 * i;
    d[0] = f.f_alpha; d[1] = f.f_sval; d[2] = f.f_logsval; d[3] = f.b_alpha; d[4] = f.b_sval; d[5] = f.b_logsval;
    d[6] = f.prob_front; d[7] = f.prob_notback;
    info_idx[2 * i] = f.f_idx; info_idx[2 * i + 1] = f.b_idx;
  }
  return 0;
  NCAMD_CATCH
  return -1;
}

void ncamd_setrngseed(ncamd_process_t h, uint64_t seed) {
  NCAMD_TRY Proc* p = extract(h); p->seed = seed; p->callctr = 0; NCAMD_CATCH
}
void ncamd_getrngstate(ncamd_process_t h, uint64_t* seed, uint64_t* ctr) {
  NCAMD_TRY Proc* p = extract(h); *seed = p->seed; *ctr = p->callctr; NCAMD_CATCH
}
void ncamd_setrngstate(ncamd_process_t h, uint64_t seed, uint64_t ctr) {
  NCAMD_TRY Proc* p = extract(h); p->seed = seed; p->callctr = ctr; NCAMD_CATCH
}

// ---- host pointer API
static void xs_many_impl(ncamd_process_t h, const double* ekin, unsigned long n, unsigned long repeat, double* out,
                         bool fp32) {
  bool ok = false;
  NCAMD_TRY
  Proc* p = extract(h);  // an invalid handle is a LogicError before any device check
  ensure_device();
  p->st_in.ensure(sizeof(double) * n);
  p->st_out1.ensure(sizeof(double) * n * repeat);
  if (n) NCAMD_HIP(hipMemcpy(p->st_in.p, ekin, sizeof(double) * n, hipMemcpyHostToDevice));
  xs_dispatch(p, p->st_in.as<double>(), (int64_t)n, (int64_t)repeat, p->st_out1.as<double>(), fp32, 0);
  NCAMD_HIP(hipGetLastError());
  if (n * repeat) NCAMD_HIP(hipMemcpy(out, p->st_out1.p, sizeof(double) * n * repeat, hipMemcpyDeviceToHost));
  ok = true;
  NCAMD_CATCH
  if (!ok)
    for (unsigned long i = 0; i < n * repeat; ++i) out[i] = -1.0;
}
void ncamd_crosssection_nonoriented_many(ncamd_process_t h, const double* ekin, unsigned long n, unsigned long repeat,
                                         double* out) {
  xs_many_impl(h, ekin, n, repeat, out, false);
}
void ncamd_crosssection_nonoriented_many_fp32(ncamd_process_t h, const double* ekin, unsigned long n,
                                              unsigned long repeat, double* out) {
  xs_many_impl(h, ekin, n, repeat, out, true);
}

void ncamd_samplescatterisotropic_many(ncamd_process_t h, const double* ekin, unsigned long n, unsigned long repeat,
                                       double* oE, double* oMu) {
  bool ok = false;
  NCAMD_TRY
  Proc* p = extract(h);  // an invalid handle is a LogicError before any device check
  ensure_device();
  const size_t tot = (size_t)n * repeat;
  p->st_in.ensure(sizeof(double) * n);
  p->st_out1.ensure(sizeof(double) * tot);
  p->st_out2.ensure(sizeof(double) * tot);
  if (n) NCAMD_HIP(hipMemcpy(p->st_in.p, ekin, sizeof(double) * n, hipMemcpyHostToDevice));
  sample_iso_dispatch(p, p->st_in.as<double>(), (int64_t)n, (int64_t)repeat, p->st_out1.as<double>(),
                      p->st_out2.as<double>(), 0);
  NCAMD_HIP(hipGetLastError());
  if (tot) {
    NCAMD_HIP(hipMemcpy(oE, p->st_out1.p, sizeof(double) * tot, hipMemcpyDeviceToHost));
    NCAMD_HIP(hipMemcpy(oMu, p->st_out2.p, sizeof(double) * tot, hipMemcpyDeviceToHost));
  }
  const unsigned long long nerr = read_reset_err(p, 0);
  if (nerr)
    raise("CalcError", std::to_string(nerr) + " neutron(s) failed in scatter sampling (rejection loop limit)");
  ok = true;
  NCAMD_CATCH
  if (!ok)
    for (size_t i = 0; i < (size_t)n * repeat; ++i) { oE[i] = -1.0; oMu[i] = -999.0; }
}

void ncamd_samplescatter_many_dirs(ncamd_process_t h, const double* ekin, const double* dirs, unsigned long n,
                                   double* oE, double* oD) {
  bool ok = false;
  NCAMD_TRY
  Proc* p = extract(h);  // an invalid handle is a LogicError before any device check
  ensure_device();
  const bool scb = p->kind == kKindSCB;
  p->st_in.ensure(sizeof(double) * n);
  p->st_in2.ensure(sizeof(double) * 3 * n);
  p->st_out1.ensure(sizeof(double) * n);
  p->st_out2.ensure(sizeof(double) * 3 * n);
  if (n) {
    NCAMD_HIP(hipMemcpy(p->st_in.p, ekin, sizeof(double) * n, hipMemcpyHostToDevice));
    NCAMD_HIP(hipMemcpy(p->st_in2.p, dirs, sizeof(double) * 3 * n, hipMemcpyHostToDevice));
  }
  if (scb) {
    const ScbWork w = scb_work(p, (int64_t)n, 0);
    launch_scb_sample(p->scb, p->st_in.as<double>(), p->st_in2.as<double>(), (int64_t)n, p->seed, next_stream(p),
                      p->st_out1.as<double>(), p->st_out2.as<double>(), w, 0);
  } else if (p->kind == kKindLC) {
    lc_sample(p, p->st_in.as<double>(), p->st_in2.as<double>(), (int64_t)n, p->st_out1.as<double>(),
              p->st_out2.as<double>(), 0);
  } else {
    sample_dirs_dispatch(p, p->st_in.as<double>(), p->st_in2.as<double>(), (int64_t)n, p->st_out1.as<double>(),
                         p->st_out2.as<double>(), 0);
  }
  NCAMD_HIP(hipGetLastError());
  if (n) {
    NCAMD_HIP(hipMemcpy(oE, p->st_out1.p, sizeof(double) * n, hipMemcpyDeviceToHost));
    NCAMD_HIP(hipMemcpy(oD, p->st_out2.p, sizeof(double) * 3 * n, hipMemcpyDeviceToHost));
  }
  const unsigned long long nerr = read_reset_err(p, 0);
  if (nerr)
    raise("CalcError", std::to_string(nerr) + " neutron(s) failed in scatter sampling (rejection loop limit)");
  ok = true;
  NCAMD_CATCH
  if (!ok)
    for (unsigned long i = 0; i < n; ++i) { oE[i] = -1.0; oD[3 * i] = oD[3 * i + 1] = oD[3 * i + 2] = 0.0; }
}

void ncamd_samplescatter_many(ncamd_process_t h, double ekin, const double (*direction)[3], unsigned long repeat,
                              double* oE, double* ox, double* oy, double* oz) {
  std::vector<double> e(repeat, ekin), d(3 * repeat), od(3 * repeat);
  for (unsigned long i = 0; i < repeat; ++i) {
    d[3 * i] = (*direction)[0]; d[3 * i + 1] = (*direction)[1]; d[3 * i + 2] = (*direction)[2];
  }
  ncamd_samplescatter_many_dirs(h, e.data(), d.data(), repeat, oE, od.data());
  for (unsigned long i = 0; i < repeat; ++i) { ox[i] = od[3 * i]; oy[i] = od[3 * i + 1]; oz[i] = od[3 * i + 2]; }
}

void ncamd_crosssection_many_dirs(ncamd_process_t h, const double* ekin, const double* dirs, unsigned long n,
                                  double* out) {
  bool scb = false, oc = false;
  NCAMD_TRY scb = extract(h)->kind == kKindSCB || extract(h)->kind == kKindLC; oc = extract(h)->oriented_comp;
  NCAMD_CATCH
  if (!scb && !oc) {
    xs_many_impl(h, ekin, n, 1, out, false);  // isotropic materials: the direction does not enter
    return;
  }
  bool ok = false;
  NCAMD_TRY
  Proc* p = extract(h);  // an invalid handle is a LogicError before any device check
  ensure_device();
  p->st_in.ensure(sizeof(double) * n);
  p->st_in2.ensure(sizeof(double) * 3 * n);
  p->st_out1.ensure(sizeof(double) * n);
  if (n) {
    NCAMD_HIP(hipMemcpy(p->st_in.p, ekin, sizeof(double) * n, hipMemcpyHostToDevice));
    NCAMD_HIP(hipMemcpy(p->st_in2.p, dirs, sizeof(double) * 3 * n, hipMemcpyHostToDevice));
  }
  if (p->oriented_comp) {
    oriented_xs(p, p->st_in.as<double>(), p->st_in2.as<double>(), (int64_t)n, p->st_out1.as<double>(), 0);
  } else if (p->kind == kKindLC) {
    lc_xs(p, p->st_in.as<double>(), p->st_in2.as<double>(), (int64_t)n, p->st_out1.as<double>(), 0);
  } else {
    const ScbWork w = scb_work(p, (int64_t)n, 0);
    launch_scb_xs(p->scb, p->st_in.as<double>(), p->st_in2.as<double>(), (int64_t)n, p->st_out1.as<double>(), w, 0);
  }
  NCAMD_HIP(hipGetLastError());
  if (n) NCAMD_HIP(hipMemcpy(out, p->st_out1.p, sizeof(double) * n, hipMemcpyDeviceToHost));
  ok = true;
  NCAMD_CATCH
  if (!ok)
    for (unsigned long i = 0; i < n; ++i) out[i] = -1.0;
}

// ---- device pointer API
void ncamd_dev_crosssection_nonoriented_many(ncamd_process_t h, const double* d_ekin, unsigned long n,
                                             unsigned long repeat, double* d_out, void* stream) {
  NCAMD_TRY
  Proc* p = extract(h);
  xs_dispatch(p, d_ekin, (int64_t)n, (int64_t)repeat, d_out, false, as_stream(stream));
  NCAMD_HIP(hipGetLastError());
  NCAMD_CATCH
}
void ncamd_dev_crosssection_nonoriented_many_fp32(ncamd_process_t h, const double* d_ekin, unsigned long n,
                                                  unsigned long repeat, double* d_out, void* stream) {
  NCAMD_TRY
  Proc* p = extract(h);
  xs_dispatch(p, d_ekin, (int64_t)n, (int64_t)repeat, d_out, true, as_stream(stream));
  NCAMD_HIP(hipGetLastError());
  NCAMD_CATCH
}
void ncamd_dev_samplescatterisotropic_many(ncamd_process_t h, const double* d_ekin, unsigned long n,
                                           unsigned long repeat, double* d_oE, double* d_oMu, void* stream) {
  NCAMD_TRY
  Proc* p = extract(h);
  sample_iso_dispatch(p, d_ekin, (int64_t)n, (int64_t)repeat, d_oE, d_oMu, as_stream(stream));
  NCAMD_HIP(hipGetLastError());
  NCAMD_CATCH
}
void ncamd_dev_samplescatterisotropic_many_noalloc(ncamd_process_t h, const double* d_ekin, unsigned long n,
                                                   unsigned long repeat, double* d_oE, double* d_oMu,
                                                   void* stream) {
  NCAMD_TRY
  Proc* p = extract(h);
  ensure_device();
  // queue entries the call needs: n x repeat per staged component plus the
  // high-energy queue (two queues for a single S(alpha,beta) stage)
  const uint64_t nq = is_absorption(p) || p->kind == kKindFG ? 2 : (uint64_t)make_view(p).nstaged + 1;
  const uint64_t cap = 2 * (uint64_t)kNoAllocMaxNeutrons;
  if (repeat && (n > cap / repeat || (uint64_t)n * repeat > cap / std::max<uint64_t>(nq, 2)))
    raise("BadInput", "the allocation-free sampling call takes at most " + std::to_string(kNoAllocMaxNeutrons) +
                          " neutrons (n x repeat) per call, proportionally fewer for compositions with more than "
                          "one S(alpha,beta) or FreeGas stage");
  unsigned char* base = noalloc_scratch();
  sample_iso_dispatch(p, d_ekin, (int64_t)n, (int64_t)repeat, d_oE, d_oMu, as_stream(stream),
                      reinterpret_cast<int64_t*>(base + kNoAllocCounterBytes),
                      reinterpret_cast<unsigned long long*>(base));
  NCAMD_HIP(hipGetLastError());
  NCAMD_CATCH
}
}  // extern "C"

namespace ncamd {
namespace {
// oriented calls of any process on device arrays (direction interleaved per
// neutron): scattering and cross sections
void dev_sample_dirs(Proc* p, const double* d_ekin, const double* d_dirs, int64_t n, double* d_oE, double* d_oD,
                     hipStream_t st) {
  if (p->kind == kKindSCB) {
    const ScbWork w = scb_work(p, n, st);
    launch_scb_sample(p->scb, d_ekin, d_dirs, n, p->seed, next_stream(p), d_oE, d_oD, w, st);
  } else if (p->kind == kKindLC) {
    lc_sample(p, d_ekin, d_dirs, n, d_oE, d_oD, st);
  } else {
    sample_dirs_dispatch(p, d_ekin, d_dirs, n, d_oE, d_oD, st);
  }
}
void dev_xs_dirs(Proc* p, const double* d_ekin, const double* d_dirs, int64_t n, double* d_out, hipStream_t st) {
  if (p->kind == kKindSCB) {
    const ScbWork w = scb_work(p, n, st);
    launch_scb_xs(p->scb, d_ekin, d_dirs, n, d_out, w, st);
  } else if (p->kind == kKindLC) {
    lc_xs(p, d_ekin, d_dirs, n, d_out, st);
  } else if (p->oriented_comp) {
    oriented_xs(p, d_ekin, d_dirs, n, d_out, st);
  } else {
    xs_dispatch(p, d_ekin, n, 1, d_out, false, st);
  }
}
}  // namespace
}  // namespace ncamd

extern "C" {

void ncamd_dev_samplescatter_many_dirs(ncamd_process_t h, const double* d_ekin, const double* d_dirs,
                                       unsigned long n, double* d_oE, double* d_oD, void* stream) {
  NCAMD_TRY
  Proc* p = extract(h);
  dev_sample_dirs(p, d_ekin, d_dirs, (int64_t)n, d_oE, d_oD, as_stream(stream));
  NCAMD_HIP(hipGetLastError());
  NCAMD_CATCH
}
void ncamd_dev_crosssection_many_dirs(ncamd_process_t h, const double* d_ekin, const double* d_dirs,
                                      unsigned long n, double* d_out, void* stream) {
  NCAMD_TRY
  Proc* p = extract(h);
  dev_xs_dirs(p, d_ekin, d_dirs, (int64_t)n, d_out, as_stream(stream));
  NCAMD_HIP(hipGetLastError());
  NCAMD_CATCH
}
void ncamd_dev_sab_sample_alphabeta(ncamd_process_t h, const double* d_ekin, unsigned long n, double* d_a,
                                    double* d_b, void* stream) {
  NCAMD_TRY
  Proc* p = extract(h);
  if (p->kind != kKindSAB) raise("BadInput", "not an S(alpha,beta) process");
  launch_sab_ab(p->sab, d_ekin, (int64_t)n, p->seed, next_stream(p), d_a, d_b, p->d_err.as<unsigned long long>(),
                as_stream(stream));
  NCAMD_HIP(hipGetLastError());
  NCAMD_CATCH
}
ncamd_batch_t ncamd_batch_create(int nmat, const ncamd_process_t* procs, const int64_t* counts) {
  ncamd_batch_t out;
  out.internal = nullptr;
  NCAMD_TRY
  ensure_device();
  if (nmat < 1) raise("BadInput", "a batch needs at least one material");
  std::unique_ptr<Batch> b(new Batch());
  std::vector<BatchMat> hm((size_t)nmat);
  int64_t off = 0;
  for (int i = 0; i < nmat; ++i) {
    Proc* p = extract(procs[i]);
    if (counts[i] < 0) raise("BadInput", "negative neutron count");
    if (p->oriented_comp) raise("BadInput", "oriented compositions are not supported in batches");
    if (p->kind == kKindFG || p->kind == kKindAbs)
      raise("BadInput", "FreeGas and absorption processes are not supported in batches");
    if (p->kind == kKindNone)
      for (Proc* c : p->comps)
        if (c->kind == kKindFG || c->kind == kKindAbs)
          raise("BadInput", "compositions with FreeGas or absorption components are not supported in batches");
    const CompView v = make_view(p);  // raises for oriented processes
    if (v.nstaged > 1)
      raise("BadInput", "materials with more than one S(alpha,beta) component are not supported in batches");
    p->refcount++;
    b->mats.push_back(p);
    b->counts.push_back(counts[i]);
    b->offsets.push_back(off);
    b->cost.push_back(batch_cost(v));
    hm[(size_t)i].view = v;
    hm[(size_t)i].offset = off;
    off += counts[i];
  }
  b->total = off;
  b->h_rngs.resize((size_t)nmat);
  b->d_mats.upload(hm.data(), hm.size());
  b->d_rngs.alloc(sizeof(BatchRng) * (size_t)nmat);
  b->d_next.alloc(2 * sizeof(unsigned));
  b->d_sab.alloc(sizeof(int64_t) * (size_t)std::max<int64_t>(b->total, 1));
  b->d_highe.alloc(sizeof(int64_t) * (size_t)std::max<int64_t>(b->total, 1));
  b->d_hcount.alloc(sizeof(unsigned long long));
  b->d_err.alloc(sizeof(unsigned long long));
  NCAMD_HIP(hipMemset(b->d_err.p, 0, sizeof(unsigned long long)));
  out.internal = b.release();
  NCAMD_CATCH
  return out;
}
void ncamd_batch_destroy(ncamd_batch_t* h) {
  NCAMD_TRY
  delete extract_batch(*h);
  h->internal = nullptr;
  NCAMD_CATCH
}
int64_t ncamd_batch_total(ncamd_batch_t h) {
  NCAMD_TRY return extract_batch(h)->total; NCAMD_CATCH
  return -1;
}
void ncamd_batch_get_tuning(ncamd_batch_t h, int* tile, int* nblocks, int* ordering) {
  NCAMD_TRY
  Batch* b = extract_batch(h);
  *tile = b->tile; *nblocks = b->nblocks; *ordering = b->ordering;
  NCAMD_CATCH
}
void ncamd_batch_graph_stats(ncamd_batch_t h, int mode, int64_t* captures, int64_t* replays) {
  NCAMD_TRY
  Batch* b = extract_batch(h);
  if (mode != 0 && mode != 1) raise("BadInput", "mode must be 0 (cross sections) or 1 (sampling)");
  *captures = b->graphs[mode].captures;
  *replays = b->graphs[mode].replays;
  NCAMD_CATCH
}
void ncamd_batch_set_tuning(ncamd_batch_t h, int tile, int nblocks, int ordering) {
  NCAMD_TRY
  Batch* b = extract_batch(h);
  if (tile < 64 || tile > (1 << 24) || nblocks < 1 || nblocks > 65536 || (ordering != 0 && ordering != 1))
    raise("BadInput", "invalid batch tuning parameters");
  if (tile != b->tile || ordering != b->ordering) b->tiles_valid = false;
  b->tile = tile; b->nblocks = nblocks; b->ordering = ordering;
  NCAMD_CATCH
}
void ncamd_dev_batch_crosssection(ncamd_batch_t h, const double* d_ekin, double* d_xs, void* stream) {
  NCAMD_TRY
  Batch* b = extract_batch(h);
  batch_upload_rngs(b, false, as_stream(stream));
  batch_run(b, 0, d_ekin, d_xs, nullptr, as_stream(stream));
  NCAMD_CATCH
}
void ncamd_dev_batch_samplescatterisotropic(ncamd_batch_t h, const double* d_ekin, double* d_oE, double* d_oMu,
                                            void* stream) {
  NCAMD_TRY
  Batch* b = extract_batch(h);
  batch_upload_rngs(b, true, as_stream(stream));
  batch_run(b, 1, d_ekin, d_oE, d_oMu, as_stream(stream));
  NCAMD_CATCH
}
// Times every candidate (tile size x resident workgroups x tile order) on
// the given inputs with sampling (outputs are scratch; no material's RNG
// state advances) and keeps the fastest.  Returns its time in ms.
double ncamd_dev_batch_autotune(ncamd_batch_t h, const double* d_ekin, double* d_o1, double* d_o2, void* stream) {
  NCAMD_TRY
  Batch* b = extract_batch(h);
  hipStream_t st = as_stream(stream);
  batch_upload_rngs(b, false, st);
  int cu = 256;
  {
    int dev = 0;
    NCAMD_HIP(hipGetDevice(&dev));
    NCAMD_HIP(hipDeviceGetAttribute(&cu, hipDeviceAttributeMultiprocessorCount, dev));
  }
  hipEvent_t e0, e1;
  NCAMD_HIP(hipEventCreate(&e0));
  NCAMD_HIP(hipEventCreate(&e1));
  double best = __builtin_huge_val();
  int bt = b->tile, bn = b->nblocks, bo = b->ordering;
  const int tiles[] = {512, 1024, 2048, 4096, 8192, 16384};
  const int per_cu[] = {2, 4, 8};
  for (int ord = 0; ord < 2; ++ord)
    for (int t : tiles)
      for (int pc : per_cu) {
        b->tile = t; b->nblocks = pc * cu; b->ordering = ord; b->tiles_valid = false;
        float ms = 0.f;
        for (int rep = 0; rep < 2; ++rep) {  // first run warms the graph and caches
          NCAMD_HIP(hipEventRecord(e0, st));
          batch_run(b, 1, d_ekin, d_o1, d_o2, st);
          NCAMD_HIP(hipEventRecord(e1, st));
          NCAMD_HIP(hipEventSynchronize(e1));
          NCAMD_HIP(hipEventElapsedTime(&ms, e0, e1));
        }
        if (ms < best) { best = ms; bt = t; bn = pc * cu; bo = ord; }
      }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  b->tile = bt; b->nblocks = bn; b->ordering = bo; b->tiles_valid = false;
  NCAMD_HIP(hipMemsetAsync(b->d_err.p, 0, sizeof(unsigned long long), st));
  return best;
  NCAMD_CATCH
  return -1.0;
}
int64_t ncamd_dev_batch_check_errors(ncamd_batch_t h, void* stream) {
  NCAMD_TRY
  Batch* b = extract_batch(h);
  unsigned long long n = 0;
  NCAMD_HIP(hipMemcpyAsync(&n, b->d_err.p, sizeof(n), hipMemcpyDeviceToHost, as_stream(stream)));
  NCAMD_HIP(hipStreamSynchronize(as_stream(stream)));
  if (n) NCAMD_HIP(hipMemset(b->d_err.p, 0, sizeof(n)));
  return (int64_t)n;
  NCAMD_CATCH
  return -1;
}

int ncamd_sab_use_envelope_sampler(ncamd_process_t h, int nintervals, double emin, double emax, int nbox_w,
                                   int nbox_k, int maxtries) {
  NCAMD_TRY
  ensure_device();
  check_tables_mutable(extract(h));
  build_kb(extract(h), nintervals, emin, emax, nbox_w, nbox_k, maxtries, nullptr);
  return 0;
  NCAMD_CATCH
  return -1;
}
int ncamd_sab_use_plain_sampler(ncamd_process_t h) {
  NCAMD_TRY
  Proc* p = extract(h);
  check_tables_mutable(p);
  if (p->kind != kKindSAB) raise("BadInput", "not an S(alpha,beta) process");
  p->use_kb = false;
  return 0;
  NCAMD_CATCH
  return -1;
}
int ncamd_sab_use_learned_sampler(ncamd_process_t h, ncamd_mlp_t model, int nintervals, double emin, double emax,
                                  int nbox_w, int nbox_k, int maxtries) {
  NCAMD_TRY
  ensure_device();
  check_tables_mutable(extract(h));
  build_learned(extract(h), extract_mlp(model), nintervals, emin, emax, nbox_w, nbox_k, maxtries);
  return 0;
  NCAMD_CATCH
  return -1;
}
int ncamd_sab_envelope_stats(ncamd_process_t h, double* out) {
  NCAMD_TRY
  Proc* p = extract(h);
  if (p->kind != kKindSAB || p->kb_stats_h.empty()) raise("BadInput", "no envelope sampler tables");
  std::copy(p->kb_stats_h.begin(), p->kb_stats_h.end(), out);
  return (int)(p->kb_stats_h.size() / 4);
  NCAMD_CATCH
  return -1;
}
int ncamd_sab_envelope_squeeze(ncamd_process_t h, int on) {
  NCAMD_TRY
  Proc* p = extract(h);
  if (p->kind != kKindSAB || p->kb_stats_h.empty()) raise("BadInput", "no envelope sampler tables");
  p->kb.squeeze = on ? 1 : 0;
  return 0;
  NCAMD_CATCH
  return -1;
}
int ncamd_sab_envelope_tables(ncamd_process_t h, double* node_e, double* boxes) {
  NCAMD_TRY
  Proc* p = extract(h);
  if (p->kind != kKindSAB || p->kb_node_h.empty()) raise("BadInput", "no envelope sampler tables");
  std::copy(p->kb_node_h.begin(), p->kb_node_h.end(), node_e);
  for (size_t i = 0; i < p->kb_box_h.size(); ++i) {
    const KBBox& b = p->kb_box_h[i];
    double* o = boxes + 6 * i;
    o[0] = b.w0; o[1] = b.w1; o[2] = b.k0; o[3] = b.k1; o[4] = b.env; o[5] = b.tlow;
  }
  return 0;
  NCAMD_CATCH
  return -1;
}
int64_t ncamd_dev_sab_sample_alphabeta_envelope(ncamd_process_t h, const double* d_ekin, unsigned long n,
                                                double* d_a, double* d_b, void* stream) {
  NCAMD_TRY
  Proc* p = extract(h);
  if (p->kind != kKindSAB || p->kb_node_h.empty()) raise("BadInput", "no envelope sampler tables");
  if (!p->q_count.p) p->q_count.alloc(2 * sizeof(unsigned long long));
  hipStream_t st = as_stream(stream);
  const size_t need = sizeof(int64_t) * (size_t)std::max<unsigned long>(n, 1);
  if (p->q_idx.bytes < need) {
    NCAMD_HIP(hipStreamSynchronize(st));  // the old queue may still be in use on this stream
    p->q_idx.alloc(need);
  }
  launch_kb_sample_ab(p->sab, p->kb, d_ekin, (int64_t)n, p->seed, next_stream(p), d_a, d_b,
                      p->q_idx.as<int64_t>(), p->q_count.as<unsigned long long>(), p->kb_maxtries, st);
  NCAMD_HIP(hipGetLastError());
  unsigned long long nf = 0;
  NCAMD_HIP(hipMemcpyAsync(&nf, p->q_count.p, sizeof(nf), hipMemcpyDeviceToHost, st));
  NCAMD_HIP(hipStreamSynchronize(st));
  return (int64_t)nf;
  NCAMD_CATCH
  return -1;
}

ncamd_mlp_t ncamd_mlp_create(uint64_t seed, double emin, double emax) {
  ncamd_mlp_t out;
  out.internal = nullptr;
  NCAMD_TRY out.internal = create_mlp(seed, emin, emax); NCAMD_CATCH
  return out;
}
void ncamd_mlp_destroy(ncamd_mlp_t* h) {
  NCAMD_TRY
  delete extract_mlp(*h);
  h->internal = nullptr;
  NCAMD_CATCH
}
int ncamd_mlp_nparams(void) { return kMlpParams; }
int ncamd_debug_gemm_bf16(const float* A, const float* B, float* C, int M, int N, int K, int ta, int tb, int accum) {
  NCAMD_TRY
  ensure_device();
  if (M < 1 || N < 1 || K < 1) raise("BadInput", "empty matrix");
  DevBuf dA, dB, dC;
  dA.upload(A, (size_t)M * K);
  dB.upload(B, (size_t)K * N);
  dC.upload(C, (size_t)M * N);
  launch_gemm_bf16(dA.as<float>(), ta ? M : K, ta != 0, dB.as<float>(), tb ? K : N, tb != 0, dC.as<float>(), N,
                   nullptr, M, N, K, accum != 0, 0);
  NCAMD_HIP(hipGetLastError());
  NCAMD_HIP(hipMemcpy(C, dC.p, sizeof(float) * (size_t)M * N, hipMemcpyDeviceToHost));
  return 0;
  NCAMD_CATCH
  return -1;
}
int ncamd_mlp_get_params(ncamd_mlp_t h, float* out) {
  NCAMD_TRY
  Mlp* m = extract_mlp(h);
  NCAMD_HIP(hipMemcpy(out, m->w.p, sizeof(float) * kMlpParams, hipMemcpyDeviceToHost));
  return kMlpParams;
  NCAMD_CATCH
  return -1;
}
int ncamd_mlp_set_params(ncamd_mlp_t h, const float* in) {
  NCAMD_TRY
  Mlp* m = extract_mlp(h);
  NCAMD_HIP(hipMemcpy(m->w.p, in, sizeof(float) * kMlpParams, hipMemcpyHostToDevice));
  return 0;
  NCAMD_CATCH
  return -1;
}
int ncamd_mlp_forward(ncamd_mlp_t h, const double* ekin, int n, float* out) {
  NCAMD_TRY
  Mlp* m = extract_mlp(h);
  if (n < 1) return 0;
  m->ensure(n);
  NCAMD_HIP(hipMemcpy(m->E.p, ekin, sizeof(double) * n, hipMemcpyHostToDevice));
  launch_mlp_features(m->E.as<double>(), n, m->lelo, m->linv, m->X.as<float>(), 0);
  mlp_forward(m, n, 0);
  NCAMD_HIP(hipGetLastError());
  NCAMD_HIP(hipMemcpy(out, m->Y.p, sizeof(float) * n * kMlpOut, hipMemcpyDeviceToHost));
  return n;
  NCAMD_CATCH
  return -1;
}
// features of ekin and targets (w, kappa) into m->X, m->T
void mlp_batch(Mlp* m, const double* ekin, const double* w, const double* kappa, int n) {
  if (n < 1) raise("BadInput", "empty batch");
  m->ensure(n);
  std::vector<float> t(2 * (size_t)n);
  for (int i = 0; i < n; ++i) { t[2 * i] = (float)w[i]; t[2 * i + 1] = (float)kappa[i]; }
  NCAMD_HIP(hipMemcpy(m->E.p, ekin, sizeof(double) * n, hipMemcpyHostToDevice));
  NCAMD_HIP(hipMemcpy(m->T.p, t.data(), sizeof(float) * 2 * n, hipMemcpyHostToDevice));
  launch_mlp_features(m->E.as<double>(), n, m->lelo, m->linv, m->X.as<float>(), 0);
}
// one fused step on the batch in m->X, m->T; returns the mean loss
float mlp_step_batch(Mlp* m, int n, bool update, double lr) {
  NCAMD_HIP(hipMemset(m->loss.p, 0, sizeof(float)));
  if (!update) NCAMD_HIP(hipMemset(m->g.p, 0, sizeof(float) * kMlpParams));
  if (update) ++m->step;
  launch_mlp_step(m->w.as<float>(), m->g.as<float>(), m->m.as<float>(), m->v.as<float>(), nullptr, nullptr, nullptr,
                  m->X.as<float>(), m->T.as<float>(), n, 1.0, 0.0, m->lelo, m->linv, m->loss.as<float>(),
                  static_cast<unsigned*>(m->counter.p), update, (float)lr, m->step, 0);
  NCAMD_HIP(hipGetLastError());
  float loss = 0.f;
  NCAMD_HIP(hipMemcpy(&loss, m->loss.p, sizeof(float), hipMemcpyDeviceToHost));
  return loss;
}
double ncamd_mlp_loss_grad(ncamd_mlp_t h, const double* ekin, const double* w, const double* kappa, int n,
                           float* grad) {
  NCAMD_TRY
  Mlp* m = extract_mlp(h);
  mlp_batch(m, ekin, w, kappa, n);
  const float loss = mlp_step_batch(m, n, false, 0.0);
  if (grad) NCAMD_HIP(hipMemcpy(grad, m->g.p, sizeof(float) * kMlpParams, hipMemcpyDeviceToHost));
  NCAMD_HIP(hipMemset(m->g.p, 0, sizeof(float) * kMlpParams));
  return loss;
  NCAMD_CATCH
  return -1.0;
}
double ncamd_mlp_train_step(ncamd_mlp_t h, const double* ekin, const double* w, const double* kappa, int n,
                            double lr) {
  NCAMD_TRY
  Mlp* m = extract_mlp(h);
  if (!(lr > 0.0)) raise("BadInput", "invalid learning rate");
  mlp_batch(m, ekin, w, kappa, n);
  return mlp_step_batch(m, n, true, lr);
  NCAMD_CATCH
  return -1.0;
}
// Trains on fresh batches of the plain S(alpha,beta) sampler (energies
// log-uniform over the model's range), drawn for runs of up to kRun steps in
// one launch; one fused kernel per step.  Returns the mean loss of the last
// 10% of the steps.
double ncamd_mlp_train_sab(ncamd_mlp_t h, ncamd_process_t sab, int steps, int batch, double lr, uint64_t seed) {
  NCAMD_TRY
  Mlp* m = extract_mlp(h);
  Proc* p = extract(sab);
  if (p->kind != kKindSAB) raise("BadInput", "training needs an S(alpha,beta) process");
  if (steps < 1 || batch < 64 || !(lr > 0.0)) raise("BadInput", "invalid training parameters");
  constexpr int kRun = 64;
  const int run = std::min(steps, kRun);
  m->ensure_train((size_t)run * batch, (size_t)steps);
  const double emin = std::exp(m->lelo), emax = std::exp(m->lelo + 1.0 / m->linv);
  const double bmax = p->sab_beta_h.back();
  hipStream_t st = 0;
  NCAMD_HIP(hipMemsetAsync(m->tloss.p, 0, sizeof(float) * steps, st));
  for (int it0 = 0, r = 0; it0 < steps; it0 += run, ++r) {
    const int k = std::min(run, steps - it0);
    const int64_t ns = (int64_t)k * batch;
    launch_fill_loguniform(m->tE.as<double>(), ns, std::log(emin), std::log(emax), seed + 7919ull * r, st);
    launch_sab_ab(p->sab, m->tE.as<double>(), ns, seed ^ 0x5ab5ab5abull, (uint32_t)r, m->ta.as<double>(),
                  m->tb.as<double>(), p->d_err.as<unsigned long long>(), st);
    for (int i = 0; i < k; ++i) {
      const size_t off = (size_t)i * batch;
      ++m->step;
      launch_mlp_step(m->w.as<float>(), m->g.as<float>(), m->m.as<float>(), m->v.as<float>(),
                      m->tE.as<double>() + off, m->ta.as<double>() + off, m->tb.as<double>() + off, nullptr, nullptr,
                      batch, p->sab.kT, bmax, m->lelo, m->linv, m->tloss.as<float>() + it0 + i,
                      static_cast<unsigned*>(m->counter.p), true, (float)lr, m->step, st);
    }
  }
  NCAMD_HIP(hipGetLastError());
  std::vector<float> losses((size_t)steps);
  NCAMD_HIP(hipMemcpyAsync(losses.data(), m->tloss.p, sizeof(float) * steps, hipMemcpyDeviceToHost, st));
  NCAMD_HIP(hipStreamSynchronize(st));
  const int nacc = std::max(1, steps / 10);
  double acc = 0.0;
  for (int i = steps - nacc; i < steps; ++i) acc += losses[(size_t)i];
  return acc / nacc;
  NCAMD_CATCH
  return -1.0;
}

int64_t ncamd_dev_check_errors(ncamd_process_t h, void* stream) {
  NCAMD_TRY
  Proc* p = extract(h);
  return (int64_t)read_reset_err(p, as_stream(stream));
  NCAMD_CATCH
  return -1;
}

void* ncamd_dev_malloc(uint64_t bytes) {
  NCAMD_TRY
  ensure_device();
  void* p = nullptr;
  NCAMD_HIP(hipMalloc(&p, bytes ? bytes : 1));
  return p;
  NCAMD_CATCH
  return nullptr;
}
void ncamd_dev_free(void* p) {
  NCAMD_TRY if (p) NCAMD_HIP(hipFree(p)); NCAMD_CATCH
}
void ncamd_memcpy_htod(void* dst, const void* src, uint64_t bytes) {
  NCAMD_TRY if (bytes) NCAMD_HIP(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice)); NCAMD_CATCH
}
void ncamd_memcpy_dtoh(void* dst, const void* src, uint64_t bytes) {
  NCAMD_TRY if (bytes) NCAMD_HIP(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost)); NCAMD_CATCH
}
void ncamd_memcpy_dtod(void* dst, const void* src, uint64_t bytes, void* s) {
  NCAMD_TRY
  if (bytes) NCAMD_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, (hipStream_t)s));
  NCAMD_CATCH
}
void* ncamd_stream_create(void) {
  NCAMD_TRY
  ensure_device();
  hipStream_t s;
  NCAMD_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
  return (void*)s;
  NCAMD_CATCH
  return nullptr;
}
void ncamd_stream_destroy(void* s) {
  NCAMD_TRY if (s) NCAMD_HIP(hipStreamDestroy((hipStream_t)s)); NCAMD_CATCH
}
void ncamd_stream_sync(void* s) {
  NCAMD_TRY NCAMD_HIP(hipStreamSynchronize((hipStream_t)s)); NCAMD_CATCH
}
void* ncamd_event_create(void) {
  NCAMD_TRY
  ensure_device();
  hipEvent_t e;
  NCAMD_HIP(hipEventCreate(&e));
  return (void*)e;
  NCAMD_CATCH
  return nullptr;
}
void ncamd_event_destroy(void* e) {
  NCAMD_TRY if (e) NCAMD_HIP(hipEventDestroy((hipEvent_t)e)); NCAMD_CATCH
}
void ncamd_event_record(void* e, void* s) {
  NCAMD_TRY NCAMD_HIP(hipEventRecord((hipEvent_t)e, (hipStream_t)s)); NCAMD_CATCH
}
float ncamd_event_elapsed_ms(void* a, void* b) {
  NCAMD_TRY
  float ms = 0.0f;
  NCAMD_HIP(hipEventSynchronize((hipEvent_t)b));
  NCAMD_HIP(hipEventElapsedTime(&ms, (hipEvent_t)a, (hipEvent_t)b));
  return ms;
  NCAMD_CATCH
  return -1.0f;
}
void ncamd_dev_fill_loguniform(double* d, uint64_t n, double emin, double emax, uint64_t seed, void* s) {
  NCAMD_TRY
  if (!(emin > 0.0 && emax > emin)) raise("BadInput", "need 0 < emin < emax");
  launch_fill_loguniform(d, (int64_t)n, log(emin), log(emax), seed, (hipStream_t)s);
  NCAMD_HIP(hipGetLastError());
  NCAMD_CATCH
}


int ncamd_sab_std_format(int knltype, int beta_grid_optimised, int nalpha, int nbeta, const double* alpha,
                         const double* beta, const double* sab, int* out_nalpha, int* out_nbeta, double* out_alpha,
                         double* out_beta, double* out_sab, int64_t* ntrimmed) {
  NCAMD_TRY
  const SabStd t = sab_std_format(knltype, beta_grid_optimised != 0, nalpha, nbeta, alpha, beta, sab, 500);
  *out_nalpha = (int)t.alpha.size();
  *out_nbeta = (int)t.beta.size();
  if (ntrimmed) *ntrimmed = t.ntrimmed;
  if (out_alpha) std::copy(t.alpha.begin(), t.alpha.end(), out_alpha);
  if (out_beta) std::copy(t.beta.begin(), t.beta.end(), out_beta);
  if (out_sab) std::copy(t.sab.begin(), t.sab.end(), out_sab);
  return 0;
  NCAMD_CATCH
  return -1;
}

// ------------------------------------------------ VDOS -> S(alpha,beta) ---
namespace {
// the VDOS of the entry points: n_egrid == 2 (range) or == n_density (grid);
// n_egrid == 1 with n_density == 0 is createVDOSDebye(egrid[0] = Debye
// temperature [K])
RegularVdos vdos_input(int n_egrid, const double* egrid, int n_density, const double* density) {
  if (n_egrid == 1 && n_density == 0) {
    if (!egrid || !(egrid[0] > 0.0 && egrid[0] < 1e5)) raise("BadInput", "Debye temperature out of range");
    double lo, hi;
    std::vector<double> d;
    vdos_debye(egrid[0], &lo, &hi, d);
    return vdos_regularise(std::vector<double>{lo, hi}, d);
  }
  if (!egrid || !density || n_egrid < 2 || n_density < 1) raise("BadInput", "missing VDOS arrays");
  return vdos_regularise(std::vector<double>(egrid, egrid + n_egrid), std::vector<double>(density, density + n_density));
}
// the last kernel built, so that the two calls of ncamd_vdos_kernel (sizes,
// then arrays) build it once
struct VdosKey {
  std::vector<double> in;
  bool operator==(const VdosKey& o) const { return in == o.in; }
};
std::mutex g_vdos_mtx;
VdosKey g_vdos_key;
std::shared_ptr<VdosKernel> g_vdos_last;
std::shared_ptr<VdosKernel> vdos_kernel_cached(int n_egrid, const double* egrid, int n_density, const double* density,
                                               double T, double mass, int lux, double target) {
  VdosKey key;
  key.in.assign(egrid, egrid + n_egrid);
  key.in.insert(key.in.end(), density, density + n_density);
  key.in.insert(key.in.end(), {(double)n_egrid, (double)n_density, T, mass, (double)lux, target});
  {
    std::lock_guard<std::mutex> lk(g_vdos_mtx);
    if (g_vdos_last && g_vdos_key == key) return g_vdos_last;
  }
  ensure_device();
  const RegularVdos v = vdos_input(n_egrid, egrid, n_density, density);
  if (lux < 0 || lux > 5) raise("BadInput", "vdoslux must be in 0..5");
  auto k = std::make_shared<VdosKernel>(vdos_kernel(v, T, mass, (unsigned)lux, target, nullptr));
  std::lock_guard<std::mutex> lk(g_vdos_mtx);
  g_vdos_key = key;
  g_vdos_last = k;
  return k;
}
}  // namespace

#define NCAMD_VDOS_CATCH                                              \
  }                                                                   \
  catch (const VdosError& e) { set_error(e.what(), e.type.c_str()); } \
  catch (const Error& e) { set_error(e.what(), e.type.c_str()); }     \
  catch (const std::exception& e) { set_error(e.what(), "std::exception"); }

int ncamd_vdos_analyse(int n_egrid, const double* egrid, int n_density, const double* density, double temperature,
                       double mass_amu, double* out4) {
  NCAMD_TRY
  const RegularVdos v = vdos_input(n_egrid, egrid, n_density, density);
  const VdosAnalysis a = vdos_analyse(v, temperature, mass_amu);
  out4[0] = a.gamma0;
  out4[1] = a.msd;
  out4[2] = a.teff;
  out4[3] = a.integral;
  return 0;
  NCAMD_VDOS_CATCH
  return -1;
}

int ncamd_vdos_gn(int n_egrid, const double* egrid, int n_density, const double* density, double temperature,
                  double mass_amu, int order, double* lo, double* binwidth, double* out, int capacity) {
  NCAMD_TRY
  ensure_device();
  const RegularVdos v = vdos_input(n_egrid, egrid, n_density, density);
  double l = 0.0, bw = 0.0;
  const std::vector<double> g = vdos_gn(v, temperature, mass_amu, (unsigned)std::max(order, 0), l, bw, nullptr);
  if (lo) *lo = l;
  if (binwidth) *binwidth = bw;
  if (out && capacity >= (int)g.size()) std::copy(g.begin(), g.end(), out);
  return (int)g.size();
  NCAMD_VDOS_CATCH
  return -1;
}

int ncamd_vdos_regularise(int n_egrid, const double* egrid, int n_density, const double* density, double* emin,
                          double* emax, double* out_density, int capacity) {
  NCAMD_TRY
  const RegularVdos v = vdos_input(n_egrid, egrid, n_density, density);
  const int n = (int)v.density.size();
  if (out_density && capacity >= n) {
    *emin = v.emin;
    *emax = v.emax;
    std::copy(v.density.begin(), v.density.end(), out_density);
  }
  return n;
  NCAMD_VDOS_CATCH
  return -1;
}

int ncamd_vdos_kernel(int n_egrid, const double* egrid, int n_density, const double* density, double temperature,
                      double mass_amu, int vdoslux, double target_emax, int* nalpha, int* nbeta, double* alpha,
                      double* beta, double* sab, double* info) {
  NCAMD_TRY
  const auto k = vdos_kernel_cached(n_egrid, egrid, n_density, density, temperature, mass_amu, vdoslux, target_emax);
  *nalpha = (int)k->alpha.size();
  *nbeta = (int)k->beta.size();
  if (alpha) std::copy(k->alpha.begin(), k->alpha.end(), alpha);
  if (beta) std::copy(k->beta.begin(), k->beta.end(), beta);
  if (sab) std::copy(k->sab.begin(), k->sab.end(), sab);
  if (info) {
    info[0] = k->suggested_emax;
    info[1] = k->msd;
    info[2] = k->gamma0;
    info[3] = k->norders;
    info[4] = k->orders_summed;
    info[5] = k->rows_summed;
    info[6] = k->gpu_ms;
  }
  return 0;
  NCAMD_VDOS_CATCH
  return -1;
}

ncamd_process_t ncamd_create_sabscatter_from_vdos(int n_egrid, const double* egrid, int n_density,
                                                  const double* density, double temperature, double mass_amu,
                                                  double sigma_bound, int vdoslux, double target_emax) {
  NCAMD_TRY
  const auto k = vdos_kernel_cached(n_egrid, egrid, n_density, density, temperature, mass_amu, vdoslux, target_emax);
  // ScatKnlData of knltype SAB with an optimised beta grid (no thickening),
  // suggestedEmax from the expansion (NCVDOSToScatKnl.cc:930-943)
  const SabStd t = sab_std_format(0, true, (int)k->alpha.size(), (int)k->beta.size(), k->alpha.data(),
                                  k->beta.data(), k->sab.data(), 500);
  return wrap(create_sab((int)t.alpha.size(), (int)t.beta.size(), t.alpha.data(), t.beta.data(), t.sab.data(),
                         temperature, mass_amu, sigma_bound, k->suggested_emax, 0, nullptr));
  NCAMD_VDOS_CATCH
  return wrap(nullptr);
}

ncamd_process_t ncamd_clone(ncamd_process_t h, uint64_t seed) {
  NCAMD_TRY return wrap(clone_proc(extract(h), seed)); NCAMD_CATCH
  return wrap(nullptr);
}

// =================================================== NCrystal C API names ===
// The ncrystal.h functions of the per-neutron path (declared in
// include/ncrystal_dropin.h) on the same process objects and error state.

namespace {
void*& handle_internal(void* addr) { return *static_cast<void**>(addr); }
ncamd_process_t as_ncamd(void* internal) {
  ncamd_process_t h;
  h.internal = internal;
  return h;
}
Proc* proc_of(void* internal) { return extract(as_ncamd(internal)); }
thread_local unsigned long t_thread_stream = 0;
std::atomic<unsigned long> g_thread_streams{0};
std::atomic<uint64_t> g_clone_counter{0};
}  // namespace

ncrystal_scatter_t ncamd_to_ncrystal_scatter(ncamd_process_t h) {
  NCAMD_TRY
  Proc* p = extract(h);
  if (is_absorption(p)) raise("BadInput", "an absorption process is not a scatter process");
  p->refcount++;
  return ncrystal_scatter_t{p};
  NCAMD_CATCH
  return ncrystal_scatter_t{nullptr};
}
ncrystal_absorption_t ncamd_to_ncrystal_absorption(ncamd_process_t h) {
  NCAMD_TRY
  Proc* p = extract(h);
  if (!is_absorption(p)) raise("BadInput", "a scatter process is not an absorption process");
  p->refcount++;
  return ncrystal_absorption_t{p};
  NCAMD_CATCH
  return ncrystal_absorption_t{nullptr};
}
ncamd_process_t ncamd_from_ncrystal(ncrystal_process_t h) { return as_ncamd(h.internal); }

// version and errors (ncrystal.cc:279-323, ncrystal.h:1228-1234)
int ncrystal_version(void) { return 4004002; }
const char* ncrystal_version_str(void) { return "4.4.2"; }
const char* ncrystal_namespace(void) { return ""; }
int ncrystal_sethaltonerror(int h) { return ncamd_sethaltonerror(h); }
int ncrystal_setquietonerror(int q) { return ncamd_setquietonerror(q); }
int ncrystal_error(void) { return g_waserror; }
const char* ncrystal_lasterror(void) { return g_waserror ? g_errmsg : nullptr; }
const char* ncrystal_lasterrortype(void) { return g_waserror ? g_errtype : nullptr; }
void ncrystal_clearerror(void) { ncamd_clearerror(); }
void ncrystal_seterrhandler(void (*handler)(char*, char*)) { g_errhandler = handler; }

// reference counting and casts (ncrystal.cc:392-503)
int ncrystal_refcount(void* addr) {
  NCAMD_TRY return proc_of(handle_internal(addr))->refcount; NCAMD_CATCH
  return -999;
}
void ncrystal_ref(void* addr) {
  NCAMD_TRY proc_of(handle_internal(addr))->refcount++; NCAMD_CATCH
}
void ncrystal_unref(void* addr) {
  NCAMD_TRY
  void*& in = handle_internal(addr);
  Proc* p = proc_of(in);
  if (--p->refcount == 0) {
    delete p;
    in = nullptr;
  }
  NCAMD_CATCH
}
int ncrystal_valid(void* addr) {
  if (!addr) return 0;
  Proc* p = static_cast<Proc*>(handle_internal(addr));
  return (p && p->magic == kMagic) ? 1 : 0;
}
void ncrystal_invalidate(void* addr) {
  if (ncrystal_valid(addr)) handle_internal(addr) = nullptr;
}
ncrystal_process_t ncrystal_cast_scat2proc(ncrystal_scatter_t s) {
  NCAMD_TRY
  if (is_absorption(proc_of(s.internal))) raise("LogicError", "handle is not a scatter handle");
  return ncrystal_process_t{s.internal};
  NCAMD_CATCH
  return ncrystal_process_t{nullptr};
}
ncrystal_process_t ncrystal_cast_abs2proc(ncrystal_absorption_t a) {
  NCAMD_TRY
  if (!is_absorption(proc_of(a.internal))) raise("LogicError", "handle is not an absorption handle");
  return ncrystal_process_t{a.internal};
  NCAMD_CATCH
  return ncrystal_process_t{nullptr};
}
ncrystal_scatter_t ncrystal_cast_proc2scat(ncrystal_process_t p) {
  NCAMD_TRY return ncrystal_scatter_t{is_absorption(proc_of(p.internal)) ? nullptr : p.internal}; NCAMD_CATCH
  return ncrystal_scatter_t{nullptr};
}
ncrystal_absorption_t ncrystal_cast_proc2abs(ncrystal_process_t p) {
  NCAMD_TRY return ncrystal_absorption_t{is_absorption(proc_of(p.internal)) ? p.internal : nullptr}; NCAMD_CATCH
  return ncrystal_absorption_t{nullptr};
}

// clones (ncrystal.cc:1527-1563)
ncrystal_scatter_t ncrystal_clone_scatter(ncrystal_scatter_t s) {
  NCAMD_TRY
  Proc* p = proc_of(s.internal);
  if (is_absorption(p)) raise("LogicError", "handle is not a scatter handle");
  return ncrystal_scatter_t{clone_proc(p, clone_seed(p->seed, 0x100000000ull + g_clone_counter++))};
  NCAMD_CATCH
  return ncrystal_scatter_t{nullptr};
}
ncrystal_scatter_t ncrystal_clone_scatter_rngbyidx(ncrystal_scatter_t s, unsigned long idx) {
  NCAMD_TRY
  Proc* p = proc_of(s.internal);
  if (is_absorption(p)) raise("LogicError", "handle is not a scatter handle");
  return ncrystal_scatter_t{clone_proc(p, clone_seed(p->seed, (uint64_t)idx))};
  NCAMD_CATCH
  return ncrystal_scatter_t{nullptr};
}
ncrystal_scatter_t ncrystal_clone_scatter_rngforcurrentthread(ncrystal_scatter_t s) {
  if (!t_thread_stream) t_thread_stream = ++g_thread_streams;
  return ncrystal_clone_scatter_rngbyidx(s, 0x80000000ul + t_thread_stream);
}
ncrystal_absorption_t ncrystal_clone_absorption(ncrystal_absorption_t a) {
  NCAMD_TRY
  Proc* p = proc_of(a.internal);
  if (!is_absorption(p)) raise("LogicError", "handle is not an absorption handle");
  return ncrystal_absorption_t{clone_proc(p, p->seed)};
  NCAMD_CATCH
  return ncrystal_absorption_t{nullptr};
}

// queries (ncrystal.cc:1061-1091)
const char* ncrystal_name(ncrystal_process_t p) {
  NCAMD_TRY return proc_name(proc_of(p.internal)); NCAMD_CATCH
  return nullptr;
}
int ncrystal_isnonoriented(ncrystal_process_t p) {
  NCAMD_TRY
  const Proc* q = proc_of(p.internal);
  return (q->kind == kKindSCB || q->kind == kKindLC || q->oriented_comp) ? 0 : 1;
  NCAMD_CATCH
  return 0;
}
void ncrystal_domain(ncrystal_process_t p, double* lo, double* hi) {
  NCAMD_TRY
  proc_domain(proc_of(p.internal), *lo, *hi);
  return;
  NCAMD_CATCH
  *lo = *hi = -1.0;
}

// single neutrons: one-element batched calls (ncrystal.cc:1093-1200)
void ncrystal_crosssection_nonoriented(ncrystal_process_t p, double ekin, double* result) {
  ncamd_crosssection_nonoriented_many(as_ncamd(p.internal), &ekin, 1, 1, result);
}
void ncrystal_crosssection(ncrystal_process_t p, double ekin, const double (*direction)[3], double* result) {
  ncamd_crosssection_many_dirs(as_ncamd(p.internal), &ekin, &(*direction)[0], 1, result);
}
void ncrystal_samplescatterisotropic(ncrystal_scatter_t s, double ekin, double* ekin_final, double* mu) {
  ncamd_samplescatterisotropic_many(as_ncamd(s.internal), &ekin, 1, 1, ekin_final, mu);
}
void ncrystal_samplescatter(ncrystal_scatter_t s, double ekin, const double (*direction)[3], double* ekin_final,
                            double (*direction_final)[3]) {
  ncamd_samplescatter_many_dirs(as_ncamd(s.internal), &ekin, &(*direction)[0], 1, ekin_final,
                                &(*direction_final)[0]);
}
void ncrystal_samplescatter_rs(double (*rngfct)(void*), void* state, ncrystal_scatter_t s, double ekin,
                               const double (*direction)[3], double* ekin_final, double (*direction_final)[3]) {
  bool ok = false;
  NCAMD_TRY
  Proc* p = proc_of(s.internal);
  if (!rngfct) raise("BadInput", "no random number function");
  // two numbers of the caller's stream key this neutron's Philox stream
  const double r1 = rngfct(state), r2 = rngfct(state);
  const uint64_t k = ((uint64_t)(r1 * 9007199254740992.0) << 11) ^ (uint64_t)(r2 * 9007199254740992.0);
  const uint64_t seed0 = p->seed, ctr0 = p->callctr;
  p->seed = k;
  p->callctr = 0;
  ncamd_samplescatter_many_dirs(wrap(p), &ekin, &(*direction)[0], 1, ekin_final, &(*direction_final)[0]);
  p->seed = seed0;
  p->callctr = ctr0;
  ok = !g_waserror;
  NCAMD_CATCH
  if (!ok) {
    *ekin_final = -1.0;
    (*direction_final)[0] = (*direction_final)[1] = (*direction_final)[2] = 0.0;
  }
}
void ncrystal_genscatter_nonoriented(ncrystal_scatter_t s, double ekin, double* angle, double* dekin) {
  double ef = -1.0, mu = -999.0;
  ncamd_samplescatterisotropic_many(as_ncamd(s.internal), &ekin, 1, 1, &ef, &mu);
  if (mu == -999.0 && ef == -1.0) { *angle = *dekin = -999.0; return; }
  *dekin = ef - ekin;
  *angle = std::acos(mu);
}

// batched calls (ncrystal.cc:1107-1330)
void ncrystal_crosssection_nonoriented_many(ncrystal_process_t p, const double* ekin, unsigned long n,
                                            unsigned long repeat, double* results) {
  ncamd_crosssection_nonoriented_many(as_ncamd(p.internal), ekin, n, repeat, results);
}
void ncrystal_samplescatterisotropic_many(ncrystal_scatter_t s, const double* ekin, unsigned long n,
                                          unsigned long repeat, double* results_ekin, double* results_mu) {
  ncamd_samplescatterisotropic_many(as_ncamd(s.internal), ekin, n, repeat, results_ekin, results_mu);
}
void ncrystal_samplescatter_many(ncrystal_scatter_t s, double ekin, const double (*direction)[3],
                                 unsigned long repeat, double* results_ekin, double* ux, double* uy, double* uz) {
  ncamd_samplescatter_many(as_ncamd(s.internal), ekin, direction, repeat, results_ekin, ux, uy, uz);
}
void ncrystal_genscatter_nonoriented_many(ncrystal_scatter_t s, const double* ekin, unsigned long n,
                                          unsigned long repeat, double* angle, double* dekin) {
  // results_dekin receives E' and results_angle mu first, then both are
  // converted in place
  ncamd_samplescatterisotropic_many(as_ncamd(s.internal), ekin, n, repeat, dekin, angle);
  const size_t tot = (size_t)n * repeat;
  for (size_t i = 0; i < tot; ++i) {
    if (angle[i] == -999.0 && dekin[i] == -1.0) { angle[i] = dekin[i] = -999.0; continue; }
    dekin[i] -= ekin[i % n];
    angle[i] = std::acos(angle[i]);
  }
}

// RNG state of a scatter handle: "seed:callcounter"
int ncrystal_rngsupportsstatemanip_ofscatter(ncrystal_scatter_t s) {
  NCAMD_TRY proc_of(s.internal); return 1; NCAMD_CATCH
  return 0;
}
char* ncrystal_getrngstate_ofscatter(ncrystal_scatter_t s) {
  NCAMD_TRY
  const Proc* p = proc_of(s.internal);
  char buf[64];
  snprintf(buf, sizeof(buf), "%llu:%llu", (unsigned long long)p->seed, (unsigned long long)p->callctr);
  char* out = static_cast<char*>(malloc(strlen(buf) + 1));
  if (!out) raise("CalcError", "out of memory");
  strcpy(out, buf);
  return out;
  NCAMD_CATCH
  return nullptr;
}
void ncrystal_setrngstate_ofscatter(ncrystal_scatter_t s, const char* state) {
  NCAMD_TRY
  Proc* p = proc_of(s.internal);
  unsigned long long seed = 0, ctr = 0;
  char tail = 0;
  if (!state || sscanf(state, "%llu:%llu%c", &seed, &ctr, &tail) != 2)
    raise("BadInput", "RNG state must read seed:callcounter");
  p->seed = seed;
  p->callctr = ctr;
  NCAMD_CATCH
}
void ncrystal_dealloc_string(char* s) { free(s); }

}  // extern "C"

// ----------------------------------------------------------------- MiniMC ---
namespace ncamd {
namespace {
struct MmcBufs {
  DevBuf x, y, z, d, E, w, id, nscat, ninel;
  void alloc(size_t n) {
    const size_t f = sizeof(double) * n;
    x.alloc(f); y.alloc(f); z.alloc(f); d.alloc(3 * f); E.alloc(f); w.alloc(f);
    id.alloc(sizeof(int64_t) * n); nscat.alloc(sizeof(int) * n); ninel.alloc(sizeof(int) * n);
  }
  MmcBasket view() {
    return MmcBasket{x.as<double>(), y.as<double>(), z.as<double>(), d.as<double>(), E.as<double>(),
                     w.as<double>(), id.as<int64_t>(), nscat.as<int>(), ninel.as<int>()};
  }
};

unsigned long long read_count(DevBuf& c, hipStream_t st) {
  unsigned long long v = 0;
  NCAMD_HIP(hipMemcpyAsync(&v, c.p, sizeof(v), hipMemcpyDeviceToHost, st));
  NCAMD_HIP(hipStreamSynchronize(st));
  return v;
}

// engine buffers of one device, kept between calls (grown to the largest
// batch seen, released when the process ends): two sets of neutron arrays,
// cross sections, sampled energies and directions, about 210 B per neutron
// of a batch, and the tally's per-workgroup partial sums
struct MmcWork {
  MmcBufs bufs[2];
  DevBuf xs_s, xs_a, oE, oD, cnt, hist, part;
  int64_t cap = 0;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
};

// The engine loop (NCMMC_SimEngine.cc:170-449, NCMMC_SimMgr.cc:106-130):
// source neutrons in batches, each batch stepped until no neutron is left.
// Calls on one device are serialised (they share that device's buffers).
void minimc_run(Proc* ps, Proc* pa, double numdens, double radius, const MmcSource& src, uint64_t nsrc, MmcTally t,
                const MmcEngine& e, int ignoremiss, uint64_t seed, double* hist_out, double* info) {
  const hipStream_t st = 0;
  const double macro = 100.0 * numdens;  // barn x atoms/Aa^3 -> 1/m
  const double r2 = radius * radius;
  const int64_t batch = (int64_t)std::min<uint64_t>(nsrc, (uint64_t)1 << 24);
  int dev = 0;
  NCAMD_HIP(hipGetDevice(&dev));
  static std::mutex mtx;
  static std::map<int, MmcWork*>* works = new std::map<int, MmcWork*>();  // never destroyed: no frees at exit
  std::lock_guard<std::mutex> lk(mtx);
  MmcWork*& wp = (*works)[dev];
  if (!wp) wp = new MmcWork();
  MmcWork& W = *wp;
  const size_t nh = (size_t)kMmcCats * 2 * (t.nbins + 2);
  if (W.cap < batch) {
    NCAMD_HIP(hipStreamSynchronize(st));  // a failed earlier call may have left work queued
    W.cap = 0;
    W.bufs[0].alloc((size_t)batch);
    W.bufs[1].alloc((size_t)batch);
    W.xs_s.alloc(sizeof(double) * batch);
    W.xs_a.alloc(sizeof(double) * batch);
    W.oE.alloc(sizeof(double) * batch);
    W.oD.alloc(3 * sizeof(double) * batch);
    W.cap = batch;
  }
  if (!W.cnt.p) W.cnt.alloc(sizeof(unsigned long long));
  if (!W.ev0) NCAMD_HIP(hipEventCreate(&W.ev0));
  if (!W.ev1) NCAMD_HIP(hipEventCreate(&W.ev1));
  W.hist.ensure(sizeof(double) * (size_t)kMmcCats * 2 * (kMmcMaxBins + 2));
  W.part.ensure(sizeof(double) * (size_t)kMmcCats * 2 * (kMmcMaxBins + 2) * kMmcBlocks);
  MmcBufs* bufs = W.bufs;
  DevBuf &xs_s = W.xs_s, &xs_a = W.xs_a, &oE = W.oE, &oD = W.oD, &cnt = W.cnt, &hist = W.hist, &part = W.part;
  int cur = 0;  // the step's input arrays; the other set receives the scattered neutrons
  NCAMD_HIP(hipMemsetAsync(part.p, 0, sizeof(double) * nh * kMmcBlocks, st));
  t.hist = hist.as<double>();
  t.part = part.as<double>();
  const bool scat = ps && macro > 0.0;
  const bool absn = pa && macro > 0.0;
  double steps = 0.0, nsteps = 0.0, maxstep = 0.0, nappend = 0.0, step_ms = 0.0;
  const hipEvent_t ev0 = W.ev0, ev1 = W.ev1;
  for (uint64_t base = 0; base < nsrc; base += (uint64_t)batch) {
    const int64_t nb = (int64_t)std::min<uint64_t>((uint64_t)batch, nsrc - base);
    NCAMD_HIP(hipMemsetAsync(cnt.p, 0, sizeof(unsigned long long), st));
    cur = 0;
    launch_mmc_source(src, r2, t, ignoremiss, seed, (int64_t)base, nb, bufs[cur].view(), cnt.as<unsigned long long>(),
                      st);
    int64_t na = (int64_t)read_count(cnt, st);
    uint32_t step = 0;
    while (na > 0) {
      if (step >= 1000000u) raise("CalcError", "MiniMC: neutrons still scattering after 1e6 steps");
      MmcBasket a = bufs[cur].view(), b = bufs[1 - cur].view();
      if (scat) dev_xs_dirs(ps, a.E, a.d, na, xs_s.as<double>(), st);
      else NCAMD_HIP(hipMemsetAsync(xs_s.p, 0, sizeof(double) * na, st));
      if (absn) dev_xs_dirs(pa, a.E, a.d, na, xs_a.as<double>(), st);
      NCAMD_HIP(hipMemsetAsync(cnt.p, 0, sizeof(unsigned long long), st));
      NCAMD_HIP(hipEventRecord(ev0, st));
      launch_mmc_step(a, na, xs_s.as<double>(), absn ? xs_a.as<double>() : nullptr, macro, r2, t, e, seed, step, b,
                      cnt.as<unsigned long long>(), st);
      NCAMD_HIP(hipEventRecord(ev1, st));
      const int64_t nbn = (int64_t)read_count(cnt, st);
      float ms = 0.0f;
      NCAMD_HIP(hipEventElapsedTime(&ms, ev0, ev1));
      step_ms += ms;
      nappend += (double)nbn;
      if (nbn > 0) {
        dev_sample_dirs(ps, b.E, b.d, nbn, oE.as<double>(), oD.as<double>(), st);
        launch_mmc_post(b, nbn, oE.as<double>(), oD.as<double>(), st);
      }
      NCAMD_HIP(hipGetLastError());
      nsteps += (double)na;
      cur = 1 - cur;
      na = nbn;
      ++step;
    }
    steps += step;
    maxstep = std::max(maxstep, (double)step);
  }
  launch_mmc_reduce(t, st);
  NCAMD_HIP(hipGetLastError());
  NCAMD_HIP(hipMemcpyAsync(hist_out, hist.p, sizeof(double) * nh, hipMemcpyDeviceToHost, st));
  NCAMD_HIP(hipStreamSynchronize(st));
  if (info) {
    info[0] = maxstep;
    info[1] = nsteps;
    info[2] = steps;
    info[3] = nappend;
    info[4] = step_ms;
  }
}
}  // namespace
}  // namespace ncamd

extern "C" {

int ncamd_minimc_sphere(ncamd_process_t scatter, ncamd_process_t absorption, double numdens, double radius,
                        int src_kind, const double* src, uint64_t nsrc, int tally, int nbins, double xmin,
                        double xmax, int breakdown, const double* engine, uint64_t seed, double* hist, double* info) {
  NCAMD_TRY
  ensure_device();
  Proc* ps = scatter.internal ? extract(scatter) : nullptr;
  Proc* pa = absorption.internal ? extract(absorption) : nullptr;
  if (ps && is_absorption(ps)) raise("BadInput", "MiniMC: the scattering process is an absorption process");
  if (pa && !is_absorption(pa)) raise("BadInput", "MiniMC: the absorption process is not an absorption process");
  if (!(numdens >= 0.0) || !std::isfinite(numdens)) raise("BadInput", "MiniMC: invalid number density");
  if (!(radius > 0.0) || !std::isfinite(radius)) raise("BadInput", "MiniMC: sphere radius must be > 0");
  if (src_kind != kMmcSrcConstant && src_kind != kMmcSrcCircular) raise("BadInput", "MiniMC: unknown source");
  if (nsrc < 1) raise("BadInput", "MiniMC: no source neutrons");
  if (tally < kMmcTallyTheta || tally > kMmcTallyQ) raise("BadInput", "MiniMC: unknown tally");
  if (nbins < 1 || nbins > kMmcMaxBins || !(xmax > xmin)) raise("BadInput", "MiniMC: invalid tally binning");
  MmcSource s{};
  s.kind = src_kind;
  s.ekin = src[0];
  s.x = src[1]; s.y = src[2]; s.z = src[3];
  const double un = std::sqrt(src[4] * src[4] + src[5] * src[5] + src[6] * src[6]);
  if (!(s.ekin > 0.0) || !std::isfinite(s.ekin)) raise("BadInput", "MiniMC: source energy must be > 0");
  if (!(un > 0.0) || !std::isfinite(un)) raise("BadInput", "MiniMC: invalid source direction");
  s.ux = src[4] / un; s.uy = src[5] / un; s.uz = src[6] / un;
  s.r = src_kind == kMmcSrcCircular ? src[7] : 0.0;
  if (!(s.r >= 0.0) || !std::isfinite(s.r)) raise("BadInput", "MiniMC: invalid beam radius");
  if (s.r > 0.0) {  // unit vectors spanning the disk normal to the beam
    double a[3] = {1.0, 0.0, 0.0};
    if (s.ux > 0.8) { a[0] = 0.0; a[1] = 1.0; }
    if (a[1] * s.uy > 0.8) { a[1] = 0.0; a[2] = 1.0; }
    double c[3] = {s.uy * a[2] - s.uz * a[1], s.uz * a[0] - s.ux * a[2], s.ux * a[1] - s.uy * a[0]};
    const double cn = std::sqrt(c[0] * c[0] + c[1] * c[1] + c[2] * c[2]);
    s.ax = c[0] / cn; s.ay = c[1] / cn; s.az = c[2] / cn;
    double b[3] = {s.uy * s.az - s.uz * s.ay, s.uz * s.ax - s.ux * s.az, s.ux * s.ay - s.uy * s.ax};
    const double bn = std::sqrt(b[0] * b[0] + b[1] * b[1] + b[2] * b[2]);
    s.bx = b[0] / bn; s.by = b[1] / bn; s.bz = b[2] / bn;
  }
  MmcTally t{};
  t.kind = tally; t.nbins = nbins; t.breakdown = breakdown ? 1 : 0;
  t.xmin = xmin; t.xmax = xmax; t.inv_bw = nbins / (xmax - xmin);
  t.u0x = s.ux; t.u0y = s.uy; t.u0z = s.uz;
  t.e0 = s.ekin;
  MmcEngine e{};
  e.roul_psurv = engine[0];
  e.roul_wthr = engine[1];
  e.roul_nscat = (int)engine[2];
  e.nscatlimit = engine[3] < 0.0 ? -1 : (int)engine[3];
  const int ignoremiss = engine[4] != 0.0 ? 1 : 0;
  const bool include_abs = engine[5] != 0.0;
  if (!(e.roul_psurv > 1e-20 && e.roul_psurv < 1.0)) raise("BadInput", "MiniMC: roulette survival probability must be in (1e-20, 1)");
  if (!(e.roul_wthr > 0.0)) raise("BadInput", "MiniMC: roulette weight threshold must be > 0");
  if (engine[3] > 32000.0) raise("BadInput", "MiniMC: nscatlimit must be in 0..32000");
  minimc_run(ps, include_abs ? pa : nullptr, numdens, radius, s, nsrc, t, e, ignoremiss, seed, hist, info);
  return 0;
  NCAMD_CATCH
  return -1;
}

}  // extern "C"
