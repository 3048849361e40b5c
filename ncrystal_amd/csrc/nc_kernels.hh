// Shared declarations between the kernel translation units and the host
// layer of libncrystal_amd.
#pragma once
#include "nc_dev.hh"
#include "nc_scb.hh"
#include "nc_kb.hh"

namespace ncamd {

enum : int { kKindNone = 0, kKindPB = 1, kKindElInc = 2, kKindSAB = 3, kKindSCB = 4, kKindFG = 5, kKindAbs = 6,
             kKindLC = 7 };
constexpr int kMaxComp = 8;
// components of a composition with an S(alpha,beta) part that are sampled
// by their own kernels after the pick: every S(alpha,beta) component (one
// per element of a compound, NCStdScatFact.cc:256-300) and FreeGas
// components beside them
constexpr int kMaxStaged = 4;

// Device view of a process composition (passed by value as kernel argument).
struct CompView {
  int n;
  int kind[kMaxComp];
  double scale[kMaxComp];
  // compositions answer 0 xs / no scattering outside the union of their
  // components' energy domains (NCProcImpl.cc:344-389); a single process is
  // evaluated as is (use_domain = 0)
  int use_domain;
  double domain_lo, domain_hi;
  PBView pb;
  ElIncView ei;
  // staged components: slot s serves component stage_comp[s]; stage_of[i] is
  // the slot of component i (-1: sampled in place by the pick)
  int nstaged;
  int stage_comp[kMaxStaged];
  int stage_of[kMaxComp];
  SABView sab[kMaxStaged];  // the S(alpha,beta) view of each S(alpha,beta) slot
  double scb_threshold;  // domain of an SCBragg component (oriented compositions)
  int has_scb;           // 1: oriented composition with an SCBragg component
  // FreeGas components (gas mixtures: one per element) and the 1/v
  // absorption constant of an absorption component, by component index
  FGView fg[kMaxComp];
  double absc[kMaxComp];
};

// Composition semantics: ncrystal_core/src/interfaces/NCProcImpl.cc:166-204
// (cumulative component xs), :353-362 (xs), :379-389 (component pick +
// sampling).
NCD bool comp_outside(const CompView& c, double E) {
  return c.use_domain && !(E >= c.domain_lo && E <= c.domain_hi);
}

template <bool FP32PB>
NCD double pb_xs_sel(const PBView& pb, double E) {
  if constexpr (FP32PB) {
    if (!pb_active(pb, E)) return 0.0;
    const int idx = pb_last_valid(pb, E);
    const double inv = 1.0 / E;
    return (double)pb.fdm32[idx] * inv;
  } else {
    return pb_xs(pb, E);
  }
}

// ext: the cross section of an SCBragg component for this neutron and
// direction (oriented compositions; computed by the SCBragg kernels).
// Gas: the composition may hold FreeGas or absorption components (only the
// kernels of such compositions carry their code).
template <bool FP32PB, bool Gas = false>
NCD double comp_xs(const CompView& c, double E, double* commul, double ext = 0.0) {
  if (comp_outside(c, E)) return 0.0;
  double tot = 0.0;
  for (int i = 0; i < c.n; ++i) {
    double xs = 0.0;
    switch (c.kind[i]) {
      case kKindPB: xs = pb_active(c.pb, E) ? pb_xs_sel<FP32PB>(c.pb, E) : 0.0; break;
      case kKindElInc: xs = elinc_xs(c.ei, E, nullptr); break;
      case kKindSAB: xs = sab_xs(c.sab[c.stage_of[i]], E); break;
      case kKindSCB: xs = E >= c.scb_threshold ? ext : 0.0; break;
      case kKindFG:
        if constexpr (Gas) xs = fg_xs(c.fg[i], E);
        break;
      case kKindAbs:
        if constexpr (Gas) xs = absoov_xs(c.absc[i], E);
        break;
      default: break;
    }
    tot += c.scale[i] * xs;
    if (commul) commul[i] = tot;
  }
  return tot;
}

// Kernel variants by the kinds a composition holds, so that a kernel carries
// only the samplers it can reach: PowderBragg only; no S(alpha,beta) part
// (one kernel); with an S(alpha,beta) part (the staged kernels below, since
// the S(alpha,beta) sampler's registers would hold every other sampler at
// one wave per SIMD); with FreeGas components (gas mixtures, whose kernel
// alone carries the free-gas sampler).
enum : int { kCompGeneral = 0, kCompNoSAB = 1, kCompPBOnly = 2, kCompFG = 3 };

template <int Variant>
NCD int comp_sample_iso(const CompView& c, Rng& rng, double E, double& Ef, double& mu) {
  static_assert(Variant == kCompNoSAB || Variant == kCompPBOnly || Variant == kCompFG,
                "compositions with S(alpha,beta) are staged");
  if (comp_outside(c, E)) { Ef = E; mu = 1.0; return kOK; }
  if constexpr (Variant == kCompPBOnly) {
    // a single PowderBragg component: no pick (NCProcImpl.cc:379-389 draws
    // nothing for one component)
    Ef = E;
    mu = pb_sample_mu(c.pb, rng, E);
    return kOK;
  }
  double commul[kMaxComp];
  comp_xs<false, Variant == kCompFG>(c, E, commul);
  const int ic = pick_idx(rng, commul, c.n);
  switch (c.kind[ic]) {
    case kKindPB: Ef = E; mu = pb_sample_mu(c.pb, rng, E); return kOK;
    case kKindElInc: Ef = E; mu = elinc_sample_mu(c.ei, rng, E); return kOK;
    case kKindFG:
      if constexpr (Variant == kCompFG) return fg_sample_demu(c.fg[ic], rng, E, Ef, mu);
      return kErrBadSampler;
    default: return kErrBadSampler;
  }
}

// First stage of a composition with an S(alpha,beta) part: component cross
// sections, the pick, and PowderBragg / ElInc sampling in place.  Returns
// kCompToStage with the component's slot when a staged component
// (S(alpha,beta), or FreeGas beside it) is picked; its sampling then
// continues the neutron's stream after the pick, and the stage kernel
// replays the pick's draw (one draw when there are several components).
// kCompToSCB likewise for the SCBragg component of an oriented composition
// (its cross section for the neutron given as ext).
// Gas: the composition holds FreeGas components (their cross sections enter
// the pick).
constexpr int kCompToStage = 1;
constexpr int kCompToSCB = 2;
template <bool Gas>
NCD int comp_pick_sample(const CompView& c, Rng& rng, double E, double& Ef, double& mu, int& slot,
                         double ext = 0.0) {
  if (comp_outside(c, E)) { Ef = E; mu = 1.0; return kOK; }
  double commul[kMaxComp];
  const double tot = comp_xs<false, Gas>(c, E, commul, ext);
  // No component scatters this neutron (an SCBragg component off every
  // reflection, the others outside their domains): the reference
  // (NCProcImpl.cc:363-377) still picks the last component, whose sampling
  // then leaves the neutron as it is; the pick is skipped here.
  if (!(tot > 0.0)) { Ef = E; mu = 1.0; return kOK; }
  const int ic = pick_idx(rng, commul, c.n);
  switch (c.kind[ic]) {
    case kKindPB: Ef = E; mu = pb_sample_mu(c.pb, rng, E); return kOK;
    case kKindElInc: Ef = E; mu = elinc_sample_mu(c.ei, rng, E); return kOK;
    case kKindSAB: slot = c.stage_of[ic]; return kCompToStage;
    case kKindFG:
      if constexpr (Gas) { slot = c.stage_of[ic]; return kCompToStage; }
      return kErrBadSampler;
    case kKindSCB: return kCompToSCB;
    default: return kErrBadSampler;
  }
}

NCHD int comp_variant(const CompView& c) {
  if (c.n == 1 && c.kind[0] == kKindPB) return kCompPBOnly;
  for (int i = 0; i < c.n; ++i)
    if (c.kind[i] == kKindSAB || c.kind[i] == kKindSCB) return kCompGeneral;
  for (int i = 0; i < c.n; ++i)
    if (c.kind[i] == kKindFG || c.kind[i] == kKindAbs) return kCompFG;
  return kCompNoSAB;
}
NCHD bool comp_has_gas(const CompView& c) {
  for (int i = 0; i < c.n; ++i)
    if (c.kind[i] == kKindFG || c.kind[i] == kKindAbs) return true;
  return false;
}

// per-neutron kernels (nc_kernels.hip)
void launch_xs_iso(const CompView&, const double* ekin, int64_t n, int64_t repeat, double* out, bool fp32,
                   hipStream_t);
// queues: [(nstaged + 1) x n*repeat] neutron indices (one queue per staged
// component and the high-energy queue), qcount: [kMaxStaged + 1] counters
// (used by compositions with an S(alpha,beta) part)
void launch_sample_iso(const CompView&, const double* ekin, int64_t n, int64_t repeat, uint64_t seed,
                       uint32_t stream, double* oE, double* oMu, unsigned long long* nerr, int64_t* queues,
                       unsigned long long* qcount, hipStream_t);
// oriented compositions with an SCBragg component: ext[n] its cross
// sections; the neutrons that pick it are listed in cq (count *cqc, cleared
// here) for the SCBragg sampling kernels
void launch_sample_dirs(const CompView&, const double* ekin, const double* dirs, int64_t n, uint64_t seed,
                        uint32_t stream, double* oE, double* oD, unsigned long long* nerr, int64_t* queues,
                        unsigned long long* qcount, hipStream_t, const double* ext = nullptr,
                        int64_t* cq = nullptr, unsigned long long* cqc = nullptr);
// cross sections of an oriented composition given its SCBragg component's
void launch_xs_ext(const CompView&, const double* ekin, const double* ext, int64_t n, double* out, hipStream_t);
void launch_sab_ab(const SABView&, const double* ekin, int64_t n, uint64_t seed, uint32_t stream, double* oa,
                   double* ob, unsigned long long* nerr, hipStream_t);
void launch_fill_loguniform(double* out, int64_t n, double lemin, double lemax, uint64_t seed, hipStream_t);
// S(alpha,beta)-only sampling: main kernel + high-energy queue kernel
void launch_sab_sample_iso(const SABView&, const double* ekin, int64_t n, int64_t repeat, uint64_t seed,
                           uint32_t stream, double* oE, double* oMu, unsigned long long* nerr, int64_t* queue,
                           unsigned long long* qcount, int domain_check, hipStream_t);
void launch_sab_guides(const SABView&, unsigned short* cdf_guide, unsigned short* row_guide, double* row_scale,
                       hipStream_t);

// box-envelope sampler of S(alpha,beta) (nc_kb.hip)
void launch_kb_bounds(const SABView&, const double4* node_par, int nn, int nbox, int64_t total, KBBox* box,
                      hipStream_t);
void launch_kb_sample_iso(const SABView&, const KBView&, const double* ekin, int64_t n, int64_t repeat,
                          uint64_t seed, uint32_t stream, double* oE, double* oMu, unsigned long long* nerr,
                          int64_t* queue, unsigned long long* qcount, int maxtries, hipStream_t);
void launch_kb_sample_ab(const SABView&, const KBView&, const double* ekin, int64_t n, uint64_t seed,
                         uint32_t stream, double* oa, double* ob, int64_t* queue, unsigned long long* qcount,
                         int maxtries, hipStream_t);
// helpers shared with the S(alpha,beta) kernels (nc_kernels.hip)
void launch_zero_u64(unsigned long long* p, hipStream_t);
// drains a queue of neutron indices with the plain sampler (from draw 0)
void launch_sab_queue(const SABView&, const double* ekin, int64_t n, int64_t total, uint64_t seed, uint32_t stream,
                      double* oE, double* oMu, unsigned long long* nerr, const int64_t* queue,
                      const unsigned long long* qcount, hipStream_t);

// MLP proposal model (nc_mlp.hip): 8 input features, 3 hidden layers of 64,
// 48 outputs = 8 bivariate Gaussian components over (w, kappa)
constexpr int kMlpIn = 8, kMlpHidden = 64, kMlpComp = 8, kMlpOut = 6 * kMlpComp;
void launch_gemm_bf16(const float* A, int lda, bool ta, const float* B, int ldb, bool tb, float* C, int ldc,
                      const float* bias, int M, int N, int K, bool accum, hipStream_t);
void launch_silu(const float* z, float* h, int64_t n, hipStream_t);
void launch_mlp_features(const double* E, int n, double lelo, double linv, float* X, hipStream_t);
// parameter offsets of the MLP (FP32 master copy): W1 [in][hidden], b1, W2,
// b2, W3, b3, W4 [hidden][out], b4
constexpr int kOffW1 = 0, kOffB1 = kOffW1 + kMlpIn * kMlpHidden, kOffW2 = kOffB1 + kMlpHidden,
              kOffB2 = kOffW2 + kMlpHidden * kMlpHidden, kOffW3 = kOffB2 + kMlpHidden,
              kOffB3 = kOffW3 + kMlpHidden * kMlpHidden, kOffW4 = kOffB3 + kMlpHidden,
              kOffB4 = kOffW4 + kMlpHidden * kMlpOut, kMlpParams = kOffB4 + kMlpOut;
// one fused training step over n samples: either (E, alpha, beta) of the
// plain sampler or precomputed features X and targets T (E null); adds the
// step's mean loss to *loss, leaves the gradient in grad (update false) or
// applies Adam step `step` and clears it (update true)
void launch_mlp_step(float* w, float* grad, float* m, float* v, const double* E, const double* sa, const double* sb,
                     const float* X, const float* T, int n, double kT, double bmax, double lelo, double linv,
                     float* loss, unsigned* counter, bool update, float lr, int step, hipStream_t st);

// MiniMC on the device (nc_mmc.hip): neutrons of one step as a structure of
// arrays (directions interleaved as (ux, uy, uz) per neutron, as the
// scattering calls take them)
struct MmcBasket {
  double *x, *y, *z, *d, *E, *w;
  int64_t* id;                 // source index (the key of the neutron's engine draws)
  int *nscat, *ninel;          // scatterings, of which inelastic
};
constexpr int kMmcSrcConstant = 0, kMmcSrcCircular = 1;
struct MmcSource {
  int kind;
  double ekin, x, y, z, ux, uy, uz, r;
  double ax, ay, az, bx, by, bz;  // unit vectors spanning the circular beam's disk
};
// tallies (NCMMC_StdTallies.cc): theta [deg], mu, nscat, w (unweighted),
// nscat_uw (unweighted), e [eV], l [Aa], de = E0 - E [eV], q [1/Aa]
constexpr int kMmcTallyTheta = 0, kMmcTallyMu = 1, kMmcTallyNscat = 2, kMmcTallyW = 3, kMmcTallyNscatUW = 4,
              kMmcTallyE = 5, kMmcTallyL = 6, kMmcTallyDE = 7, kMmcTallyQ = 8;
constexpr int kMmcCats = 6;       // the total and the five scattering-history categories
constexpr int kMmcMaxBins = 400;
struct MmcTally {
  int kind, nbins, breakdown;
  double xmin, xmax, inv_bw;
  double u0x, u0y, u0z;        // reference direction (the beam's)
  double e0;                   // reference energy (the beam's) [eV]
  double* hist;                // [kMmcCats][sum w, sum w^2][underflow, nbins, overflow]
  double* part;                // [kMmcBlocks][the same]: per-workgroup partial sums
};
constexpr int kMmcBlocks = 1024;  // most workgroups of an engine launch
struct MmcEngine {
  double roul_psurv, roul_wthr;
  int roul_nscat, nscatlimit;  // nscatlimit < 0: none
};
void launch_mmc_source(const MmcSource& src, double r2, const MmcTally& t, int ignoremiss, uint64_t seed, int64_t id0,
                       int64_t n, const MmcBasket& out, unsigned long long* count, hipStream_t);
void launch_mmc_step(const MmcBasket& in, int64_t n, const double* xs_s, const double* xs_a, double macro, double r2,
                     const MmcTally& t, const MmcEngine& e, uint64_t seed, uint32_t step, const MmcBasket& out,
                     unsigned long long* count, hipStream_t);
void launch_mmc_post(const MmcBasket& b, int64_t n, const double* oE, const double* oD, hipStream_t);
void launch_mmc_reduce(const MmcTally& t, hipStream_t);  // hist = sum of the partial sums

// many materials in one persistent launch (nc_batch.hip)
struct BatchMat {
  CompView view;
  int64_t offset;     // first neutron of this material in the concatenated arrays
};
struct BatchRng {
  uint64_t seed;      // the material's RNG seed and call counter value for this call
  uint64_t stream;
};
struct BatchTile {
  int mat;
  int64_t begin, end;  // neutron range within the material
};
// work buffers of the staged sampling launch
struct BatchWork {
  unsigned* next;              // [2] tile counters of the two persistent stages
  int64_t* sab;                // [total] per tile, its neutrons that picked S(alpha,beta)
  int* tcount;                 // [ntiles] their number per tile
  int64_t* highe;              // [total] (tile << 40 | neutron) above the table's Emax
  unsigned long long* hcount;  // [1]
};
// mode 0: cross sections into o1; mode 1: isotropic sampling into (o1=E', o2=mu)
void launch_batch(const BatchMat* mats, const BatchRng* rngs, const BatchTile* tiles, int ntiles, int nblocks,
                  int mode, const double* ekin, double* o1, double* o2, const BatchWork& w,
                  unsigned long long* nerr, hipStream_t);

// single-crystal Bragg, oriented (nc_scb.hip): counting sort of the batch by
// scan length, main kernel, Romberg queue kernel
constexpr int kScbBuckets = 1024;
struct ScbWork {
  unsigned* keys;              // [n] bucket of each neutron
  int64_t* perm;               // [n] neutrons in bucket order
  unsigned* hist;              // [kScbBuckets]
  unsigned long long* cursor;  // [kScbBuckets]
  int64_t* queue;              // [n] neutrons for the Romberg kernels
  unsigned long long* qcount;  // [1]
  int* cnt;                    // [n] sampling: contributions at each sorted position
  double* tot;                 // [n] sampling: their cumulative total
  double4* sel;                // [n] sampling: selected normal and rounded 1/(2d)
};
void launch_scb_xs(const SCBView&, const double* ekin, const double* dirs, int64_t n, double* out,
                   const ScbWork&, hipStream_t);
// idmap (optional): neutron i of the (compact) inputs writes slot idmap[i]
// of oE/oD and draws from Philox index idmap[i], after skipping `skip`
// draws (an oriented composition's pick)
void launch_scb_sample(const SCBView&, const double* ekin, const double* dirs, int64_t n, uint64_t seed,
                       uint32_t stream, double* oE, double* oD, const ScbWork&, hipStream_t,
                       const int64_t* idmap = nullptr, int skip = 0);
// layered crystal by the reference models (nc_lc.hip): the SCBragg kernels
// over rotations of each neutron around the layer axis (unit, lab frame)
struct LCView {
  double ax, ay, az;
};
// n x N rotated neutrons (lane i * N + j): rotation j from the table cs, or
// random (then its (cos, sin) goes to ocs)
void launch_lc_expand(const double* ekin, const double* dirs, int64_t n, int N, const LCView&, const double2* cs,
                      uint64_t seed, uint32_t stream, double* oE, double* oD, double2* ocs, hipStream_t);
void launch_lc_mean(const double* xs, int64_t n, int N, double* out, hipStream_t);
void launch_lc_pick(const double* xs, const double* rE, const double* rD, int64_t n, int N, uint64_t seed,
                    uint32_t stream, double* sE, double* sD, int* sel, hipStream_t);
void launch_lc_unrotate(const double* scD, const int* sel, const double2* cs, const double* ekin, const double* dirs,
                        int64_t n, int N, const LCView&, double* oE, double* oD, hipStream_t);
// the layered-crystal model proper (lcmode=0): plane sets of equal (d,
// angle to the layer axis), sorted by decreasing (d, angle)
struct LC0Set {
  double twodsp, inv2d, c1, s1, c1minus, c1plus, fsq, pad;
};
struct LC0View {
  const LC0Set* sets;
  int nsets;
  double ax, ay, az;        // layer axis, lab frame
  double cta, sta, acc;     // cos/sin of the mosaic truncation angle, Romberg accuracy
  double xsfact;            // 1 / (V natoms)
};
void launch_lc0_xs(const LC0View&, const SCBView&, const double* ekin, const double* dirs, int64_t n, double* out,
                   hipStream_t);
void launch_lc0_sample(const LC0View&, const SCBView&, const double* ekin, const double* dirs, int64_t n, uint64_t seed,
                       uint32_t stream, double* oE, double* oD, hipStream_t);

// compacts the neutrons listed in q (count *qc <= nmax): energies, directions
// and original indices
void launch_scb_gather(const int64_t* q, const unsigned long long* qc, int64_t nmax, const double* ekin,
                       const double* dirs, double* cE, double* cD, int64_t* idmap, hipStream_t);

// S(alpha,beta) table construction (nc_sab_build.hip)
struct SABBuildBufs {
  int na, nb, ne;
  const double* alpha;
  const double* beta;
  const double* sab;
  double2* sl;        // out [nb*na]
  double* cumul;      // out [nb*na]
  double kT, sigma_bound;
};
void launch_sab_derive(const SABBuildBufs&, hipStream_t);
// xs only at n energies
void launch_sab_xs_points(const SABBuildBufs&, const double* energies, int n, double* xs_out, hipStream_t);
// full analysis on the final egrid (xs + Alg1 sampler data)
void launch_sab_analyse(const SABBuildBufs&, const double* egrid, int ne, double* xs_out, int4* shdr,
                        double* sfirst, double* scdf, double2* sxy, AInfo* sinfo, hipStream_t);
// free-gas xs at n energies (sigma_free, ca)
void launch_fg_xs(double sigma_free, double ca, const double* e, int n, double* out, hipStream_t);

// FreeGas and 1/v absorption processes (nc_kernels.hip): cross sections
// over the repeat layout; FreeGas sampling, with new directions when dirs
// is given (oX = 3 doubles per neutron) or mu otherwise
void launch_simple_xs(int kind, const FGView& g, double abs_c, const double* ekin, int64_t n, int64_t repeat,
                      double* out, hipStream_t);
void launch_fg_sample(const FGView& g, const double* ekin, const double* dirs, int64_t n, int64_t repeat,
                      uint64_t seed, uint32_t stream, double* oE, double* oX, unsigned long long* nerr, hipStream_t);

// hkl |F|^2 evaluation (nc_hkl.hip)
struct HKLBuild {
  double rm[9];                 // reciprocal lattice (row-major; diagonal 2pi/a, 2pi/b, 2pi/c for 90/90/90)
  int max_h, max_k, max_l;
  int nspecies;
  const double* pos;            // [3*natoms_total]
  const int* pos_off;           // [nspecies+1]
  const double* csl;            // [nspecies]
  const double* msd;            // [nspecies]
  const double* thr;            // [nspecies] whkl thresholds
  double fsquarecut;
  double ksq_lo, ksq_hi, dcut_lo, dcut_hi;
};
// evaluates a contiguous range [first, first+count) of the loop domain;
// writes fsq/dsp and a validity flag per point
void launch_hkl_eval(const HKLBuild&, int64_t first, int64_t count, double* fsq, double* dsp,
                     unsigned char* valid, hipStream_t);
// space-group path: evaluates the listed representatives (hkl[3*count])
// with the reciprocal lattice rm (row-major 3x3)
struct HKLMatrix {
  double m[9];
};
void launch_hkl_eval_keys(const HKLBuild&, const HKLMatrix& rm, const int* hkl, int64_t count, double* fsq,
                          double* dsp, unsigned char* valid, hipStream_t);

}  // namespace ncamd
