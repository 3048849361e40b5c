// Shared declarations between the kernel translation units and the host
// layer of libncrystal_amd.
#pragma once
#include "nc_dev.hh"
#include "nc_scb.hh"

namespace ncamd {

enum : int { kKindNone = 0, kKindPB = 1, kKindElInc = 2, kKindSAB = 3, kKindSCB = 4 };
constexpr int kMaxComp = 4;

// Device view of a process composition (passed by value as kernel argument).
struct CompView {
  int n;
  int kind[kMaxComp];
  double scale[kMaxComp];
  double domain_lo, domain_hi;
  PBView pb;
  ElIncView ei;
  SABView sab;
};

// per-neutron kernels (nc_kernels.hip)
void launch_xs_iso(const CompView&, const double* ekin, int64_t n, int64_t repeat, double* out, bool fp32,
                   hipStream_t);
void launch_sample_iso(const CompView&, const double* ekin, int64_t n, int64_t repeat, uint64_t seed,
                       uint32_t stream, double* oE, double* oMu, unsigned long long* nerr, hipStream_t);
void launch_sample_dirs(const CompView&, const double* ekin, const double* dirs, int64_t n, uint64_t seed,
                        uint32_t stream, double* oE, double* oD, unsigned long long* nerr, hipStream_t);
void launch_sab_ab(const SABView&, const double* ekin, int64_t n, uint64_t seed, uint32_t stream, double* oa,
                   double* ob, unsigned long long* nerr, hipStream_t);
void launch_fill_loguniform(double* out, int64_t n, double lemin, double lemax, uint64_t seed, hipStream_t);
// S(alpha,beta)-only sampling: main kernel + high-energy queue kernel
void launch_sab_sample_iso(const SABView&, const double* ekin, int64_t n, int64_t repeat, uint64_t seed,
                           uint32_t stream, double* oE, double* oMu, unsigned long long* nerr, int64_t* queue,
                           unsigned long long* qcount, hipStream_t);
void launch_sab_guides(const SABView&, unsigned short* cdf_guide, unsigned short* row_guide, double* row_scale,
                       hipStream_t);

// single-crystal Bragg, oriented (nc_scb.hip): main kernel + Romberg queue kernel
void launch_scb_xs(const SCBView&, const double* ekin, const double* dirs, int64_t n, double* out,
                   int64_t* queue, unsigned long long* qcount, hipStream_t);
void launch_scb_sample(const SCBView&, const double* ekin, const double* dirs, int64_t n, uint64_t seed,
                       uint32_t stream, double* oE, double* oD, int64_t* queue, unsigned long long* qcount,
                       hipStream_t);

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

// hkl |F|^2 evaluation (nc_hkl.hip)
struct HKLBuild {
  double rl0, rl1, rl2;         // diagonal reciprocal lattice 2pi/a, 2pi/b, 2pi/c
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

}  // namespace ncamd
