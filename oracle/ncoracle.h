/*
 * ncoracle.h -- TEST INFRASTRUCTURE ONLY.
 *
 * Plain-C restatement of the per-neutron physics of NCrystal v4.4.2
 * (PowderBragg, Debye-model incoherent elastic, S(alpha,beta) inelastic with
 * its free-gas high-energy extender, hkl enumeration) used as the CPU checker
 * for the HIP implementation in ncrystal_amd/.  Only tests/, the smoke() entry
 * and the CPU-baseline leg of bench.py may load this library.  The product
 * never links or calls it.
 *
 * Every function cites the reference file:line whose algorithm it restates
 * (paths relative to the reference repository root).
 *
 * Parity status of each function (how it is pinned) is listed in the header
 * comment of ncoracle.c.
 */
#ifndef NCORACLE_H
#define NCORACLE_H
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

/* ---- counter-based RNG (Philox4x32-10, Salmon et al. SC'11) ---- */
void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
double orc_u64_to_fp01(uint64_t x);
void orc_rng_fill(uint64_t seed, uint32_t stream, uint64_t idx, int n, double* out);

/* ---- numerics helpers exposed for known-answer tests ---- */
double orc_integrate_alpha_interval_fast(double a1, double s1, double a2, double s2,
                                         double logs1, double logs2);
void orc_sincos_2pix(double x, double* s, double* c);
double orc_debye_msd(double debye_temp, double temp, double mass_amu);
void orc_alpha_limits(double ekin_div_kT, double beta, double* alow, double* aupp);
double orc_freegas_xs_shape(double a_squared);

/* ---- hkl enumeration for cells with all angles 90 degrees, no space group ---- */
/* positions: 3*sum(natoms) fractional coords, species-major. csl in sqrt(barn)
   (10 fm).  Returns number of families written (sorted like the reference's
   InfoBuilder), -1 if maxout too small, -2 if the 1e6-family guard fires. */
int orc_hkl_orthogonal(double a, double b, double c,
                       int nspecies, const int* natoms, const double* positions,
                       const double* csl, const double* msd,
                       double dcutoff, double fsquarecut, double merge_tolerance,
                       int maxout, int* out_hkl, int* out_mult,
                       double* out_dsp, double* out_fsq);
/* same, plus the explicit hkl list of every family (members sorted like the
   reference, first = representative): out_mem_off[nfam+1], out_mem_hkl[3*nmem] */
int orc_hkl_orthogonal_members(double a, double b, double c,
                               int nspecies, const int* natoms, const double* positions,
                               const double* csl, const double* msd,
                               double dcutoff, double fsquarecut, double merge_tolerance,
                               int maxout, int* out_hkl, int* out_mult,
                               double* out_dsp, double* out_fsq,
                               int64_t* out_mem_off, int* out_mem_hkl, int64_t maxmem);

/* ---- hkl enumeration with a space group (1..230): reflections grouped by
   symmetry equivalence, NCFillHKL.cc:545-709 with NCEqRefl.cc:25-150.  Cells
   with all angles 90, alpha = beta = 90 and gamma = 120, or alpha = gamma =
   90 (angles in degrees).  Returns the number of families (sorted like the
   reference's InfoBuilder; members = the equivalent reflections, one of each
   (x, -x) pair), -1 if maxout/maxmem are too small, -2 if the 1e6-family
   guard fires, -3 for an unsupported cell, -4 for a bad space group. */
int orc_hkl_spacegroup_members(int sg, double a, double b, double c, double alpha_deg, double beta_deg,
                               double gamma_deg, int nspecies, const int* natoms, const double* positions,
                               const double* csl, const double* msd, double dcutoff, double fsquarecut,
                               int maxout, int* out_hkl, int* out_mult, double* out_dsp, double* out_fsq,
                               int64_t* out_mem_off, int* out_mem_hkl, int64_t maxmem);
/* reciprocal lattice (row-major 3x3) of such a cell; -3 for other cells */
int orc_reci_lattice(double a, double b, double c, double alpha_deg, double beta_deg, double gamma_deg,
                     double* m);

/* the hkl family sort alone (test hook): perm[i] = input index of the i-th
   family in the reference's order */
void orc_debug_sort_fams(int n, const double* dsp, const double* fsq, const int* mult, const int* hkl,
                         int* perm);

/* ---- processes ---- */
typedef struct orc_proc orc_proc;
/* FreeGas scatter process (ncrystal_core/src/freegas/NCFreeGas.cc:25-77):
   temperature [K], target mass [amu], free scattering cross section [barn];
   xs and isotropic sampling through the generic orc_xs_many /
   orc_sample_iso_many.  NULL on invalid input. */
orc_proc* orc_create_freegas(double temperature, double mass_amu, double sigma_free);
/* Procedure of every free-gas draw (FreeGas processes and the S(alpha,beta)
   high-energy extender): 0 (default) = two-body collision with a Maxwellian
   target, the device's procedure; 1 = restatement of the reference's
   FreeGasSampler (NCFreeGasUtils.cc:492-930).  Both sample the same
   distribution; process-wide, set it before sampling. */
void orc_set_freegas_procedure(int which);
/* AbsOOV 1/v absorption (ncrystal_core/src/absoov/NCAbsOOV.cc:30-44):
   xs = sqrt(E_2200)*sigma_abs/sqrt(E) (inf at E = 0); xs only. */
orc_proc* orc_create_absoov(double sigma_abs);

orc_proc* orc_create_powderbragg(double v0_times_natoms, const double* dsp,
                                 const double* fsq_x_mult, int n);
int orc_pb_ntables(const orc_proc*);
void orc_pb_tables(const orc_proc*, double* e2d, double* fdm_commul);

orc_proc* orc_create_elinc(int n, const double* msd, const double* bixs,
                           const double* scale);

/* S(alpha,beta) table: sab[ibeta*nalpha+ialpha].  egrid: NULL/0 => automatic,
   3 values => {emin,emax,npts} (0 = auto), >3 => explicit grid. */
orc_proc* orc_create_sab(int nalpha, int nbeta, const double* alpha,
                         const double* beta, const double* sab, double temperature,
                         double mass_amu, double sigma_bound, double suggested_emax,
                         int negrid_in, const double* egrid_in);
int orc_sab_negrid(const orc_proc*);
void orc_sab_egrid(const orc_proc*, double* egrid, double* xs);
/* per egrid point sampler summary: ibeta_low, nvals, kind, first-bin value */
void orc_sab_sampler_info(const orc_proc*, int iegrid, int* ibeta_low, int* nvals,
                          int* kind, double* firstbin);
void orc_sab_sampler_arrays(const orc_proc*, int iegrid, double* bx, double* by,
                            double* bcdf, double* info_dbl /*8 per beta*/,
                            int* info_idx /*2 per beta*/);
void orc_sab_derived(const orc_proc*, double* logsab, double* cumul);
double orc_sab_analyse_xs(const orc_proc*, double ekin);

orc_proc* orc_create_composition(int n, orc_proc* const* comps, const double* scales);
int orc_comp_ncomponents(const orc_proc*);
int orc_comp_component(const orc_proc*, int i, double* scale);
const orc_proc* orc_comp_component_ptr(const orc_proc*, int i);
int orc_elinc_n(const orc_proc*);
void orc_elinc_data(const orc_proc*, double* msd, double* bixs);
void orc_free(orc_proc*);

/* batched evaluation (nthreads<=1: single thread) */
void orc_xs_many(const orc_proc*, const double* ekin, int64_t n, double* out, int nthreads);
/* out index = irepeat*n + i; RNG stream = (seed, stream, idx_offset + out index)
   returns number of neutrons whose sampling raised an error (poisoned outputs) */
int64_t orc_sample_iso_many(const orc_proc*, uint64_t seed, uint32_t stream,
                            uint64_t idx_offset, const double* ekin, int64_t n,
                            int64_t repeat, double* out_ekin, double* out_mu,
                            int nthreads);
/* sample (alpha,beta) from the SAB sampler only (no conversion) */
/* Planted changes of the Alg1 procedure, for tests that show the sampler
   pins detect them: 0 = none (the reference's procedure), 1 = no first-bin
   rule, 2 = egrid margin 1.0, 3 = nearer beta row instead of quantile
   interpolation.  Process-wide; set before creating the table. */
void orc_set_alg1_variant(int variant);
/* createTailedBreakdown (NCSABUtils.cc:538-633) on one alpha slice with
   cumulative integrals `cumul`: out = (xs_front, xs_middle, xs_back) */
void orc_tailed_breakdown(const double* ag, int na, const double* sab, const double* lsab, const double* cumul,
                          double alow, double aupp, unsigned ilo, unsigned iup, double* out, unsigned* imid);
/* SABSamplerAtE_Alg1::sampleAlpha (NCSABSamplerModels.cc:157-234) of beta
   row ibeta of the sampler at egrid point ie, at the m percentiles u */
void orc_sab_sample_alpha(const orc_proc* p, int ie, int ibeta, int m, const double* u, double* out);
/* SABSampler::sampleHighE alone (NCSABSampler.cc:59-156) for E >= Emax:
   oa = -1 where the table at Emax must draw; returns the error count */
int64_t orc_sab_sample_highe_many(const orc_proc* p, uint64_t seed, uint32_t stream, const double* ekin,
                                  int64_t n, double* oa, double* ob);
/* transformKernelToStdFormat (NCSABUtils.cc:134-325): kernel type 0 S,
   1 scaled S, 2 scaled symmetric S; NULL outputs give the sizes only */
int orc_sab_std_format(int type, int beta_opt, int na, int nb, const double* a, const double* b,
                       const double* s, int* out_na, int* out_nb, double* oa, double* ob, double* os,
                       int64_t* ntrim);
/* PointwiseDist percentiles (NCPointwiseDist.cc:32-107) of the piecewise
   linear density through (x, y) at the m values p; 0 or -1 (bad input). */
int orc_pwdist_percentiles(int n, const double* x, const double* y, int m, const double* p, double* out,
                           unsigned* idx);
int64_t orc_sab_sample_ab_many(const orc_proc*, uint64_t seed, uint32_t stream,
                               const double* ekin, int64_t n, double* out_alpha,
                               double* out_beta, int nthreads);
/* oriented: per-neutron directions (isotropic materials: mu + azimuth) */
int64_t orc_sample_dir_many(const orc_proc*, uint64_t seed, uint32_t stream,
                            uint64_t idx_offset, const double* ekin,
                            const double* dirs, int64_t n, double* out_ekin,
                            double* out_dirs, int nthreads);

/* SCBragg: families (d, |F|^2 and explicit hkl members, in hkl-list order);
   reciprocal lattice and cry2lab row-major 3x3; mosaicity FWHM [rad]; prec as
   NCrystal's mosprec; ntrunc 0 = derived from prec */
orc_proc* orc_create_scbragg(const double* reci_lattice, double v0_times_natoms, int nfam,
                             const double* dsp, const double* fsq, const int64_t* mem_off,
                             const int* mem_hkl, const double* cry2lab, double mos_fwhm, double prec,
                             double ntrunc);
int orc_scb_info(const orc_proc*, int* nfam, int64_t* nnormals, double* threshold, double* gos15);
void orc_scb_tables(const orc_proc*, double* inv2d, double* xsfact, int64_t* off, double* normals,
                    double* lt_evalcosx /*2*nlt*/, double* lt_sofcosd /*2*nlt*/);
void orc_xs_dir_many(const orc_proc*, const double* ekin, const double* dirs, int64_t n, double* out,
                     int nthreads);

/* VDOS -> S(alpha,beta) (ncvdos.c).  Errors: -1 / NULL with the message in
   orc_vdos_lasterror(). */
const char* orc_vdos_lasterror(void);
/* regulariseVDOSGrid: two-call (cap < npts returns npts without writing) */
int orc_vdos_regularise(const double* egrid, int negrid, const double* density, int n, double* emin_out,
                        double* emax_out, double* dens_out, int cap);
/* createVDOSDebye: 20 density points */
void orc_vdos_debye(double debye_temp, double* emin, double* emax, double* density20);
/* out = {gamma0, msd [Aa^2], effective temperature [K], integral of the given density} */
int orc_vdos_analyse(double emin, double emax, const double* density, int n, double T, double mass, double* out);
/* G_order spectrum on lo + i*bw: two-call as above */
int orc_vdos_gn(double emin, double emax, const double* density, int n, double T, double mass, int order,
                double* lo, double* bw, double* spec, int cap);
typedef struct orc_vdos_kernel orc_vdos_kernel;
orc_vdos_kernel* orc_vdos_kernel_create(double emin, double emax, const double* density, int n, double T,
                                        double mass, unsigned vdoslux, double target_emax);
/* info3 = {suggested Emax, msd, gamma0} */
int orc_vdos_kernel_dims(const orc_vdos_kernel*, int* na, int* nb, int* norders, double* info3);
void orc_vdos_kernel_copy(const orc_vdos_kernel*, double* alpha, double* beta, double* sab);
void orc_vdos_kernel_free(orc_vdos_kernel*);

#ifdef __cplusplus
}
#endif
#endif
