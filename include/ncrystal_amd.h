/*
 * ncrystal_amd.h -- C ABI of the MI355X (gfx950) implementation of NCrystal's
 * per-neutron physics hot path.
 *
 * Plain C types only (pointers, sizes, doubles).  Each batched entry point
 * mirrors the NCrystal v4.4.2 C API function named in its comment
 * (ncrystal_core/include/NCrystal/cinterface/ncrystal.h, implementation in
 * ncrystal_core/src/cinterface/ncrystal.cc) with the same argument meaning,
 * output layout and error convention:
 *   - outputs are poisoned on error (xs=-1, ekin_final=-1, mu=-999,
 *     direction=(0,0,0)), ncrystal.cc:1095-1156,1240-1245;
 *   - the error flag/message are queried with ncamd_error/ncamd_lasterror/
 *     ncamd_clearerror (ncrystal.h:1031-1045); with haltonerror set (the
 *     default, like ncrystal.cc:279-323) the process exits on error.
 *
 * Handles are opaque, reference counted, and not thread safe (one handle per
 * thread, as for NCrystal's Scatter objects).  "_dev" entry points take
 * device pointers and a hipStream_t (passed as void*) and are asynchronous;
 * all others take host pointers and return after the results are copied back.
 */
#ifndef NCRYSTAL_AMD_H
#define NCRYSTAL_AMD_H
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

typedef struct { void* internal; } ncamd_process_t;

/* ---------- library / errors (ncrystal.h:1031-1045, ncrystal.cc:279-323) ---------- */
const char* ncamd_version_str(void);
int ncamd_device_count(void);
int ncamd_error(void);
const char* ncamd_lasterror(void);
const char* ncamd_lasterrortype(void);
void ncamd_clearerror(void);
int ncamd_sethaltonerror(int);   /* returns the old value */
int ncamd_setquietonerror(int);  /* returns the old value */
/* called with (type, message) on every error, before printing/halting
 * (ncrystal.h:1047-1049) */
void ncamd_seterrhandler(void (*handler)(char* type, char* message));

/* ---------- handle lifetime (ncrystal.h:672-676) ---------- */
void ncamd_ref(ncamd_process_t);
void ncamd_unref(ncamd_process_t*);
int ncamd_refcount(ncamd_process_t);
int ncamd_valid(ncamd_process_t);
/* 0 isotropic powder-like process, 1 oriented single crystal
 * (ncrystal_isnonoriented of include/ncrystal_dropin.h is its negation, as
 * the reference's ncrystal.h:752) */
int ncamd_isoriented(ncamd_process_t);
/* a process that shares the tables of `h` with its own stream (seed) and
 * staging memory, for use from another host thread; the source lives while
 * clones exist, and neither may change its sampler mode meanwhile
 * (Scatter::clone, ncrystal_core/include/NCrystal/interfaces/NCProc.hh:118-127) */
ncamd_process_t ncamd_clone(ncamd_process_t h, uint64_t seed);

/* ---------- material compilation (init path) ---------- */

/* hkl families of a cell with all angles 90 degrees, no space group.
 * Replaces calculateHKLPlanes (ncrystal_core/src/extd_utils/NCFillHKL.cc:230-450)
 * followed by the InfoBuilder sort (ncrystal_core/src/infobld/NCInfoBuilder.cc:452).
 * The |F|^2 of every (h,k,l) is evaluated on the GPU; family grouping and
 * sorting run on the host.  positions: 3*sum(natoms) fractional coordinates,
 * species-major; csl in sqrt(barn); msd in Aa^2.  Returns the number of
 * families (<= maxout written), or -1 (maxout too small; call again with the
 * returned requirement in *nrequired), or -2 (more than 1e6 families).
 */
int64_t ncamd_hkl_orthogonal(double a, double b, double c, int nspecies, const int* natoms,
                             const double* positions, const double* csl, const double* msd,
                             double dcutoff, double fsquarecut, double merge_tolerance,
                             int64_t maxout, int* out_hkl, int* out_mult, double* out_dsp,
                             double* out_fsq, int64_t* nrequired);

/* Same, plus the explicit hkl list of every family (the members that
 * NCrystal's HKLInfo carries for ExplicitHKLs lists and that the standard
 * plane provider expands, ncrystal_core/src/extd_utils/NCPlaneProvider.cc:111-175):
 * family i owns members [out_mem_off[i], out_mem_off[i+1]) of out_mem_hkl
 * (3 ints each, one per +-pair, sorted so that the first is the
 * representative hkl).  Returns -1 if maxout or maxmem is too small
 * (*nrequired / *nmem_required say how much is needed). */
int64_t ncamd_hkl_orthogonal_members(double a, double b, double c, int nspecies, const int* natoms,
                                     const double* positions, const double* csl, const double* msd,
                                     double dcutoff, double fsquarecut, double merge_tolerance,
                                     int64_t maxout, int* out_hkl, int* out_mult, double* out_dsp,
                                     double* out_fsq, int64_t* nrequired, int64_t maxmem,
                                     int64_t* out_mem_off, int* out_mem_hkl, int64_t* nmem_required);
/* the same for any cell (angles in degrees): the reference's enumeration
 * without symmetry (NCFillHKL.cc:230-450) with the reciprocal lattice of
 * getReciprocalLatticeRot (NCLatticeUtils.cc:95-149, its general branch for
 * triclinic and rhombohedral cells); arguments as above after the angles */
int64_t ncamd_hkl_cell_members(double a, double b, double c, double alpha, double beta, double gamma, int nspecies,
                               const int* natoms, const double* positions, const double* csl, const double* msd,
                               double dcutoff, double fsquarecut, double merge_tolerance, int64_t maxout,
                               int* out_hkl, int* out_mult, double* out_dsp, double* out_fsq, int64_t* nrequired,
                               int64_t maxmem, int64_t* out_mem_off, int* out_mem_hkl, int64_t* nmem_required);

/* hkl families of a crystal with a space group (1..230): reflections grouped
 * by symmetry equivalence.  Replaces calculateHKLPlanesWithSymEqRefl
 * (ncrystal_core/src/extd_utils/NCFillHKL.cc:545-709, equivalent reflections
 * by ncrystal_core/src/phys_utils/NCEqRefl.cc:25-150) followed by the
 * InfoBuilder sort (ncrystal_core/src/infobld/NCInfoBuilder.cc:452).  Any
 * cell (angles in degrees): closed-form reciprocal lattices for 90/90/90,
 * 90/90/120 and 90/beta/90 cells, the general branch otherwise
 * (ncrystal_core/src/utils/NCLatticeUtils.cc:95-149).  The loop over (h,k,l) and the symmetry bookkeeping
 * run on the host, |F|^2 of every representative on the GPU.  Each family's
 * members are its equivalent reflections, one per +-pair, sorted so that the
 * first is the representative hkl, and mult = 2 x their number; the member
 * arrays may both be NULL.  Arguments otherwise as ncamd_hkl_orthogonal_members. */
int64_t ncamd_hkl_spacegroup_members(int spacegroup, double a, double b, double c, double alpha,
                                     double beta, double gamma, int nspecies, const int* natoms,
                                     const double* positions, const double* csl, const double* msd,
                                     double dcutoff, double fsquarecut, int64_t maxout, int* out_hkl,
                                     int* out_mult, double* out_dsp, double* out_fsq, int64_t* nrequired,
                                     int64_t maxmem, int64_t* out_mem_off, int* out_mem_hkl,
                                     int64_t* nmem_required);

/* Reciprocal lattice (row-major 3x3, 2 pi included) of such a cell, as
 * getReciprocalLatticeRot (ncrystal_core/src/utils/NCLatticeUtils.cc:95-149)
 * builds it; the input of ncamd_create_scbragg.  Returns 0, or -1 on error. */
int ncamd_reci_lattice(double a, double b, double c, double alpha, double beta, double gamma,
                       double* out9);

/* Debye-model isotropic mean-squared displacement [Aa^2]
 * (ncrystal_core/src/phys_utils/NCDebyeMSD.cc:27-36). */
double ncamd_debye_msd(double debye_temp, double temperature, double mass_amu);

/* FreeGas scatter process: replaces FreeGas(Temperature, AtomMass, SigmaFree)
 * (ncrystal_core/src/freegas/NCFreeGas.cc:25-77): cross section of
 * FreeGasXSProvider (ncrystal_core/src/phys_utils/NCFreeGasUtils.cc:28-120),
 * isotropic scattering by FreeGasSampler::sampleDeltaEMu
 * (ncrystal_core/include/NCrystal/internal/phys_utils/NCFreeGasUtils.hh:146-160).
 * temperature [K], target mass [amu], free scattering cross section [barn]
 * (NCrystal's SigmaBound::free: sigma_bound * (A/(1+A))^2).  On its own or
 * as a component of a composition (gas mixtures: one FreeGas process per
 * element; compounds: beside S(alpha,beta) components, one per element with
 * free-gas dynamics, NCStdScatFact.cc:256-300); not in many-material
 * batches.  Sampling
 * refuses energies that are not finite and >= 0 (poisoned outputs, counted
 * as errors). */
ncamd_process_t ncamd_create_freegas(double temperature, double mass_amu, double sigma_free);

/* 1/v absorption process: replaces AbsOOV(SigmaAbsorption)
 * (ncrystal_core/src/absoov/NCAbsOOV.cc:30-44): xs = sigma_abs
 * sqrt(E_2200m/s / E), infinite at E = 0.  Cross-section calls only; the
 * sampling calls raise BadInput.  Compositions of absorption processes merge
 * them into one (c = s1 c1 + s2 c2, NCAbsOOV.cc:60-70) and hold no
 * scattering process. */
ncamd_process_t ncamd_create_absoov(double sigma_abs);

/* PowderBragg from (d-spacing, |F|^2*multiplicity) pairs; replaces
 * PowderBragg(double v0_times_natoms, VectDFM&&)
 * (ncrystal_core/src/powderbragg/NCPowderBragg.cc:117-120,68-110).
 * fp32_tables != 0 additionally builds FP32 cumulative sums (compensated:
 * within one float ulp of the FP64 ones), used by
 * ncamd_crosssection_nonoriented_many_fp32. */
ncamd_process_t ncamd_create_powderbragg(double v0_times_natoms, const double* dspacings,
                                         const double* fsq_x_mult, int64_t n, int fp32_tables);

/* Incoherent elastic, replaces ElIncScatter(msd, bixs, scale)
 * (ncrystal_core/src/elincscatter/NCElIncScatter.cc:171-180). */
ncamd_process_t ncamd_create_elincscatter(int n, const double* msd, const double* bound_incoh_xs,
                                          const double* scale);

/* S(alpha,beta) scatter, replaces SABScatter(SABData&&, const VectD& egrid)
 * (ncrystal_core/src/sabscatter/NCSABScatter.cc:68-80) with the
 * SABIntegrator table build (ncrystal_core/src/sab/NCSABIntegrator.cc:283-559)
 * running on the GPU in FP64.  sab[ibeta*nalpha+ialpha].  egrid: n_egrid==0
 * automatic; 3 values {emin,emax,npts} with 0 meaning automatic; >3 explicit. */
ncamd_process_t ncamd_create_sabscatter(int nalpha, int nbeta, const double* alpha,
                                        const double* beta, const double* sab,
                                        double temperature, double mass_amu,
                                        double sigma_bound, double suggested_emax,
                                        int n_egrid, const double* egrid);

/* Scattering-kernel data to the standard S(alpha,beta) table, replaces
 * SABUtils::transformKernelToStdFormat (ncrystal_core/src/sab/NCSABUtils.cc:
 * 134-325) with validateScatKnlData's checks (NCScatKnlData.cc:26-80):
 * all-zero edges trimmed (upper alpha columns, rows at beta > 0 from the top
 * and at beta < 0 from the bottom), a symmetric table (knltype 2, beta grid
 * from 0) mirrored to negative beta, a scaled table (knltype 1 or 2)
 * unscaled by exp(-beta/2), and beta grids of fewer than 500 rows thickened
 * by linear interpolation unless beta_grid_optimised.  knltype 0 is plain
 * S(alpha,beta).  Call with NULL output arrays to get the sizes
 * (out_nalpha, out_nbeta), then with arrays of those sizes; the result
 * feeds ncamd_create_sabscatter.  Returns 0, or -1 on error. */
int ncamd_sab_std_format(int knltype, int beta_grid_optimised, int nalpha, int nbeta, const double* alpha,
                         const double* beta, const double* sab, int* out_nalpha, int* out_nbeta, double* out_alpha,
                         double* out_beta, double* out_sab, int64_t* ntrimmed);

/* ---- VDOS -> S(alpha,beta) ----
 * A phonon density of states is given as n_density values on n_egrid == 2
 * (range [emin, emax], equidistant) or n_egrid == n_density (grid) energies
 * [eV], as in NCMAT files; n_egrid == 1 with n_density == 0 is the Debye
 * spectrum of createVDOSDebye (egrid[0] = Debye temperature [K]; 20 points
 * of E^2 on [E_D/2, E_D], ncrystal_core/src/dyninfoutils/NCDynInfoUtils.cc:
 * 377-396).  The grid is first regularised as regulariseVDOSGrid does
 * (ncrystal_core/src/vdos/NCVDOSEval.cc:497-640). */

/* regulariseVDOSGrid: returns the number of points of the regular grid; with
 * out_density of at least that capacity also writes emin, emax and the
 * densities.  -1 on error. */
int ncamd_vdos_regularise(int n_egrid, const double* egrid, int n_density, const double* density, double* emin,
                          double* emax, double* out_density, int capacity);

/* VDOSEval (NCVDOSEval.cc:147-445) at temperature [K] for an element of mass
 * [amu]: out4 = {gamma0, mean squared displacement [Aa^2] (the Debye-Waller
 * MSD of the element), effective temperature [K], integral of the given
 * density}.  Host computation.  0, or -1 on error. */
int ncamd_vdos_analyse(int n_egrid, const double* egrid, int n_density, const double* density, double temperature,
                       double mass_amu, double* out4);

/* VDOSGn (ncrystal_core/src/vdos/NCVDOSGn.cc): the spectrum G_n of phonon
 * order n (1..100000) at temperature [K] for an element of mass [amu], on
 * the grid lo + i * binwidth [eV].  Returns the number of points; with out of
 * at least that capacity also writes the values.  G_1 from VDOSEval, each
 * higher order as the FFT convolution of orders n - n/2 and n/2 with the
 * reference's truncation and thinning; transforms of 32768 points or more run
 * on the GPU with the host's arithmetic (bit-identical results), orders of one
 * generation (2^k+1..2^(k+1)) on host threads.  -1 on error. */
int ncamd_vdos_gn(int n_egrid, const double* egrid, int n_density, const double* density, double temperature,
                  double mass_amu, int order, double* lo, double* binwidth, double* out, int capacity);

/* createScatteringKernel (ncrystal_core/src/vdos/NCVDOSToScatKnl.cc:787-944)
 * with the default truncation and thinning of the phonon-order functions:
 * the orders needed to reach the kinematic curve of target_emax [eV] (0: the
 * vdoslux default 0.5, 1, 3, 5, 8, 12 eV for vdoslux 0..5), the alpha and
 * beta grids (nbeta = 100 * 2^vdoslux, nalpha = nbeta/2) and
 * sab[ibeta*nalpha+ialpha] = sum over orders n of x^n exp(-x)/n! kT
 * G_n(beta kT), summed on the GPU's FP64 matrix cores.  info (7 values, or
 * NULL) = {suggested Emax [eV], msd [Aa^2], gamma0, number of orders, orders
 * summed (all, or whole blocks of at least 16 as the reference sums them),
 * rows beta <= 0 summed, device time of the sum [ms]}.  Call with NULL arrays for
 * the sizes, then with arrays of those sizes (the kernel of the last inputs
 * is kept, so it is built once).  0, or -1 on error (CalcError when the
 * expansion cannot reach target_emax). */
int ncamd_vdos_kernel(int n_egrid, const double* egrid, int n_density, const double* density, double temperature,
                      double mass_amu, int vdoslux, double target_emax, int* nalpha, int* nbeta, double* alpha,
                      double* beta, double* sab, double* info);

/* SABScatter of the kernel above, as NCrystal's standard scatter factory
 * creates one per element with VDOS dynamics (NCStdScatFact.cc:256-275,
 * NCDynInfoUtils.cc:290-340): the kernel in standard form (zero edges
 * trimmed, its beta grid kept) with its suggested Emax as the top of the
 * energy grid; sigma_bound [barn] is the element's bound scattering cross
 * section. */
ncamd_process_t ncamd_create_sabscatter_from_vdos(int n_egrid, const double* egrid, int n_density,
                                                  const double* density, double temperature, double mass_amu,
                                                  double sigma_bound, int vdoslux, double target_emax);

/* Composition of processes with scales, replaces ProcComposition
 * (ncrystal_core/src/interfaces/NCProcImpl.cc:110-389).  Components are
 * added as ProcComposition::addComponent does (NCProcImpl.cc:263-338):
 * zero scales and PowderBragg processes without planes are dropped, a nested
 * composition adds its components with multiplied scales, a process already
 * present adds to its scale, and PowderBragg + PowderBragg or ElInc + ElInc
 * are merged into one process at scale 1 (createMerged:
 * ncrystal_core/src/powderbragg/NCPowderBragg.cc:218-331,
 * ncrystal_core/src/phys_utils/NCElIncXS.cc:204-233), so a multi-phase
 * material is the composition of its phases.  After merging: at most 8
 * components, at most one SCBragg process, and at most 4 S(alpha,beta)
 * components and FreeGas components beside them (compounds carry one per
 * element, NCStdScatFact.cc:256-300; each is sampled by its own kernel after
 * the pick, with a queue of n entries); order is kept (it fixes the
 * component pick).  With
 * an SCBragg component the composition is oriented (as NCrystal's
 * single-crystal materials: SCBragg for d >= sccutoff, PowderBragg below it,
 * ElInc, S(alpha,beta)): only the calls with neutron directions apply, the
 * isotropic components scatter through randDirectionGivenScatterMu
 * (NCProcImpl.cc:29-37), and the sampling call waits for the number of
 * neutrons that pick the SCBragg component before sampling them.  A
 * composition holds either scattering or absorption processes. */
ncamd_process_t ncamd_create_composition(int n, const ncamd_process_t* procs, const double* scales);

/* Components of a composition after merging, replaces
 * ProcComposition::components() (ncrystal_core/include/NCrystal/interfaces/
 * NCProcImpl.hh:328,473).  Returns the number of components; with
 * 0 <= i < that number, *component receives a new reference to component i
 * (release it with ncamd_unref) and *scale its scale. */
int ncamd_composition_component(ncamd_process_t h, int i, ncamd_process_t* component, double* scale);

/* read back tables (for tests and tools) */
int64_t ncamd_powderbragg_tables(ncamd_process_t, double* e2d, double* fdm, float* fdm32, int64_t maxn);
int ncamd_sab_egrid(ncamd_process_t, double* egrid, double* xs, int maxn);
int ncamd_sab_derived(ncamd_process_t, double* logsab, double* cumul);
/* per energy-grid point: kind, ibeta_low, nvals, first-bin value; arrays of
 * the beta distribution (x, y, cdf; nvals each) and alpha infos
 * (8 doubles + 2 ints per beta, nvals-1 entries). */
int ncamd_sab_sampler(ncamd_process_t, int iegrid, int* kind, int* ibeta_low, int* nvals,
                      double* firstbin, double* bx, double* by, double* bcdf,
                      double* info_dbl, int* info_idx);

/* Mosaic single crystal (oriented), replaces
 * SCBragg(const Info&, const SCOrientation&, MosaicityFWHM, 0.0, nullptr, prec, ntrunc)
 * (ncrystal_core/src/scbragg/NCSCBragg.cc:90-231) with the standard plane
 * provider: families (d-spacing [Aa], |F|^2 [barn]) with explicit hkl members
 * (mem_off[nfam+1], mem_hkl[3*nmem], as ncamd_hkl_orthogonal_members returns),
 * the reciprocal lattice (row-major 3x3, getReciprocalLatticeRot of
 * ncrystal_core/src/utils/NCLatticeUtils.cc:95-117) and the crystal-to-lab
 * rotation (row-major 3x3, getCrystal2LabRot of
 * ncrystal_core/src/extd_utils/NCOrientUtils.cc:28-35).  Mosaicity is the
 * Gaussian FWHM in radians; prec and ntrunc as NCrystal's mosprec and
 * truncation (ntrunc 0 = derived from prec).  Evaluated with
 * ncamd_crosssection_many_dirs / ncamd_samplescatter_many(_dirs) and their
 * device-pointer forms; the non-oriented calls raise BadInput. */
ncamd_process_t ncamd_create_scbragg(double v0_times_natoms, int64_t nfam, const double* dspacing,
                                     const double* fsquared, const int64_t* mem_off, const int* mem_hkl,
                                     const double* reci_lattice, const double* cry2lab,
                                     double mosaicity_fwhm, double prec, double ntrunc);
/* Layered crystal by one of the reference models, replaces
 * LCBragg(const Info&, const SCOrientation&, MosaicityFWHM, const LCAxis&,
 * mode, 0.0, nullptr, prec, ntrunc) for mode != 0
 * (ncrystal_core/src/lcbragg/NCLCBragg.cc:62-73): the single crystal
 * (an SCBragg process of the same orientation, mosaicity and precision)
 * rotated around the layer axis (lab frame; the reference converts the
 * crystal-frame lcaxis with the crystal-to-lab rotation).  mode > 0:
 * LCBraggRef (ncrystal_core/src/extd_utils/NCLCRefModels.cc:25-80), cross
 * sections averaged over N equally spaced rotations phi_i = 2 pi i / N - pi
 * (N the smallest prime >= mode), scattering over `mode` random rotations
 * picked by cross section; mode < 0: LCBraggRndmRot (NCLCRefModels.cc:82-171),
 * -mode random rotations for both.  Mode 0 (the LCHelper model) raises
 * BadInput here; ncamd_create_lcbragg builds it.  The process holds a reference to the SCBragg process; oriented
 * calls only; it cannot be a component of a composition. */
ncamd_process_t ncamd_create_lcbragg_ref(ncamd_process_t scbragg, const double* lcaxis_lab, int mode);
/* Layered crystal, replaces LCBragg(const Info&, const SCOrientation&,
 * MosaicityFWHM, const LCAxis&, mode, 0.0, nullptr, prec, ntrunc)
 * (ncrystal_core/src/lcbragg/NCLCBragg.cc:36-73) with the standard plane
 * provider: the SCBragg inputs of ncamd_create_scbragg and the layer axis in
 * the crystal frame.  mode 0: the layered-crystal model (LCHelper,
 * ncrystal_core/src/extd_utils/NCLCUtils.cc:58-516: plane sets of equal
 * (d, angle to the axis), rotation ranges reaching the mosaic band, Romberg
 * averages over them; scattering: a range by its cross section, the rotation
 * by rejection under the reference's 8-bin envelope, NCLCUtils.cc:518-690).
 * mode != 0: as ncamd_create_lcbragg_ref with the axis taken to the lab
 * frame by cry2lab. */
ncamd_process_t ncamd_create_lcbragg(double v0_times_natoms, int64_t nfam, const double* dspacing,
                                     const double* fsquared, const int64_t* mem_off, const int* mem_hkl,
                                     const double* reci_lattice, const double* cry2lab,
                                     const double* lcaxis_crystal, double mosaicity_fwhm, double prec,
                                     double ntrunc, int mode);
/* read-back: family count, demi-normal count, threshold energy [eV] and 15
 * GaussOnSphere parameters (sigma, truncangle, prec, cos/sin truncangle,
 * expfact, norm, k1, k2, numacc, table size, first argument and inverse step
 * of both spline tables) */
int ncamd_scbragg_info(ncamd_process_t, int64_t* nfam, int64_t* nnormals, double* threshold, double* gos);
/* families (1/(2d), |F|^2/(V0 n), demi-normal offsets[nfam+1]), lab-frame
 * demi-normals (3 per) and the two spline tables ((y, y'') pairs) */
int ncamd_scbragg_tables(ncamd_process_t, double* inv2d, double* xsfact, int64_t* off, double* normals,
                         double* lt_evalcosx, double* lt_sofcosd);

/* ---------- exact envelope sampler of S(alpha,beta) (no counterpart in
 * NCrystal; an alternative to the plain sampler of NCSABSampler.cc that
 * samples the interpolated table density itself) ----------
 * Switches an S(alpha,beta) process to rejection sampling in kinematic
 * coordinates (w = sqrt(beta+eps)/wmax, kappa = position between the
 * kinematic alpha limits) with a box envelope per energy interval:
 * nintervals log-spaced intervals over [emin, emax), clipped to the
 * process's energy grid (0 = its whole range; outside the grid the reference
 * does not sample the table density itself, so the plain sampler is used
 * there), nbox_w strips in w with nbox_k boxes each, at most
 * maxtries proposals per neutron.  Every box carries exact upper and lower
 * bounds of the target; proposals below the lower bound are accepted
 * without evaluating the table (squeeze), with the same decisions as the
 * full test.  Neutrons outside [emin, emax) or without an accepted proposal
 * use the plain sampler (from the first draw of their stream).  The
 * accepted (alpha, beta) follow the interpolated table density exactly. */
int ncamd_sab_use_envelope_sampler(ncamd_process_t, int nintervals, double emin, double emax, int nbox_w,
                                   int nbox_k, int maxtries);
int ncamd_sab_use_plain_sampler(ncamd_process_t);
/* interval edges [nintervals+1] and per box (interval-major, strip-major)
 * w0, w1, kappa0, kappa1, upper bound, lower bound [6*nintervals*nbox_w*nbox_k] */
int ncamd_sab_envelope_tables(ncamd_process_t, double* node_e, double* boxes);
/* per interval: envelope mass of the grid in use, envelope mass of the
 * uniform grid, blend weight of the uniform grid in the edges of the grid in
 * use (1: uniform), mass under the lower bounds; the expected acceptance is
 * (integral of the target) / (envelope mass).  Returns nintervals. */
int ncamd_sab_envelope_stats(ncamd_process_t, double* out /* 4 per interval */);
/* test hook: 0 switches the squeeze off (every proposal evaluates the table) */
int ncamd_sab_envelope_squeeze(ncamd_process_t, int on);
/* raw (alpha, beta) of the envelope (or learned) sampler; returns the number
 * of neutrons that fell back to the plain sampler (synchronises the stream) */
int64_t ncamd_dev_sab_sample_alphabeta_envelope(ncamd_process_t, const double* d_ekin, unsigned long n,
                                                double* d_alpha, double* d_beta, void* stream);

/* ---------- learned proposal model (no counterpart in NCrystal) ----------
 * MLP: 8 energy features -> 3 x 64 (SiLU) -> 48 outputs = 8 bivariate
 * Gaussian components over the kinematic coordinates (w, kappa) of an
 * S(alpha,beta) sample.  BF16 matrix-core products with FP32 accumulation,
 * FP32 Adam.  Energies are scaled over [emin, emax]. */
typedef struct { void* internal; } ncamd_mlp_t;
ncamd_mlp_t ncamd_mlp_create(uint64_t seed, double emin, double emax);
void ncamd_mlp_destroy(ncamd_mlp_t*);
int ncamd_mlp_nparams(void);
int ncamd_mlp_get_params(ncamd_mlp_t, float* out);
int ncamd_mlp_set_params(ncamd_mlp_t, const float* in);
/* test hook of the matrix-core product used by the model: C (+)= op(A) op(B)
 * with BF16 operands and FP32 accumulation; A stored M x K (K x M if ta),
 * B stored K x N (N x K if tb), C M x N, all row-major FP32 host arrays */
int ncamd_debug_gemm_bf16(const float* A, const float* B, float* C, int M, int N, int K, int ta, int tb, int accum);
/* outputs [n*48] at the given energies */
int ncamd_mlp_forward(ncamd_mlp_t, const double* ekin, int n, float* out);
/* mean mixture negative log-likelihood of targets (w, kappa) and its
 * gradient with respect to all parameters [ncamd_mlp_nparams()], from the
 * fused training-step kernel without its update */
double ncamd_mlp_loss_grad(ncamd_mlp_t, const double* ekin, const double* w, const double* kappa, int n,
                           float* grad);
/* one fused training step (forward, likelihood, backward, Adam with step
 * size lr) on the given batch; returns its mean loss */
double ncamd_mlp_train_step(ncamd_mlp_t, const double* ekin, const double* w, const double* kappa, int n,
                            double lr);
/* trains on fresh batches drawn by the plain sampler of an S(alpha,beta)
 * process, one fused kernel per step; returns the mean loss over the last
 * 10% of the steps */
double ncamd_mlp_train_sab(ncamd_mlp_t, ncamd_process_t sab, int steps, int batch, double lr, uint64_t seed);

/* ---------- MiniMC on the device ----------
 * A beam of nsrc neutrons through a sphere of radius [m] centred at the
 * origin, filled with a material of number density numdens [atoms/Aa^3]
 * whose scattering is `scatter` and absorption `absorption` (either may be a
 * null handle), simulated with the forced-collision engine of NCrystal's
 * MiniMC (ncrystal_core/src/minimc/NCMMC_SimEngine.cc:170-449: per step the
 * unscattered part of each neutron is tallied at the surface, the rest
 * scatters at a distance drawn from the truncated exponential, absorption by
 * weight, Russian roulette), sources "constant" (src_kind 0) and "circular"
 * (1) of NCMMC_Source.cc with src = {ekin [eV], x, y, z [m], ux, uy, uz, r
 * [m]}, missing neutrons tallied as in the reference unless ignored.
 * Tallies (NCMMC_StdTallies.cc): 0 theta [deg] (angle to the beam), 1 mu,
 * 2 nscat (-1 for neutrons missing the sphere), 3 w (unweighted), 4
 * nscat_uw (unweighted), 5 e [eV], 6 l [Aa], 7 de = E0 - E [eV], 8 q [1/Aa],
 * binned in nbins (<= 400) over [xmin, xmax), with the
 * breakdown by scattering history when `breakdown`.  engine = {roulette
 * survival probability, weight threshold, nscat threshold, nscatlimit (< 0:
 * none), ignoremiss, include absorption}.  hist[6][2][nbins + 2] receives,
 * for the total and the categories NOSCAT, SINGLESCAT_ELAS,
 * SINGLESCAT_INELAS, MULTISCAT_PUREELAS, MULTISCAT_OTHER, the sums of weights
 * and of squared weights in (underflow, bins, overflow); info[5] = {most
 * steps of a batch, neutron-steps, steps summed over batches, neutrons
 * passed on to scatter, device time of the step kernels [ms]}.  Replaces
 * NCrystal.minimc.run's engine for this geometry (ncrystal_python
 * minimc.py run(), NCMMC_RunSim.cc).  0, or -1 on error. */
int ncamd_minimc_sphere(ncamd_process_t scatter, ncamd_process_t absorption, double numdens, double radius,
                        int src_kind, const double* src, uint64_t nsrc, int tally, int nbins, double xmin,
                        double xmax, int breakdown, const double* engine, uint64_t seed, double* hist, double* info);

/* ---------- learned-proposal sampler of S(alpha,beta) (no counterpart in
 * NCrystal) ----------
 * The envelope sampler above on grids placed by the model: for each energy
 * interval the model's mixture at the interval's upper energy gives the
 * w-edges (quantiles of its w-marginal) and the kappa-edges of each strip
 * (quantiles of its kappa distribution in the strip), blended with the
 * uniform grid; the blend with the smallest exact envelope mass is kept,
 * the uniform grid included, so the learned grid is never worse than the
 * uniform one.  Accept/reject stays exact against the table; statistics
 * through ncamd_sab_envelope_stats. */
int ncamd_sab_use_learned_sampler(ncamd_process_t, ncamd_mlp_t, int nintervals, double emin, double emax,
                                  int nbox_w, int nbox_k, int maxtries);

/* ---------- RNG (ncrystal.h:1067-1087) ----------
 * Each process handle owns a counter-based stream: (seed, call counter).
 * Every sampling call consumes one call-counter value; neutron k of a call
 * uses Philox4x32-10 with key=seed, counter=(draw, call, k). */
void ncamd_setrngseed(ncamd_process_t, uint64_t seed);
void ncamd_getrngstate(ncamd_process_t, uint64_t* seed, uint64_t* callcounter);
void ncamd_setrngstate(ncamd_process_t, uint64_t seed, uint64_t callcounter);

/* ---------- batched per-neutron calls (host pointers) ---------- */

/* ncrystal_crosssection_nonoriented_many (ncrystal.h:1289-1293,
 * ncrystal.cc:1107-1141): results[irepeat*n_ekin+i] = xs(ekin[i]) [barn]. */
void ncamd_crosssection_nonoriented_many(ncamd_process_t, const double* ekin,
                                         unsigned long n_ekin, unsigned long repeat,
                                         double* results);
/* same with FP32 PowderBragg cumulative tables (FP64 thresholds/searches) */
void ncamd_crosssection_nonoriented_many_fp32(ncamd_process_t, const double* ekin,
                                              unsigned long n_ekin, unsigned long repeat,
                                              double* results);

/* ncrystal_samplescatterisotropic_many (ncrystal.h:1295-1301,
 * ncrystal.cc:1201-1246): out index irepeat*n_ekin+i. */
void ncamd_samplescatterisotropic_many(ncamd_process_t, const double* ekin,
                                       unsigned long n_ekin, unsigned long repeat,
                                       double* results_ekin, double* results_mu);

/* ncrystal_samplescatter_many (ncrystal.h:1303-1309, ncrystal.cc:1248-1282):
 * one energy and direction, repeated. */
void ncamd_samplescatter_many(ncamd_process_t, double ekin, const double (*direction)[3],
                              unsigned long repeat, double* results_ekin, double* results_dirx,
                              double* results_diry, double* results_dirz);

/* New: per-neutron directions (NCrystal's Python vectorises this in Python,
 * ncrystal_python/src/NCrystal/_chooks.py:579-593).  dirs/out_dirs: 3*n. */
void ncamd_crosssection_many_dirs(ncamd_process_t, const double* ekin, const double* dirs,
                                  unsigned long n, double* results);
void ncamd_samplescatter_many_dirs(ncamd_process_t, const double* ekin, const double* dirs,
                                   unsigned long n, double* results_ekin, double* results_dirs);

/* ---------- batched per-neutron calls (device pointers, async on stream) ----------
 * Same semantics; RNG call counter advanced at enqueue time.  Errors raised
 * inside kernels are counted on the device and reported by
 * ncamd_dev_check_errors (which synchronises the stream). */
void ncamd_dev_crosssection_nonoriented_many(ncamd_process_t, const double* d_ekin,
                                             unsigned long n_ekin, unsigned long repeat,
                                             double* d_results, void* stream);
void ncamd_dev_crosssection_nonoriented_many_fp32(ncamd_process_t, const double* d_ekin,
                                                  unsigned long n_ekin, unsigned long repeat,
                                                  double* d_results, void* stream);
void ncamd_dev_samplescatterisotropic_many(ncamd_process_t, const double* d_ekin,
                                           unsigned long n_ekin, unsigned long repeat,
                                           double* d_results_ekin, double* d_results_mu,
                                           void* stream);
/* Allocation-free form of ncamd_dev_samplescatterisotropic_many (same
 * arguments, results and RNG stream use).  No counterpart in NCrystal: it is
 * for callers that cannot allocate device memory during a call (HIP graph
 * capture, tight serving loops).  Its queues and counters live in a fixed
 * per-GPU scratch area of the library (32 MiB + 256 B, held until the
 * process ends), so a call takes at most 2097152 neutrons (n x repeat;
 * larger calls raise BadInput), and calls that overlap on one GPU (on
 * different streams or host threads) race on that area. */
void ncamd_dev_samplescatterisotropic_many_noalloc(ncamd_process_t, const double* d_ekin,
                                                   unsigned long n_ekin, unsigned long repeat,
                                                   double* d_results_ekin, double* d_results_mu,
                                                   void* stream);
void ncamd_dev_samplescatter_many_dirs(ncamd_process_t, const double* d_ekin,
                                       const double* d_dirs, unsigned long n,
                                       double* d_results_ekin, double* d_results_dirs,
                                       void* stream);
/* per-neutron directions (oriented processes; isotropic ones ignore d_dirs) */
void ncamd_dev_crosssection_many_dirs(ncamd_process_t, const double* d_ekin, const double* d_dirs,
                                      unsigned long n, double* d_results, void* stream);
/* S(alpha,beta) only: raw (alpha,beta) samples of the plain sampler */
void ncamd_dev_sab_sample_alphabeta(ncamd_process_t, const double* d_ekin, unsigned long n,
                                    double* d_alpha, double* d_beta, void* stream);
/* number of neutrons whose sampling failed since the last check (syncs stream) */
int64_t ncamd_dev_check_errors(ncamd_process_t, void* stream);

/* ---------- many materials in one launch (no counterpart in NCrystal, whose
 * C API evaluates one material per call) ----------
 * A batch references nmat processes (isotropic: PowderBragg, ElInc,
 * S(alpha,beta) or compositions of them) and the number of neutrons of each.
 * Inputs/outputs are concatenated per material in batch order (material i
 * owns entries [sum(counts[0..i)), +counts[i])).  One call runs all materials
 * in one persistent launch (replayed from a captured HIP graph) and returns
 * exactly what one ncamd_dev_* call per material would return: every
 * material keeps its own seed and call counter, and a sampling call advances
 * each counter by one. */
typedef struct { void* internal; } ncamd_batch_t;
ncamd_batch_t ncamd_batch_create(int nmat, const ncamd_process_t* procs, const int64_t* counts);
void ncamd_batch_destroy(ncamd_batch_t*);
int64_t ncamd_batch_total(ncamd_batch_t);
void ncamd_dev_batch_crosssection(ncamd_batch_t, const double* d_ekin, double* d_xs, void* stream);
void ncamd_dev_batch_samplescatterisotropic(ncamd_batch_t, const double* d_ekin, double* d_results_ekin,
                                            double* d_results_mu, void* stream);
int64_t ncamd_dev_batch_check_errors(ncamd_batch_t, void* stream);
/* launch parameters: neutrons per tile, resident workgroups, tile order
 * (0: material order, 1: most expensive materials first) */
void ncamd_batch_get_tuning(ncamd_batch_t, int* tile, int* nblocks, int* ordering);
void ncamd_batch_set_tuning(ncamd_batch_t, int tile, int nblocks, int ordering);
/* times the candidate tunings on these inputs (outputs are scratch, RNG
 * states unchanged), keeps the fastest and returns its time [ms] */
double ncamd_dev_batch_autotune(ncamd_batch_t, const double* d_ekin, double* d_o1, double* d_o2, void* stream);
/* how often the graph of a mode (0 cross sections, 1 sampling) was captured
 * and how often it was replayed unchanged */
void ncamd_batch_graph_stats(ncamd_batch_t, int mode, int64_t* captures, int64_t* replays);

/* ---------- device memory, streams and events (plumbing for callers that
 * do not bring their own HIP runtime objects; no counterpart in NCrystal) ---- */
void* ncamd_dev_malloc(uint64_t bytes);
void ncamd_dev_free(void* p);
void ncamd_memcpy_htod(void* dst, const void* src, uint64_t bytes);
void ncamd_memcpy_dtoh(void* dst, const void* src, uint64_t bytes);
void ncamd_memcpy_dtod(void* dst, const void* src, uint64_t bytes, void* stream);
void* ncamd_stream_create(void);
void ncamd_stream_destroy(void* stream);
void ncamd_stream_sync(void* stream);
void* ncamd_event_create(void);
void ncamd_event_destroy(void* ev);
void ncamd_event_record(void* ev, void* stream);
float ncamd_event_elapsed_ms(void* ev_start, void* ev_stop);
/* fill d_out[0..n) with E = exp(U(log emin, log emax)) from Philox(seed) */
void ncamd_dev_fill_loguniform(double* d_out, uint64_t n, double emin, double emax, uint64_t seed, void* stream);

#ifdef __cplusplus
}
#endif
#endif
