// VDOS -> S(alpha,beta) (nc_vdos.hip): host part and GPU fill.
#pragma once
#include <hip/hip_runtime.h>

#include <stdexcept>
#include <string>
#include <vector>

namespace ncamd {

// errors carry the NCrystal error type (BadInput, CalcError, ...)
struct VdosError : std::runtime_error {
  std::string type;
  VdosError(const std::string& t, const std::string& m) : std::runtime_error(m), type(t) {}
};

// a VDOS on the regular grid emin + i*(emax-emin)/(n-1) that extends down to
// zero in whole bins (regulariseVDOSGrid, NCVDOSEval.cc:497-640)
struct RegularVdos {
  double emin = 0.0, emax = 0.0;
  std::vector<double> density;
};
RegularVdos vdos_regularise(const std::vector<double>& egrid, const std::vector<double>& density);
// createVDOSDebye (NCDynInfoUtils.cc:377-396): 20 points, density E^2/E_D^2
void vdos_debye(double debye_temp, double* emin, double* emax, std::vector<double>& density);

struct VdosAnalysis {
  double gamma0, msd, teff, integral;
};
VdosAnalysis vdos_analyse(const RegularVdos&, double temperature, double mass_amu);

// createScatteringKernel (NCVDOSToScatKnl.cc:787-944): sab[ibeta*nalpha+ialpha]
struct VdosKernel {
  std::vector<double> alpha, beta, sab;
  int norders = 0, orders_summed = 0, rows_summed = 0;
  double suggested_emax = 0.0, msd = 0.0, gamma0 = 0.0;
  double gpu_ms = 0.0;  // device time of the operand and order-sum kernels
};
VdosKernel vdos_kernel(const RegularVdos&, double temperature, double mass_amu, unsigned vdoslux,
                       double target_emax, hipStream_t stream);

// the phonon-order-n spectrum G_n of the expansion (VDOSGn): values on the
// grid lo + i * binwidth
std::vector<double> vdos_gn(const RegularVdos&, double temperature, double mass_amu, unsigned order, double& lo,
                            double& binwidth, hipStream_t stream);

}  // namespace ncamd
