// Polynomial trigonometry shared by host (table construction) and device
// code.  Operation order follows the reference so that results are
// bit-identical with -ffp-contract=off:
//  - sincos_mpi2pi2: ncrystal_core/src/utils/NCMath.cc:96-125 (Taylor series
//    at A/2 and the double-angle step)
//  - sincos_mpi8pi8: NCMath.cc:124-148 (short series for |A| <= pi/8)
//  - cos_mpipi / cos_mpi2pi2: NCMath.cc:150-195
//  - sincos_0pi32: ncrystal_core/include/NCrystal/internal/utils/NCMath.hh:368-374
#pragma once
#include "nc_dev.hh"

namespace ncamd {

NCHD void sincos_mpi2pi2(double A, double& cosA, double& sinA) {
  const double x = 0.5 * A;
  const double mx2 = -x * x;
  const double s2 = x * (1.0 + mx2 * (1.66666666666666666666666666666666666666666667e-1
                   + mx2 * (8.33333333333333333333333333333333333333333333e-3
                   + mx2 * (1.98412698412698412698412698412698412698412698e-4
                   + mx2 * (2.75573192239858906525573192239858906525573192e-6
                   + mx2 * (2.50521083854417187750521083854417187750521084e-8
                   + mx2 * (1.60590438368216145993923771701549479327257105e-10
                   + mx2 * (7.64716373181981647590113198578807044415510024e-13))))))));
  const double c2m1 = mx2 * (0.5
                + mx2 * (4.16666666666666666666666666666666666666666667e-2
                + mx2 * (1.38888888888888888888888888888888888888888889e-3
                + mx2 * (2.48015873015873015873015873015873015873015873e-5
                + mx2 * (2.75573192239858906525573192239858906525573192e-7
                + mx2 * (2.08767569878680989792100903212014323125434237e-9
                + mx2 * (1.14707455977297247138516979786821056662326504e-11
                + mx2 * (4.77947733238738529743820749111754402759693765e-14))))))));
  const double k = 2.0 * c2m1;
  sinA = (k + 2.0) * s2;
  cosA = 1.0 + k * (c2m1 + 2.0);
}

// cos(x) for |x| <= pi/2: Taylor series in -x^2 to order 22
NCHD double cos_mpi2pi2(double x) {
  const double mx2 = -x * x;
  return 1.0 + mx2 * (0.5
         + mx2 * (4.16666666666666666666666666666666666666666667e-2
         + mx2 * (1.38888888888888888888888888888888888888888889e-3
         + mx2 * (2.48015873015873015873015873015873015873015873e-5
         + mx2 * (2.75573192239858906525573192239858906525573192e-7
         + mx2 * (2.08767569878680989792100903212014323125434237e-9
         + mx2 * (1.14707455977297247138516979786821056662326504e-11
         + mx2 * (4.77947733238738529743820749111754402759693765e-14
         + mx2 * (1.56192069685862264622163643500573334235194041e-16
         + mx2 * (4.1103176233121648584779906184361403746103695e-19
         + mx2 * (8.89679139245057328674889744250246834331248809e-22)))))))))));
}

// cos(A) for |A| <= pi via the reflection about pi/2
NCHD double cos_mpipi(double A) {
  const double Aabs = fabs(A);
  const double x = dmin(Aabs, kPi - Aabs);
  return copysign(cos_mpi2pi2(x), kPiHalf - Aabs);
}

NCHD void sincos_mpi8pi8(double A, double& cosA, double& sinA) {
  const double x = 0.5 * A;
  const double mx2 = -x * x;
  const double s2 = x * (1.0 + mx2 * (1.66666666666666666666666666666666666666666667e-1
                   + mx2 * (8.33333333333333333333333333333333333333333333e-3
                   + mx2 * (1.98412698412698412698412698412698412698412698e-4
                   + mx2 * (2.75573192239858906525573192239858906525573192e-6
                   + mx2 * (2.50521083854417187750521083854417187750521084e-8))))));
  const double c2m1 = mx2 * (0.5 + mx2 * (4.16666666666666666666666666666666666666666667e-2
                + mx2 * (1.38888888888888888888888888888888888888888889e-3
                + mx2 * (2.48015873015873015873015873015873015873015873e-5
                + mx2 * (2.75573192239858906525573192239858906525573192e-7)))));
  const double k = 2.0 * c2m1;
  sinA = (k + 2.0) * s2;
  cosA = 1.0 + k * (c2m1 + 2.0);
}

// A in [0, 3pi/2]: reflection about pi/2 for the cosine sign
NCHD void sincos_0pi32(double A, double& c, double& s) {
  sincos_mpi2pi2(dmin(A, kPi - A), c, s);
  c = copysign(c, kPiHalf - A);
}

}  // namespace ncamd
