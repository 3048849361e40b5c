/* Compiled against include/ncrystal_dropin.h and linked with
 * libncrystal_amd.so; runs without a GPU.  Checks the parts of the NCrystal C
 * API convention that need no device: version strings, halt/quiet setters
 * returning the old value, the error handler's (type, message) arguments,
 * poisoned outputs and the error state on invalid handles, handle validity
 * and invalidation.  Prints one line per check; exit status 0 when all pass. */
#include <math.h>
#include <stdio.h>
#include <string.h>

#include "ncrystal_dropin.h"

static int nfail = 0;
static int nhandler = 0;
static char last_type[64];

static void check(int ok, const char* what) {
  printf("%s %s\n", ok ? "ok  " : "FAIL", what);
  if (!ok) ++nfail;
}

static void handler(char* type, char* msg) {
  ++nhandler;
  strncpy(last_type, type, sizeof(last_type) - 1);
  (void)msg;
}

int main(void) {
  check(ncrystal_version() == 4004002, "ncrystal_version");
  check(strcmp(ncrystal_version_str(), "4.4.2") == 0, "ncrystal_version_str");
  check(strcmp(ncrystal_namespace(), "") == 0, "ncrystal_namespace");
  check(ncrystal_sethaltonerror(0) == 1, "halt-on-error defaults to 1, setter returns the old value");
  check(ncrystal_setquietonerror(1) == 0, "quiet defaults to 0, setter returns the old value");
  check(ncrystal_sethaltonerror(0) == 0, "halt setter returns the value just set");
  ncrystal_seterrhandler(handler);
  check(ncrystal_error() == 0 && ncrystal_lasterror() == NULL && ncrystal_lasterrortype() == NULL,
        "no error state at start (NULL message and type)");

  ncrystal_process_t bad = {NULL};
  ncrystal_scatter_t bads = {NULL};
  double xs = 0.0;
  ncrystal_crosssection_nonoriented(bad, 0.025, &xs);
  check(xs == -1.0 && ncrystal_error() == 1, "xs of an invalid handle: -1 and the error flag");
  check(nhandler == 1 && strcmp(last_type, "LogicError") == 0, "error handler called with the type");
  check(ncrystal_lasterror() != NULL && strcmp(ncrystal_lasterrortype(), "LogicError") == 0,
        "lasterror / lasterrortype");
  ncrystal_clearerror();
  check(ncrystal_error() == 0, "clearerror");

  double ekin[3] = {0.01, 0.02, 0.03}, out[6];
  ncrystal_crosssection_nonoriented_many(bad, ekin, 3, 2, out);
  int allm1 = 1;
  for (int i = 0; i < 6; ++i) allm1 &= out[i] == -1.0;
  check(allm1 && ncrystal_error(), "_many xs poisoned to -1 for all n*repeat outputs");
  ncrystal_clearerror();

  double ef[6], mu[6];
  ncrystal_samplescatterisotropic_many(bads, ekin, 3, 2, ef, mu);
  int pois = 1;
  for (int i = 0; i < 6; ++i) pois &= ef[i] == -1.0 && mu[i] == -999.0;
  check(pois, "_many isotropic sampling poisoned (E'=-1, mu=-999)");
  ncrystal_clearerror();

  const double dir[3] = {0.0, 0.0, 1.0};
  double e1, d1[3] = {5, 5, 5};
  ncrystal_samplescatter(bads, 0.025, &dir, &e1, &d1);
  check(e1 == -1.0 && d1[0] == 0.0 && d1[1] == 0.0 && d1[2] == 0.0, "oriented sampling poisoned (dir 0,0,0)");
  ncrystal_clearerror();

  double ux[2], uy[2], uz[2], em[2];
  ncrystal_samplescatter_many(bads, 0.025, &dir, 2, em, ux, uy, uz);
  check(em[0] == -1.0 && em[1] == -1.0 && ux[1] == 0.0 && uz[0] == 0.0, "_many oriented sampling poisoned");
  ncrystal_clearerror();

  double lo = 0, hi = 0;
  ncrystal_domain(bad, &lo, &hi);
  check(lo == -1.0 && hi == -1.0, "domain of an invalid handle: (-1, -1)");
  ncrystal_clearerror();

  double ang, dek;
  ncrystal_genscatter_nonoriented(bads, 0.025, &ang, &dek);
  check(ang == -999.0 && dek == -999.0, "obsolete genscatter poisoned to -999");
  ncrystal_clearerror();

  check(ncrystal_refcount(&bad) == -999 && ncrystal_error(), "refcount of an invalid handle: -999 and error");
  ncrystal_clearerror();
  check(ncrystal_valid(&bad) == 0 && ncrystal_valid(NULL) == 0, "invalid handles are not valid");
  check(ncrystal_name(bad) == NULL && ncrystal_error(), "name of an invalid handle: NULL and error");
  ncrystal_clearerror();
  ncrystal_scatter_t c = ncrystal_cast_proc2scat(bad);
  check(c.internal == NULL && ncrystal_error(), "cast of an invalid handle: null handle and error");
  ncrystal_clearerror();
  check(nhandler >= 10, "the handler saw every error");
  printf("%d failures\n", nfail);
  return nfail ? 1 : 0;
}
