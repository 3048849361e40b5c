"""The free-gas model of the CPU checker (FreeGas process, and the sampler the
S(alpha,beta) high-energy extender shares) pinned by closed forms:

* cross section: sigma_free * [(1 + 1/(2a^2)) erf(a) + exp(-a^2)/(a sqrt(pi))]
  with a^2 = A E / kT (ncrystal_core/src/phys_utils/NCFreeGasUtils.cc:28-120
  evaluates this with series and asymptotic forms);
* sampling: (alpha, beta) follow the free-gas scattering function
  S(alpha, beta) = exp(-(alpha/A + beta)^2 / (4 alpha/A)) / sqrt(4 pi alpha/A)
  on the kinematic region.  With u = sqrt(alpha/A), the alpha integral is
  closed: the integrand is exp(-g^2)/sqrt(pi) du with g = u/2 + beta/(2u), and
  d/du [erf(g) - exp(-g^2) erfcx(h)]/2 = exp(-g^2)/sqrt(pi) for
  h = u/2 - beta/(2u).  This gives the beta marginal (by one-dimensional
  quadrature over beta) and the conditional CDF of alpha given beta, both
  compared with the samples by Kolmogorov-Smirnov tests;
* 1/v absorption: sigma_abs sqrt(E_2200 / E).

The sampling checks run on both free-gas procedures of the checker: the
two-body collision draw that the device performs (the default) and the
restatement of the reference's FreeGasSampler; a two-sample test compares
the two directly.
"""
import math

import numpy as np
import pytest
from scipy import special, stats

KB = 8.6173303e-5          # eV/K, NCrystal's constant_boltzmann
MN = 1.00866491588         # neutron mass [amu]
EKIN2200 = 0.5 * (MN * 931494095.17 / (299792458e10 ** 2)) * 2200.0e10 * 2200.0e10


def _shape(a2):
    a = np.sqrt(a2)
    return (1 + 0.5 / a2) * special.erf(a) + np.exp(-a2) / (a * math.sqrt(math.pi))


def test_freegas_xs_closed_form(oracle):
    for T, M, sf in ((293.15, 4.0026, 0.76), (20.0, 1.00794, 20.5), (600.0, 39.95, 0.68)):
        g = oracle.freegas(T, M, sf)
        e = np.geomspace(1e-7, 10.0, 4000)
        a2 = (M / MN) * e / (KB * T)
        ref = sf * _shape(a2)
        assert np.max(np.abs(g.xs(e) - ref) / ref) < 1e-9, (T, M)


def _I(u, beta):
    g = u / 2 + beta / (2 * u)
    h = u / 2 - beta / (2 * u)
    return 0.5 * (special.erf(g) - np.exp(-g * g) * special.erfcx(h))


def _alpha_lims(c, beta):
    r = np.sqrt(np.maximum(c + beta, 0.0))
    return (math.sqrt(c) - r) ** 2, (math.sqrt(c) + r) ** 2


def _sample(oracle, procedure, T, M, E, n, seed=7):
    g = oracle.freegas(T, M, 1.0)
    with oracle.freegas_procedure(procedure):
        return g.sample_iso(np.full(n, E), seed=seed, nthreads=8)


CASES = [(293.15, 1.00794, 0.025), (293.15, 4.0026, 0.001), (293.15, 4.0026, 0.5),
         (100.0, 39.95, 0.05), (600.0, 12.011, 2.0), (293.15, 26.98, 1e-6)]


@pytest.mark.parametrize("procedure", ["collision", "reference"])
@pytest.mark.parametrize("T,M,E", CASES)
def test_freegas_sampling_distribution(oracle, procedure, T, M, E):
    n = 200000
    kT = KB * T
    A = M / MN
    c = E / kT
    ef, mu, nerr = _sample(oracle, procedure, T, M, E, n)
    assert nerr == 0 and np.all(np.abs(mu) <= 1.0) and np.all(ef >= 0.0)
    beta = (ef - E) / kT
    alpha = (E + ef - 2 * mu * np.sqrt(E * ef)) / kT
    # beta marginal: P(beta) ~ I(u+) - I(u-) over [-c, bmax]
    bmax = beta.max() + 5.0
    bg = np.concatenate([-c + (c + 1e-12) * np.linspace(0, 1, 4001) ** 2, np.linspace(1e-12, bmax, 8001)[1:]])
    bg = bg[bg > -c]
    lo, hi = _alpha_lims(c, bg)
    p = _I(np.sqrt(hi / A), bg) - _I(np.sqrt(np.maximum(lo, 1e-300) / A), bg)
    cdf = np.concatenate([[0.0], np.cumsum(0.5 * (p[1:] + p[:-1]) * np.diff(bg))])
    cdf /= cdf[-1]
    d_beta = stats.kstest(beta, lambda x: np.interp(x, bg, cdf)).statistic
    # alpha given beta: conditional CDF values are uniform
    lo, hi = _alpha_lims(c, beta)
    ulo, uhi = np.sqrt(lo / A), np.sqrt(hi / A)
    ua = np.sqrt(np.clip(alpha, lo, hi) / A)
    den = _I(uhi, beta) - _I(ulo, beta)
    ok = den > 1e-300
    f = (_I(ua[ok], beta[ok]) - _I(ulo[ok], beta[ok])) / den[ok]
    d_alpha = stats.kstest(f, "uniform").statistic
    lim = 1.8 / math.sqrt(n)  # KS critical value at about 0.3% significance
    assert d_beta < lim and d_alpha < lim, (d_beta, d_alpha, lim)


@pytest.mark.parametrize("procedure", ["collision", "reference"])
def test_freegas_sampler_fails_a_wrong_mass(oracle, procedure):
    """the check above has the power to see a 10% change of the target mass"""
    n = 200000
    T, M, E = 293.15, 4.0026, 0.025
    kT, c = KB * T, 0.025 / (KB * 293.15)
    ef, mu, _ = _sample(oracle, procedure, T, M * 1.1, E, n)
    beta = (ef - E) / kT
    alpha = (E + ef - 2 * mu * np.sqrt(E * ef)) / kT
    A = M / MN
    lo, hi = _alpha_lims(c, beta)
    den = _I(np.sqrt(hi / A), beta) - _I(np.sqrt(lo / A), beta)
    f = (_I(np.sqrt(np.clip(alpha, lo, hi) / A), beta) - _I(np.sqrt(lo / A), beta)) / den
    assert stats.kstest(f, "uniform").statistic > 2 * 1.8 / math.sqrt(n)


@pytest.mark.parametrize("T,M,E", CASES + [(293.15, 26.98, 50.0), (20.0, 2.014, 1e-4)])
def test_freegas_procedures_agree(oracle, T, M, E):
    """collision draw and the reference's sampler: two-sample KS on E' and mu
    (different streams, 3e5 draws each)"""
    n = 300000
    e1, m1, n1 = _sample(oracle, "collision", T, M, E, n, seed=3)
    e2, m2, n2 = _sample(oracle, "reference", T, M, E, n, seed=4)
    assert n1 == 0 and n2 == 0
    lim = 1.95 * math.sqrt(2.0 / n)  # about 0.1% significance
    assert stats.ks_2samp(e1, e2).statistic < lim
    assert stats.ks_2samp(m1, m2).statistic < lim


def test_absoov_closed_form(oracle):
    a = oracle.absoov(0.231)
    e = np.concatenate([[0.0], np.geomspace(1e-9, 100.0, 1000)])
    x = a.xs(e)
    assert np.isinf(x[0])
    ref = 0.231 * np.sqrt(EKIN2200 / e[1:])
    assert np.max(np.abs(x[1:] - ref) / ref) < 1e-15
