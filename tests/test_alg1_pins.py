"""Pins of the checker's S(alpha,beta) sampling procedure that do not rest on
the checker's own code.

* PointwiseDist percentiles: the reference's own test
  (tests/src/app_pointwisedist: a uniform density written as a piecewise
  linear one, 117 percentiles, fixture `pointwisedist_uniform.json`) and, for
  random piecewise-linear densities, the exact distribution function
  evaluated at the returned value (the percentile inverts a quadratic per
  bin, NCPointwiseDist.cc:76-107).
* createTailedBreakdown: the reference's own test (tests/src/app_sabsliceinteg,
  S = exp(alpha) on a 20-point grid, 16 alpha ranges: bin indices exact,
  printed front/middle/back integrals to their 6 digits, the sum within
  1e-15 of the exact integral; fixture `sabsliceinteg.json`).
* Alg1 (SABSampler::sampleAlphaBeta NCSABSampler.cc:158-227 with
  SABSamplerAtE_Alg1::sampleAlphaBeta NCSABSamplerModels.cc:48-155): the
  density of the whole procedure is computed here from the table alone --
  the sampler energy picked with the 1.05 egrid margin, beta weights as the
  exact integrals of the log-linearly interpolated S over the kinematic
  alpha range at the sampler energy, a piecewise-linear beta density whose
  first bin is widened by 1/3 (the 4/3 rule) and resampled with the
  sqrt-shaped endpoint rule, alpha as the linear interpolation in beta of the
  two rows' quantiles at one shared percentile (its distribution function
  found by a monotone root solve), and rejection at the neutron's own
  kinematic limits.  The checker's samples must follow it (KS of beta; KS of
  the conditional percentile of alpha given beta), and each of three planted
  changes of the procedure (no first-bin rule, margin 1.0, nearer row instead
  of the quantile interpolation) must fail it.
"""
import json
import math

import numpy as np
import pytest
from scipy import stats

from conftest import GOLDEN

KB = 8.6173303e-5


# ---------------------------------------------------------------- PointwiseDist
def test_pointwisedist_reference_log(oracle):
    g = json.loads((GOLDEN / "pointwisedist_uniform.json").read_text())
    r = np.linspace(0.0, 1.0, g["npoints"])
    p, idx = oracle.pwdist_percentiles(g["x"], g["w"], r)
    assert len(g["printed"]) == r.size
    for (rp, pp), rv, pv in zip(g["printed"], r, p):
        assert float(f"{rv:.6g}") == rp
        assert float(f"{pv:.6g}") == pp, (rv, pv, pp)
    # the reference test's own assertion: p = 4 r within 1e-12
    assert np.all(np.abs(p - 4.0 * r) <= 1e-12 * np.maximum(1.0, 4.0 * r))


def _pw_cdf(x, y, v):
    """exact distribution function of the piecewise-linear density (x, y) at v"""
    area = 0.5 * np.diff(x) * (y[1:] + y[:-1])
    tot = area.sum()
    cum = np.concatenate([[0.0], np.cumsum(area)])
    i = np.clip(np.searchsorted(x, v, side="right") - 1, 0, x.size - 2)
    dx = v - x[i]
    slope = (y[i + 1] - y[i]) / (x[i + 1] - x[i])
    return (cum[i] + y[i] * dx + 0.5 * slope * dx * dx) / tot, i


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_pointwisedist_inverts_exact_cdf(oracle, seed):
    rng = np.random.default_rng(seed)
    n = 25
    x = np.cumsum(rng.uniform(0.01, 1.0, n))
    y = rng.uniform(0.0, 2.0, n)
    y[rng.integers(0, n, 4)] = 0.0  # zero weights: triangles and gaps
    y[0] = 0.0
    p = np.concatenate([rng.uniform(0, 1, 20000), [0.0, 1.0, 1e-300, 0.5]])
    v, idx = oracle.pwdist_percentiles(x, y, p)
    F, i = _pw_cdf(x, y, v)
    assert np.all((v >= x[0]) & (v <= x[-1]))
    assert np.max(np.abs(F - p)) < 1e-12
    # the reported bin holds the value (bins of zero mass are never reported)
    assert np.all((x[idx] <= v) & (v <= x[idx + 1]))


# -------------------------------------------------------- createTailedBreakdown
def _tb_indices(ag, alow, aupp):
    """bin indices as the reference test selects them
    (tests/src/app_sabsliceinteg/main.cc:46-58)"""
    it = int(np.searchsorted(ag, alow, side="left"))
    if it > 0:
        it -= 1
    while it + 1 < ag.size and ag[it + 1] <= alow:
        it += 1
    lo = it
    it = int(np.searchsorted(ag[:-1], aupp, side="left"))
    while it > 0 and ag[it - 1] >= aupp:
        it -= 1
    return lo, it


def test_tailed_breakdown_reference_log(oracle):
    g = json.loads((GOLDEN / "sabsliceinteg.json").read_text())
    ag = np.array(g["alpha"])
    s = np.exp(ag)
    ls = np.log(s)

    def exact(a, b):
        if a >= ag[-1] or b <= ag[0]:
            return 0.0
        a, b = max(a, ag[0]), min(b, ag[-1])
        return math.exp(a) * math.expm1(b - a)

    cumul = np.concatenate([[0.0], np.cumsum([exact(ag[i - 1], ag[i]) for i in range(1, ag.size)])])
    assert len(g["cases"]) == 16
    for c in g["cases"]:
        lo, up = _tb_indices(ag, c["alow"], c["aupp"])
        assert [lo, up] == c["aidx"], c
        f, m, b = oracle.tailed_breakdown(ag, s, ls, cumul, c["alow"], c["aupp"], lo, up)
        for v, ref in ((f, c["front"]), (m, c["middle"]), (b, c["back"]), (f + m + b, c["total"])):
            assert float(f"{v:.6g}") == ref, (c, v, ref)
        ex = exact(c["alow"], c["aupp"])
        assert float(f"{ex:.6g}") == c["expected"]
        assert (f + m + b == ex) or abs((f + m + b) / ex - 1.0) < 2e-15, (c, f + m + b, ex)


# -------------------------------------------------------------------- Alg1
def _alpha_lims(c, beta):
    beta = np.asarray(beta, dtype=np.float64)
    r = np.sqrt(np.maximum(c + beta, 0.0))
    s = math.sqrt(c) + r
    return beta * beta / (s * s), s * s


class Row:
    """log-linearly interpolated S(alpha) of one beta row, restricted to an
    alpha interval; exact masses, distribution function and quantiles"""

    def __init__(self, ag, srow, lo, hi):
        lo, hi = max(lo, ag[0]), min(hi, ag[-1])
        self.empty = not hi > lo
        if self.empty:
            self.lo, self.hi, self.W = lo, hi, 0.0
            return
        inner = ag[(ag > lo) & (ag < hi)]
        a = np.concatenate([[lo], inner, [hi]])
        self.a = a
        self.s = self.interp(ag, srow, a)
        self.d = np.diff(a)
        self.r = np.log(self.s[1:] / self.s[:-1])
        m = self._mass(self.s[:-1], self.r, self.d, self.d)
        self.cum = np.concatenate([[0.0], np.cumsum(m)])
        self.W = self.cum[-1]
        self.lo, self.hi = lo, hi

    @staticmethod
    def interp(ag, srow, a):
        i = np.clip(np.searchsorted(ag, a, side="right") - 1, 0, ag.size - 2)
        t = (a - ag[i]) / (ag[i + 1] - ag[i])
        return srow[i] ** (1 - t) * srow[i + 1] ** t

    @staticmethod
    def _mass(s0, r, d, x):
        """integral of s0 exp(r t/d) over t in [0, x]"""
        z = r * x / d
        small = np.abs(z) < 1e-6
        with np.errstate(over="ignore", invalid="ignore", divide="ignore"):
            big = s0 * d * np.expm1(z) / r
        ser = s0 * x * (1 + z / 2 + z * z / 6)
        return np.where(small, ser, big)

    def cdf(self, alpha):
        alpha = np.clip(alpha, self.lo, self.hi)
        k = np.clip(np.searchsorted(self.a, alpha, side="right") - 1, 0, self.d.size - 1)
        part = self._mass(self.s[k], self.r[k], self.d[k], alpha - self.a[k])
        return (self.cum[k] + part) / self.W

    def quantile(self, u):
        m = np.asarray(u) * self.W
        k = np.clip(np.searchsorted(self.cum, m, side="right") - 1, 0, self.d.size - 1)
        rest = m - self.cum[k]
        s0, r, d = self.s[k], self.r[k], self.d[k]
        z = r * rest / (s0 * d)
        with np.errstate(invalid="ignore", divide="ignore"):
            x = np.where(np.abs(r) > 1e-9, np.log1p(z) * d / r, rest / s0)
        return np.clip(self.a[k] + np.clip(x, 0, d), self.lo, self.hi)


def _table():
    """10 x 14 table: a free-gas-like ridge plus a floor, positive everywhere"""
    ag = np.geomspace(0.02, 40.0, 10)
    bg = np.linspace(-12.0, 12.0, 14)
    A, B = np.meshgrid(ag, bg)
    w = 1.6
    sab = np.exp(-(A / w + B) ** 2 / (4 * A / w)) / np.sqrt(4 * math.pi * A / w) * np.exp(-B / 2) + 2e-3
    return ag, bg, sab


class Procedure:
    """distribution of the Alg1 procedure for one neutron energy"""

    def __init__(self, ag, bg, sab, kT, egrid, E, margin=1.05):
        self.ag, self.bg, self.sab = ag, bg, sab
        self.c = E / kT
        up = int(np.searchsorted(egrid, E, side="right"))
        assert 0 < up < egrid.size
        while up + 1 < egrid.size and E * margin > egrid[up]:
            up += 1
        self.cs = cs = egrid[up] / kT
        # rows with kinematically accessible alpha at the sampler energy
        ibl = None
        rows = {}
        for j, b in enumerate(bg):
            if b <= -cs:
                continue
            lo, hi = _alpha_lims(cs, b)
            if not (ag[-1] <= lo or ag[0] >= hi):
                if ibl is None:
                    ibl = j
            if ibl is not None:
                rows[j] = Row(ag, sab[j], float(lo), float(hi))
        assert ibl is not None and ibl > 0 and bg[ibl - 1] <= -cs, "the pin needs a kinematic first bin"
        assert all(not r.empty and r.W > 0 for r in rows.values())
        self.ibl, self.rows = ibl, rows
        b0, b1 = -cs, bg[ibl]
        x0 = b0 - (b1 - b0) / 3.0  # the 4/3 rule
        self.x = np.concatenate([[x0], bg[ibl:]])
        self.y = np.concatenate([[0.0], [rows[j].W for j in range(ibl, bg.size)]])
        area = 0.5 * np.diff(self.x) * (self.y[1:] + self.y[:-1])
        self.A = area.sum()
        self.P0 = area[0] / self.A
        self.b0, self.b1 = b0, b1
        # first-bin rule: beta uniform in [b0, b1], alpha from the first row
        # within the sampler's limits at beta; an attempt with beta < -c is
        # rejected outright
        r1 = rows[ibl]
        bb = np.linspace(max(b0, -self.c), b1, 4001)
        lo, hi = _alpha_lims(cs, bb)
        Ms = r1.cdf(hi) - r1.cdf(lo)
        Pc = np.trapezoid(Ms, bb) / (b1 - b0)
        Pr = max(0.0, -self.c - b0) / (b1 - b0)
        self.K0 = self.P0 / ((Pc + Pr) * (b1 - b0))
        assert (1 - Pc - Pr) ** 30 < 1e-6  # the 30-try fallback is negligible

    def rows_of(self, beta):
        k = np.searchsorted(self.bg, beta, side="left")  # beta in (bg[k-1], bg[k]]
        return k - 1, k

    def alpha_of_u(self, l, h, t, u):
        return (1 - t) * self.rows[l].quantile(u) + t * self.rows[h].quantile(u)

    def u_of_alpha(self, l, h, t, alpha):
        """percentile U with alpha(U) = alpha (monotone root solve)"""
        lo = np.zeros_like(alpha)
        hi = np.ones_like(alpha)
        for _ in range(44):
            mid = 0.5 * (lo + hi)
            below = self.alpha_of_u(l, h, t, mid) < alpha
            lo = np.where(below, mid, lo)
            hi = np.where(below, hi, mid)
        return 0.5 * (lo + hi)

    def window(self, beta):
        """for beta in the interpolated bins: (l, h, t, U-, U+) of the
        accepted percentiles"""
        l, h = self.rows_of(beta)
        res = np.zeros((5, beta.size))
        for k in np.unique(h):
            sel = h == k
            b = beta[sel]
            t = (b - self.bg[k - 1]) / (self.bg[k] - self.bg[k - 1])
            alo, ahi = _alpha_lims(self.c, b)
            ulo = self.u_of_alpha(k - 1, k, t, alo)
            uhi = self.u_of_alpha(k - 1, k, t, ahi)
            res[:, sel] = [np.full(b.size, k - 1), np.full(b.size, k), t, ulo, uhi]
        return res

    def density(self, beta):
        """unnormalised density of accepted beta"""
        g = np.zeros_like(beta)
        first = (beta >= max(self.b0, -self.c)) & (beta <= self.b1)
        r1 = self.rows[self.ibl]
        lo, hi = _alpha_lims(self.c, beta[first])
        g[first] = self.K0 * (r1.cdf(hi) - r1.cdf(lo))
        rest = (beta > self.b1) & (beta > -self.c)
        b = beta[rest]
        w = self.window(b)
        pib = np.interp(b, self.x, self.y) / self.A
        g[rest] = pib * (w[4] - w[3])
        return g

    def beta_cdf(self):
        pts = []
        if max(self.b0, -self.c) < self.b1:
            pts.append(np.linspace(max(self.b0, -self.c), self.b1, 3001))
        edges = self.bg[self.bg >= self.b1]
        for a, b in zip(edges[:-1], edges[1:]):
            if b > -self.c:
                seg = np.linspace(max(a, -self.c), b, 3001)
                pts.append(seg[1:] if pts else seg)
        xb = np.concatenate(pts)
        g = self.density(xb)
        # the accepted percentile window on this grid, for conditional_u
        rest = (xb > self.b1) & (xb > -self.c)
        self._wgrid = (xb[rest], self.window(xb[rest]))
        cdf = np.concatenate([[0.0], np.cumsum(0.5 * (g[1:] + g[:-1]) * np.diff(xb))])
        return xb, cdf / cdf[-1]

    def conditional_u(self, alpha, beta):
        """percentile of alpha within the accepted range given beta (uniform
        under the procedure)"""
        v = np.empty_like(alpha)
        first = beta <= self.b1
        r1 = self.rows[self.ibl]
        lo, hi = _alpha_lims(self.c, beta[first])
        f0, f1 = r1.cdf(lo), r1.cdf(hi)
        v[first] = (r1.cdf(alpha[first]) - f0) / (f1 - f0)
        rest = ~first
        br = beta[rest]
        l, h = self.rows_of(br)
        t = (br - self.bg[l]) / (self.bg[h] - self.bg[l])
        u = np.empty(br.size)
        for k in np.unique(h):
            sel = h == k
            u[sel] = self.u_of_alpha(k - 1, k, t[sel], alpha[rest][sel])
        # window interpolated from the fine grid of beta_cdf (smooth within
        # each bin; grid steps < 1e-3 of a bin)
        xg, wg = self._wgrid
        ulo = np.interp(br, xg, wg[3])
        uhi = np.interp(br, xg, wg[4])
        v[rest] = (u - ulo) / (uhi - ulo)
        return v


def _ks_pair(proc, a, b):
    xb, cdf = proc.beta_cdf()  # also tabulates the percentile window
    d_beta = stats.kstest(b, lambda v: np.interp(v, xb, cdf)).statistic
    d_alpha = stats.kstest(proc.conditional_u(a, b), "uniform").statistic
    return d_beta, d_alpha


N_ALG1 = 1000000
KS_LIM = 1.95 / math.sqrt(N_ALG1)  # about 0.1% significance for each of the two tests


@pytest.fixture(scope="module")
def alg1_table():
    ag, bg, sab = _table()
    T = 293.15
    kT = KB * T
    # egrid[5] = 5.5 kT: -5.5 lies inside the beta bin (-6.46, -4.62)
    egrid = 5.5 * kT * 1.6 ** (np.arange(12) - 5.0)
    return ag, bg, sab, T, kT, egrid


# Two neutron energies: "first bin" = 0.9 egrid[5] (sampler at egrid[5];
# -E/kT = -4.95 inside the sampler's first, kinematic, beta bin, so the
# first-bin rule shapes accepted draws); "margin" = egrid[5]/1.03 (the 1.05
# margin moves the sampler from egrid[5] to egrid[6]).
CASES = {"first bin": (5, 0.9), "margin": (5, 1 / 1.03)}


def _energy(table, case):
    i, f = CASES[case]
    return table[5][i] * f


def _sample(oracle, table, E, variant=0, seed=5):
    ag, bg, sab, T, kT, egrid = table
    with oracle.alg1_variant(variant):
        o = oracle.sab(ag, bg, sab, T, 26.98, 1.0, egrid=egrid)
        a, b, nerr = o.sample_ab(np.full(N_ALG1, E), seed=seed, nthreads=8)
    assert nerr == 0
    return a, b


@pytest.mark.parametrize("case", list(CASES))
def test_alg1_follows_procedure_density(oracle, alg1_table, case):
    ag, bg, sab, T, kT, egrid = alg1_table
    E = _energy(alg1_table, case)
    proc = Procedure(ag, bg, sab, kT, egrid, E)
    if case == "margin":
        assert proc.cs > E / kT * 1.5  # the margin moved the sampler
    a, b = _sample(oracle, alg1_table, E)
    lo, hi = _alpha_lims(E / kT, b)
    assert np.all((b >= -E / kT) & (a >= lo * (1 - 1e-12)) & (a <= hi * (1 + 1e-12)))
    if case == "first bin":
        assert np.mean(b <= proc.b1) > 0.002  # the first bin is populated
    d_beta, d_alpha = _ks_pair(proc, a, b)
    print(case, "Alg1 vs procedure density: KS beta", d_beta, "KS alpha|beta", d_alpha, "limit", KS_LIM)
    assert d_beta < KS_LIM and d_alpha < KS_LIM


@pytest.mark.parametrize("variant,case", [(1, "first bin"), (2, "margin"), (3, "first bin"), (3, "margin")])
def test_alg1_pins_detect_planted_changes(oracle, alg1_table, variant, case):
    ag, bg, sab, T, kT, egrid = alg1_table
    E = _energy(alg1_table, case)
    proc = Procedure(ag, bg, sab, kT, egrid, E)
    a, b = _sample(oracle, alg1_table, E, variant=variant)
    d_beta, d_alpha = _ks_pair(proc, a, b)
    print("planted change", variant, case, "KS beta", d_beta, "KS alpha|beta", d_alpha)
    assert max(d_beta, d_alpha) > 3 * KS_LIM
