"""MiniMC restatement for the checker (test infrastructure only: imported by
tests/, never by the library).

NCrystal's standard MiniMC engine (ncrystal_core/src/minimc/
NCMMC_SimEngine.cc:170-449) for a beam through a sphere, vectorised over all
neutrons of a step with numpy and the checker's processes (pyoracle.Proc
cross sections and scattering with directions):

* source: "constant" (all neutrons at one point) or "circular" (uniform on a
  disk normal to the beam) with one energy and direction
  (NCMMC_Source.cc:643-840); neutrons are moved to the sphere's surface
  (NCMMC_Sphere.hh distToVolumeEntry) and those that miss it are tallied
  unscattered unless ignored (NCMMC_SimMgr.cc:106-130);
* each step (NCMMC_SimEngine.cc:216-449): distance D to the surface,
  macroscopic cross sections 100 * numdens * sigma (NCMMC_Utils.hh:103),
  transmission P = exp(-Sigma_s D); the transmitted part w exp(-Sigma_a D) P
  is tallied; the rest scatters at a distance drawn from exp(-Sigma_s x)
  truncated to [0, D] (NCMMC_Utils.cc:152-173) with weight
  w exp(-Sigma_a x) (1 - P), after Russian roulette of neutrons with at least
  nscat_threshold scatterings and weight below weight_threshold (survival
  probability p, survivors' weight / p);
* tallies (NCMMC_StdTallies.cc:223-330, 328-434, 435-520, 590-670): theta =
  acos(u . u0) in degrees, mu, nscat (-1 for neutrons missing the volume),
  e, l = wavelength, de = E0 - E and q = |k_f - k_i|, weighted; w and
  nscat_uw unweighted (one entry per neutron and step); with the breakdown
  by scattering history (none, one elastic, one inelastic, several all
  elastic, several otherwise; a scattering is elastic when the energy is
  unchanged).

The random numbers come from numpy's PCG64, so a run agrees with the
reference's and the library's only statistically.
"""
import numpy as np

TALLIES = ("theta", "mu", "nscat", "w", "nscat_uw", "e", "l", "de", "q")
_WL2EKIN = 0.081804209605330899
_EKIN2KSQ = 4.0 * np.pi ** 2 / _WL2EKIN
CATEGORIES = ("NOSCAT", "SINGLESCAT_ELAS", "SINGLESCAT_INELAS", "MULTISCAT_PUREELAS", "MULTISCAT_OTHER")


def _sphere_entry(r2, p, u):
    b = np.einsum("ij,ij->i", p, u)
    c = np.einsum("ij,ij->i", p, p) - r2
    out = np.full(len(p), -1.0)
    inside = c <= 0.0
    out[inside] = np.where(np.minimum(c[inside], b[inside]) < 0.0, 0.0, -1.0)
    o = ~inside
    disc = b[o] * b[o] - c[o]
    t = np.full(o.sum(), -1.0)
    ok = disc >= 0.0
    t[ok] = -b[o][ok] - np.sqrt(disc[ok])
    t[t < 0.0] = -1.0
    out[o] = t
    return out


def _sphere_exit(r2, p, u):
    b = np.einsum("ij,ij->i", p, u)
    c = np.einsum("ij,ij->i", p, p) - r2
    d = -b + np.sqrt(np.maximum(b * b - c, 0.0))
    return np.maximum(d, 0.0)


class _Hist:
    def __init__(self, nbins, xmin, xmax, breakdown):
        self.nbins, self.xmin, self.xmax = nbins, xmin, xmax
        self.h = np.zeros((6, 2, nbins + 2))
        self.breakdown = breakdown

    def fill(self, v, w, cat):
        # NCHists.hh valueToBin with under/overflow bins (the upper edge goes
        # to the last bin); weighted fills skip weights that are not positive
        keep = w > 0.0
        v, w, cat = v[keep], w[keep], cat[keep]
        f = (v - self.xmin) * (self.nbins / (self.xmax - self.xmin))
        inner = 1 + np.minimum(np.floor(np.where(np.isfinite(f), f, 0.0)), self.nbins).astype(np.int64)
        s = np.where(~(v >= self.xmin), 0,
                     np.where(v >= self.xmax, np.where(v == self.xmax, self.nbins, self.nbins + 1), inner))
        s = s.astype(np.int64)
        n = self.nbins + 2
        self.h[0, 0] += np.bincount(s, weights=w, minlength=n)
        self.h[0, 1] += np.bincount(s, weights=w * w, minlength=n)
        if self.breakdown:
            for k in range(5):
                m = cat == k
                if m.any():
                    self.h[1 + k, 0] += np.bincount(s[m], weights=w[m], minlength=n)
                    self.h[1 + k, 1] += np.bincount(s[m], weights=w[m] * w[m], minlength=n)


def _value(tally, u, u0, w, nscat, E, E0):
    mu = np.clip(u @ u0, -1.0, 1.0)
    if tally == "theta":
        return np.degrees(np.arccos(mu))
    if tally == "mu":
        return mu
    if tally in ("nscat", "nscat_uw"):
        return nscat.astype(np.float64)
    if tally == "e":
        return E
    if tally == "l":
        return np.sqrt(_WL2EKIN / np.maximum(E, 5e-324))
    if tally == "de":
        return E0 - E
    if tally == "q":
        return np.sqrt(np.maximum((E + E0 - 2.0 * mu * np.sqrt(E * E0)) * _EKIN2KSQ, 0.0))
    return w


def _fill_weight(tally, w):
    return np.ones_like(w) if tally in ("w", "nscat_uw") else w


def _category(nscat, ninel):
    return np.where(nscat > 1, np.where(ninel > 0, 4, 3), np.where(nscat == 1, np.where(ninel > 0, 2, 1), 0))


def minimc_sphere(scatter, absorption, numdens, radius, ekin, n, source="constant", pos=(0.0, 0.0, 0.0),
                  direction=(0.0, 0.0, 1.0), beam_radius=0.0, tally="theta", nbins=90, xrange=(0.0, 180.0),
                  breakdown=True, roulette=(0.1, 1e-2, 2), nscatlimit=None, ignoremiss=False,
                  include_absorption=True, seed=0, nthreads=8, sample_seed=1):
    """Same arguments and result layout as ncrystal_amd.core.minimc_sphere."""
    rng = np.random.default_rng(seed)
    macro = 100.0 * numdens
    r2 = radius * radius
    u0 = np.asarray(direction, dtype=np.float64)
    u0 = u0 / np.linalg.norm(u0)
    hist = _Hist(nbins, xrange[0], xrange[1], breakdown)
    p = np.tile(np.asarray(pos, dtype=np.float64), (n, 1))
    if source == "circular" and beam_radius > 0.0:
        a = np.array([1.0, 0.0, 0.0])
        if a @ u0 > 0.8:
            a = np.array([0.0, 1.0, 0.0])
        if a @ u0 > 0.8:
            a = np.array([0.0, 0.0, 1.0])
        a = np.cross(u0, a)
        a /= np.linalg.norm(a)
        b = np.cross(u0, a)
        b /= np.linalg.norm(b)
        rad = beam_radius * np.sqrt(rng.random(n))
        phi = 2.0 * np.pi * rng.random(n)
        p += (rad * np.cos(phi))[:, None] * a + (rad * np.sin(phi))[:, None] * b
    u = np.tile(u0, (n, 1))
    d = _sphere_entry(r2, p, u)
    miss = d < 0.0
    if miss.any() and not ignoremiss:
        zero = np.zeros(miss.sum(), dtype=np.int64)
        w1 = np.ones(miss.sum())
        hist.fill(_value(tally, u[miss], u0, w1, zero - 1, np.full(len(w1), float(ekin)), float(ekin)), w1, zero)
    keep = ~miss
    p = p[keep] + d[keep, None] * u[keep]
    u = u[keep]
    m = len(p)
    E = np.full(m, float(ekin))
    w = np.ones(m)
    nscat = np.zeros(m, dtype=np.int64)
    ninel = np.zeros(m, dtype=np.int64)
    step = 0
    while m > 0:
        D = _sphere_exit(r2, p, u)
        if scatter is not None and macro > 0.0:
            ss = macro * scatter.xs_dir(E, u, nthreads=nthreads)
        else:
            ss = np.zeros(m)
        if nscatlimit is not None:
            ss[nscat >= nscatlimit] = 0.0
        if absorption is not None and include_absorption and macro > 0.0:
            sa = macro * absorption.xs(E, nthreads=nthreads)
        else:
            sa = np.zeros(m)
        ptr = np.where(ss > 0.0, np.exp(-ss * D), 1.0)
        wt = w * np.exp(-sa * D) * ptr
        hist.fill(_value(tally, u, u0, wt, nscat, E, float(ekin)), _fill_weight(tally, wt), _category(nscat, ninel))
        go = (w > 0.0) & (ss > 0.0)
        u_roul = rng.random(m)
        u_dist = 1.0 - rng.random(m)  # (0, 1]
        rf = np.ones(m)
        roul = go & (nscat >= roulette[2]) & (w < roulette[1])
        dead = roul & (u_roul > roulette[0])
        rf[roul & ~dead] = 1.0 / roulette[0]
        go &= ~dead
        with np.errstate(divide="ignore", invalid="ignore"):
            x = -np.log1p(u_dist * np.expm1(-ss * D)) / ss
        x = np.clip(np.where(go, x, 0.0), 0.0, D)
        ws = w * rf * np.exp(-sa * x) * (1.0 - ptr)
        go &= ws > 0.0
        if not go.any():
            break
        p = p[go] + x[go, None] * u[go]
        E0 = E[go]
        E1, u1, nerr = scatter.sample_dir(E0, np.ascontiguousarray(u[go]), seed=sample_seed, stream=step,
                                          nthreads=nthreads)
        if nerr:
            raise RuntimeError("checker sampling reported %d errors" % nerr)
        w = ws[go]
        nscat = nscat[go] + 1
        ninel = ninel[go] + (E1 != E0)
        E, u = np.asarray(E1), np.asarray(u1).reshape(-1, 3)
        m = len(E)
        step += 1

    def h(k):
        return dict(content=hist.h[k, 0, 1:-1].copy(), errorsq=hist.h[k, 1, 1:-1].copy(),
                    underflow=hist.h[k, 0, 0], underflow_errorsq=hist.h[k, 1, 0], overflow=hist.h[k, 0, -1],
                    overflow_errorsq=hist.h[k, 1, -1])
    out = {"total": h(0), "info": dict(max_steps=step)}
    if breakdown:
        for i, name in enumerate(CATEGORIES):
            out[name] = h(1 + i)
    return out
