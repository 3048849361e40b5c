"""Synthetic materials and neutron batches for the benchmark and test workloads.

Everything here is input preparation (host-side numpy): crystal cells, atom
data, Debye temperatures, a three-Gaussian vibrational density of states and
its phonon expansion into an S(alpha,beta) table, and seeded neutron energies
and directions.  None of it is on the per-neutron path.

Physical conventions follow NCrystal (ncrystal_core/include/NCrystal/core/
NCDefs.hh:79-128): energies in eV, lengths in Angstrom, cross sections in barn,
coherent scattering lengths in sqrt(barn) (= 10 fm), alpha/beta dimensionless
with alpha = hbar^2 q^2 / (2 m_neutron kT) and beta = (E_final - E)/kT
(ncrystal_core/include/NCrystal/internal/phys_utils/NCKinUtils.hh:108-121).
"""
from dataclasses import dataclass, field

import numpy as np

K_BOLTZMANN = 8.6173303e-5
K_NEUTRON_MASS_AMU = 1.00866491588
WL2EKIN = 0.081804209605330899


def wl2ekin(wl):
    wl = np.asarray(wl, dtype=np.float64)
    return WL2EKIN / (wl * wl)


def ekin2wl(e):
    return np.sqrt(WL2EKIN / np.asarray(e, dtype=np.float64))


@dataclass
class Species:
    name: str
    csl: float            # coherent scattering length [sqrt(barn)]
    sigma_inc: float      # bound incoherent xs [barn]
    sigma_abs: float      # absorption xs at 2200 m/s [barn]
    mass: float           # [amu]
    debye_temp: float     # [K]
    positions: np.ndarray  # (n,3) fractional, mapped to [0,1) and sorted

    @property
    def natoms(self):
        return len(self.positions)


@dataclass
class Crystal:
    a: float
    b: float
    c: float
    species: list
    temperature: float
    dcutoff: float
    name: str = "crystal"
    extra: dict = field(default_factory=dict)

    @property
    def volume(self):
        # ncrystal_core/src/infobld/NCInfoBuilder.cc:237-244 (diagonal lattice)
        return (self.a * self.b) * self.c

    @property
    def n_atoms(self):
        return sum(s.natoms for s in self.species)


def _norm_positions(pos):
    """Map to [0,1) and sort lexicographically like the reference's
    InfoBuilder (ncrystal_core/src/infobld/NCInfoBuilder.cc:118-123,180-192)."""
    p = np.asarray(pos, dtype=np.float64).reshape(-1, 3).copy()
    p = np.where(p < 0.0, p + 1.0, p)
    p = np.where(p >= 1.0, p - 1.0, p)
    order = sorted(range(len(p)), key=lambda i: (p[i, 0], p[i, 1], p[i, 2]))
    return p[order]


def fcc_powder(a=4.05, mass=26.98, csl=0.3449, sigma_inc=0.0082, sigma_abs=0.231,
               debye_temp=410.0, temperature=293.0, dcutoff=0.5):
    """fcc single-species powder (a=4.05 Aa, 4 atoms/cell, b_coh 3.449 fm,
    sigma_inc 0.0082 b, sigma_abs 0.231 b, T_Debye 410 K, T 293 K)."""
    pos = _norm_positions([(0, .5, .5), (0, 0, 0), (.5, .5, 0), (.5, 0, .5)])
    sp = Species("Al", csl, sigma_inc, sigma_abs, mass, debye_temp, pos)
    return Crystal(a, a, a, [sp], temperature, dcutoff, name="fcc_a4.05")


def ortho_random(seed=1, dcutoff=0.1, temperature=293.0, nspecies=3, atoms_per_species=8):
    """Orthorhombic cell a,b,c ~ U(5,15) Aa with 3 species x 8 atoms at
    seeded fractional positions; b_coh ~ U(2,8) fm; T_Debye ~ U(200,600) K.
    Masses (not fixed by the workload definition) ~ U(20,60) amu, sigma_inc ~
    U(0,2) b, sigma_abs ~ U(0,1) b."""
    rng = np.random.default_rng(seed)
    a, b, c = rng.uniform(5.0, 15.0, size=3)
    species = []
    for s in range(nspecies):
        csl = rng.uniform(2.0, 8.0) / 10.0
        pos = _norm_positions(rng.uniform(0.0, 1.0, size=(atoms_per_species, 3)))
        species.append(Species(f"X{s}", csl, rng.uniform(0.0, 2.0), rng.uniform(0.0, 1.0),
                               rng.uniform(20.0, 60.0), rng.uniform(200.0, 600.0), pos))
    return Crystal(float(a), float(b), float(c), species, temperature, dcutoff, name=f"ortho_seed{seed}")


# ----------------------------------------------------------------------------
# S(alpha,beta) from a vibrational density of states (incoherent Gaussian
# approximation, phonon expansion; Sjolander 1958).  Standard textbook method,
# written independently of the reference's VDOS expansion code.

def vdos_3gauss(seed=1, emax=0.15, npts=500):
    rng = np.random.default_rng(seed)
    centres = rng.uniform(0.01, 0.12, size=3)
    widths = rng.uniform(0.003, 0.02, size=3)
    amps = rng.uniform(0.5, 1.5, size=3)
    e = np.linspace(0.0, emax, npts)
    rho = np.zeros_like(e)
    for c0, w, A in zip(centres, widths, amps):
        rho += A * np.exp(-0.5 * ((e - c0) / w) ** 2)
    rho[0] = 0.0
    # Debye-like low-energy behaviour rho ~ E^2 below the first grid point region
    rho *= np.clip(e / 0.004, 0.0, 1.0) ** 2
    return e, rho


def phonon_expand_sab(vdos_e, vdos_rho, temperature, mass_amu, alpha, beta, nmax=None):
    """S(alpha,beta) (non-symmetric, S(a,-b) = e^b S(a,b) detailed balance)
    for an incoherent harmonic solid with the given VDOS, excluding the
    elastic (n=0) term.  alpha uses the neutron mass (NCrystal convention);
    the phonon expansion uses alpha/A with A = M/m_n."""
    kT = K_BOLTZMANN * temperature
    A = mass_amu / K_NEUTRON_MASS_AMU
    alpha = np.asarray(alpha, dtype=np.float64)
    beta = np.asarray(beta, dtype=np.float64)
    amax = alpha[-1] / A
    # Debye-Waller exponent lam from the VDOS (needed for the number of terms)
    eb = vdos_e / kT
    bpos_grid = np.linspace(1e-6, eb[-1], 4000)
    rho_pos = np.interp(bpos_grid, eb, vdos_rho)
    norm = np.trapezoid(rho_pos, bpos_grid)
    lam_est = np.trapezoid(rho_pos / norm / (bpos_grid * np.tanh(bpos_grid / 2.0)), bpos_grid)
    if nmax is None:
        x = amax * lam_est
        nmax = int(min(600, max(30, x + 12.0 * np.sqrt(x + 1.0) + 10)))
    # fine uniform beta grid; the margin keeps the truncated linear convolutions
    # exact inside the window that is finally used
    db = 0.02
    bneed = max(abs(beta[0]), abs(beta[-1])) + 2.0
    bmax_fine = bneed + 60.0
    nfine = int(np.ceil(bmax_fine / db))
    bf = np.arange(-nfine, nfine + 1) * db
    rho_b = np.interp(np.abs(bf), eb, vdos_rho, right=0.0) / norm
    T1 = np.zeros_like(bf)
    m = (rho_b > 0.0) & (bf != 0.0)
    # P(b) e^{-b/2} with P = rho/(2 b sinh(b/2)); e^{-b/2}/sinh(b/2) = 2/(e^b - 1)
    T1[m] = rho_b[m] / (bf[m] * np.expm1(bf[m]))
    i0 = nfine
    T1[i0] = 0.5 * (T1[i0 - 1] + T1[i0 + 1])
    lam = np.trapezoid(T1, bf)
    T1 = T1 / lam
    L = len(bf)
    nfft = 1 << int(np.ceil(np.log2(2 * L)))
    F1 = np.fft.rfft(T1 * db, nfft)
    Tn = T1.copy()
    Tlist = [Tn]
    for n in range(2, nmax + 1):
        full = np.fft.irfft(np.fft.rfft(Tn, nfft) * F1, nfft)
        Tn = np.maximum(full[nfine:nfine + L], 0.0)
        Tlist.append(Tn)
    Tarr = np.array(Tlist)  # (nmax, L)
    # interpolate each T_n onto the requested beta grid
    Tb = np.array([np.interp(beta, bf, t) for t in Tarr])  # (nmax, nb)
    sab = np.zeros((beta.size, alpha.size))
    from math import lgamma
    for ia, al in enumerate(alpha):
        x = al * lam / A
        ln = np.array([n * np.log(x) - lgamma(n + 1) for n in range(1, nmax + 1)]) - x
        w = np.exp(ln)
        sab[:, ia] = w @ Tb
    sab[sab < 1e-300] = 0.0
    return sab


def sab_table_c3(seed=1, temperature=293.0, mass_amu=26.98, nalpha=200, nbeta=800):
    """200 x 800 S(alpha,beta) table from the 3-Gaussian VDOS (500 points up
    to 0.15 eV) at T=293 K."""
    e, rho = vdos_3gauss(seed)
    kT = K_BOLTZMANN * temperature
    emax_cover = 1.2  # eV, kinematic coverage of the alpha/beta grids
    amax = 4.6 * emax_cover / kT
    alpha = np.geomspace(1e-3, amax, nalpha)
    bneg = -(emax_cover / kT) * 1.02
    bpos = 25.0
    u = np.linspace(-1.0, 1.0, nbeta)
    # denser near beta=0
    beta = np.where(u < 0, -np.abs(bneg) * np.abs(u) ** 1.5, bpos * np.abs(u) ** 1.5)
    beta = np.unique(beta)
    assert beta.size == nbeta
    sab = phonon_expand_sab(e, rho, temperature, mass_amu, alpha, beta)
    return dict(alpha=alpha, beta=beta, sab=sab, temperature=temperature, mass=mass_amu,
                sigma_bound=1.5, vdos=(e, rho))


def vdos_debye(debye_temp, npts=400):
    """Debye VDOS: rho ~ E^2 up to k_B T_Debye, zero above."""
    ed = K_BOLTZMANN * debye_temp
    e = np.linspace(0.0, ed, npts)
    return e, e * e


def sab_table_c1(temperature=293.0, debye_temp=410.0, mass_amu=26.98, sigma_bound=None,
                 nalpha=100, nbeta=240, emax_cover=0.6):
    """S(alpha,beta) of the fcc powder's Debye-model inelastic scattering:
    Debye VDOS to k_B 410 K expanded at 293 K on an alpha/beta grid covering
    energies to `emax_cover` eV (the free-gas extender takes over above the
    table's energy grid).  sigma_bound defaults to the fcc material's total
    bound scattering xs 4 pi b_coh^2 + sigma_inc (b_coh 3.449 fm, 0.0082 b)."""
    if sigma_bound is None:
        sigma_bound = 4 * np.pi * 0.3449 ** 2 + 0.0082
    e, rho = vdos_debye(debye_temp)
    kT = K_BOLTZMANN * temperature
    amax = 4.6 * emax_cover / kT
    alpha = np.geomspace(1e-3, amax, nalpha)
    bneg = -(emax_cover / kT) * 1.02
    bpos = 15.0
    u = np.linspace(-1.0, 1.0, nbeta)
    beta = np.unique(np.where(u < 0, -np.abs(bneg) * np.abs(u) ** 1.5, bpos * np.abs(u) ** 1.5))
    sab = phonon_expand_sab(e, rho, temperature, mass_amu, alpha, beta)
    return dict(alpha=alpha, beta=beta, sab=sab, temperature=temperature, mass=mass_amu,
                sigma_bound=float(sigma_bound), vdos=(e, rho))


def small_sab_table(seed=3, temperature=293.0, mass_amu=12.0, nalpha=40, nbeta=60, amax=40.0,
                    bmin=-20.0, bmax=10.0):
    """Small table for fast tests (high-E extender path is reached above Emax)."""
    e, rho = vdos_3gauss(seed)
    alpha = np.geomspace(2e-3, amax, nalpha)
    beta = np.linspace(bmin, bmax, nbeta)
    sab = phonon_expand_sab(e, rho, temperature, mass_amu, alpha, beta)
    return dict(alpha=alpha, beta=beta, sab=sab, temperature=temperature, mass=mass_amu,
                sigma_bound=4.0)


# ----------------------------------------------------------------------------
# neutron batches

def energies_logu(n, emin, emax, seed=1):
    rng = np.random.default_rng(seed)
    return np.exp(rng.uniform(np.log(emin), np.log(emax), size=n))


def isotropic_dirs(n, seed=2):
    rng = np.random.default_rng(seed)
    uz = rng.uniform(-1.0, 1.0, size=n)
    ph = rng.uniform(0.0, 2 * np.pi, size=n)
    s = np.sqrt(1.0 - uz * uz)
    return np.stack([s * np.cos(ph), s * np.sin(ph), uz], axis=1)


def cone_dirs(n, axis, half_angle_rad, seed=3):
    rng = np.random.default_rng(seed)
    axis = np.asarray(axis, dtype=np.float64)
    axis = axis / np.linalg.norm(axis)
    cmin = np.cos(half_angle_rad)
    ct = rng.uniform(cmin, 1.0, size=n)
    ph = rng.uniform(0.0, 2 * np.pi, size=n)
    st = np.sqrt(1.0 - ct * ct)
    tmp = np.array([1.0, 0.0, 0.0]) if abs(axis[0]) < 0.9 else np.array([0.0, 1.0, 0.0])
    e1 = np.cross(axis, tmp)
    e1 /= np.linalg.norm(e1)
    e2 = np.cross(axis, e1)
    return (ct[:, None] * axis[None, :] + (st * np.cos(ph))[:, None] * e1[None, :]
            + (st * np.sin(ph))[:, None] * e2[None, :])


def random_rotation(seed=4):
    """Uniformly random proper rotation (3x3, det +1) from a seeded Gaussian
    matrix (QR with sign fix)."""
    rng = np.random.default_rng(seed)
    q, r = np.linalg.qr(rng.normal(size=(3, 3)))
    q = q * np.sign(np.diag(r))[None, :]
    if np.linalg.det(q) < 0:
        q[:, 0] = -q[:, 0]
    return q


def random_unit_vector(seed=5):
    rng = np.random.default_rng(seed)
    v = rng.normal(size=3)
    return v / np.linalg.norm(v)


def wavelengths_uniform(n, wl_lo, wl_hi, seed=6):
    rng = np.random.default_rng(seed)
    return rng.uniform(wl_lo, wl_hi, size=n)



def freegas_like_sab(nalpha, nbeta, mass_amu, temperature=293.0, alpha_max=None, beta_min=None, beta_max=None):
    """S(alpha,beta) of a free gas of mass mass_amu (non-symmetric form,
    S(a,-b) = e^b S(a,b)): sqrt(A/(4 pi a)) exp(-(A b + a)^2 / (4 A a)),
    on a log-spaced alpha grid and a beta grid dense around 0."""
    A = mass_amu / 1.00866491588
    amax = alpha_max if alpha_max is not None else 40.0 * A
    alpha = np.geomspace(1e-3, amax, nalpha)
    bmin = beta_min if beta_min is not None else -40.0
    bmax = beta_max if beta_max is not None else 25.0
    u = np.linspace(-1.0, 1.0, nbeta)
    beta = np.where(u < 0, -bmin * u ** 3, bmax * u ** 3)
    beta = np.sort(beta)
    a = alpha[None, :]
    b = beta[:, None]
    sab = np.sqrt(A / (4 * np.pi * a)) * np.exp(-(A * b + a) ** 2 / (4 * A * a))
    sab = np.maximum(sab, 1e-300)
    return dict(alpha=alpha, beta=beta, sab=sab, temperature=temperature, mass=mass_amu,
                sigma_bound=float(4.0 * (A + 1) ** 2 / A ** 2))


def batch_materials(nmat=512, seed=11, planes=(1e2, 1e5), grids=((50, 100), (300, 1200)), counts=(1e4, 1e6)):
    """Inputs of the many-materials workload: per material a synthetic plane
    list (N_planes log-uniform in `planes`, d with density ~ d^-4 in
    [0.25, 8] Aa, |F|^2 x multiplicity with a Debye-Waller fall-off), one to
    three incoherent species, a free-gas-like S(alpha,beta) table with a grid
    between grids[0] and grids[1] (log-uniform per axis) and a neutron count
    log-uniform in `counts`."""
    rng = np.random.default_rng(seed)
    mats = []
    for i in range(nmat):
        npl = int(round(np.exp(rng.uniform(np.log(planes[0]), np.log(planes[1])))))
        u = rng.uniform(size=npl)
        dmin, dmax = 0.25, 8.0
        d = 1.0 / ((dmin ** -3 - u * (dmin ** -3 - dmax ** -3)) ** (1.0 / 3.0))
        d = np.sort(d)[::-1].copy()
        msd = rng.uniform(0.005, 0.02)
        mult = 2 * rng.integers(1, 25, size=npl)
        fsq = rng.uniform(0.01, 5.0, size=npl) * np.exp(-msd * (2 * np.pi / d) ** 2)
        v0n = rng.uniform(50.0, 2000.0) * rng.integers(1, 25)
        ns = int(rng.integers(1, 4))
        ei = dict(msd=rng.uniform(0.005, 0.03, size=ns), bixs=rng.uniform(0.0, 5.0, size=ns),
                  scale=rng.dirichlet(np.ones(ns)))
        na = int(round(np.exp(rng.uniform(np.log(grids[0][0]), np.log(grids[1][0])))))
        nb = int(round(np.exp(rng.uniform(np.log(grids[0][1]), np.log(grids[1][1])))))
        sab = freegas_like_sab(na, nb, mass_amu=float(rng.uniform(5.0, 200.0)))
        n = int(round(np.exp(rng.uniform(np.log(counts[0]), np.log(counts[1])))))
        mats.append(dict(pb=(v0n, d, fsq * mult), elinc=ei, sab=sab, count=n))
    return mats
