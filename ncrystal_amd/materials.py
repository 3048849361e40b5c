"""Material assembly: from a synthetic crystal description to a process
composition on the GPU, following the component wiring of NCrystal's
standard scatter factory (ncrystal_core/src/stdscatfactory/NCStdScatFact.cc:
137-146 incoherent elastic first, :148-185 Bragg, :263-330 inelastic; the
composition keeps that order).
"""
import numpy as np

from . import core


def species_msd(crystal):
    """Debye-model MSD per species (ncrystal_core/src/phys_utils/NCDebyeMSD.cc)."""
    return [core.debye_msd(s.debye_temp, crystal.temperature, s.mass) for s in crystal.species]


def hkl_list(crystal, msd=None):
    msd = species_msd(crystal) if msd is None else msd
    sp = [(s.csl, m, s.positions) for s, m in zip(crystal.species, msd)]
    return core.hkl_orthogonal(crystal.a, crystal.b, crystal.c, sp, crystal.dcutoff)


def powderbragg_inputs(crystal, msd=None):
    hkl, mult, dsp, fsq = hkl_list(crystal, msd)
    return crystal.volume * crystal.n_atoms, dsp, fsq * mult


def elinc_inputs(crystal, msd=None):
    """(msd, bound incoherent xs, scale) per species; scale = fraction of atoms
    (ncrystal_core/src/elincscatter/NCElIncScatter.cc:70-97)."""
    msd = species_msd(crystal) if msd is None else msd
    ntot = crystal.n_atoms
    scale = [float(s.natoms) * 1.0 / ntot for s in crystal.species]
    return np.asarray(msd), np.asarray([s.sigma_inc for s in crystal.species]), np.asarray(scale)


class Material:
    """Scatter of a one-phase material: ElInc, PowderBragg, then the inelastic
    components.  sab: one S(alpha,beta) table (dict of alpha, beta, sab,
    temperature, mass, sigma_bound, optional egrid) at scale 1.  dynamics:
    the inelastic components per element instead, as the standard scatter
    factory adds them (NCStdScatFact.cc:256-300): a list of (fraction, entry),
    entry an S(alpha,beta) table dict, {"freegas": (mass_amu, sigma_free)} or
    {"vdos": (egrid, density) or Debye temperature, "mass": amu,
    "sigma_bound": barn, "vdoslux": 0..5} at the crystal's temperature, each a
    component scaled by the element's number fraction."""

    def __init__(self, crystal, sab=None, with_elinc=True, with_bragg=True, fp32_tables=True, dynamics=None):
        if sab is not None and dynamics is not None:
            raise ValueError("give either sab or dynamics")
        self.crystal = crystal
        msd = species_msd(crystal)
        self.msd = msd
        comps = []
        self.components = {}
        if with_elinc:
            m, b, s = elinc_inputs(crystal, msd)
            ei = core.create_elincscatter(m, b, s)
            comps.append(ei)
            self.components["elinc"] = ei
        if with_bragg:
            v0n, dsp, fm = powderbragg_inputs(crystal, msd)
            self.pb_inputs = (v0n, dsp, fm)
            pb = core.create_powderbragg(v0n, dsp, fm, fp32_tables=fp32_tables)
            comps.append(pb)
            self.components["bragg"] = pb
        scales = [1.0] * len(comps)
        if sab is not None:
            dynamics = [(1.0, sab)]
        self.inelastic = []
        for frac, d in dynamics or []:
            if "freegas" in d:
                mass, sigma = d["freegas"]
                c = core.create_freegas(crystal.temperature, mass, sigma)
            elif "vdos" in d and "sab" not in d:
                # an S(alpha,beta) expanded from the element's VDOS on the GPU
                # ((egrid, density) or a Debye temperature)
                c = core.create_sabscatter_from_vdos(d["vdos"], crystal.temperature, d["mass"], d["sigma_bound"],
                                                     vdoslux=d.get("vdoslux", 3))
            else:
                c = core.create_sabscatter(d["alpha"], d["beta"], d["sab"], d["temperature"], d["mass"],
                                           d["sigma_bound"], egrid=d.get("egrid"))
            comps.append(c)
            scales.append(float(frac))
            self.inelastic.append(c)
        if sab is not None:
            self.components["inelas"] = self.inelastic[0]
        self.scatter = core.create_composition(comps, scales)


def _stable_sum(values):
    # Neumaier summation (StableSum, ncrystal_core/include/NCrystal/internal/utils/NCMath.hh:194)
    s = c = 0.0
    for x in values:
        t = s + x
        c += (s - t) + x if abs(s) >= abs(x) else (x - t) + s
        s = t
    return s + c


def phase_weights(phases):
    """Scales of the phases of a multi-phase material: volume fraction times
    number density, normalised by their sum, in ascending order of weight
    (ncrystal_core/src/factories/NCFactImpl.cc:500-540; phases of equal
    weight keep the given order).  phases: [(volume fraction, Material)].
    Returns [(weight, Material)]."""
    w = [frac * m.crystal.n_atoms / m.crystal.volume for frac, m in phases]
    tot = _stable_sum(w)
    if not tot > 0.0:
        raise ValueError("phases have no atoms")
    out = [(x / tot, m) for x, (_, m) in zip(w, phases)]
    return sorted(out, key=lambda e: e[0])


class MultiPhase:
    """Scatter of a multi-phase material: the composition of the phases'
    scatters with their weights (NCFactImpl.cc:615-642).  The composition
    flattens the phases and merges their PowderBragg and ElInc components."""

    def __init__(self, phases):
        self.phases = phase_weights(phases)
        self.scatter = core.create_composition([m.scatter for _, m in self.phases],
                                               [w for w, _ in self.phases])
