"""Compounds: one inelastic component per element, as NCrystal's standard
scatter factory adds them (ncrystal_core/src/stdscatfactory/NCStdScatFact.cc:
256-300: an S(alpha,beta) scatter per element with scattering-kernel
dynamics, a FreeGas scatter per element with free-gas dynamics, each scaled
by the element's number fraction), beside ElInc and PowderBragg, through the
C ABI on the GPU against the CPU checker's composition of the same inputs
(ProcComposition, NCProcImpl.cc:340-389): cross sections to 1e-12 and
samples on the same Philox streams (the pick, then each staged component's
sampler continuing the stream), isotropic and with directions, on two calls
in a row.  Also: multi-phase materials whose phases each hold an
S(alpha,beta) component, and the limits (four staged components per
composition; one staged component per material in many-material batches)."""
import numpy as np
import pytest

from ncrystal_amd import synth
from conftest import assert_streams_match

pytestmark = pytest.mark.gpu

T = 293.0
# element tables: (mass [amu], table seed, alpha max)
ELEMENTS = [(26.98, 3, 40.0), (15.999, 5, 60.0), (1.008, 7, 200.0)]
HELIUM = (4.0026, 0.7596)


@pytest.fixture(scope="module")
def amd():
    from ncrystal_amd import _capi, core, materials
    if _capi.lib().ncamd_device_count() < 1:
        pytest.fail("no GPU visible to libncrystal_amd")
    return _capi, core, materials


def _rel(a, b):
    den = np.maximum(np.abs(a), np.abs(b))
    with np.errstate(invalid="ignore", divide="ignore"):
        r = np.where(den > 0, np.abs(a - b) / np.where(den > 0, den, 1.0), 0.0)
    return np.where(a == b, 0.0, r)


def _table(mass, seed, amax):
    return synth.small_sab_table(seed=seed, temperature=T, mass_amu=mass, amax=amax)


def _checker(oracle, materials, m, dynamics):
    parts, scales = [], []
    mm, bb, ss = materials.elinc_inputs(m.crystal, m.msd)
    parts.append(oracle.elinc(mm, bb, ss))
    parts.append(oracle.powderbragg(*m.pb_inputs))
    scales += [1.0, 1.0]
    for frac, d in dynamics:
        if "freegas" in d:
            parts.append(oracle.freegas(m.crystal.temperature, *d["freegas"]))
        else:
            parts.append(oracle.sab(d["alpha"], d["beta"], d["sab"], d["temperature"], d["mass"], d["sigma_bound"]))
        scales.append(frac)
    return oracle.composition(parts, scales)


CASES = {
    "three S(alpha,beta)": [(0.4, ELEMENTS[0]), (0.4, ELEMENTS[1]), (0.2, ELEMENTS[2])],
    "S(alpha,beta) + FreeGas": [(0.5, ELEMENTS[0]), (0.5, "gas")],
    "four staged": [(0.25, ELEMENTS[0]), (0.25, ELEMENTS[1]), (0.25, ELEMENTS[2]), (0.25, "gas")],
}


def _dynamics(case):
    return [(f, {"freegas": HELIUM} if e == "gas" else _table(*e)) for f, e in CASES[case]]


@pytest.mark.parametrize("case", list(CASES))
def test_compound_matches_checker(amd, oracle, case):
    C, core, materials = amd
    dyn = _dynamics(case)
    m = materials.Material(synth.fcc_powder(dcutoff=0.3), dynamics=dyn)
    o = _checker(oracle, materials, m, dyn)
    g = m.scatter
    parts = core.composition_components(g)
    assert [s for _, s in parts] == [1.0, 1.0] + [f for f, _ in dyn]
    n = 150000
    e = synth.energies_logu(n, 1e-5, 10.0, seed=5)
    assert np.max(_rel(g.crossSectionIsotropic(e), o.xs(e, nthreads=8))) <= 1e-12
    g.setRandomSeed(77)
    for call in range(2):
        ge, gm = g.sampleScatterIsotropic(e)
        oe, om, nerr = o.sample_iso(e, seed=77, stream=call, nthreads=8)
        assert nerr == 0
        assert_streams_match((_rel(ge, oe) > 1e-9) | (np.abs(gm - om) > 1e-9), f"{case}: iso call {call}")
        assert np.all(ge >= 0.0) and np.all(np.abs(gm) <= 1.0)
    dirs = synth.isotropic_dirs(n, seed=9)
    g.setRandomSeed(78)
    ge, gd = g.sampleScatter(e, dirs)
    oe, od, nerr = o.sample_dir(e, dirs, seed=78, nthreads=8)
    assert nerr == 0
    assert_streams_match((_rel(ge, oe) > 1e-9) | (np.max(np.abs(gd - od), axis=1) > 1e-9), f"{case}: dirs")
    assert np.abs(np.linalg.norm(gd, axis=1) - 1).max() < 1e-12
    # every staged component carries a share of the picks
    tot = g.crossSectionIsotropic(e)
    for c, (f, _) in zip(m.inelastic, dyn):
        assert np.mean(f * c.crossSectionIsotropic(e) / tot) > 1e-2


def test_repeat_matches_tiled_checker(amd, oracle):
    """repeat > 1 (n x repeat queue entries per staged component)"""
    C, core, materials = amd
    dyn = _dynamics("three S(alpha,beta)")
    m = materials.Material(synth.fcc_powder(dcutoff=0.3), dynamics=dyn)
    o = _checker(oracle, materials, m, dyn)
    e = synth.energies_logu(20001, 1e-4, 3.0, seed=6)
    m.scatter.setRandomSeed(5)
    ge, gm = m.scatter.sampleScatterIsotropic(e, repeat=3)
    oe, om, nerr = o.sample_iso(e, repeat=3, seed=5, nthreads=8)
    assert nerr == 0
    assert_streams_match((_rel(ge, oe) > 1e-9) | (np.abs(gm - om) > 1e-9), "three S(alpha,beta), repeat 3")


def test_multiphase_with_sab_phases(amd, oracle):
    """two phases with an S(alpha,beta) component each: the merged material
    keeps both S(alpha,beta) components (NCProcImpl.cc:263-338 adds them
    side by side), with the phases' PowderBragg and ElInc merged"""
    C, core, materials = amd
    ta, tb = _table(*ELEMENTS[0]), _table(*ELEMENTS[1])
    a = materials.Material(synth.fcc_powder(dcutoff=0.3), sab=ta)
    b = materials.Material(synth.fcc_powder(a=4.2, dcutoff=0.3), sab=tb)
    mp = materials.MultiPhase([(0.3, a), (0.7, b)])
    ph = {id(a): _checker(oracle, materials, a, [(1.0, ta)]), id(b): _checker(oracle, materials, b, [(1.0, tb)])}
    o = oracle.composition([ph[id(x)] for _, x in mp.phases], [w for w, _ in mp.phases])
    g = mp.scatter
    gparts = core.composition_components(g)
    oparts = oracle.components(o)
    assert len(gparts) == len(oparts) == 4
    assert [s for _, s in gparts] == [s for _, s, _ in oparts]
    e = synth.energies_logu(100000, 1e-5, 10.0, seed=8)
    assert np.max(_rel(g.crossSectionIsotropic(e), o.xs(e, nthreads=8))) <= 1e-12
    g.setRandomSeed(19)
    ge, gm = g.sampleScatterIsotropic(e)
    oe, om, nerr = o.sample_iso(e, seed=19, nthreads=8)
    assert nerr == 0
    assert_streams_match((_rel(ge, oe) > 1e-9) | (np.abs(gm - om) > 1e-9), "two S(alpha,beta) phases")


def test_staged_limits(amd):
    C, core, materials = amd
    t = _table(*ELEMENTS[0])
    args = (t["alpha"], t["beta"], t["sab"], t["temperature"], t["mass"], t["sigma_bound"])
    sabs = [core.create_sabscatter(*args) for _ in range(5)]
    gas = core.create_freegas(T, *HELIUM)
    with pytest.raises(C.NCBadInput):
        core.create_composition(sabs, [0.2] * 5)
    with pytest.raises(C.NCBadInput):
        core.create_composition(sabs[:4] + [gas], [0.2] * 5)
    four = core.create_composition(sabs[:4], [0.25] * 4)
    # a batch material holds at most one staged component
    with pytest.raises(C.NCBadInput):
        core.Batch([four], [10])
    with pytest.raises(C.NCBadInput):
        core.Batch([core.create_composition([sabs[0], gas], [0.5, 0.5])], [10])
    one = core.create_composition([sabs[0]], [1.0])
    core.Batch([one], [10])
