"""Multi-phase materials through the C ABI on the GPU against the CPU checker:
the composition of two phases' scatters with their number-density weights
(ncrystal_core/src/factories/NCFactImpl.cc:500-540,615-642), flattened and
merged as ProcComposition::addComponent does (NCProcImpl.cc:263-338), which
tests/test_composition_merge.py pins in the checker.  Components and scales
identical, merged PowderBragg tables bit-identical, cross sections to 1e-12,
samples on the same Philox streams, FP32 Bragg tables of the merged process
within 1e-6 of FP64; two S(alpha,beta) phases keep both S(alpha,beta)
components."""
import numpy as np
import pytest

from ncrystal_amd import synth
from conftest import assert_streams_match

pytestmark = pytest.mark.gpu


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


def _checker_phase(oracle, materials, m, sab):
    parts = []
    mm, bb, ss = materials.elinc_inputs(m.crystal, m.msd)
    parts.append(oracle.elinc(mm, bb, ss))
    v0n, dsp, fm = m.pb_inputs
    parts.append(oracle.powderbragg(v0n, dsp, fm))
    if sab is not None:
        parts.append(oracle.sab(sab["alpha"], sab["beta"], sab["sab"], sab["temperature"], sab["mass"],
                                sab["sigma_bound"]))
    return oracle.composition(parts, [1.0] * len(parts))


def _two_phases(materials, oracle):
    sab = synth.small_sab_table(mass_amu=26.98)
    a = materials.Material(synth.fcc_powder(dcutoff=0.3), sab=sab)
    b = materials.Material(synth.ortho_random(seed=3, dcutoff=0.3))
    mp = materials.MultiPhase([(0.4, a), (0.6, b)])
    ochk = {id(a): _checker_phase(oracle, materials, a, sab), id(b): _checker_phase(oracle, materials, b, None)}
    o = oracle.composition([ochk[id(m)] for _, m in mp.phases], [w for w, _ in mp.phases])
    return mp, o


def test_multiphase_matches_checker(amd, oracle):
    C, core, materials = amd
    mp, o = _two_phases(materials, oracle)
    g = mp.scatter
    w = [x for x, _ in mp.phases]
    assert abs(sum(w) - 1.0) < 1e-15 and w == sorted(w)
    gparts = core.composition_components(g)
    oparts = oracle.components(o)
    kinds = {"PowderBragg": "PowderBragg", "ElInc": "ElInc", "SAB": "SAB"}
    assert len(gparts) == len(oparts) == 3
    for (gp, gs), (ot, os_, od) in zip(gparts, oparts):
        assert gs == os_, (ot, gs, os_)
        if kinds[ot] == "PowderBragg":
            e2d, fdm, f32 = core.powderbragg_tables(gp)
            assert np.array_equal(e2d, od[0])
            assert np.max(np.abs(fdm - od[1]) / od[1][-1]) <= 1e-15
    n = 100000
    e = synth.energies_logu(n, 1e-5, 10.0, seed=3)
    assert np.max(_rel(g.crossSectionIsotropic(e), o.xs(e))) <= 1e-12
    g.setRandomSeed(17)
    ge, gm = g.sampleScatterIsotropic(e)
    oe, om, nerr = o.sample_iso(e, seed=17, nthreads=8)
    assert nerr == 0
    assert_streams_match(np.maximum(_rel(ge, oe), _rel(gm, om)) > 1e-9, "multiphase iso")
    dirs = synth.isotropic_dirs(n, seed=8)
    g.setRandomSeed(5)
    ge2, gd = g.sampleScatter(e, dirs)
    oe2, od2, nerr = o.sample_dir(e, dirs, seed=5, nthreads=8)
    assert nerr == 0
    assert_streams_match(np.max(np.abs(gd - od2), axis=1) > 1e-9, "multiphase dirs")
    # the merged Bragg component carries FP32 tables since both phases had them
    x64 = g.crossSectionIsotropic(e)
    x32 = g.crossSectionIsotropic_fp32(e)
    nz = x64 > 0
    assert np.max(_rel(x32[nz], x64[nz])) <= 1e-6


def test_multiphase_refusals(amd, oracle):
    C, core, materials = amd
    sab = synth.small_sab_table(mass_amu=26.98)
    a = materials.Material(synth.fcc_powder(dcutoff=0.4), sab=sab)
    b = materials.Material(synth.fcc_powder(a=4.2, dcutoff=0.4), sab=synth.small_sab_table(mass_amu=40.0))
    # two S(alpha,beta) phases keep both S(alpha,beta) components (parity in
    # tests/test_gpu_compound.py)
    assert len(core.composition_components(materials.MultiPhase([(0.5, a), (0.5, b)]).scatter)) == 4
    # the same phase twice adds its weights
    mp = materials.MultiPhase([(0.5, a), (0.5, a)])
    parts = core.composition_components(mp.scatter)
    assert [s for _, s in parts] == [1.0, 1.0, 1.0]
