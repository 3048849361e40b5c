"""Compositions with FreeGas components (gas mixtures: one FreeGas process
per element, scaled by its number fraction, as NCrystal's gas-mixture
materials compose them) and compositions of absorption processes, through
the C ABI on the GPU against the CPU checker: cross sections to 1e-12,
samples on the same Philox streams (the component pick, then the free-gas
sampler of tests/test_freegas_pins.py), absorption compositions merged into
one 1/v process (ncrystal_core/src/absoov/NCAbsOOV.cc:60-70) and refused
for sampling; such compositions are refused in many-material batches."""
import numpy as np
import pytest

from ncrystal_amd import synth
from conftest import assert_streams_match

pytestmark = pytest.mark.gpu

# (mass [amu], free-atom scattering cross section [barn], number fraction)
HE_AR_XE = [(4.0026, 0.7596, 0.6), (39.948, 0.6660, 0.3), (131.293, 3.6963, 0.1)]


@pytest.fixture(scope="module")
def amd():
    from ncrystal_amd import _capi, core
    if _capi.lib().ncamd_device_count() < 1:
        pytest.fail("no GPU visible to libncrystal_amd")
    return _capi, core


def _rel(a, b):
    den = np.maximum(np.abs(a), np.abs(b))
    with np.errstate(invalid="ignore", divide="ignore"):
        r = np.where(den > 0, np.abs(a - b) / np.where(den > 0, den, 1.0), 0.0)
    return np.where(a == b, 0.0, r)


@pytest.mark.parametrize("ngas,T", [(2, 293.15), (3, 77.0), (1, 600.0)])
def test_gas_mixture_matches_checker(amd, oracle, ngas, T):
    C, core = amd
    gases = HE_AR_XE[:ngas]
    g = core.create_composition([core.create_freegas(T, m, s) for m, s, _ in gases], [f for _, _, f in gases])
    o = oracle.composition([oracle.freegas(T, m, s) for m, s, _ in gases], [f for _, _, f in gases])
    e = synth.energies_logu(60000, 1e-5, 10.0, seed=ngas)
    xg = g.crossSectionIsotropic(e)
    assert np.max(_rel(xg, o.xs(e))) <= 1e-12
    ref = sum(f * oracle.freegas(T, m, s).xs(e) for m, s, f in gases)
    assert np.max(_rel(xg, ref)) <= 1e-13
    g.setRandomSeed(31)
    ge, gm = g.sampleScatterIsotropic(e)
    oe, om, nerr = o.sample_iso(e, seed=31, nthreads=8)
    assert nerr == 0
    assert_streams_match((np.abs(ge - oe) > 1e-9 * np.abs(oe)) | (np.abs(gm - om) > 1e-9), f"gas mixture {ngas} iso")
    assert np.all(np.abs(gm) <= 1.0) and np.all(ge >= 0.0)
    dirs = synth.isotropic_dirs(len(e), seed=7)
    g.setRandomSeed(8)
    de, dd = g.sampleScatter(e, dirs)
    re_, rd, nerr = o.sample_dir(e, dirs, seed=8, nthreads=8)
    assert nerr == 0
    assert_streams_match(np.max(np.abs(dd - rd), axis=1) > 1e-9, f"gas mixture {ngas} dirs")
    assert np.abs(np.linalg.norm(dd, axis=1) - 1).max() < 1e-12


def test_gas_mixture_not_in_batches(amd):
    C, core = amd
    g = core.create_composition([core.create_freegas(293.15, m, s) for m, s, _ in HE_AR_XE[:2]], [0.5, 0.5])
    with pytest.raises(C.NCBadInput):
        core.Batch([g], [10])


def test_absorption_composition(amd, oracle):
    C, core = amd
    sig = [(0.231, 0.25), (2.56, 0.5), (5.08, 0.25)]
    g = core.create_composition([core.create_absoov(s) for s, _ in sig], [f for _, f in sig])
    o = oracle.composition([oracle.absoov(s) for s, _ in sig], [f for _, f in sig])
    parts = core.composition_components(g)
    assert len(parts) == 1 and parts[0][1] == 1.0
    e = np.concatenate([synth.energies_logu(20000, 1e-6, 100.0, seed=4), [0.0]])
    xg = g.crossSectionIsotropic(e)
    assert np.array_equal(xg, o.xs(e))
    assert np.isinf(xg[-1])
    ref = sum(f * oracle.absoov(s).xs(e[:-1]) for s, f in sig)
    assert np.max(_rel(xg[:-1], ref)) < 1e-15
    with pytest.raises(C.NCBadInput):
        g.sampleScatterIsotropic(e)
    with pytest.raises(C.NCBadInput):
        g.sampleScatter(e, synth.isotropic_dirs(len(e), seed=1))
