"""FreeGas and 1/v absorption processes through the C ABI on the GPU against
the CPU checker (whose free-gas model tests/test_freegas_pins.py pins to the
closed-form scattering function): cross sections to 1e-12, sampled (E', mu)
and directions on the same Philox streams, the repeat layout, refused
energies, and the calls these processes do not support."""
import numpy as np
import pytest

from ncrystal_amd import synth
from conftest import assert_streams_match

pytestmark = pytest.mark.gpu

GASES = [(293.15, 4.0026, 0.7596), (20.0, 1.00794, 20.49), (600.0, 39.948, 0.683)]


@pytest.fixture(scope="module")
def amd():
    from ncrystal_amd import _capi, core
    if _capi.lib().ncamd_device_count() < 1:
        pytest.fail("no GPU visible to libncrystal_amd")
    return _capi, core


def _rel(a, b):
    """relative difference; equal values (infinities included) count 0"""
    same = (a == b) | (np.isnan(a) & np.isnan(b))
    den = np.maximum(np.abs(a), np.abs(b))
    with np.errstate(invalid="ignore"):
        r = np.where(den > 0, np.abs(a - b) / np.where(den > 0, den, 1.0), 0.0)
    return np.where(same, 0.0, r)


@pytest.mark.parametrize("T,M,sf", GASES)
def test_freegas_matches_checker(amd, oracle, T, M, sf):
    C, core = amd
    g = core.create_freegas(T, M, sf)
    o = oracle.freegas(T, M, sf)
    e = np.concatenate([synth.energies_logu(30000, 1e-5, 10.0, seed=2), [0.0, 1e-300, 1e4, 1e6]])
    assert np.max(_rel(g.crossSectionIsotropic(e), o.xs(e))) <= 1e-12
    g.setRandomSeed(5)
    ge, gm = g.sampleScatterIsotropic(e, repeat=2)
    oe, om, nerr = o.sample_iso(e, repeat=2, seed=5, nthreads=8)
    assert nerr == 0 and ge.shape == (2 * len(e),)
    assert_streams_match((np.abs(ge - oe) > 1e-9 * np.abs(oe)) | (np.abs(gm - om) > 1e-9), "freegas iso")
    # oriented call: new directions from mu on the same streams
    dirs = synth.isotropic_dirs(len(e), seed=3)
    g.setRandomSeed(9)
    de, dd = g.sampleScatter(e, dirs)
    re_, rd, nerr = o.sample_dir(e, dirs, seed=9, nthreads=8)
    assert nerr == 0
    assert_streams_match(np.max(np.abs(dd - rd), axis=1) > 1e-9, "freegas dirs")
    assert np.abs(np.linalg.norm(dd, axis=1) - 1).max() < 1e-12
    xd = g.crossSection(e, dirs)
    assert np.array_equal(xd, g.crossSectionIsotropic(e))


def test_freegas_refused_energies(amd, oracle):
    C, core = amd
    g = core.create_freegas(293.15, 4.0026, 0.7596)
    o = oracle.freegas(293.15, 4.0026, 0.7596)
    e = np.array([0.025, -1.0, np.nan, np.inf, -np.inf, 0.0, 1.0])
    st = C.Stream()
    de = C.DeviceArray.from_numpy(e)
    d1, d2 = C.DeviceArray(len(e)), C.DeviceArray(len(e))
    g.setRandomSeed(1)
    g.dev_sampleScatterIsotropic(de.ptr, len(e), d1.ptr, d2.ptr, stream=st.ptr)
    assert g.dev_check_errors(stream=st.ptr) == 4
    ge, gm = d1.numpy(), d2.numpy()
    oe, om, nerr = o.sample_iso(e, seed=1)
    assert nerr == 4
    bad = np.array([False, True, True, True, True, False, False])
    assert np.all(ge[bad] == -1.0) and np.all(gm[bad] == -999.0)
    assert np.allclose(ge[~bad], oe[~bad], rtol=1e-9) and np.allclose(gm[~bad], om[~bad], atol=1e-9)
    with pytest.raises(C.NCCalcError):
        g.sampleScatterIsotropic(e)


def test_absoov(amd, oracle):
    C, core = amd
    a = core.create_absoov(0.231)
    o = oracle.absoov(0.231)
    e = np.concatenate([[0.0, 1e-300], synth.energies_logu(10000, 1e-9, 100.0, seed=4)])
    xg, xo = a.crossSectionIsotropic(e), o.xs(e)
    assert np.isinf(xg[0]) and np.max(_rel(xg[1:], xo[1:])) <= 1e-15
    assert np.array_equal(a.crossSectionIsotropic(e, repeat=3).reshape(3, -1), np.tile(xg, (3, 1)))
    st = C.Stream()
    d = C.DeviceArray.from_numpy(e)
    d1, d2 = C.DeviceArray(len(e)), C.DeviceArray(len(e))
    with pytest.raises(C.NCBadInput):
        C.call("ncamd_dev_samplescatterisotropic_many", a._h, d.ptr, len(e), 1, d1.ptr, d2.ptr, st.ptr)


def test_composition_rules(amd):
    C, core = amd
    g = core.create_freegas(293.15, 4.0026, 0.7596)
    a = core.create_absoov(0.231)
    e = core.create_elincscatter([0.01], [1.0], [1.0])
    # scattering and absorption processes do not mix (NCProcImpl.cc:269-272)
    for x in (e, g):
        with pytest.raises(C.NCBadInput):
            core.create_composition([x, a], [1.0, 1.0])
    # FreeGas beside S(alpha,beta) is a composition of its own kind (a staged
    # component; parity in tests/test_gpu_compound.py)
    t = synth.small_sab_table(mass_amu=4.0026)
    sab = core.create_sabscatter(t["alpha"], t["beta"], t["sab"], t["temperature"], t["mass"], t["sigma_bound"])
    assert len(core.composition_components(core.create_composition([g, sab], [1.0, 1.0]))) == 2
    with pytest.raises(C.NCBadInput):
        a.crossSectionIsotropic_fp32(np.array([0.025]))


def test_batch_refuses_freegas_and_absorption(amd):
    """standalone FreeGas and 1/v absorption processes are not batch materials:
    BadInput, as for compositions holding them"""
    C, core = amd
    for p in (core.create_freegas(293.15, 4.0026, 0.7596), core.create_absoov(0.231)):
        with pytest.raises(C.NCBadInput):
            core.Batch([p], [10])
