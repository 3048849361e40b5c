"""Edge cases through the C ABI on the GPU: empty batches for every entry
point, special energies (0, negative, +-inf, NaN, subnormal, far above every
table) against the CPU checker on the same Philox streams, the repeat layout
of the staged composition kernels, and many-material batches with an empty
material or a single material."""
import numpy as np
import pytest

from ncrystal_amd import synth
from conftest import assert_streams_match

pytestmark = pytest.mark.gpu

SPECIAL = np.array([0.0, -1.0, -0.0, np.inf, -np.inf, np.nan, 5e-324, 1e-300, 1e-12, 1e3, 1e6])


@pytest.fixture(scope="module")
def amd():
    from ncrystal_amd import _capi, core, materials
    if _capi.lib().ncamd_device_count() < 1:
        pytest.fail("no GPU visible to libncrystal_amd")
    return _capi, core, materials


def _same(a, b, rtol=1e-12):
    """equal including NaN positions, finite values within rtol"""
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    if a.shape != b.shape or not np.array_equal(np.isnan(a), np.isnan(b)):
        return False
    ok = ~np.isnan(a)
    if not np.array_equal(np.isinf(a[ok]), np.isinf(b[ok])):
        return False
    fin = ok & np.isfinite(a)
    den = np.maximum(np.abs(a[fin]), np.abs(b[fin]))
    return bool(np.all(np.abs(a[fin] - b[fin]) <= rtol * np.where(den > 0, den, 1.0)))


def _c1(amd, oracle):
    _, core, materials = amd
    cr = synth.fcc_powder()
    sabd = synth.small_sab_table(mass_amu=26.98)
    m = materials.Material(cr, sab=sabd)
    mm, bb, ss = materials.elinc_inputs(cr, m.msd)
    v0n, dsp, fm = m.pb_inputs
    o_ei = oracle.elinc(mm, bb, ss)
    o_pb = oracle.powderbragg(v0n, dsp, fm)
    o_sab = oracle.sab(sabd["alpha"], sabd["beta"], sabd["sab"], sabd["temperature"], sabd["mass"],
                       sabd["sigma_bound"])
    g_ei = core.create_elincscatter(mm, bb, ss)
    g_pb = core.create_powderbragg(v0n, dsp, fm)
    g_sab = core.create_sabscatter(sabd["alpha"], sabd["beta"], sabd["sab"], sabd["temperature"], sabd["mass"],
                                   sabd["sigma_bound"])
    comp = oracle.composition([o_ei, o_pb, o_sab], [1.0, 1.0, 1.0])
    return dict(comp=(m.scatter, comp), elinc=(g_ei, o_ei), pb=(g_pb, o_pb), sab=(g_sab, o_sab))


@pytest.mark.parametrize("which", ["comp", "elinc", "pb", "sab"])
def test_special_energies_match_checker(amd, oracle, which):
    C = amd[0]
    g, o = _c1(amd, oracle)[which]
    e = np.concatenate([SPECIAL, synth.energies_logu(1000, 1e-5, 10.0, seed=3)])
    xg = g.crossSectionIsotropic(e)
    xo = o.xs(e)
    assert _same(xg, xo), (which, xg[:len(SPECIAL)], xo[:len(SPECIAL)])
    # sampling through the device-pointer call: neutrons whose sampling fails
    # carry the poison values (E' = -1, mu = -999) and are counted, as in the
    # checker (the host-pointer call raises instead)
    st = C.Stream()
    de = C.DeviceArray.from_numpy(e)
    d1, d2 = C.DeviceArray(len(e)), C.DeviceArray(len(e))
    g.setRandomSeed(31)
    g.dev_sampleScatterIsotropic(de.ptr, len(e), d1.ptr, d2.ptr, stream=st.ptr)
    nerr_g = g.dev_check_errors(stream=st.ptr)
    ge, gm = d1.numpy(), d2.numpy()
    oe, om, nerr_o = o.sample_iso(e, seed=31)
    assert nerr_g == nerr_o, (which, nerr_g, nerr_o)
    # special entries exactly as the checker treats them
    n = len(SPECIAL)
    assert _same(ge[:n], oe[:n]) and _same(gm[:n], om[:n]), (which, ge[:n], oe[:n], gm[:n], om[:n])
    # ordinary entries: same Philox streams
    assert_streams_match((np.abs(ge[n:] - oe[n:]) > 1e-9 * np.abs(oe[n:])) | (np.abs(gm[n:] - om[n:]) > 1e-9),
                         f"special-energy batch {which}")


def test_empty_batches(amd, oracle):
    C, core, _ = amd
    procs = _c1(amd, oracle)
    empty = np.zeros(0)
    for name, (g, _) in procs.items():
        assert g.crossSectionIsotropic(empty).shape == (0,), name
        oe, om = g.sampleScatterIsotropic(empty)
        assert oe.shape == (0,) and om.shape == (0,), name
        st = C.Stream()
        d = C.DeviceArray(1)
        g.dev_crossSectionIsotropic(d.ptr, 0, d.ptr, stream=st.ptr)
        g.dev_sampleScatterIsotropic(d.ptr, 0, d.ptr, d.ptr, stream=st.ptr)
        assert g.dev_check_errors(stream=st.ptr) == 0, name


def test_composition_repeat_layout(amd, oracle):
    g, o = _c1(amd, oracle)["comp"]
    e = synth.energies_logu(20000, 1e-5, 10.0, seed=5)
    g.setRandomSeed(9)
    ge, gm = g.sampleScatterIsotropic(e, repeat=3)
    oe, om, nerr = o.sample_iso(e, repeat=3, seed=9)
    assert nerr == 0 and ge.shape == (3 * len(e),)
    assert_streams_match((np.abs(ge - oe) > 1e-9 * np.abs(oe)) | (np.abs(gm - om) > 1e-9), "repeat layout")
    xr = g.crossSectionIsotropic(e[:50], repeat=4)
    assert np.array_equal(xr.reshape(4, 50), np.tile(g.crossSectionIsotropic(e[:50]), (4, 1)))


def test_batch_empty_and_single_material(amd, oracle):
    C, core, _ = amd
    procs = _c1(amd, oracle)
    g = procs["comp"][0]
    e = synth.energies_logu(30000, 1e-4, 1.0, seed=6)
    st = C.Stream()
    de = C.DeviceArray.from_numpy(e)
    # reference: one device call
    g.setRandomSeed(44)
    a, b = C.DeviceArray(len(e)), C.DeviceArray(len(e))
    g.dev_sampleScatterIsotropic(de.ptr, len(e), a.ptr, b.ptr, stream=st.ptr)
    st.sync()
    ref_e, ref_m = a.numpy(), b.numpy()
    # same neutrons as the only non-empty material of a batch
    pb = procs["pb"][0]
    batch = core.Batch([pb, g, pb], [0, len(e), 0])
    assert batch.total == len(e)
    g.setRandomSeed(44)
    oe, om = C.DeviceArray(len(e)), C.DeviceArray(len(e))
    batch.dev_sampleScatterIsotropic(de.ptr, oe.ptr, om.ptr, stream=st.ptr)
    assert batch.dev_check_errors(stream=st.ptr) == 0
    st.sync()
    assert np.array_equal(oe.numpy(), ref_e) and np.array_equal(om.numpy(), ref_m)
    xs = C.DeviceArray(len(e))
    batch.dev_crossSection(de.ptr, xs.ptr, stream=st.ptr)
    st.sync()
    assert np.array_equal(xs.numpy(), g.crossSectionIsotropic(e))


def test_scbragg_special_inputs(amd, oracle):
    """special energies and direction vectors that are not unit length (the
    reference normalises them) against the checker"""
    from test_gpu_scbragg import _fcc
    C, core, _ = amd
    g, o = _fcc(core, oracle)
    rng = np.random.default_rng(12)
    e = np.concatenate([SPECIAL, synth.energies_logu(3000, 1e-3, 1.0, seed=8)])
    dirs = synth.isotropic_dirs(len(e), seed=9) * rng.uniform(0.5, 4.0, size=(len(e), 1))
    assert _same(g.crossSection(e, dirs), o.xs_dir(e, dirs))
    st = C.Stream()
    de, dd = C.DeviceArray.from_numpy(e), C.DeviceArray.from_numpy(dirs.ravel())
    oe, od = C.DeviceArray(len(e)), C.DeviceArray(3 * len(e))
    g.setRandomSeed(77)
    g.dev_sampleScatter(de.ptr, dd.ptr, len(e), oe.ptr, od.ptr, stream=st.ptr)
    nerr_g = g.dev_check_errors(stream=st.ptr)
    re_, rd, nerr_o = o.sample_dir(e, dirs, seed=77)
    assert nerr_g == nerr_o
    ge, gd = oe.numpy(), od.numpy().reshape(-1, 3)
    n = len(SPECIAL)
    assert _same(ge[:n], re_[:n]) and _same(gd[:n].ravel(), rd[:n].ravel())
    assert_streams_match(np.max(np.abs(gd[n:] - rd[n:]), axis=1) > 1e-9, "scbragg special inputs")
