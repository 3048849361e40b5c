"""Allocation-free sampling call (ncamd_dev_samplescatterisotropic_many_noalloc)
against the allocating one (ncamd_dev_samplescatterisotropic_many), through
the C ABI: the same kernels on the same Philox streams, so outputs must be
identical, on two calls in a row (the second smaller, with repeat), for an
S(alpha,beta) process reaching its free-gas extender, the same process on its
envelope sampler, a full composition (ElInc + PowderBragg + S(alpha,beta)),
a PowderBragg process, a FreeGas process and a compound (two S(alpha,beta)
components and a FreeGas one, four queues); oversized calls raise
BadInput."""
import numpy as np
import pytest

from ncrystal_amd import synth

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def env():
    from ncrystal_amd import _capi, core, materials
    if _capi.lib().ncamd_device_count() < 1:
        pytest.fail("no GPU visible to libncrystal_amd")
    L = _capi.lib()
    f = L.ncamd_dev_samplescatterisotropic_many_noalloc
    f.restype = None
    f.argtypes = [_capi.Handle, _capi._vp, _capi.ctypes.c_ulong, _capi.ctypes.c_ulong, _capi._vp, _capi._vp,
                  _capi._vp]
    return _capi, core, materials


def _processes(core, materials):
    t = synth.small_sab_table(mass_amu=26.98)
    args = (t["alpha"], t["beta"], t["sab"], t["temperature"], t["mass"], t["sigma_bound"])
    sab = core.create_sabscatter(*args)
    env = core.create_sabscatter(*args)
    env.use_envelope_sampler(nintervals=64, emin=1e-3, emax=0.2, nbox_w=32, nbox_k=16)
    m = materials.Material(synth.fcc_powder(), sab=t)
    pb = m.components["bragg"]
    fg = core.create_freegas(293.15, 4.0026, 0.7596)
    t2 = synth.small_sab_table(seed=5, mass_amu=15.999, amax=60.0)
    cpd = materials.Material(synth.fcc_powder(), dynamics=[(0.4, t), (0.4, t2), (0.2, {"freegas": (4.0026, 0.7596)})])
    return {"sab": sab, "envelope": env, "composition": m.scatter, "powderbragg": pb, "freegas": fg,
            "compound": cpd.scatter}


def _run(C, proc, fn, de, n, repeat, st):
    tot = n * repeat
    oe, om = C.DeviceArray(tot), C.DeviceArray(tot)
    proc.setRNGState(4242, 7)
    if fn == "noalloc":
        C.lib().ncamd_dev_samplescatterisotropic_many_noalloc(proc.handle, de.ptr, n, repeat, oe.ptr, om.ptr,
                                                              st.ptr)
        C.check()
    else:
        proc.dev_sampleScatterIsotropic(de.ptr, n, oe.ptr, om.ptr, stream=st.ptr, repeat=repeat)
    assert proc.dev_check_errors(stream=st.ptr) == 0
    assert proc.getRNGState() == (4242, 8)
    return oe.numpy(), om.numpy()


def test_noalloc_matches_allocating_call(env):
    C, core, materials = env
    st = C.Stream()
    for name, proc in _processes(core, materials).items():
        for n, repeat in ((300000, 1), (70001, 3)):
            e = synth.energies_logu(n, 1e-5, 5.0, seed=n)
            de = C.DeviceArray.from_numpy(e)
            print(f"{name}: n={n} repeat={repeat}", flush=True)
            a_e, a_m = _run(C, proc, "alloc", de, n, repeat, st)
            b_e, b_m = _run(C, proc, "noalloc", de, n, repeat, st)
            assert np.array_equal(a_e, b_e) and np.array_equal(a_m, b_m), (name, n, repeat)
            assert np.all(np.abs(b_m) <= 1.0) and np.all(b_e >= 0.0)


def test_noalloc_refuses_oversized_calls(env):
    C, core, materials = env
    pb = _processes(core, materials)["powderbragg"]
    st = C.Stream()
    d = C.DeviceArray(1)
    before = pb.getRNGState()
    C.lib().ncamd_dev_samplescatterisotropic_many_noalloc(pb.handle, d.ptr, (1 << 21) + 1, 1, d.ptr, d.ptr, st.ptr)
    with pytest.raises(C.NCBadInput):
        C.check()
    C.lib().ncamd_dev_samplescatterisotropic_many_noalloc(pb.handle, d.ptr, 1 << 20, 3, d.ptr, d.ptr, st.ptr)
    with pytest.raises(C.NCBadInput):
        C.check()
    assert pb.getRNGState() == before  # refused before any launch
