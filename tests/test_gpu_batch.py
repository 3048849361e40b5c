"""Many materials in one persistent launch vs one call per material: the
batched launch must return bit-identical cross sections and samples (every
material keeps its own Philox streams), for every tuning, and advance each
material's call counter exactly like a per-material call."""
import numpy as np
import pytest

from ncrystal_amd import synth

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def env():
    from ncrystal_amd import _capi, core
    if _capi.lib().ncamd_device_count() < 1:
        pytest.fail("no GPU visible to libncrystal_amd")
    return _capi, core


def _materials(core, nmat=24, seed=5):
    mats = synth.batch_materials(nmat=nmat, seed=seed, planes=(1e2, 3e4), grids=((50, 100), (120, 400)),
                                 counts=(1e3, 4e4))
    procs = []
    for i, m in enumerate(mats):
        v0n, d, fm = m["pb"]
        ei = core.create_elincscatter(m["elinc"]["msd"], m["elinc"]["bixs"], m["elinc"]["scale"])
        pb = core.create_powderbragg(v0n, d, fm)
        s = m["sab"]
        sab = core.create_sabscatter(s["alpha"], s["beta"], s["sab"], s["temperature"], s["mass"], s["sigma_bound"])
        # a mix of full compositions and single processes
        if i % 4 == 1:
            p = sab
        elif i % 4 == 2:
            p = pb
        else:
            p = core.create_composition([ei, pb, sab], [1.0, 1.0, 1.0])
        p.setRandomSeed(1000 + i)
        procs.append(p)
    return mats, procs


def test_batch_matches_per_material_calls(env):
    C, core = env
    mats, procs = _materials(core)
    counts = [m["count"] for m in mats]
    batch = core.Batch(procs, counts)
    tot = batch.total
    assert tot == sum(counts)
    e = synth.energies_logu(tot, 1e-4, 1.0, seed=7)
    st = C.Stream()
    de = C.DeviceArray.from_numpy(e)
    xs, oe, om = C.DeviceArray(tot), C.DeviceArray(tot), C.DeviceArray(tot)
    # reference: one device call per material on slices of the same arrays
    ref_xs, ref_e, ref_m = np.empty(tot), np.empty(tot), np.empty(tot)
    off = batch.offsets
    for i, p in enumerate(procs):
        n = counts[i]
        sub_e = C.DeviceArray.from_numpy(e[off[i]:off[i + 1]])
        a, b, c = C.DeviceArray(n), C.DeviceArray(n), C.DeviceArray(n)
        p.dev_crossSectionIsotropic(sub_e.ptr, n, a.ptr, stream=st.ptr)
        p.dev_sampleScatterIsotropic(sub_e.ptr, n, b.ptr, c.ptr, stream=st.ptr)
        st.sync()
        ref_xs[off[i]:off[i + 1]] = a.numpy()
        ref_e[off[i]:off[i + 1]] = b.numpy()
        ref_m[off[i]:off[i + 1]] = c.numpy()
    states = [p.getRNGState() for p in procs]
    for p, (sd, cc) in zip(procs, states):
        p.setRNGState(sd, cc - 1)  # replay the sampling call in the batch
    for tuning in [(512, 256, 0), (4096, 2048, 1), (1024, 1024, 1)]:
        batch.set_tuning(*tuning)
        for p, (sd, cc) in zip(procs, states):
            p.setRNGState(sd, cc - 1)
        batch.dev_crossSection(de.ptr, xs.ptr, stream=st.ptr)
        batch.dev_sampleScatterIsotropic(de.ptr, oe.ptr, om.ptr, stream=st.ptr)
        assert batch.dev_check_errors(stream=st.ptr) == 0
        st.sync()
        assert np.array_equal(xs.numpy(), ref_xs), tuning
        assert np.array_equal(oe.numpy(), ref_e), tuning
        assert np.array_equal(om.numpy(), ref_m), tuning
        assert [p.getRNGState() for p in procs] == states
    # autotuning leaves every RNG state alone and picks one of the candidates
    ms = batch.dev_autotune(de.ptr, oe.ptr, om.ptr, stream=st.ptr)
    print("autotune best ms", ms, batch.tuning())
    assert ms > 0 and [p.getRNGState() for p in procs] == states
    for p, (sd, cc) in zip(procs, states):
        p.setRNGState(sd, cc - 1)
    batch.dev_sampleScatterIsotropic(de.ptr, oe.ptr, om.ptr, stream=st.ptr)
    st.sync()
    assert np.array_equal(oe.numpy(), ref_e) and np.array_equal(om.numpy(), ref_m)
    # alternating cross-section and sampling calls replay one graph per mode,
    # with eager work in between, and give the same result every time
    c0, r0 = batch.graph_stats(0)
    c1, r1 = batch.graph_stats(1)
    for rep in range(3):
        for p, (sd, cc) in zip(procs, states):
            p.setRNGState(sd, cc - 1)
        batch.dev_crossSection(de.ptr, xs.ptr, stream=st.ptr)
        batch.dev_sampleScatterIsotropic(de.ptr, oe.ptr, om.ptr, stream=st.ptr)
        st.sync()
        assert np.array_equal(xs.numpy(), ref_xs), rep
        assert np.array_equal(om.numpy(), ref_m), rep
        C.call("ncamd_dev_fill_loguniform", xs.ptr, tot, 1e-3, 1.0, rep, st.ptr)
        st.sync()
    c0b, r0b = batch.graph_stats(0)
    c1b, r1b = batch.graph_stats(1)
    print("graph captures/replays: xs", (c0, r0), "->", (c0b, r0b), "sampling", (c1, r1), "->", (c1b, r1b))
    assert c1b == c1 and r1b == r1 + 3  # sampling graph replayed each time
    assert c0b <= c0 + 1 and r0b >= r0 + 2  # xs graph captured at most once more (new tuning), then replayed


def test_batch_rejects_oriented_and_bad_input(env):
    C, core = env
    mats, procs = _materials(core, nmat=2)
    with pytest.raises(C.NCAMDError):
        core.Batch(procs, [-1, 5])
