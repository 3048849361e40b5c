"""The many-materials workload at full size against the CPU checker: 512
materials (ElInc + PowderBragg + S(alpha,beta) each; plane counts
log-uniform 1e2..1e5, grids 50x100..300x1200, 1e4..1e6 neutrons each,
~1.1e8 neutrons, E log-uniform 1e-4..1 eV) in one batched launch.  Every
37th material and the extremes (most and fewest planes, largest and smallest
grid, most neutrons) are compared with the checker's composition of the same
inputs: cross sections to 1e-12 and samples on the same Philox streams
(seed, call 0, neutron index within the material); over all neutrons,
cross sections are finite and >= 0, E' >= 0 and |mu| <= 1, with no errors."""
import numpy as np
import pytest

from conftest import assert_streams_match
from ncrystal_amd import synth

pytestmark = pytest.mark.gpu


def _rel(a, b):
    den = np.maximum(np.abs(a), np.abs(b))
    with np.errstate(invalid="ignore", divide="ignore"):
        r = np.where(den > 0, np.abs(a - b) / np.where(den > 0, den, 1.0), 0.0)
    return np.where(a == b, 0.0, r)


def test_batch_512_materials_against_checker(oracle):
    from ncrystal_amd import _capi as C
    from ncrystal_amd import core
    if C.lib().ncamd_device_count() < 1:
        pytest.fail("no GPU visible to libncrystal_amd")
    mats = synth.batch_materials(nmat=512, seed=11)
    procs = []
    for i, m in enumerate(mats):
        v0n, d, fm = m["pb"]
        t = m["sab"]
        parts = [core.create_elincscatter(m["elinc"]["msd"], m["elinc"]["bixs"], m["elinc"]["scale"]),
                 core.create_powderbragg(v0n, d, fm),
                 core.create_sabscatter(t["alpha"], t["beta"], t["sab"], t["temperature"], t["mass"],
                                        t["sigma_bound"])]
        p = core.create_composition(parts, [1.0, 1.0, 1.0])
        p.setRandomSeed(1000 + i)
        procs.append(p)
    counts = [m["count"] for m in mats]
    batch = core.Batch(procs, counts)
    tot = batch.total
    off = batch.offsets
    print("neutrons", tot, flush=True)
    st = C.Stream()
    de = C.DeviceArray(tot)
    C.call("ncamd_dev_fill_loguniform", de.ptr, tot, 1e-4, 1.0, 3, st.ptr)
    xs, oe, om = C.DeviceArray(tot), C.DeviceArray(tot), C.DeviceArray(tot)
    batch.dev_crossSection(de.ptr, xs.ptr, stream=st.ptr)
    batch.dev_sampleScatterIsotropic(de.ptr, oe.ptr, om.ptr, stream=st.ptr)
    assert batch.dev_check_errors(stream=st.ptr) == 0
    st.sync()
    e, x, ef, mu = de.numpy(), xs.numpy(), oe.numpy(), om.numpy()
    assert np.all(np.isfinite(x) & (x >= 0.0))
    assert np.all(ef >= 0.0) and np.all(np.abs(mu) <= 1.0)
    nplanes = [len(m["pb"][1]) for m in mats]
    cells = [m["sab"]["sab"].size for m in mats]
    pick = set(range(0, 512, 37))
    pick |= {int(np.argmax(nplanes)), int(np.argmin(nplanes)), int(np.argmax(cells)), int(np.argmin(cells)),
             int(np.argmax(counts))}
    for i in sorted(pick):
        m = mats[i]
        v0n, d, fm = m["pb"]
        t = m["sab"]
        comp = oracle.composition(
            [oracle.elinc(m["elinc"]["msd"], m["elinc"]["bixs"], m["elinc"]["scale"]),
             oracle.powderbragg(v0n, d, fm),
             oracle.sab(t["alpha"], t["beta"], t["sab"], t["temperature"], t["mass"], t["sigma_bound"])],
            [1.0, 1.0, 1.0])
        sl = slice(int(off[i]), int(off[i + 1]))
        ei = e[sl]
        assert _rel(x[sl], comp.xs(ei, nthreads=8)).max() <= 1e-12, i
        re_, rm, nerr = comp.sample_iso(ei, seed=1000 + i, stream=0, nthreads=8)
        assert nerr == 0
        assert_streams_match((_rel(ef[sl], re_) > 1e-9) | (np.abs(mu[sl] - rm) > 1e-9),
                             f"material {i}: {nplanes[i]} planes, grid {t['alpha'].size}x{t['beta'].size}")
