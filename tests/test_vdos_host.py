"""The library's host-side VDOS functions (ncamd_vdos_regularise,
ncamd_vdos_analyse: regulariseVDOSGrid and VDOSEval,
ncrystal_core/src/vdos/NCVDOSEval.cc:147-640) against the CPU checker, whose
values tests/test_vdos_pins.py pins to the reference's logs: identical
results on every density of tests/golden/vdos_reference.json, including the
Debye spectrum given by its temperature, and BadInput for what the reference
refuses.  No GPU is used by these calls."""
import json
import pathlib

import numpy as np
import pytest

FIX = json.loads((pathlib.Path(__file__).parent / "golden" / "vdos_reference.json").read_text())


@pytest.fixture(scope="module")
def lib():
    from ncrystal_amd import _capi, core
    return _capi, core


def _inputs(oracle, key):
    v = FIX["vdos"][key]
    if v["kind"] == "debye":
        return v["debye_temp"], oracle.vdos_debye(v["debye_temp"]), v["mass"]
    eg, d = np.array(v["egrid"]), np.array(v["density"])
    return (eg, d), (eg, d), v["mass"]


def test_regularise_identical(lib, oracle):
    C, core = lib
    for key in FIX["vdos"]:
        vin, _, _ = _inputs(oracle, key)
        if np.isscalar(vin):
            lo, hi, d = oracle.vdos_debye(vin)
            ref = oracle.vdos_regularise(np.array([lo, hi]), d)
        else:
            ref = oracle.vdos_regularise(*vin)
        got = core.vdos_regularise(vin)
        assert got[0] == ref[0] and got[1] == ref[1] and np.array_equal(got[2], ref[2]), key


def test_analyse_identical(lib, oracle):
    C, core = lib
    for r in FIX["analyse"]:
        vin, _, mass = _inputs(oracle, r["vdos"])
        if np.isscalar(vin):
            lo, hi, d = oracle.vdos_debye(vin)
            lo, hi, d = oracle.vdos_regularise(np.array([lo, hi]), d)
        else:
            lo, hi, d = oracle.vdos_regularise(*vin)
        ref = oracle.vdos_analyse(lo, hi, d, r["temperature"], mass)
        got = core.vdos_analyse(vin, r["temperature"], mass)
        assert got == ref, (r["vdos"], got, ref)


def test_refusals(lib):
    C, core = lib
    with pytest.raises(C.NCBadInput):  # below 1e-5 eV
        core.vdos_regularise((np.array([5e-6, 0.03]), np.linspace(0.1, 1.0, 20)))
    with pytest.raises(C.NCBadInput):  # too few points
        core.vdos_regularise((np.array([0.01, 0.03]), np.ones(4)))
    with pytest.raises(C.NCBadInput):  # grid and density sizes
        core.vdos_regularise((np.linspace(0.01, 0.03, 7), np.ones(8)))
    with pytest.raises(C.NCBadInput):  # mass
        core.vdos_analyse(400.0, 300.0, 0.1)
