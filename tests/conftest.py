import pathlib
import sys

import pytest

ROOT = pathlib.Path(__file__).resolve().parents[1]
for p in (ROOT, ROOT / "oracle"):
    if str(p) not in sys.path:
        sys.path.insert(0, str(p))

GOLDEN = ROOT / "tests" / "golden"

# Per-neutron sample parity on the same Philox streams: device and host libm
# may differ in the last bit, which can flip a rejection test, so at most this
# fraction of neutrons may differ beyond 1e-9 (measured: none).
MISMATCH_FRAC = 1e-4


def assert_streams_match(bad, label, limit=MISMATCH_FRAC):
    """bad: per-neutron mask of outputs that differ beyond 1e-9; prints the
    count and the first differing indices, fails above `limit`."""
    import numpy as np
    bad = np.asarray(bad, dtype=bool).ravel()
    idx = np.flatnonzero(bad)
    frac = idx.size / max(bad.size, 1)
    print(f"{label}: {idx.size} of {bad.size} neutrons differ beyond 1e-9"
          + (f", first indices {idx[:20].tolist()}" if idx.size else ""))
    assert frac <= limit, (label, idx.size, bad.size, idx[:20].tolist())
    return frac


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


@pytest.fixture(scope="session")
def oracle():
    import pyoracle
    pyoracle.build()
    return pyoracle
