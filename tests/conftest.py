import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")
for p in (ROOT, os.path.join(ROOT, "oracle"), os.path.join(ROOT, "ppopp26-artifact_amd")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device); runs through the C-ABI")


@pytest.fixture(scope="session")
def golden():
    def load(name):
        return np.load(os.path.join(GOLDEN, name), allow_pickle=False)
    return load


def assert_fr_equal(got, want, what=""):
    """Element-for-element equality of two (n, 4) uint64 arrays; on a
    mismatch, names the first differing element and the count."""
    got = np.asarray(got)
    want = np.asarray(want)
    assert got.shape == want.shape, "%s: shape %s vs %s" % (what, got.shape, want.shape)
    bad = np.any(got != want, axis=-1) if got.ndim > 1 else (got != want)
    if bad.any():
        i = int(np.argmax(bad))
        raise AssertionError("%s: %d of %d elements differ, first at %d: got %s want %s"
                             % (what, int(bad.sum()), bad.size, i, [hex(int(v)) for v in np.atleast_1d(got[i])],
                                [hex(int(v)) for v in np.atleast_1d(want[i])]))
