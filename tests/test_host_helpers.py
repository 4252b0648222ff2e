"""The host bit-reverse helpers of the operator API (calculate_map1,
bit_rev_step1_parallel; PLONK/utils/function.cuh:157-244, function.cu:503-509),
run through `api_check --host-helpers`, which makes no GPU call.  Expected
values come from a numpy bit reversal."""
import os
import subprocess

import numpy as np
import pytest

from conftest import ROOT

BIN = os.path.join(ROOT, "ppopp26-artifact_amd", "bin", "api_check")


def bitrev(i, bits):
    i = np.asarray(i, dtype=np.uint64)
    r = np.zeros_like(i)
    for b in range(bits):
        r |= ((i >> np.uint64(b)) & np.uint64(1)) << np.uint64(bits - 1 - b)
    return r


@pytest.mark.skipif(not os.path.exists(BIN), reason="api_check not built")
@pytest.mark.parametrize("lg", [1, 5, 12])
def test_calculate_map1_and_bit_rev_step1(tmp_path, lg):
    p = subprocess.run([BIN, "--host-helpers", str(tmp_path), str(lg)], capture_output=True, text=True, timeout=60)
    assert p.returncode == 0, p.stderr
    n = 1 << lg
    x = np.fromfile(tmp_path / "in.bin", dtype=np.uint64).reshape(-1, 4)
    assert x.shape == (n, 4) and len(np.unique(x, axis=0)) == n
    m = np.fromfile(tmp_path / "map.bin", dtype=np.uint32)
    np.testing.assert_array_equal(m, bitrev(np.arange(n), lg))
    want = np.zeros_like(x)
    want[bitrev(np.arange(n), lg)] = x
    for t in (1, 3, 8):
        np.testing.assert_array_equal(np.fromfile(tmp_path / ("brev_t%d.bin" % t), dtype=np.uint64).reshape(-1, 4),
                                      want)
    want2 = np.zeros((2 * n, 4), dtype=np.uint64)
    want2[bitrev(n + np.arange(n), lg + 1)] = x
    np.testing.assert_array_equal(np.fromfile(tmp_path / "brev_idx.bin", dtype=np.uint64).reshape(-1, 4), want2)
