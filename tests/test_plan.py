"""Pass-split plans and the autotuner's cache, host side (no device)."""
import os

import pytest

from conftest import ROOT  # noqa: F401


@pytest.fixture()
def pl():
    import plffi
    if not os.path.exists(plffi.LIB_PATH):
        pytest.skip("libpl.so not built")
    plffi.lib()
    yield plffi
    for lg in range(4, 29):  # back to the default splits
        plffi.plan_set(lg, [])


def test_default_plans(pl):
    # ceil(lg/6) passes of 2^3..2^6-point columns, sizes within one, larger first
    assert pl.plan_get(24) == [6, 6, 6, 6]
    assert pl.plan_get(20) == [5, 5, 5, 5]
    assert pl.plan_get(13) == [5, 4, 4]
    assert pl.plan_get(12) == [6, 6]
    assert pl.plan_get(9) == [5, 4]
    assert pl.plan_get(6) == [6]
    assert pl.plan_get(3) == []  # tiny kernel


def test_candidates_are_valid_splits(pl):
    for lg in range(4, 29):
        cands = pl.plan_candidates(lg)
        assert cands, lg
        assert pl.plan_get(lg) in cands  # the default is among them
        for m in cands:
            assert sum(m) == lg and all(3 <= x <= 6 for x in m), (lg, m)
        assert len(set(map(tuple, cands))) == len(cands)


def test_plan_set_and_errors(pl):
    pl.plan_set(24, [6, 6, 4, 4, 4])
    assert pl.plan_get(24) == [6, 6, 4, 4, 4]
    assert pl.lib().pl_ntt_num_passes(24) == 5
    pl.plan_set(24, [])
    assert pl.plan_get(24) == [6, 6, 6, 6]
    import ctypes
    bad = (ctypes.c_int * 8)(6, 6, 6, 5)
    assert pl.lib().pl_plan_set(24, bad, 4) == -1          # wrong sum
    bad = (ctypes.c_int * 8)(8, 8, 8)
    assert pl.lib().pl_plan_set(24, bad, 3) == -1          # m > 6
    bad = (ctypes.c_int * 8)(2, 4)
    assert pl.lib().pl_plan_set(6, bad, 2) == -1           # m < 3


def test_plan_cache_file(pl, tmp_path):
    p = tmp_path / "plans.txt"
    p.write_text("# comment\n20 4 6 6 4 4 0.1234\n13 3 5 5 3 0.01\n12 3 4 4 4 0.5\n14 2 8 6 1.0\nbad line\n")
    assert pl.lib().pl_plan_load(str(p).encode()) == 3     # the 8-point split is rejected
    assert pl.plan_get(20) == [6, 6, 4, 4]
    assert pl.plan_get(13) == [5, 5, 3]
    assert pl.plan_get(12) == [4, 4, 4]
    assert pl.plan_get(14) == [5, 5, 4]
    assert pl.lib().pl_plan_load(str(tmp_path / "missing.txt").encode()) == -1


def test_autotune_argument_errors(pl):
    L = pl.lib()
    assert L.pl_autotune(3, 10, None, 3, 0) == -1
    assert L.pl_autotune(12, 11, None, 3, 0) == -1
    assert L.pl_autotune(12, 13, None, 0, 0) == -1
    assert L.pl_autotune(12, 13, None, 3, 4) == -1
    try:
        import torch
        if torch.cuda.is_available():
            return
    except ImportError:
        pass
    assert L.pl_autotune(12, 13, None, 3, 0) == -5
