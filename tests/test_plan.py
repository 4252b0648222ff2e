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


def test_host_groups_cache_and_set(pl, tmp_path):
    # copy-group counts of the host-to-host NTT: default 2, set/restore,
    # and "h <lg> <lg_k> <ms>" cache lines (malformed or out-of-range ones ignored)
    L = pl.lib()
    assert L.pl_ntt_host_groups(24) == 2
    assert L.pl_ntt_host_groups(0) == -1 and L.pl_ntt_host_groups(29) == -1
    assert L.pl_ntt_host_groups_set(24, 5) == 0 and L.pl_ntt_host_groups(24) == 5
    assert L.pl_ntt_host_groups_set(24, -1) == 0 and L.pl_ntt_host_groups(24) == 2
    assert L.pl_ntt_host_groups_set(24, 21) == -1 and L.pl_ntt_host_groups_set(30, 1) == -1
    p = tmp_path / "plans.txt"
    p.write_text("16 3 6 5 5 0.02\n# host\nh 20 3 4.1\nh 22 0 9.9\nh 23 99 1.0\nh x 1\nh 24\n")
    assert L.pl_plan_load(str(p).encode()) == 3  # one split, two copy-group counts
    assert pl.host_groups(20) == 3 and pl.host_groups(22) == 0
    assert pl.host_groups(23) == 2 and pl.host_groups(24) == 2
    for lg in (20, 22):
        L.pl_ntt_host_groups_set(lg, -1)
    L.pl_plan_set(16, None, 0)
    assert L.pl_autotune(12, 13, None, 3, 8) == -1  # unknown flag
