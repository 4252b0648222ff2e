"""GPU checks of the pass splits: every candidate split computes the same
exact transform, and the autotuner's choices are applied and cached."""
import numpy as np
import pytest

import pyoracle as po

pytestmark = pytest.mark.gpu


@pytest.fixture()
def pl():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a HIP device"
    import plffi
    plffi.lib()
    yield plffi
    for lg in range(4, 29):
        plffi.plan_set(lg, [])


@pytest.mark.parametrize("lg", [8, 9, 13, 16])
def test_every_split_is_exact(pl, lg):
    x = po.uniform(1 << lg, po.seed(po.OP_NTT, lg, 77))
    ref = po.ntt(x, lg)
    refi = po.ntt(x, lg, inverse=True)
    xd = pl.to_dev(x)
    for m in pl.plan_candidates(lg):
        print("split", lg, m, flush=True)
        pl.plan_set(lg, m)
        dom = pl.Domain(lg)
        assert pl.ntt_num_passes(lg) == len(m)
        np.testing.assert_array_equal(pl.to_host(pl.ntt(xd, dom)), ref)
        np.testing.assert_array_equal(pl.to_host(pl.ntt(xd, dom, inverse=True)), refi)


def test_autotune_applies_and_caches(pl, tmp_path):
    import time
    cache = str(tmp_path / "plans.txt")
    pl.autotune(10, 14, cache, reps=2)
    lines = [l.split() for l in open(cache) if not l.startswith("#")]
    got = {int(l[0]): [int(v) for v in l[2:2 + int(l[1])]] for l in lines}
    assert sorted(got) == [10, 11, 12, 13, 14]
    for lg, m in got.items():
        assert pl.plan_get(lg) == m
        assert m in pl.plan_candidates(lg)
    # a tuned domain is still exact
    x = po.uniform(1 << 14, 5)
    np.testing.assert_array_equal(pl.to_host(pl.ntt(pl.to_dev(x), pl.Domain(14))), po.ntt(x, 14))
    # second call reads the cache instead of measuring
    for lg in range(4, 29):
        pl.plan_set(lg, [])
    t0 = time.time()
    pl.autotune(10, 14, cache, reps=2)
    assert time.time() - t0 < 5.0
    for lg, m in got.items():
        assert pl.plan_get(lg) == m


def test_domain_keeps_its_split(pl):
    """A domain built under one split keeps it after the process-wide split
    changes: its pass count, the workspace size and a call with no workspace
    (the library allocates by the domain's own split) stay consistent."""
    lg = 6
    x = po.uniform(1 << lg, po.seed(po.OP_NTT, lg, 5))
    pl.plan_set(lg, [3, 3])
    dom = pl.Domain(lg)
    pl.plan_set(lg, [])
    assert pl.ntt_num_passes(lg) == 1 and pl.domain_num_passes(dom) == 2
    assert pl.ntt_work_bytes(lg) == 32 << lg
    print("split kept lg=6", flush=True)
    np.testing.assert_array_equal(pl.to_host(pl.ntt(pl.to_dev(x), dom)), po.ntt(x, lg))
    np.testing.assert_array_equal(pl.to_host(pl.ntt(pl.to_dev(x), dom, inverse=True)), po.ntt(x, lg, inverse=True))
    # passes 0 .. domain_num_passes - 1 equal the whole transform
    import torch
    xd = pl.to_dev(x)
    mid = torch.empty_like(xd)
    out = torch.empty_like(xd)
    pl.ntt_pass(xd, mid, dom, 0)
    pl.ntt_pass(mid, out, dom, 1)
    np.testing.assert_array_equal(pl.to_host(out), po.ntt(x, lg))
