"""pl_grand_product_noalloc (library-held scratch, no allocation) through the
C-ABI, against pl_grand_product with caller workspace: two calls in a row,
the second on a smaller input (the shared array keeps the first call's
partials), then the size limit."""
import ctypes

import numpy as np
import pytest

import pyoracle as po

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def pl():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a HIP device"
    import plffi
    plffi.lib()
    return plffi


def test_noalloc_grand_product_matches_the_workspace_form(pl):
    import torch
    for n, seed in ((3 << 20, 1), (5000, 2), (1 << 22, 3), (77, 4)):
        x = po.uniform(n, po.seed(po.OP_GP, 0, 900 + seed))
        print("grand product noalloc n=%d" % n, flush=True)
        a = pl.to_dev(x)
        inc_a = torch.zeros_like(a)
        pl.grand_product_noalloc(a, incl_out=inc_a)
        b = pl.to_dev(x)
        inc_b = torch.zeros_like(b)
        work = torch.empty(((pl.lib().pl_scan_work_bytes(n) + 31) // 32, 4), dtype=torch.int64, device="cuda")
        pl.grand_product(b, incl_out=inc_b, work=work)
        torch.cuda.synchronize()
        assert torch.equal(a, b) and torch.equal(inc_a, inc_b)
        if n <= 5000:
            np.testing.assert_array_equal(pl.to_host(a), po.grand_product(x))


def test_noalloc_rejects_sizes_beyond_the_array(pl):
    import torch
    t = torch.zeros((16, 4), dtype=torch.int64, device="cuda")
    L = pl.lib()
    # the size check comes before any device access
    assert L.pl_grand_product_noalloc(ctypes.c_void_p(t.data_ptr()), (1 << 28) + 1, None, None) == -3
    assert L.pl_grand_product_noalloc(None, 16, None, None) == -1
    assert L.pl_grand_product_noalloc(ctypes.c_void_p(t.data_ptr()), 0, None, None) == -1
