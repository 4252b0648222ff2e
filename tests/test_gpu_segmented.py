"""GPU parity of the segmented NTT step (pl_ntt_stage_no_rotate, the
reference's Ntt::stage_no_rotate) against the oracle's restatement, through
the C-ABI, bit-exact; and the driver's composition of such calls on the GPU
equal to the NTT."""
import numpy as np
import pytest

import pyoracle as po
from test_segmented import composed_ntt

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def pl():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a HIP device"
    import plffi
    plffi.lib()
    return plffi


@pytest.fixture(scope="module")
def params(pl):
    return pl.params()


# (lg_domain, iterations, stage, lg_chunk, chunk_id): the three factor rules,
# blocks smaller than a radix block, iterations 1..8 and 10, chunk ids > 0
CASES = [
    (12, 6, 0, 12, 0), (12, 6, 0, 9, 3), (12, 6, 6, 10, 2), (20, 6, 6, 12, 5),
    (12, 5, 6, 8, 1), (18, 6, 12, 10, 7), (14, 7, 7, 12, 1), (16, 8, 8, 12, 9),
    (16, 10, 6, 11, 0), (22, 10, 12, 12, 100), (13, 7, 3, 8, 4), (12, 3, 7, 4, 17),
    (10, 1, 9, 1, 300), (10, 2, 0, 2, 5), (24, 8, 16, 14, 777), (28, 10, 18, 12, 2000),
]


@pytest.mark.parametrize("lg,it,stage,lgc,cid", CASES)
def test_stage_vs_oracle(pl, params, lg, it, stage, lgc, cid):
    x = po.uniform(1 << lgc, po.seed(po.OP_NTT, lg, stage * 16 + it))
    print("stage_no_rotate lg=%d it=%d stage=%d chunk 2^%d id %d" % (lg, it, stage, lgc, cid), flush=True)
    d = pl.to_dev(x)
    pl.stage_no_rotate(d, lg, it, stage, cid, params)
    np.testing.assert_array_equal(pl.to_host(d), po.stage_no_rotate(x, lg, it, stage, cid))


@pytest.mark.parametrize("lg,schedule,lgc", [(12, [(0, 6), (6, 6)], 9), (16, [(0, 8), (8, 8)], 12),
                                             (18, [(0, 6), (6, 6), (12, 6)], 14), (20, [(0, 10), (10, 10)], 16)])
def test_driver_composition_on_gpu(pl, params, lg, schedule, lgc):
    x = po.uniform(1 << lg, po.seed(po.OP_NTT, lg, 99))
    print("composed lg=%d %s" % (lg, schedule), flush=True)

    def step(chunk, lg_, it, stage, cid):
        d = pl.to_dev(chunk)
        pl.stage_no_rotate(d, lg_, it, stage, cid, params)
        return pl.to_host(d)

    np.testing.assert_array_equal(composed_ntt(x, lg, schedule, lgc, step), po.ntt(x, lg))


def test_argument_errors(pl, params):
    import ctypes
    L = pl.lib()
    d = pl.to_dev(po.uniform(64, 1))
    ptr = ctypes.c_void_p(d.data_ptr())
    pp = ctypes.c_void_p(params.data_ptr())
    assert L.pl_ntt_stage_no_rotate(ptr, 64, 20, 9, 0, 0, pp, None) == -3   # radix 9
    assert L.pl_ntt_stage_no_rotate(ptr, 64, 10, 6, 6, 0, pp, None) == -1   # stage + it > lg
    assert L.pl_ntt_stage_no_rotate(ptr, 32, 10, 6, 0, 0, pp, None) == -1   # chunk < 2^it
    assert L.pl_ntt_stage_no_rotate(ptr, 48, 10, 4, 0, 0, pp, None) == -1   # not a power of two
    assert L.pl_ntt_stage_no_rotate(ptr, 64, 10, 6, 0, 0, None, None) == -1
