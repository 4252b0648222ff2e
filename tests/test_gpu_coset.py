"""GPU parity of the coset transform (the reference's Ntt_coset /
Intt_coset surface: pl_coset_ntt, pl_coset_scale, pl_coset_ntt_step)
against the oracle, through the C-ABI.  Bit-exact."""
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


@pytest.mark.parametrize("lg", [1, 3, 5, 6, 7, 12, 13, 16, 20])
def test_coset_ntt_vs_oracle(pl, lg):
    x = po.uniform(1 << lg, po.seed(po.OP_LDE, lg, 600))
    g = po.from_u64(7)
    print("coset ntt lg=%d" % lg, flush=True)
    dom = pl.CosetDomain(lg)
    xd = pl.to_dev(x)
    y = pl.to_host(pl.coset_ntt(xd, dom))
    np.testing.assert_array_equal(y, po.coset_ntt(x, lg, g))
    np.testing.assert_array_equal(pl.to_host(pl.coset_ntt(xd, dom, inverse=True)), po.coset_intt(x, lg, g))
    # in place, and the round trip
    t = pl.to_dev(x)
    pl.coset_ntt(t, dom, out=t)
    pl.coset_ntt(t, dom, inverse=True, out=t)
    np.testing.assert_array_equal(pl.to_host(t), x)


def test_coset_ntt_other_shift_and_extremes(pl):
    lg = 14
    s = po.uniform(1, 77)
    x = np.zeros((1 << lg, 4), dtype=np.uint64)
    x[:] = [0xFFFFFFFF00000000, 0x53BDA402FFFE5BFE, 0x3339D80809A1D805, 0x73EDA753299D7D48]
    dom = pl.CosetDomain(lg, shift=s)
    print("coset ntt shift, r-1", flush=True)
    np.testing.assert_array_equal(pl.to_host(pl.coset_ntt(pl.to_dev(x), dom)), po.coset_ntt(x, lg, s))
    np.testing.assert_array_equal(pl.to_host(pl.coset_ntt(pl.to_dev(x), dom, inverse=True)), po.coset_intt(x, lg, s))


@pytest.mark.parametrize("lg", [4, 6, 11, 18])
def test_coset_three_steps_compose(pl, lg):
    """init-free three-call form (forward1/2/3): steps 1..3 equal the one call"""
    import torch
    x = po.uniform(1 << lg, po.seed(po.OP_LDE, lg, 601))
    dom = pl.CosetDomain(lg)
    for inverse in (False, True):
        print("coset steps lg=%d inv=%s" % (lg, inverse), flush=True)
        a = pl.to_dev(x)
        out = torch.empty_like(a)
        pl.coset_ntt_step(a, a, 1, dom, inverse)
        pl.coset_ntt_step(a, None, 2, dom, inverse)
        pl.coset_ntt_step(a, out, 3, dom, inverse)
        want = po.coset_intt(x, lg, po.from_u64(7)) if inverse else po.coset_ntt(x, lg, po.from_u64(7))
        np.testing.assert_array_equal(pl.to_host(out), want)


def test_coset_scale_chunks(pl):
    """init on chunks: chunk c of the vector times g^(c C + i); applied to
    every chunk it equals the coset powers on the whole vector"""
    lg, lgc = 12, 9
    x = po.uniform(1 << lg, 602)
    dom = pl.CosetDomain(lg)
    t = pl.to_dev(x)
    C = 1 << lgc
    for c in range(1 << (lg - lgc)):
        pl.coset_scale(t[c * C:(c + 1) * C], c, dom)
    g = po.from_u64(7)
    pw = np.concatenate([po.pow_(g, i) for i in range(1 << lg)])
    np.testing.assert_array_equal(pl.to_host(t), po.mul(x, pw))
    for c in range(1 << (lg - lgc)):
        pl.coset_scale(t[c * C:(c + 1) * C], c, dom, inverse=True)
    np.testing.assert_array_equal(pl.to_host(t), x)
    # the scaled vector's plain NTT is the coset NTT
    pl.coset_scale(t, 0, dom)
    np.testing.assert_array_equal(pl.to_host(pl.ntt(t, pl.Domain(lg))), po.coset_ntt(x, lg, g))


def test_coset_argument_errors(pl):
    import ctypes
    import torch
    L = pl.lib()
    dom = pl.CosetDomain(8)
    t = torch.zeros((256, 4), dtype=torch.int64, device="cuda")
    p = ctypes.c_void_p(t.data_ptr())
    assert L.pl_coset_scale(p, 512, 0, 0, dom.h, None) == -1   # chunk beyond the vector
    assert L.pl_coset_scale(p, 128, 2, 0, dom.h, None) == -1
    assert L.pl_coset_ntt_step(p, p, 3, 0, dom.h, None) == -1   # last step must not be in place
    assert L.pl_coset_ntt_step(p, p, 4, 0, dom.h, None) == -1
    h = ctypes.c_void_p()
    z = np.zeros(4, dtype=np.uint64)
    assert L.pl_coset_domain_create(8, ctypes.c_void_p(z.ctypes.data), ctypes.byref(h)) == -1  # shift 0
