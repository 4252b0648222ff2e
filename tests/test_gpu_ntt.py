"""GPU parity of the NTT / iNTT against the oracle and the golden vectors,
through the C-ABI.  Bit-exact (integer arithmetic)."""
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


def test_ntt_golden_small(pl, golden):
    k = golden("ntt_small.npz")
    for lg in range(1, 9):
        dom = pl.Domain(lg)
        x = pl.to_dev(k["in_%d" % lg])
        y = pl.to_host(pl.ntt(x, dom))
        np.testing.assert_array_equal(y, k["ntt_%d" % lg], err_msg="ntt lg=%d" % lg)
        z = pl.to_host(pl.ntt(x, dom, inverse=True))
        np.testing.assert_array_equal(z, k["intt_%d" % lg], err_msg="intt lg=%d" % lg)


@pytest.mark.parametrize("lg", [9, 10, 12, 13, 16, 17, 20])
def test_ntt_vs_oracle(pl, lg):
    dom = pl.Domain(lg)
    x = po.uniform(1 << lg, po.seed(po.OP_NTT, lg))
    print("ntt lg=%d" % lg, flush=True)
    y = pl.to_host(pl.ntt(pl.to_dev(x), dom))
    np.testing.assert_array_equal(y, po.ntt(x, lg))
    z = pl.to_host(pl.ntt(pl.to_dev(x), dom, inverse=True))
    np.testing.assert_array_equal(z, po.ntt(x, lg, inverse=True))


R_MINUS_1 = [0xFFFFFFFF00000000, 0x53BDA402FFFE5BFE, 0x3339D80809A1D805, 0x73EDA753299D7D48]


@pytest.mark.parametrize("lg", [6, 9, 12, 16, 20])
@pytest.mark.parametrize("pattern", ["r-1", "alternating", "one-hot"])
def test_ntt_extreme_inputs(pl, lg, pattern):
    """Inputs at the top of the canonical range drive the largest sums through
    the passes' signed limbs (every DC sum of r - 1), against the oracle."""
    n = 1 << lg
    x = np.zeros((n, 4), dtype=np.uint64)
    if pattern == "r-1":
        x[:] = R_MINUS_1
    elif pattern == "alternating":
        x[0::2] = R_MINUS_1
    else:
        x[n // 3] = R_MINUS_1
    dom = pl.Domain(lg)
    print("ntt extreme lg=%d %s" % (lg, pattern), flush=True)
    np.testing.assert_array_equal(pl.to_host(pl.ntt(pl.to_dev(x), dom)), po.ntt(x, lg))
    np.testing.assert_array_equal(pl.to_host(pl.ntt(pl.to_dev(x), dom, inverse=True)), po.ntt(x, lg, inverse=True))


@pytest.mark.parametrize("lg", [22, 24])
def test_ntt_roundtrip_large(pl, lg):
    # size-independent properties at the benchmark sizes:
    # iNTT(NTT(x)) == x, and NTT at the first 64 outputs vs direct evaluation
    import torch
    dom = pl.Domain(lg)
    n = 1 << lg
    x = po.uniform(n, po.seed(po.OP_NTT, lg))
    xd = pl.to_dev(x)
    print("ntt roundtrip lg=%d" % lg, flush=True)
    y = pl.ntt(xd, dom)
    z = pl.ntt(y, dom, inverse=True)
    torch.cuda.synchronize()
    assert torch.equal(z, xd)
    yh = pl.to_host(y)
    # X[k] = x(w^k): check k = 0, 1, n-1 by Horner on the CPU
    w = po.root_of_unity(lg)
    for k in (0, 1, n - 1):
        pt = po.pow_(w, k)
        np.testing.assert_array_equal(yh[k:k + 1], po.eval_(x, pt))


@pytest.mark.parametrize("lg", [22, 24])
def test_ntt_full_vs_oracle_large(pl, lg):
    """Bit-exact over the whole vector at the benchmark sizes, forward and
    inverse (the oracle's radix-2 transform runs on the host's threads)."""
    from conftest import assert_fr_equal
    import torch
    dom = pl.Domain(lg)
    x = po.uniform(1 << lg, po.seed(po.OP_NTT, lg))
    xd = pl.to_dev(x)
    print("ntt full lg=%d: GPU" % lg, flush=True)
    y = pl.to_host(pl.ntt(xd, dom))
    yi = pl.to_host(pl.ntt(xd, dom, inverse=True))
    del xd
    torch.cuda.empty_cache()
    print("ntt full lg=%d: oracle on %d threads" % (lg, po.get_threads()), flush=True)
    assert_fr_equal(y, po.ntt(x, lg), "ntt lg=%d" % lg)
    assert_fr_equal(yi, po.ntt(x, lg, inverse=True), "intt lg=%d" % lg)


def test_ntt_in_place(pl):
    lg = 14
    dom = pl.Domain(lg)
    x = po.uniform(1 << lg, 3)
    xd = pl.to_dev(x)
    pl.ntt(xd, dom, out=xd)
    np.testing.assert_array_equal(pl.to_host(xd), po.ntt(x, lg))


def test_ntt_batch(pl):
    lg, npoly = 11, 5
    dom = pl.Domain(lg)
    x = po.uniform(npoly << lg, 4)
    y = pl.to_host(pl.ntt_batch(pl.to_dev(x), dom, npoly))
    for p in range(npoly):
        np.testing.assert_array_equal(y[p << lg:(p + 1) << lg], po.ntt(x[p << lg:(p + 1) << lg], lg))


def test_reference_input_prefix(pl, golden):
    # the reference's own input data (first 4096 elements of input.bin)
    import os
    from conftest import GOLDEN
    raw = np.fromfile(os.path.join(GOLDEN, "input_prefix.bin"), dtype=np.uint64).reshape(-1, 4)
    dom = pl.Domain(12)
    y = pl.to_host(pl.ntt(pl.to_dev(raw), dom))
    np.testing.assert_array_equal(y, po.ntt(raw, 12))


def test_ntt_roundtrip_2p26(pl):
    # the reference's largest NTT scale: round trip, and X[k] = x(w^k) at three
    # k by Horner on the CPU
    import torch
    lg = 26
    n = 1 << lg
    dom = pl.Domain(lg)
    x = po.uniform(n, po.seed(po.OP_NTT, lg))
    xd = pl.to_dev(x)
    print("ntt roundtrip lg=26", flush=True)
    y = pl.ntt(xd, dom)
    assert torch.equal(pl.ntt(y, dom, inverse=True), xd)
    w = po.root_of_unity(lg)
    for k in (1, 12345, n - 1):
        np.testing.assert_array_equal(pl.to_host(y[k:k + 1]), po.eval_(x, po.pow_(w, k)))
    # every output weighs in: sum_k X[k] w^(-jk) = N x_j (the inverse DFT at
    # one point, by the device evaluation, which the oracle pins up to 2^24)
    nn = po.from_u64(n)
    for j in (0, 1, 777, n - 5):
        pt = pl.to_dev(po.inv(po.pow_(w, j)))
        np.testing.assert_array_equal(pl.to_host(pl.evaluate(y, pt)), po.mul(nn, x[j:j + 1]))


def test_ntt_roundtrip_2p28(pl):
    # the largest supported size (8 GiB per vector): round trip, and
    # X[k] = x(w^k) through the device evaluation (itself checked against the
    # oracle up to 2^24)
    import torch
    lg = 28
    n = 1 << lg
    dom = pl.Domain(lg)
    xd = torch.empty((n, 4), dtype=torch.int64, device="cuda")
    pl.fill_uniform(xd, po.seed(po.OP_NTT, lg))
    print("ntt roundtrip lg=28", flush=True)
    y = pl.ntt(xd, dom)
    assert torch.equal(pl.ntt(y, dom, inverse=True), xd)
    w = po.root_of_unity(lg)
    for k in (1, n // 2 + 3):
        pt = pl.to_dev(po.pow_(w, k))
        np.testing.assert_array_equal(pl.to_host(pl.evaluate(xd, pt)), pl.to_host(y[k:k + 1]))
    # the inverse DFT at one point weighs every output: sum_k X[k] w^(-jk) = N x_j
    for j in (0, 5):
        pt = pl.to_dev(po.inv(po.pow_(w, j)))
        np.testing.assert_array_equal(pl.to_host(pl.evaluate(y, pt)), po.mul(po.from_u64(n), pl.to_host(xd[j:j + 1])))
    del xd, y
    torch.cuda.empty_cache()


def test_concurrent_calls_on_distinct_streams(pl):
    # the C-ABI is thread-safe for distinct streams: four host threads, each
    # with its own stream, domain and buffers, run transforms and scans at
    # once (ctypes releases the interpreter lock during the calls)
    import threading
    import torch
    lgs = [12, 14, 16, 18]
    xs = [po.uniform(1 << lg, 700 + lg) for lg in lgs]
    pts = [po.uniform(1, 800 + lg) for lg in lgs]
    got, err = [None] * 4, []

    def work(i):
        try:
            s = torch.cuda.Stream()
            lg = lgs[i]
            with torch.cuda.stream(s):
                dom = pl.Domain(lg)
                d = pl.to_dev(xs[i])
                outs = []
                for _ in range(5):
                    y = pl.ntt(d, dom, stream=s)
                    e = pl.evaluate(d, pl.to_dev(pts[i]), stream=s)
                    outs = [y, e]
                s.synchronize()
                got[i] = [pl.to_host(t) for t in outs]
        except Exception as ex:  # reported below
            err.append(repr(ex))

    ts = [threading.Thread(target=work, args=(i,)) for i in range(4)]
    for t in ts:
        t.start()
    for t in ts:
        t.join(timeout=120)
    assert not err, err
    for i, lg in enumerate(lgs):
        np.testing.assert_array_equal(got[i][0], po.ntt(xs[i], lg))
        np.testing.assert_array_equal(got[i][1], po.eval_(xs[i], pts[i]))
