"""Entry points captured in a hipGraph (torch.cuda.graph) with caller
workspaces, replayed with new inputs copied into the captured buffers and
with eager work between the replays; every replay equals the oracle.  The
library records no memset nodes (zeroing is a kernel), so replays after the
first see the same state as the first."""
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


def _fr_buf(n):
    import torch
    return torch.empty((n, 4), dtype=torch.int64, device="cuda")


def _work(nbytes):
    return _fr_buf(max(1, (nbytes + 31) // 32))


def test_graph_replays_transforms_and_scans(pl):
    import torch
    lg, lb = 14, 2
    n = 1 << lg
    dom = pl.Domain(lg)
    ldom = pl.LdeDomain(lg, lb)
    ks = [po.from_u64(1), po.from_u64(7), po.from_u64(13)]
    karr = np.concatenate(ks)
    # captured buffers
    x, y = _fr_buf(n), _fr_buf(n)
    c, lde_out = _fr_buf(n), _fr_buf(n << lb)
    cols = [_fr_buf(n) for _ in range(6)]
    beta, gamma, z = _fr_buf(1), _fr_buf(1), _fr_buf(n)
    a, dc, rem = _fr_buf(n), _fr_buf(1), _fr_buf(1)
    w_ntt = _work(pl.ntt_work_bytes(lg))
    w_lde = _work(ldom.work_bytes)
    w_gp = _work(pl.lib().pl_vec_work_bytes(n))
    w_div = _work(pl.lib().pl_scan_work_bytes(n))

    def inputs(rep):
        h = {"x": po.uniform(n, 500 + rep), "c": po.uniform(n, 600 + rep),
             "cols": [po.uniform(n, 700 + 10 * rep + j) for j in range(6)],
             "beta": po.uniform(1, 800 + rep), "gamma": po.uniform(1, 900 + rep),
             "a": po.uniform(n, 1000 + rep), "dc": po.uniform(1, 1100 + rep)}
        x.copy_(pl.to_dev(h["x"]))
        c.copy_(pl.to_dev(h["c"]))
        for j in range(6):
            cols[j].copy_(pl.to_dev(h["cols"][j]))
        beta.copy_(pl.to_dev(h["beta"]))
        gamma.copy_(pl.to_dev(h["gamma"]))
        a.copy_(pl.to_dev(h["a"]))  # the division works in place
        dc.copy_(pl.to_dev(h["dc"]))
        return h

    def calls(s):
        pl.ntt(x, dom, out=y, work=w_ntt, stream=s)
        pl.lde(c, ldom, out=lde_out, work=w_lde, stream=s)
        pl.perm_ratio(cols[:3], cols[3:], beta, gamma, karr, dom, out=z, grand_product=True, work=w_gp, stream=s)
        pl.div_linear(a, dc, rem=rem, work=w_div, stream=s)

    s = torch.cuda.Stream()
    inputs(0)
    torch.cuda.synchronize()
    s.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(s):
        calls(s)  # warm-up outside the capture
    torch.cuda.current_stream().wait_stream(s)
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g, stream=s):
        calls(s)
    for rep in range(1, 4):
        h = inputs(rep)
        torch.cuda.synchronize()
        print("graph replay", rep, flush=True)
        g.replay()
        torch.cuda.synchronize()
        np.testing.assert_array_equal(pl.to_host(y), po.ntt(h["x"], lg))
        np.testing.assert_array_equal(pl.to_host(lde_out), po.lde(h["c"], lg, lb, po.ref_lde_shift(lb)))
        np.testing.assert_array_equal(
            pl.to_host(z), po.grand_product(po.perm_ratio(h["cols"], h["beta"], h["gamma"], ks, lg)))
        q, r = po.div(h["a"], h["dc"])
        np.testing.assert_array_equal(pl.to_host(a), q)
        np.testing.assert_array_equal(pl.to_host(rem), r)
        # eager work between replays, on the default stream and its own buffers
        e = pl.ntt(pl.to_dev(po.uniform(n, 77)), dom)
        pl.div_linear(e, pl.to_dev(po.uniform(1, 78)))
        torch.cuda.synchronize()
