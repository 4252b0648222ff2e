"""GPU parity of the LDE, grand product, evaluation and division against the
oracle and the golden vectors, through the C-ABI.  Bit-exact."""
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


def test_lde_golden_small(pl, golden):
    k = golden("lde_small.npz")
    for lg_n in range(1, 6):
        c = pl.to_dev(k["in_%d" % lg_n])
        for lb in range(1, 4):
            print("lde golden", lg_n, lb, flush=True)
            ref = pl.LdeDomain(lg_n, lb)
            np.testing.assert_array_equal(pl.to_host(pl.lde(c, ref)), k["lde_ref_%d_%d" % (lg_n, lb)])
            g = pl.LdeDomain(lg_n, lb, shift=po.from_u64(7))
            np.testing.assert_array_equal(pl.to_host(pl.lde(c, g)), k["lde_g_%d_%d" % (lg_n, lb)])


@pytest.mark.parametrize("lg_n,lb", [(8, 1), (9, 2), (12, 2), (12, 3), (12, 4), (16, 2), (17, 1), (18, 4),
                                     (20, 2), (20, 3), (22, 2), (22, 3), (24, 2), (24, 3)])
def test_lde_vs_oracle(pl, lg_n, lb):
    """our_lde bit-exact over the whole output, up to the benchmark sizes
    (2^24 -> 2^26 and 2^27; the oracle's coset transforms run on the host's
    threads)."""
    from conftest import assert_fr_equal
    import torch
    print("lde %d %d: GPU" % (lg_n, lb), flush=True)
    c = po.uniform(1 << lg_n, po.seed(po.OP_LDE, lg_n, lb))
    dom = pl.LdeDomain(lg_n, lb)
    out = pl.to_host(pl.lde(pl.to_dev(c), dom))
    dom.close()
    torch.cuda.empty_cache()
    print("lde %d %d: oracle" % (lg_n, lb), flush=True)
    assert_fr_equal(out, po.lde(c, lg_n, lb, po.ref_lde_shift(lb)), "lde %d %d" % (lg_n, lb))


@pytest.mark.parametrize("lg_n", [1, 3, 6, 7, 12, 16, 20, 22])
def test_lde_blowup3_vs_oracle(pl, lg_n):
    """The blowup-3 LDE (3 coset NTTs of size N with shifts 7 w_3N^t, output
    interleaved to the natural order of the 3N-point domain) against the
    oracle, whose closed form is pinned by brute force; at 2^20 and 2^22 the
    evaluation-domain pipeline (iNTT, times 7^i, coset NTTs) as well."""
    from conftest import assert_fr_equal
    import torch
    c = po.uniform(1 << lg_n, po.seed(po.OP_LDE, lg_n, 3))
    print("lde blowup 3 lg=%d" % lg_n, flush=True)
    dom = pl.LdeDomain(lg_n, 0, blowup3=True)
    assert dom.blowup == 3
    out = pl.to_host(pl.lde(pl.to_dev(c), dom))
    assert_fr_equal(out, po.lde3(c, lg_n, po.from_u64(7)), "lde3 %d" % lg_n)
    if lg_n >= 20:
        coeffs = pl.ntt(pl.to_dev(c), pl.Domain(lg_n), inverse=True)
        out = pl.to_host(pl.lde(coeffs, dom))
        assert_fr_equal(out, po.lde3(po.ntt(c, lg_n, inverse=True), lg_n, po.from_u64(7)), "evals->lde3")
    dom.close()
    torch.cuda.empty_cache()


@pytest.mark.parametrize("lg_n,lb", [(10, 1), (12, 3), (16, 2), (20, 2)])
def test_lde_naive_vs_oracle(pl, lg_n, lb):
    """naive_lde (lde.cu:1367-1490): out[k] = NTT_N(c)[k mod N]"""
    c = po.uniform(1 << lg_n, po.seed(po.OP_LDE, lg_n, 100 + lb))
    print("lde naive %d %d" % (lg_n, lb), flush=True)
    out = pl.to_host(pl.lde_naive(pl.to_dev(c), pl.Domain(lg_n), lb))
    np.testing.assert_array_equal(out, po.naive_lde(c, lg_n, lb))


@pytest.mark.parametrize("lg_n,lb", [(12, 1), (20, 1), (20, 2), (22, 1), (22, 2), (24, 1)])
def test_evals_to_coset_lde_pipeline(pl, lg_n, lb):
    """The prover pipeline on evaluations: iNTT -> multiply coefficient i by
    7^i -> zero-pad to 2^lb N -> NTT of size 2^lb N, i.e. pl_ntt(inverse)
    then pl_lde with shift 7, against the oracle's composition."""
    from conftest import assert_fr_equal
    import torch
    e = po.uniform(1 << lg_n, po.seed(po.OP_LDE, lg_n, 200 + lb))
    print("evals -> coset lde %d %d" % (lg_n, lb), flush=True)
    coeffs = pl.ntt(pl.to_dev(e), pl.Domain(lg_n), inverse=True)
    out = pl.to_host(pl.lde(coeffs, pl.LdeDomain(lg_n, lb, shift=po.from_u64(7))))
    torch.cuda.empty_cache()
    want = po.lde(po.ntt(e, lg_n, inverse=True), lg_n, lb, po.from_u64(7))
    assert_fr_equal(out, want, "evals->lde %d %d" % (lg_n, lb))


@pytest.mark.parametrize("lg_n", [12, 20, 22, 24])
def test_evals_to_blowup3_pipeline(pl, lg_n):
    """The same pipeline with blowup 3: iNTT -> the three cosets 7 w_3N^t of
    size N (pl_lde3_domain_create, shift 7), against the oracle's composition."""
    from conftest import assert_fr_equal
    import torch
    e = po.uniform(1 << lg_n, po.seed(po.OP_LDE, lg_n, 300))
    print("evals -> blowup-3 lde %d" % lg_n, flush=True)
    coeffs = pl.ntt(pl.to_dev(e), pl.Domain(lg_n), inverse=True)
    dom = pl.LdeDomain(lg_n, 0, shift=po.from_u64(7), blowup3=True)
    out = pl.to_host(pl.lde(coeffs, dom))
    dom.close()
    torch.cuda.empty_cache()
    want = po.lde3(po.ntt(e, lg_n, inverse=True), lg_n, po.from_u64(7))
    assert_fr_equal(out, want, "evals->lde3 %d" % lg_n)


@pytest.mark.parametrize("lg_n,lb", [(12, 2), (16, 3)])
def test_lde_extreme_inputs(pl, lg_n, lb):
    """all coefficients r - 1: the largest sums through the coset passes"""
    c = np.zeros((1 << lg_n, 4), dtype=np.uint64)
    c[:] = [0xFFFFFFFF00000000, 0x53BDA402FFFE5BFE, 0x3339D80809A1D805, 0x73EDA753299D7D48]
    print("lde extreme", lg_n, lb, flush=True)
    dom = pl.LdeDomain(lg_n, lb)
    out = pl.to_host(pl.lde(pl.to_dev(c), dom))
    np.testing.assert_array_equal(out, po.lde(c, lg_n, lb, po.ref_lde_shift(lb)))


def test_lde_in_place(pl):
    import torch
    lg_n, lb = 10, 2
    c = po.uniform(1 << lg_n, 99)
    buf = torch.zeros(((1 << lg_n) << lb, 4), dtype=torch.int64, device="cuda")
    buf[: 1 << lg_n] = pl.to_dev(c)
    dom = pl.LdeDomain(lg_n, lb)
    pl.lde(buf[: 1 << lg_n], dom, out=buf)
    np.testing.assert_array_equal(pl.to_host(buf), po.lde(c, lg_n, lb, po.ref_lde_shift(lb)))


@pytest.mark.parametrize("lg_n,lb", [(22, 2), (24, 2), (24, 3), (24, 4)])
def test_lde_large_properties(pl, lg_n, lb):
    # the benchmark sizes: out[j*B + t] is the size-N coset NTT with shift
    # 7^B * w_M^t (B = 2^lb); the inverse NTT of one coset column gives
    # c_i * shift_t^i.  Check the first and the last coset at five positions.
    import torch
    n = 1 << lg_n
    B = 1 << lb
    c = po.uniform(n, po.seed(po.OP_LDE, lg_n, lb))
    cd = pl.to_dev(c)
    print("lde large", lg_n, flush=True)
    out = pl.lde(cd, pl.LdeDomain(lg_n, lb))
    dom = pl.Domain(lg_n)
    for t in (0, B - 1):
        col = out[t::B].contiguous()
        back = pl.to_host(pl.ntt(col, dom, inverse=True))
        # back[i] = c_i * s_t^i ; compare at a few positions against the oracle
        s = po.mul(po.ref_lde_shift(lb), po.pow_(po.root_of_unity(lg_n + lb), t))
        for i in (0, 1, 2, n // 3, n - 1):
            np.testing.assert_array_equal(back[i:i + 1], po.mul(c[i:i + 1], po.pow_(s, i)))
    del out
    torch.cuda.synchronize()
    torch.cuda.empty_cache()


def test_grand_product_golden(pl, golden):
    k = golden("scan_small.npz")
    for n in (1, 2, 3, 64, 257):
        p = pl.to_dev(k["gp_in_%d" % n])
        pl.grand_product(p)
        np.testing.assert_array_equal(pl.to_host(p), k["gp_out_%d" % n])


@pytest.mark.parametrize("n", [4095, 4096, 4097, 100000, 1 << 20, (1 << 24) + 5])
def test_grand_product_vs_oracle(pl, n):
    import torch
    print("gp", n, flush=True)
    p = po.uniform(n, po.seed(po.OP_GP, 0, n & 0xff))
    pd = pl.to_dev(p)
    incl = torch.empty_like(pd)
    pl.grand_product(pd, incl_out=incl)
    exp = po.grand_product(p)
    got = pl.to_host(pd)
    np.testing.assert_array_equal(got, exp)
    # inclusive = exclusive shifted by one, last = total product
    gi = pl.to_host(incl)
    np.testing.assert_array_equal(gi[:-1], exp[1:])
    np.testing.assert_array_equal(gi[-1:], po.mul(exp[-1:], p[-1:]))


def test_grand_product_zeros(pl):
    # zeros in the input (a quarter of the reference's input.bin is zero)
    n = 9000
    p = po.uniform(n, 5)
    p[100] = 0
    p[5000] = 0
    pd = pl.to_dev(p)
    pl.grand_product(pd)
    np.testing.assert_array_equal(pl.to_host(pd), po.grand_product(p))


def test_eval_golden(pl, golden):
    k = golden("scan_small.npz")
    for n in (1, 2, 3, 64, 100):
        c = pl.to_dev(k["eval_in_%d" % n])
        for tag in ("r", "0", "1"):
            y = pl.evaluate(c, pl.to_dev(k["eval_x_%s_%d" % (tag, n)]))
            np.testing.assert_array_equal(pl.to_host(y), k["eval_y_%s_%d" % (tag, n)])


@pytest.mark.parametrize("n", [4095, 4096, 4097, 65281, 123457, 255 * 256 * 17 + 1, 1 << 20, (1 << 20) + 5,
                               255 * 256 * 28, 255 * 256 * 28 + 1, 1 << 22, 1 << 24])
def test_eval_vs_oracle(pl, n):
    print("eval", n, flush=True)
    c = po.uniform(n, po.seed(po.OP_EVAL, 0, n & 0xff))
    x = po.uniform(1, 7)
    y = pl.evaluate(pl.to_dev(c), pl.to_dev(x))
    np.testing.assert_array_equal(pl.to_host(y), po.eval_(c, x))


def test_eval_div_random_mid_sizes_one_workspace(pl):
    """seeded random sizes over the whole mid-size range (every run length K =
    4..28, ragged last blocks), evaluation and division with remainder, one
    workspace sized for the largest reused by every call, against the oracle"""
    import random
    import torch
    rng = random.Random(31337)
    sizes = [rng.randrange(4096, 255 * 256 * 28 + 1) for _ in range(14)]
    work = torch.empty((pl.lib().pl_scan_work_bytes(max(sizes)) // 32 + 1, 4), dtype=torch.int64, device="cuda")
    for j, n in enumerate(sizes):
        print("mid-size eval/div n=%d" % n, flush=True)
        a = po.uniform(n, po.seed(po.OP_DIV, 1, j))
        x = po.uniform(1, 50 + j)
        ad = pl.to_dev(a)
        y = pl.evaluate(ad, pl.to_dev(x), work=work)
        np.testing.assert_array_equal(pl.to_host(y), po.eval_(a, x), "eval n=%d" % n)
        rem = torch.empty((1, 4), dtype=torch.int64, device="cuda")
        pl.div_linear(ad, pl.to_dev(x), rem=rem, work=work)
        q, r = po.div(a, x)
        np.testing.assert_array_equal(pl.to_host(ad), q, "div n=%d" % n)
        np.testing.assert_array_equal(pl.to_host(rem), r, "rem n=%d" % n)


def test_eval_reused_work_and_zero_point(pl):
    # one work buffer across calls with different points (the mid-size form
    # keeps a flag there that each call must clear), x = 0 and x = 1 included
    import torch
    n = (1 << 20) + 3
    c = po.uniform(n, 31)
    cd = pl.to_dev(c)
    work = torch.empty((pl.lib().pl_scan_work_bytes(n) // 32 + 1, 4), dtype=torch.int64, device="cuda")
    for x in (po.uniform(1, 1), po.from_u64(0), po.one(), po.uniform(1, 2)):
        y = pl.evaluate(cd, pl.to_dev(x), work=work)
        np.testing.assert_array_equal(pl.to_host(y), po.eval_(c, x))


def test_eval_graph_replays(pl):
    # captured once, replayed with a new point each time and eager work in
    # between; every replay equals the oracle
    import torch
    n = 1 << 18
    c = po.uniform(n, 32)
    cd = pl.to_dev(c)
    xd = pl.to_dev(po.uniform(1, 3))
    y = torch.zeros((1, 4), dtype=torch.int64, device="cuda")
    work = torch.empty((pl.lib().pl_scan_work_bytes(n) // 32 + 1, 4), dtype=torch.int64, device="cuda")
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(s):
        pl.evaluate(cd, xd, y=y, work=work, stream=s)
    torch.cuda.current_stream().wait_stream(s)
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g, stream=s):
        pl.evaluate(cd, xd, y=y, work=work, stream=s)
    for rep in range(4):
        x = po.uniform(1, 100 + rep)
        xd.copy_(pl.to_dev(x))
        y.zero_()
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        np.testing.assert_array_equal(pl.to_host(y), po.eval_(c, x))
        pl.evaluate(cd, pl.to_dev(po.uniform(1, 7)), work=torch.empty_like(work))  # eager work between replays


def test_eval_empty(pl):
    import torch
    y = torch.full((1, 4), 7, dtype=torch.int64, device="cuda")
    x = pl.to_dev(po.uniform(1, 3))
    pl.lib().pl_eval(None, 0, pl._ptr(x), pl._ptr(y), None, pl._stream(None))
    assert pl.to_host(y).sum() == 0


def test_div_golden(pl, golden):
    import torch
    k = golden("scan_small.npz")
    for n in (1, 2, 3, 64, 100):
        a = pl.to_dev(k["div_in_%d" % n])
        rem = torch.empty((1, 4), dtype=torch.int64, device="cuda")
        pl.div_linear(a, pl.to_dev(k["div_c_%d" % n]), rem=rem)
        np.testing.assert_array_equal(pl.to_host(a), k["div_q_%d" % n])
        np.testing.assert_array_equal(pl.to_host(rem), k["div_rem_%d" % n])


@pytest.mark.parametrize("n", [4095, 4096, 4097, 65281, 77777, 255 * 256 * 17 + 1, 1 << 20, (1 << 20) + 5,
                               255 * 256 * 28, 255 * 256 * 28 + 1, (1 << 24) + 3])
def test_div_vs_oracle(pl, n):
    import torch
    print("div", n, flush=True)
    a = po.uniform(n, po.seed(po.OP_DIV, 0, n & 0xff))
    c = po.uniform(1, 11)
    ad = pl.to_dev(a)
    cd = pl.to_dev(c)
    rem = torch.empty((1, 4), dtype=torch.int64, device="cuda")
    pl.div_linear(ad, cd, rem=rem)
    q, r = po.div(a, c)
    np.testing.assert_array_equal(pl.to_host(ad), q)
    np.testing.assert_array_equal(pl.to_host(rem), r)
    np.testing.assert_array_equal(pl.to_host(cd), c)  # divisor untouched


def test_div_reused_work_and_special_divisors(pl):
    # one work buffer across calls; c = 0 (division by X), c = -1 (X - 1)
    import torch
    n = (1 << 20) + 3
    a = po.uniform(n, 41)
    work = torch.empty((pl.lib().pl_scan_work_bytes(n) // 32 + 1, 4), dtype=torch.int64, device="cuda")
    for c in (po.uniform(1, 1), po.from_u64(0), po.neg(po.one()), po.uniform(1, 2)):
        ad = pl.to_dev(a)
        rem = torch.empty((1, 4), dtype=torch.int64, device="cuda")
        pl.div_linear(ad, pl.to_dev(c), rem=rem, work=work)
        q, r = po.div(a, c)
        np.testing.assert_array_equal(pl.to_host(ad), q)
        np.testing.assert_array_equal(pl.to_host(rem), r)


def test_miss_ops_literals(pl, golden):
    import json, os
    from conftest import GOLDEN
    consts = json.load(open(os.path.join(GOLDEN, "ref_constants.json")))
    k = golden("scan_small.npz")
    pt = int(consts["miss_ops_point"], 16)
    x = np.array([[(pt >> (64 * i)) & 0xFFFFFFFFFFFFFFFF for i in range(4)]], dtype=np.uint64)
    poly = pl.to_dev(k["mo_poly"])
    np.testing.assert_array_equal(pl.to_host(pl.evaluate(poly, pl.to_dev(x))), k["mo_eval"])
    cv = int(consts["miss_ops_divisor_gpu"], 16)
    c = np.array([[(cv >> (64 * i)) & 0xFFFFFFFFFFFFFFFF for i in range(4)]], dtype=np.uint64)
    pl.div_linear(poly, pl.to_dev(c))
    np.testing.assert_array_equal(pl.to_host(poly), k["mo_div_q"])
