"""GPU parity of the element-wise operators, the zero-safe batch inversion and
the PLONK permutation ratio / grand product, through the C-ABI.  Bit-exact."""
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


def host_op(op, a, b):
    if op == "add":
        return po.add(a, b)
    if op == "sub":
        return po.sub(a, b)
    if op == "mul":
        return po.mul(a, b)
    if op == "div":
        return po.mul(a, po.inv(b))
    raise ValueError(op)


@pytest.mark.parametrize("op", ["add", "sub", "mul", "div"])
@pytest.mark.parametrize("n", [1, 17, 4096, 4097, 70000])
def test_binary_ops(pl, op, n):
    a = po.uniform(n, 100 + n)
    b = po.uniform(n, 200 + n)
    if n > 3:
        b[2] = 0  # x/0 = 0
    out = pl.to_host(pl.vec_op(op, pl.to_dev(a), pl.to_dev(b)))
    np.testing.assert_array_equal(out, host_op(op, a, b))


@pytest.mark.parametrize("op", ["add_scalar", "sub_scalar", "mul_scalar", "div_scalar"])
def test_scalar_ops(pl, op):
    n = 5000
    a = po.uniform(n, 300)
    s = po.uniform(1, 301)
    out = pl.to_host(pl.vec_op(op, pl.to_dev(a), pl.to_dev(s)))
    np.testing.assert_array_equal(out, host_op(op.split("_")[0], a, np.repeat(s, n, axis=0)))


def test_unary_ops(pl):
    n = 3000
    a = po.uniform(n, 400)
    a[5] = 0
    ad = pl.to_dev(a)
    np.testing.assert_array_equal(pl.to_host(pl.vec_op("neg", ad)), po.neg(a))
    np.testing.assert_array_equal(pl.to_host(pl.vec_op("inv", ad)), po.inv(a))
    tm = np.zeros((n, 4), dtype=np.uint64)
    for i in range(n):
        po.lib().or_to_mont(po.P(tm[i:i + 1]), po.P(np.ascontiguousarray(a[i:i + 1])))
    np.testing.assert_array_equal(pl.to_host(pl.vec_op("to_mont", ad)), tm)
    np.testing.assert_array_equal(pl.to_host(pl.vec_op("from_mont", pl.to_dev(tm))), a)


@pytest.mark.parametrize("n", [1, 2, 4096, 4097, (1 << 20) + 1, 3 << 22])
def test_batch_inverse_large(pl, n):
    # x * x^-1 = 1 except zeros, at sizes spanning one to three block levels
    import torch
    print("inv", n, flush=True)
    a = po.uniform(n, 500 + (n & 0xfff))
    a[n // 2] = 0
    ad = pl.to_dev(a)
    inv = pl.vec_op("inv", ad)
    prod = pl.to_host(pl.vec_op("mul", ad, inv))
    one = po.one()[0]
    expect = np.repeat(po.one(), n, axis=0)
    expect[n // 2] = 0
    np.testing.assert_array_equal(prod, expect)
    # spot values against the oracle
    idx = [0, n - 1, n // 3]
    ih = pl.to_host(inv)
    for i in idx:
        np.testing.assert_array_equal(ih[i:i + 1], po.inv(a[i:i + 1]))
    torch.cuda.synchronize()


def test_perm_ratio_golden(pl, golden):
    k = golden("scan_small.npz")
    ks = np.concatenate([po.from_u64(1), po.from_u64(7), po.from_u64(13)])
    for lg in (3, 6):
        dom = pl.Domain(lg)
        w = [pl.to_dev(k["perm_%s_%d" % (nm, lg)]) for nm in ("a", "b", "c")]
        s = [pl.to_dev(k["perm_%s_%d" % (nm, lg)]) for nm in ("s1", "s2", "s3")]
        out = pl.perm_ratio(w, s, pl.to_dev(k["perm_beta_%d" % lg]), pl.to_dev(k["perm_gamma_%d" % lg]), ks, dom)
        np.testing.assert_array_equal(pl.to_host(out), k["perm_ratio_%d" % lg])


def test_perm_ratio_extreme_columns(pl):
    """columns, beta and gamma at r - 1 (and zeros mixed in), against the oracle"""
    lg = 12
    n = 1 << lg
    rm1 = [0xFFFFFFFF00000000, 0x53BDA402FFFE5BFE, 0x3339D80809A1D805, 0x73EDA753299D7D48]
    cols = []
    for j in range(6):
        c = np.zeros((n, 4), dtype=np.uint64)
        c[j % 2::2] = rm1
        cols.append(c)
    beta = np.array([rm1], dtype=np.uint64)
    gamma = np.array([rm1], dtype=np.uint64)
    ks = [po.from_u64(1), po.from_u64(7), po.from_u64(13)]
    dom = pl.Domain(lg)
    dev = [pl.to_dev(c) for c in cols]
    ratio = pl.perm_ratio(dev[:3], dev[3:], pl.to_dev(beta), pl.to_dev(gamma), np.concatenate(ks), dom)
    np.testing.assert_array_equal(pl.to_host(ratio), po.perm_ratio(cols, beta, gamma, ks, lg))


@pytest.mark.parametrize("lg", [12, 20])
def test_perm_grand_product_vs_oracle(pl, lg):
    n = 1 << lg
    print("perm gp", lg, flush=True)
    cols = [po.uniform(n, po.seed(po.OP_GP, lg, 10 + j)) for j in range(6)]
    beta, gamma = po.uniform(1, po.seed(po.OP_GP, lg, 20)), po.uniform(1, po.seed(po.OP_GP, lg, 21))
    ks = [po.from_u64(1), po.from_u64(7), po.from_u64(13)]
    dom = pl.Domain(lg)
    dev = [pl.to_dev(c) for c in cols]
    ratio = pl.perm_ratio(dev[:3], dev[3:], pl.to_dev(beta), pl.to_dev(gamma), np.concatenate(ks), dom)
    exp_ratio = po.perm_ratio(cols, beta, gamma, ks, lg)
    np.testing.assert_array_equal(pl.to_host(ratio), exp_ratio)
    z = pl.perm_ratio(dev[:3], dev[3:], pl.to_dev(beta), pl.to_dev(gamma), np.concatenate(ks), dom,
                      grand_product=True)
    np.testing.assert_array_equal(pl.to_host(z), po.grand_product(exp_ratio))


def _perm_inputs(n, lg, salt):
    cols = [po.uniform(n, po.seed(po.OP_GP, lg, salt + j)) for j in range(6)]
    beta, gamma = po.uniform(1, po.seed(po.OP_GP, lg, salt + 10)), po.uniform(1, po.seed(po.OP_GP, lg, salt + 11))
    return cols, beta, gamma


@pytest.mark.parametrize("lg", [1, 3, 8, 11, 12, 13, 14, 15, 18, 20, 21, 22])
def test_perm_grand_product_levels(pl, lg):
    """every top-level geometry (a single block of 256 x 1, 2, 4, 8 run
    products, directly above the runs of 8 elements or above block products,
    whose count ranges from 2 to 256 blocks here) against the oracle's ratio
    and grand product"""
    n = 1 << lg
    print("perm gp levels", lg, flush=True)
    cols, beta, gamma = _perm_inputs(n, lg, 40)
    ks = [po.from_u64(1), po.from_u64(7), po.from_u64(13)]
    dom = pl.Domain(lg)
    dev = [pl.to_dev(c) for c in cols]
    z = pl.perm_ratio(dev[:3], dev[3:], pl.to_dev(beta), pl.to_dev(gamma), np.concatenate(ks), dom,
                      grand_product=True)
    np.testing.assert_array_equal(pl.to_host(z), po.grand_product(po.perm_ratio(cols, beta, gamma, ks, lg)))


@pytest.mark.parametrize("lg,i0", [(3, 0), (3, 7), (12, 4000), (13, 2047), (13, 2048), (15, 12345)])
def test_perm_grand_product_zero_den(pl, lg, i0):
    """a zero denominator (w_0 = -(beta s_0 + gamma) at i0) gives a zero ratio
    under x/0 = 0: z is zero after i0"""
    n = 1 << lg
    print("perm gp zero den", lg, i0, flush=True)
    cols, beta, gamma = _perm_inputs(n, lg, 60)
    t = po.add(po.mul(beta, cols[3][i0:i0 + 1]), gamma)
    cols[0][i0] = po.neg(t)[0]
    ks = [po.from_u64(1), po.from_u64(7), po.from_u64(13)]
    ratio = po.perm_ratio(cols, beta, gamma, ks, lg)
    assert not ratio[i0].any()
    exp = po.grand_product(ratio)
    dom = pl.Domain(lg)
    dev = [pl.to_dev(c) for c in cols]
    z = pl.perm_ratio(dev[:3], dev[3:], pl.to_dev(beta), pl.to_dev(gamma), np.concatenate(ks), dom,
                      grand_product=True)
    np.testing.assert_array_equal(pl.to_host(z), exp)


@pytest.mark.parametrize("lg", [12, 20])
@pytest.mark.parametrize("kind", ["small", "permuted", "random"])
def test_perm_general_constants(pl, lg, kind):
    """coset constants other than (1, 7, 13) take the general terms path (one
    product per k_j w^i term); (1, 7, 13) itself forms 7 t and 13 t by limb
    shifts.  Both against the oracle, ratio and grand product."""
    n = 1 << lg
    print("perm general k", lg, kind, flush=True)
    cols, beta, gamma = _perm_inputs(n, lg, 80)
    if kind == "small":
        ks = [po.from_u64(1), po.from_u64(5), po.from_u64(11)]
    elif kind == "permuted":
        ks = [po.from_u64(7), po.from_u64(1), po.from_u64(13)]
    else:
        ks = [po.uniform(1, 90 + j) for j in range(3)]
    dom = pl.Domain(lg)
    dev = [pl.to_dev(c) for c in cols]
    exp_ratio = po.perm_ratio(cols, beta, gamma, ks, lg)
    ratio = pl.perm_ratio(dev[:3], dev[3:], pl.to_dev(beta), pl.to_dev(gamma), np.concatenate(ks), dom)
    np.testing.assert_array_equal(pl.to_host(ratio), exp_ratio)
    z = pl.perm_ratio(dev[:3], dev[3:], pl.to_dev(beta), pl.to_dev(gamma), np.concatenate(ks), dom,
                      grand_product=True)
    np.testing.assert_array_equal(pl.to_host(z), po.grand_product(exp_ratio))


def test_perm_reference_constants_extreme_terms(pl):
    """k = (1, 7, 13) with beta = r - 1 and w^i near r: the limb-shift terms 7 t
    and 13 t at their largest digits, against the oracle"""
    lg = 14
    n = 1 << lg
    rm1 = [0xFFFFFFFF00000000, 0x53BDA402FFFE5BFE, 0x3339D80809A1D805, 0x73EDA753299D7D48]
    cols, _, gamma = _perm_inputs(n, lg, 120)
    beta = np.array([rm1], dtype=np.uint64)
    ks = [po.from_u64(1), po.from_u64(7), po.from_u64(13)]
    dom = pl.Domain(lg)
    dev = [pl.to_dev(c) for c in cols]
    exp_ratio = po.perm_ratio(cols, beta, gamma, ks, lg)
    z = pl.perm_ratio(dev[:3], dev[3:], pl.to_dev(beta), pl.to_dev(gamma), np.concatenate(ks), dom,
                      grand_product=True)
    np.testing.assert_array_equal(pl.to_host(z), po.grand_product(exp_ratio))


def test_perm_grand_product_deep(pl):
    """2^26 elements (runs of 8, then block products over two levels): the
    fused grand product equals the oracle's grand product of the ratio vector,
    itself checked against the oracle at smaller sizes"""
    import torch
    lg = 26
    n = 1 << lg
    print("perm gp 2^26", flush=True)
    dev = [pl.fill_uniform(torch.empty((n, 4), dtype=torch.int64, device="cuda"), 900 + j) for j in range(6)]
    beta = pl.to_dev(po.uniform(1, 991))
    gamma = pl.to_dev(po.uniform(1, 992))
    ks = np.concatenate([po.from_u64(1), po.from_u64(7), po.from_u64(13)])
    dom = pl.Domain(lg)
    ratio = pl.to_host(pl.perm_ratio(dev[:3], dev[3:], beta, gamma, ks, dom))
    z = pl.to_host(pl.perm_ratio(dev[:3], dev[3:], beta, gamma, ks, dom, grand_product=True))
    del dev
    np.testing.assert_array_equal(z, po.grand_product(ratio))


# ---- exp_mod / gen_sequence / repeat_to_poly ---------------------------------

@pytest.mark.parametrize("e", [0, 1, 2, 3, 5, 65537, 2 ** 31 - 1])
@pytest.mark.parametrize("n", [1, 4097])
def test_exp_mod(pl, e, n):
    a = po.uniform(n, 300 + n + e % 997)
    a[0] = 0
    got = pl.to_host(pl.exp_mod(pl.to_dev(a), e))
    np.testing.assert_array_equal(got, po.vec_exp(a, e))


def test_exp_mod_in_place_and_errors(pl):
    a = po.uniform(70000, 77)
    d = pl.to_dev(a)
    pl.exp_mod(d, 3, out=d)
    np.testing.assert_array_equal(pl.to_host(d), po.vec_exp(a, 3))
    with pytest.raises(pl.PlError):
        pl.exp_mod(d, -1)


@pytest.mark.parametrize("n", [1, 2, 300, 16384, 16385, 3 * 16384 + 77, 1 << 20])
def test_gen_sequence(pl, n):
    x = po.uniform(1, 400 + n % 1000)
    got = pl.to_host(pl.gen_sequence(n, pl.to_dev(x)))
    np.testing.assert_array_equal(got, po.sequence(x, n))


def test_gen_sequence_edge_values(pl):
    for x in (po.from_u64(0), po.one(), po.root_of_unity(10)):
        got = pl.to_host(pl.gen_sequence(20000, pl.to_dev(x)))
        np.testing.assert_array_equal(got, po.sequence(x, 20000))


def test_gen_sequence_2p24(pl):
    x = po.uniform(1, 24)
    got = pl.to_host(pl.gen_sequence(1 << 24, pl.to_dev(x)))
    np.testing.assert_array_equal(got, po.sequence(x, 1 << 24))


def test_gen_sequence_errors(pl):
    import torch
    x = pl.to_dev(po.uniform(1, 1))
    with pytest.raises(pl.PlError):
        pl.gen_sequence((1 << 28) + 1, x, out=torch.empty((1, 4), dtype=torch.int64, device=x.device))
    buf = pl.to_dev(po.uniform(8, 2))
    with pytest.raises(pl.PlError):  # out overlaps x
        pl.gen_sequence(4, buf[2:3], out=buf)


@pytest.mark.parametrize("n", [1, 1000, (1 << 20) + 3])
def test_repeat_to_poly(pl, n):
    x = po.uniform(1, 500 + n % 1000)
    got = pl.to_host(pl.repeat_to_poly(pl.to_dev(x), n))
    np.testing.assert_array_equal(got, np.repeat(x, n, axis=0))
