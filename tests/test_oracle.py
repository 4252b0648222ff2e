"""Pins the CPU oracle (oracle/oracle.c) before anything is checked against it.

Each oracle function is checked against values it did not compute:
Python big-int closed forms (tests/golden/*.npz, written by make_golden.py),
the reference's own constant tables (ref_constants.json), the reference's
own blst field (oracle/_ref, when built), and algebraic invariants chosen so
that a dropped term, a wrong sign or index, or a transposed operand fails.
"""
import json
import os

import numpy as np
import pytest

import pyoracle as po
from conftest import GOLDEN

R_MOD = 0x73EDA753299D7D483339D80809A1D80553BDA402FFFE5BFEFFFFFFFF00000001


def to_int(a):
    a = np.asarray(a, dtype=np.uint64).reshape(-1, 4)
    return [sum(int(x[i]) << (64 * i) for i in range(4)) for x in a]


def from_int(vals):
    return np.array([[(v >> (64 * i)) & 0xFFFFFFFFFFFFFFFF for i in range(4)] for v in vals],
                    dtype=np.uint64).reshape(-1, 4)


@pytest.fixture(scope="module")
def consts():
    with open(os.path.join(GOLDEN, "ref_constants.json")) as f:
        return json.load(f)


def test_field_kat(golden):
    k = golden("field_kat.npz")
    a, b = k["a"], k["b"]
    np.testing.assert_array_equal(po.add(a, b), k["add"])
    np.testing.assert_array_equal(po.sub(a, b), k["sub"])
    np.testing.assert_array_equal(po.mul(a, b), k["mul"])
    np.testing.assert_array_equal(po.inv(a), k["inv"])


def test_field_constants_match_reference(consts):
    assert int(consts["BLS12_381_r"], 16) == R_MOD
    assert to_int(po.one())[0] == int(consts["BLS12_381_rone"], 16)
    assert int(consts["miss_ops_one"], 16) == int(consts["BLS12_381_rone"], 16)
    assert to_int(po.from_u64(7))[0] == int(consts["group_gen"][0], 16)
    g_inv = po.inv(po.from_u64(7))
    assert to_int(g_inv)[0] == int(consts["group_gen_inverse"][0], 16)


def test_roots_match_reference_tables(consts):
    fw = [int(v, 16) for v in consts["forward_roots_of_unity"]]
    iv = [int(v, 16) for v in consts["inverse_roots_of_unity"]]
    dinv = [int(v, 16) for v in consts["domain_size_inverse"]]
    assert len(fw) == 33 and len(iv) == 33 and len(dinv) == 33
    for lg in range(33):
        w = po.root_of_unity(lg)
        assert to_int(w)[0] == fw[lg], lg
        assert to_int(po.inv(w))[0] == iv[lg], lg
        assert to_int(po.domain_size_inverse(lg))[0] == dinv[lg], lg


def test_root_orders():
    # omega_{2^lg} has order exactly 2^lg
    for lg in (1, 5, 20, 32):
        w = po.root_of_unity(lg)
        assert to_int(po.pow_(w, 1 << lg)) == to_int(po.one())
        assert to_int(po.pow_(w, 1 << (lg - 1))) != to_int(po.one())


def test_field_against_reference_blst():
    ref = po.ref_lib()
    if ref is None:
        pytest.skip("reference blst not built here (needs /root/reference)")
    a = po.uniform(300, po.seed(9, 0, 2))
    b = po.uniform(300, po.seed(9, 0, 3))
    b[:3] = 0
    b[3] = po.one()[0]
    for name, f in (("ref_mul", po.mul), ("ref_add", po.add), ("ref_sub", po.sub)):
        out = po.fr_array(len(a))
        for i in range(len(a)):
            getattr(ref, name)(po.P(out[i:i + 1]), po.P(np.ascontiguousarray(a[i:i + 1])),
                               po.P(np.ascontiguousarray(b[i:i + 1])))
        np.testing.assert_array_equal(out, f(a, b), err_msg=name)
    out = po.fr_array(len(a))
    for i in range(len(a)):
        ref.ref_inv(po.P(out[i:i + 1]), po.P(np.ascontiguousarray(a[i:i + 1])))
    np.testing.assert_array_equal(out, po.inv(a))


def test_uniform_generator_is_canonical_and_seeded():
    x = po.uniform(1000, po.seed(po.OP_NTT, 10))
    assert all(v < R_MOD for v in to_int(x))
    y = po.uniform(1000, po.seed(po.OP_NTT, 10))
    np.testing.assert_array_equal(x, y)
    z = po.uniform(1000, po.seed(po.OP_NTT, 11))
    assert not np.array_equal(x, z)


def test_generator_matches_python_restatement(golden):
    # ntt_small.npz inputs were drawn by make_golden.py's own SplitMix64
    k = golden("ntt_small.npz")
    for lg in range(0, 9):
        np.testing.assert_array_equal(po.uniform(1 << lg, po.seed(po.OP_NTT, lg)), k["in_%d" % lg])


def test_ntt_brute_force(golden):
    k = golden("ntt_small.npz")
    for lg in range(0, 9):
        x = k["in_%d" % lg]
        np.testing.assert_array_equal(po.ntt(x, lg), k["ntt_%d" % lg], err_msg="lg=%d" % lg)
        np.testing.assert_array_equal(po.ntt(x, lg, inverse=True), k["intt_%d" % lg], err_msg="lg=%d" % lg)


def test_ntt_invariants():
    lg = 12
    n = 1 << lg
    x = po.uniform(n, po.seed(po.OP_NTT, lg, 5))
    np.testing.assert_array_equal(po.ntt(po.ntt(x, lg), lg, inverse=True), x)
    delta = po.fr_array(n)
    delta[0] = po.one()[0]
    np.testing.assert_array_equal(po.ntt(delta, lg), np.repeat(po.one(), n, axis=0))
    ones = np.repeat(po.one(), n, axis=0)
    expect = po.fr_array(n)
    expect[0] = po.from_u64(n)[0]
    np.testing.assert_array_equal(po.ntt(ones, lg), expect)
    # convolution theorem: NTT(a (*) b) = NTT(a) . NTT(b), cyclic convolution
    m = 16
    a = po.uniform(m, 1)
    b = po.uniform(m, 2)
    conv = po.fr_array(m)
    for i in range(m):
        for j in range(m):
            conv[(i + j) % m] = po.add(conv[(i + j) % m:(i + j) % m + 1], po.mul(a[i:i + 1], b[j:j + 1]))[0]
    np.testing.assert_array_equal(po.ntt(conv, 4), po.mul(po.ntt(a, 4), po.ntt(b, 4)))


def test_lde_brute_force(golden):
    k = golden("lde_small.npz")
    for lg_n in range(1, 6):
        c = k["in_%d" % lg_n]
        for lb in range(1, 4):
            np.testing.assert_array_equal(po.lde(c, lg_n, lb, po.ref_lde_shift(lb)),
                                          k["lde_ref_%d_%d" % (lg_n, lb)])
            np.testing.assert_array_equal(po.lde(c, lg_n, lb, po.from_u64(7)),
                                          k["lde_g_%d_%d" % (lg_n, lb)])
            np.testing.assert_array_equal(po.naive_lde(c, lg_n, lb), k["naive_%d_%d" % (lg_n, lb)])


def test_lde_ref_shift_is_not_seven():
    # our_lde's coset factor is 7^(2^lambda), not 7 (lde.cu:642-645)
    assert to_int(po.ref_lde_shift(2)) != to_int(po.from_u64(7))
    assert to_int(po.ref_lde_shift(2)) == to_int(po.from_u64(7 ** 4))


def test_grand_product(golden):
    k = golden("scan_small.npz")
    for n in (1, 2, 3, 64, 257):
        np.testing.assert_array_equal(po.grand_product(k["gp_in_%d" % n]), k["gp_out_%d" % n])
        np.testing.assert_array_equal(po.naive_gp(k["gp_in_%d" % n]), k["naive_gp_out_%d" % n])


def test_naive_gp_against_reference_blst(golden):
    ref = po.ref_lib()
    if ref is None:
        pytest.skip("reference blst not built here")
    k = golden("scan_small.npz")
    for n in (1, 2, 3, 64, 257):
        p = k["gp_in_%d" % n]
        out = po.fr_array(n)
        ref.ref_naive_gp(po.P(np.ascontiguousarray(p)), po.P(out), n)
        np.testing.assert_array_equal(out, po.naive_gp(p))


def test_eval(golden):
    k = golden("scan_small.npz")
    for n in (1, 2, 3, 64, 100):
        for tag in ("r", "0", "1"):
            np.testing.assert_array_equal(po.eval_(k["eval_in_%d" % n], k["eval_x_%s_%d" % (tag, n)]),
                                          k["eval_y_%s_%d" % (tag, n)])


def test_div(golden):
    k = golden("scan_small.npz")
    for n in (1, 2, 3, 64, 100):
        q, rem = po.div(k["div_in_%d" % n], k["div_c_%d" % n])
        np.testing.assert_array_equal(q, k["div_q_%d" % n])
        np.testing.assert_array_equal(rem, k["div_rem_%d" % n])


def test_div_identity():
    # q(X)(X - z) + a(z) = a(X)
    n = 50
    a = po.uniform(n, 77)
    c = po.uniform(1, 78)
    q, rem = po.div(a, c)
    z = po.neg(c)
    # coefficients of q(X)*(X - z): q_{i-1} - z q_i
    recon = po.fr_array(n)
    for i in range(n):
        t = po.mul(q[i:i + 1], z)
        prev = q[i - 1:i] if i > 0 else po.fr_array(1)
        recon[i] = po.sub(prev, t)[0]
    recon[0] = po.add(recon[0:1], rem)[0]
    np.testing.assert_array_equal(recon, a)
    np.testing.assert_array_equal(po.eval_(a, z), rem)


def test_naive_div_is_gpu_quotient_without_tail(golden):
    k = golden("scan_small.npz")
    one = po.one()
    for n in (2, 3, 64, 100):
        a = k["div_in_%d" % n]
        c = k["div_c_%d" % n]
        np.testing.assert_array_equal(po.naive_div(a, c, one), k["div_q_%d" % n][:-1])


def test_naive_div_against_reference_blst(golden):
    ref = po.ref_lib()
    if ref is None:
        pytest.skip("reference blst not built here")
    k = golden("scan_small.npz")
    for n in (2, 3, 64, 100):
        a = np.ascontiguousarray(k["div_in_%d" % n])
        c = np.ascontiguousarray(k["div_c_%d" % n])
        q = po.fr_array(n - 1)
        ref.ref_naive_div(po.P(a), n, po.P(c), po.P(po.one()), po.P(q))
        np.testing.assert_array_equal(q, po.naive_div(a, c, po.one()))


def test_miss_ops_literals(golden, consts):
    k = golden("scan_small.npz")
    pt = from_int([int(consts["miss_ops_point"], 16)])
    np.testing.assert_array_equal(po.eval_(k["mo_poly"], pt), k["mo_eval"])
    c = from_int([int(consts["miss_ops_divisor_gpu"], 16)])
    q, _ = po.div(k["mo_poly"], c)
    np.testing.assert_array_equal(q, k["mo_div_q"])


def test_perm_ratio(golden):
    k = golden("scan_small.npz")
    ks = [po.from_u64(1), po.from_u64(7), po.from_u64(13)]
    for lg in (3, 6):
        cols = [k["perm_%s_%d" % (nm, lg)] for nm in ("a", "b", "c", "s1", "s2", "s3")]
        out = po.perm_ratio(cols, k["perm_beta_%d" % lg], k["perm_gamma_%d" % lg], ks, lg)
        np.testing.assert_array_equal(out, k["perm_ratio_%d" % lg])


def test_reference_input_prefix_shape():
    raw = np.fromfile(os.path.join(GOLDEN, "input_prefix.bin"), dtype=np.uint64).reshape(-1, 4)
    assert raw.shape == (4096, 4)
    vals = to_int(raw)
    assert all(v < R_MOD for v in vals)
    assert vals[0] == 0 and vals[1] == to_int(po.one())[0]


def test_reference_field_eval_and_perm_product_against_the_goldens(golden):
    """The reference-field Horner evaluation and permutation grand product
    (the timed CPU baseline of the miss_ops workload) against the Python
    big-int fixtures."""
    ref = po.ref_lib()
    if ref is None:
        pytest.skip("reference blst not built here")
    k = golden("scan_small.npz")
    for n in (1, 2, 3, 64, 100):
        for tag in ("r", "0", "1"):
            y = po.fr_array(1)
            c = np.ascontiguousarray(k["eval_in_%d" % n])
            ref.ref_eval(po.P(c), n, po.P(np.ascontiguousarray(k["eval_x_%s_%d" % (tag, n)])), po.P(y))
            np.testing.assert_array_equal(y, k["eval_y_%s_%d" % (tag, n)])
    ks = np.concatenate([po.from_u64(1), po.from_u64(7), po.from_u64(13)])
    for lg in (3, 6):
        cols = [k["perm_%s_%d" % (nm, lg)] for nm in ("a", "b", "c", "s1", "s2", "s3")]
        z = po.ref_perm_grand_product(cols, k["perm_beta_%d" % lg], k["perm_gamma_%d" % lg], ks,
                                      po.root_of_unity(lg))
        np.testing.assert_array_equal(z, po.grand_product(k["perm_ratio_%d" % lg]))


def test_coset_intt_inverts_coset_ntt(golden):
    """or_coset_intt undoes or_coset_ntt; or_coset_ntt itself is the first
    column of or_lde, which the brute-force fixtures pin (shift 7 and the
    reference's 7^(2^lg_blowup))."""
    k = golden("lde_small.npz")
    for lg in (1, 3, 5):
        c = k["in_%d" % lg]
        B = 2
        lde = k["lde_g_%d_1" % lg]
        col0 = lde[0::B]
        np.testing.assert_array_equal(po.coset_ntt(c, lg, po.from_u64(7)), col0)
    for lg, shift in ((6, po.from_u64(7)), (10, po.uniform(1, 5)), (12, po.from_u64(5))):
        x = po.uniform(1 << lg, 40 + lg)
        np.testing.assert_array_equal(po.coset_intt(po.coset_ntt(x, lg, shift), lg, shift), x)
        # and the coset powers by themselves: coset_ntt(x) = ntt(x_i shift^i)
        pw = np.concatenate([po.pow_(shift, i) for i in range(8)])
        y = x.copy()
        y[:8] = po.mul(x[:8], pw)
        rest = po.mul(x[8:], np.concatenate([po.pow_(shift, i) for i in range(8, 1 << lg)])) if lg <= 10 else None
        if rest is not None:
            y[8:] = rest
            np.testing.assert_array_equal(po.coset_ntt(x, lg, shift), po.ntt(y, lg))


def test_blowup3_lde_by_brute_force():
    """or_lde3 against the closed form out[k] = sum_i c_i (g w^k)^i over the
    3N-point domain, w = 7^((r-1)/(3N)) computed here in Python integers
    (its order is exactly 3N and its cube is the N-th root of the tables)."""
    Rinv = pow(2 ** 256, -1, R_MOD)

    def val(a):  # Montgomery limbs -> the field value
        return [x * Rinv % R_MOD for x in to_int(a)]

    for lg in (0, 1, 2, 4, 5):
        n = 1 << lg
        w = pow(7, (R_MOD - 1) // (3 * n), R_MOD)
        assert pow(w, 3 * n, R_MOD) == 1 and pow(w, n, R_MOD) != 1 and (n == 1 or pow(w, 3 * n // 2, R_MOD) != 1)
        assert val(po.root_of_unity_3n(lg)) == [w]
        if lg:
            assert val(po.root_of_unity(lg)) == [pow(w, 3, R_MOD)]
        c = po.uniform(n, 70 + lg)
        ci = val(c)
        for g in (7, 5):
            got = val(po.lde3(c, lg, po.from_u64(g)))
            want = []
            for k in range(3 * n):
                x = g * pow(w, k, R_MOD) % R_MOD
                want.append(sum(ci[i] * pow(x, i, R_MOD) for i in range(n)) % R_MOD)
            assert got == want, (lg, g)


R_MONT = 1 << 256
R_INV = pow(R_MONT, -1, R_MOD)


def mont_pow_int(s, e):
    """Montgomery word of x^e for the word s = x R mod r (Python big ints)"""
    return pow(s * R_INV % R_MOD, e, R_MOD) * R_MONT % R_MOD


@pytest.mark.parametrize("e", [0, 1, 2, 3, 4, 5, 17, 65537, 2 ** 31 - 1])
def test_vec_exp_by_python_pow(e):
    a = po.uniform(64, 11 + e % 1000)
    a[3] = 0
    a[4] = po.one()[0]
    got = to_int(po.vec_exp(a, e))
    want = [mont_pow_int(s, e) for s in to_int(a)]
    assert got == want


def test_sequence_by_python_pow():
    for x in (po.uniform(1, 5), po.from_u64(0), po.one(), po.from_u64(2)):
        s = to_int(x)[0]
        got = to_int(po.sequence(x, 300))
        assert got == [mont_pow_int(s, i) for i in range(300)]
    assert po.sequence(po.uniform(1, 6), 0).shape == (0, 4)


def test_threads_do_not_change_results():
    # the oracle's transforms split work over host threads; the results do
    # not depend on how many
    lg = 15
    x = po.uniform(1 << lg, 99)
    was = po.get_threads()
    try:
        po.set_threads(1)
        a = (po.ntt(x, lg), po.ntt(x, lg, inverse=True), po.lde(x[: 1 << 13], 13, 2, po.from_u64(7)),
             po.coset_intt(x, lg, po.from_u64(5)))
        po.set_threads(16)
        b = (po.ntt(x, lg), po.ntt(x, lg, inverse=True), po.lde(x[: 1 << 13], 13, 2, po.from_u64(7)),
             po.coset_intt(x, lg, po.from_u64(5)))
    finally:
        po.set_threads(was)
    for u, v in zip(a, b):
        np.testing.assert_array_equal(u, v)
