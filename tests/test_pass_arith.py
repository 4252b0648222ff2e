"""The signed-limb arithmetic of the NTT pass kernels (csrc/fr29.hpp,
csrc/ntt_pass.hpp) on the host, against Python integers.

* the radix-2^29 product in its four digit variants (m and output digits in
  [0, 2^29) or balanced in [-2^28, 2^28)): congruence x w 2^-261 mod r and the
  output ranges the kernel relies on, for signed input limbs up to 2^31 and
  balanced factor digits;
* pbd words (packed balanced digits): pack / unpack round trip; unpack_ext
  (balanced digits of any 256-bit value) and to_pbd;
* reduce_centered (into (-0.51 r, 0.51 r)) and reduce_below_2r (into
  [0, 1.01 r)) from any signed limbs below 2^31;
* one pass column (tests/native/pass_sim.cpp, the kernel's sequence of
  operations on the same functions) for every column size 2^3..2^6, both
  directions, intermediate and last pass, caller or pbd input: every limb
  stays a signed 32-bit value, every product column below 2^63, every packed
  output in range, and the outputs equal the column DFT computed with Python
  integers (times the pass twiddles, or canonical in the last pass).  Inputs
  include the extremes: all 2^256 - 1 (the largest DC sums), all digits at
  -2^28 or 2^28 - 1, and chains of two passes;
* the permutation-ratio terms (ops.hip perm_table_kernel's table entry and
  perm_terms_kernel) of one element in the same arithmetic: every bound holds
  (the table entry fits a pbd word) and num / den carry exactly the common
  factor 2^-10 (random and all-(2^256 - 1) inputs).
"""
import ctypes
import json
import os
import random
import subprocess

import pytest

R = 0x73EDA753299D7D483339D80809A1D80553BDA402FFFE5BFEFFFFFFFF00000001
HERE = os.path.dirname(__file__)
CSRC = os.path.join(HERE, "..", "ppopp26-artifact_amd", "csrc")
M29 = (1 << 29) - 1
U32 = ctypes.c_uint32


def limbs(v, n, b):
    return [(v >> (b * i)) & ((1 << b) - 1) for i in range(n)]


def s32(v):
    v &= 0xFFFFFFFF
    return v - (1 << 32) if v >> 31 else v


def val9(b):
    return sum(s32(x) << (29 * i) for i, x in enumerate(b))


def balanced(v):
    """balanced digits (i < 8 in [-2^28, 2^28)) of an integer |v| < 2^255"""
    d = []
    for _ in range(8):
        x = v & M29
        if x >= 1 << 28:
            x -= 1 << 29
        d.append(x)
        v = (v - x) >> 29
    d.append(v)
    return d


def arr(vals, n):
    return (U32 * n)(*[v & 0xFFFFFFFF for v in vals])


@pytest.fixture(scope="module")
def lib(tmp_path_factory):
    d = tmp_path_factory.mktemp("pass_sim")
    so = d / "libpass_sim.so"
    subprocess.check_call(["g++", "-O2", "-std=c++17", "-shared", "-fPIC", "-I", os.path.abspath(CSRC),
                           os.path.join(HERE, "native", "pass_sim.cpp"), "-o", str(so)])
    return ctypes.CDLL(str(so))


def mont(lib, a, w, bm, bt):
    A = arr(a, 9)
    W = arr(w, 9)
    T = (U32 * 9)()
    lib.t_mont29(A, W, T, int(bm), int(bt))
    return [T[i] for i in range(9)]


@pytest.mark.parametrize("bm,bt", [(False, False), (False, True), (True, False), (True, True)])
def test_product_variants(lib, bm, bt):
    rinv = pow(1 << 261, -1, R)
    rng = random.Random(7 + 2 * bm + bt)
    for trial in range(3000):
        if trial % 3 == 0:
            a = [rng.randrange(-(1 << 31), 1 << 31) for _ in range(9)]
        elif trial % 3 == 1:
            a = [rng.choice((-(1 << 31), (1 << 31) - 1)) for _ in range(9)]
        else:
            a = [rng.randrange(-(1 << 28), 1 << 28) for _ in range(9)]
        w = balanced(rng.randrange(R))
        t = mont(lib, a, w, bm, bt)
        x = val9(a)
        T = val9(t)
        assert (T - x * val9(w) * rinv) % R == 0
        xw = x * val9(w) / (1 << 261)
        if bm:
            assert xw - R / 2 - 1 < T < xw + R / 2 + 1
        else:
            assert xw - R - 1 < T <= xw
        for i in range(8):
            d = s32(t[i])
            assert (-(1 << 28) <= d < (1 << 28)) if bt else (0 <= d < (1 << 29))


def test_pbd_round_trip_and_ext(lib):
    rng = random.Random(11)
    for trial in range(2000):
        b = [rng.randrange(-(1 << 28), 1 << 28) for _ in range(8)] + [rng.randrange(-(1 << 23), 1 << 23)]
        if trial < 4:
            b = [(-(1 << 28)) if trial & 1 else (1 << 28) - 1] * 8 + [(-(1 << 23)) if trial & 2 else (1 << 23) - 1]
        X = (U32 * 8)()
        lib.t_pbd_pack(arr(b, 9), X)
        B = (U32 * 9)()
        lib.t_pbd_unpack(X, B)
        assert [s32(B[i]) for i in range(9)] == b
    cases = [0, 1, R - 1, R, (1 << 256) - 1, (1 << 256) - (1 << 232), (1 << 255), 2 * R - 1]
    cases += [rng.randrange(1 << 256) for _ in range(2000)]
    for x in cases:
        B = (U32 * 9)()
        lib.t_unpack_ext(arr(limbs(x, 8, 32), 8), B)
        d = [s32(B[i]) for i in range(9)]
        assert val9(d) == x
        assert all(-(1 << 28) <= v < (1 << 28) for v in d[:8]) and 0 <= d[8] < (1 << 25)
        if x < R:
            X = (U32 * 8)()
            lib.t_to_pbd(arr(limbs(x, 8, 32), 8), X)
            lib.t_pbd_unpack(X, B)
            assert val9([B[i] for i in range(9)]) == x


def test_reductions(lib):
    rng = random.Random(13)
    for trial in range(4000):
        if trial % 4 == 0:
            v = [rng.choice((-(1 << 31), (1 << 31) - 1)) for _ in range(9)]
        else:
            v = [rng.randrange(-(1 << 31), 1 << 31) for _ in range(9)]
        x = val9(v)
        B = (U32 * 9)()
        lib.t_reduce_centered(arr(v, 9), B)
        c = [s32(B[i]) for i in range(9)]
        y = val9(c)
        assert (y - x) % R == 0 and abs(y) < 0.51 * R
        assert all(-(1 << 28) <= d < (1 << 28) for d in c[:8])
        lib.t_reduce_below_2r(arr(v, 9), B)
        t = [B[i] for i in range(9)]
        z = sum(d << (29 * i) for i, d in enumerate(t))
        assert (z - x) % R == 0 and 0 <= z < 1.01 * R
        assert all(d <= M29 for d in t[:8])


# ---- one pass column -------------------------------------------------------
REF = json.load(open(os.path.join(HERE, "golden", "ref_constants.json")))
RINV256 = pow(1 << 256, -1, R)
W256 = int(REF["forward_roots_of_unity"][8], 16) * RINV256 % R
W256I = int(REF["inverse_roots_of_unity"][8], 16) * RINV256 % R
R261 = (1 << 261) % R


def pbd_words(vals):
    """pbd words (8 x u32 each) of field values times 2^261 mod r"""
    out = []
    for v in vals:
        b = balanced(v * R261 % R)
        assert -(1 << 23) <= b[8] < (1 << 23)
        word = sum(((x & M29) if i < 8 else (x & 0xFFFFFF)) << (29 * i) for i, x in enumerate(b))
        out += limbs(word, 8, 32)
    return out


def w16b_words(inv):
    w16 = pow(W256I if inv else W256, 16, R)
    out = []
    for e in range(8):
        out += [x & 0xFFFFFFFF for x in balanced(pow(w16, e, R) * R261 % R)]
    return out


def run_column(lib, la, inv, last, ext, xin, tw=None, scale=None):
    A = 1 << la
    Mp = 8 * A
    inner = pbd_words([pow(W256I if inv else W256, e, R) for e in range(256)])
    tww = pbd_words(tw) if tw is not None else [0] * (8 * Mp)
    sc = pbd_words([scale]) if scale is not None else [0] * 8
    out = (U32 * (8 * Mp))()
    bad = lib.sim_column(la, int(last), int(ext), arr(xin, 8 * Mp), arr(w16b_words(inv), 72), arr(inner, 2048),
                         arr(tww, 8 * Mp), int(scale is not None), arr(sc, 8), out)
    return bad, [out[i] for i in range(8 * Mp)]


def in_value(words, k, ext):
    w = words[8 * k:8 * k + 8]
    if ext:
        return sum(x << (32 * i) for i, x in enumerate(w))
    word = sum(x << (32 * i) for i, x in enumerate(w))
    d = [(word >> (29 * i)) & M29 for i in range(8)] + [word >> 232]
    d = [x - (1 << 29) if x >> 28 else x for x in d[:8]] + [d[8] - (1 << 24) if d[8] >> 23 else d[8]]
    return val9(d)


def out_value(words, k, last):
    w = words[8 * k:8 * k + 8]
    if last:
        return sum(x << (32 * i) for i, x in enumerate(w))
    return in_value(words, k, False)


def column_dft(x, la, inv):
    Mp = 8 << la
    wM = pow(W256I if inv else W256, 256 // Mp, R)
    return [sum(x[j] * pow(wM, j * k, R) for j in range(Mp)) % R for k in range(Mp)]


def pbd_input(rng, mode):
    """one pbd word: random small value, or extreme digits with a small top"""
    if mode == "random":
        v = rng.randrange(-int(0.64 * R), int(0.64 * R))
        b = balanced(v)
    else:
        lo = -(1 << 28) if mode == "low" else (1 << 28) - 1
        b = [lo] * 8 + [rng.randrange(-(1 << 22), 1 << 22)]
    word = sum(((x & M29) if i < 8 else (x & 0xFFFFFF)) << (29 * i) for i, x in enumerate(b))
    return limbs(word, 8, 32)


@pytest.mark.parametrize("la", [0, 1, 2, 3])
@pytest.mark.parametrize("inv", [False, True])
def test_column_bounds_and_values(lib, la, inv):
    rng = random.Random(100 * la + inv)
    Mp = 8 << la
    for last in (False, True):
        for ext, mode in ((True, "random"), (True, "max"), (True, "zero_r"), (False, "random"), (False, "low"),
                          (False, "high")):
            if ext:
                if mode == "random":
                    xs = [rng.randrange(1 << 256) for _ in range(Mp)]
                elif mode == "max":
                    xs = [(1 << 256) - 1] * Mp
                else:
                    xs = [rng.choice((0, R - 1, (1 << 256) - 1)) for _ in range(Mp)]
                words = sum((limbs(x, 8, 32) for x in xs), [])
            else:
                words = sum((pbd_input(rng, mode) for _ in range(Mp)), [])
            tw = [rng.randrange(R) for _ in range(Mp)]
            scale = rng.randrange(R) if last and mode == "random" else None
            bad, out = run_column(lib, la, inv, last, ext, words, None if last else tw, scale)
            assert bad == 0, (la, inv, last, ext, mode)
            x = [in_value(words, j, ext) % R for j in range(Mp)]
            X = column_dft(x, la, inv)
            for k in range(Mp):
                got = out_value(out, k, last)
                if last:
                    want = X[k] * (scale if scale is not None else 1) % R
                    assert got == want, (la, inv, mode, k)
                else:
                    assert (got - X[k] * tw[k]) % R == 0, (la, inv, mode, k)
                    assert abs(got) < 0.66 * R


def test_two_passes_chained(lib):
    """a pass's pbd output as the next pass's input (worst case: 2^256 - 1 in)"""
    rng = random.Random(5)
    for la in (3, 2):
        Mp = 8 << la
        words = sum((limbs((1 << 256) - 1, 8, 32) for _ in range(Mp)), [])
        tw = [rng.randrange(R) for _ in range(Mp)]
        bad, mid = run_column(lib, la, False, False, True, words, tw)
        assert bad == 0
        bad, out = run_column(lib, la, False, True, False, mid)
        assert bad == 0
        x = [in_value(mid, j, False) % R for j in range(Mp)]
        X = column_dft(x, la, False)
        assert [out_value(out, k, True) for k in range(Mp)] == X


def test_perm_terms_model(lib):
    rinv = pow(1 << 256, -1, R)
    rng = random.Random(21)
    for trial in range(300):
        big = trial < 20
        vals = [((1 << 256) - 1) if big else rng.randrange(R) for _ in range(10)]
        w0, w1, beta, gamma = vals[:4]
        ws, ss = vals[4:7], vals[7:10]
        k = [1, 7, 13] if trial % 2 else [rng.randrange(R) for _ in range(3)]
        b32 = 32 * beta % R
        k29 = []
        for kj in k:
            b = balanced(1024 * kj * (1 << 256) % R)  # pbd word of (32 k) R' = 1024 k R
            word = sum(((x & M29) if i < 8 else (x & 0xFFFFFF)) << (29 * i) for i, x in enumerate(b))
            k29 += limbs(word, 8, 32)
        out_n = (U32 * 8)()
        out_d = (U32 * 8)()
        bad = lib.sim_perm_terms(arr(limbs(w0, 8, 32), 8), arr(limbs(w1, 8, 32), 8), arr(limbs(b32, 8, 32), 8),
                                 arr(limbs(gamma, 8, 32), 8), arr(k29, 24),
                                 arr(sum((limbs(x, 8, 32) for x in ws), []), 24),
                                 arr(sum((limbs(x, 8, 32) for x in ss), []), 24), out_n, out_d)
        assert bad == 0, trial
        # field values of the stored-form inputs (x R): a = x R^-1
        f = lambda x: x * rinv % R  # noqa: E731
        wi = f(w0) * f(w1) % R
        num = den = 1
        for j in range(3):
            num = num * (f(ws[j]) + f(gamma) + f(beta) * k[j] * wi) % R
            den = den * (f(ws[j]) + f(gamma) + f(beta) * f(ss[j])) % R
        inv1024 = pow(1024, -1, R)
        got_n = sum(out_n[i] << (32 * i) for i in range(8))
        got_d = sum(out_d[i] << (32 * i) for i in range(8))
        assert got_n == num * inv1024 % R * (1 << 256) % R
        assert got_d == den * inv1024 % R * (1 << 256) % R
