"""The radix-2^29 Montgomery product of the NTT passes (csrc/fr29.hpp) and
its constants, on the host.

* fr29.hpp's plain-C product (signed digit, then + r), built with g++,
  against Python integers: x w 2^-261 mod r, and in (0, 1.031 r) for any
  256-bit x and w < r (the bound the lazy [0, 2r) butterflies rely on);
* the DFT constants in ntt_pass.hpp (balanced digits) against the
  reference's own 16th roots of unity (tests/golden/ref_constants.json),
  times 2^261;
* r as written in fr29.hpp;
* the permutation terms' limb shifts for k = (1, 7, 13) (7 t, 13 t);
* mont29_columns.inc is what its generator writes, every column block adds
  each product term exactly once, and the balanced-digit blocks close the
  previous column as (acc + 2^28) >> 29.

The device asm blocks are exercised bit-for-bit by every GPU NTT/LDE test.
"""
import ctypes
import json
import os
import random
import re
import subprocess

import pytest

R = 0x73EDA753299D7D483339D80809A1D80553BDA402FFFE5BFEFFFFFFFF00000001
HERE = os.path.dirname(__file__)
CSRC = os.path.join(HERE, "..", "ppopp26-artifact_amd", "csrc")
M29 = (1 << 29) - 1


def limbs(v, n, b):
    return [(v >> (b * i)) & ((1 << b) - 1) for i in range(n)]


@pytest.fixture(scope="module")
def mul29(tmp_path_factory):
    d = tmp_path_factory.mktemp("fr29")
    src = d / "fr29_host.cpp"
    src.write_text(
        '#include "%s"\n'
        'extern "C" void mul29(const uint32_t* x, const uint32_t* w9, uint32_t* y) {\n'
        "  uint32_t a[9], t[9];\n"
        "  pl::split29(x, a);\n"
        "  pl::mont29(a, w9, t);\n"
        "  pl::join29(t, y);\n"
        "  pl::add_r256(y);\n"
        "}\n" % os.path.abspath(os.path.join(CSRC, "fr29.hpp"))
    )
    lib = d / "libfr29.so"
    subprocess.check_call(["g++", "-O2", "-std=c++17", "-shared", "-fPIC", str(src), "-o", str(lib)])
    so = ctypes.CDLL(str(lib))

    def call(x, w):
        X = (ctypes.c_uint32 * 8)(*limbs(x, 8, 32))
        W = (ctypes.c_uint32 * 9)(*limbs(w, 9, 29))
        Y = (ctypes.c_uint32 * 8)()
        so.mul29(X, W, Y)
        return sum(int(Y[i]) << (32 * i) for i in range(8))

    return call


def test_product_matches_integers(mul29):
    rinv = pow(1 << 261, -1, R)
    rng = random.Random(29)
    cases = [(0, R - 1), (1, 1), (R - 1, R - 1), (2 * R - 1, R - 1), ((1 << 256) - 1, R - 1), ((1 << 256) - 1, 1)]
    cases += [((1 << k) - 1, R - 1 - k) for k in range(1, 257)]
    cases += [(rng.randrange(1 << 256), rng.randrange(R)) for _ in range(4000)]
    cases += [(2 * R - 1 - rng.randrange(1 << 40), R - 1 - rng.randrange(1 << 40)) for _ in range(500)]
    worst = 0
    for x, w in cases:
        y = mul29(x, w)
        assert y % R == x * w * rinv % R, (hex(x), hex(w))
        worst = max(worst, y)
    assert worst < 1.031 * R


def _table(text, name, rows, cols):
    blk = text[text.index(name):]
    blk = blk[: blk.index("}};") + 3]
    nums = [int(v, 16) for v in re.findall(r"0x([0-9a-f]+)u", blk)]
    assert len(nums) == 2 * rows * cols
    return [[nums[(d * rows + e) * cols:(d * rows + e + 1) * cols] for e in range(rows)] for d in range(2)]


def _s32(v):
    return v - (1 << 32) if v >> 31 else v


def test_dft16_constants_are_reference_roots_times_2_261():
    ref = json.load(open(os.path.join(HERE, "golden", "ref_constants.json")))
    rinv = pow(1 << 256, -1, R)
    w = int(ref["forward_roots_of_unity"][4], 16) * rinv % R
    wi = int(ref["inverse_roots_of_unity"][4], 16) * rinv % R
    assert pow(w, 16, R) == 1 and pow(w, 8, R) == R - 1 and w * wi % R == 1
    tab = _table(open(os.path.join(CSRC, "ntt_pass.hpp")).read(), "kW16b[2][8][9]", 8, 9)
    for d, base in ((0, w), (1, wi)):
        for e in range(8):
            digits = [_s32(v) for v in tab[d][e]]
            assert all(-(1 << 28) <= v <= (1 << 28) for v in digits)
            v = sum(l << (29 * i) for i, l in enumerate(digits))
            assert v == pow(base, e, R) * (1 << 261) % R, (d, e)


def test_modulus_constants():
    text = open(os.path.join(CSRC, "fr29.hpp")).read()
    body = text[text.index("constexpr uint32_t r29(int i)"):]
    r29 = [int(v, 16) for v in re.findall(r"0x([0-9a-f]+)u", body[: body.index(";")])]
    assert r29 == limbs(R, 9, 29)


def test_generated_columns_current_and_complete():
    gen = os.path.join(CSRC, "gen_mont29_columns.py")
    inc = os.path.join(CSRC, "mont29_columns.inc")
    out = subprocess.check_output(["python3", gen]).decode()
    assert out == open(inc).read()
    # in each of the 70 blocks every product term of its column appears once
    spec = re.findall(r"struct Mont29Col<(\d+), (true|false), (true|false)> \{.*?asm\(\"([^\"]*)\"\s*: ([^\n]*)"
                      r"\s*: ([^;]*)\);", out, re.S)
    assert len(spec) == 70
    for k, ub, rnd, text, _outs, ins in spec:
        k = int(k)
        ops = ["acc", "d"] + [o.strip() for o in re.findall(r'"[vs]"\(([^)]*\)?)\)', ins)]
        first = text.split("\\n\\t")[:2]
        if rnd == "true":
            # (prev + 2^28) >> 29: the carry of a balanced digit
            assert first[0] == "v_mad_i64_i32 %0, %1, 1, %3, %2" and ops[3] == "1u << 28", k
            assert first[1] == "v_ashrrev_i64 %0, 29, %0", k
        elif k >= 1:
            assert first[0] == "v_ashrrev_i64 %0, 29, %2", k
        pairs = []
        for t in text.split("\\n\\t"):
            m = re.match(r"v_mad_i64_i32 %0, %1, %(\d+), %(\d+), (%0|0)", t)
            if m:
                pairs.append((ops[int(m.group(1))], ops[int(m.group(2))]))
        want = set()
        if k <= 16:
            for i in range(max(0, k - 8), min(k, 8) + 1):
                want.add(("x[%d]" % i, "w[%d]" % (k - i)))
                if i < k:
                    want.add(("m[%d]" % i, "nr29(%d)" % (k - i)))
        assert sorted(pairs) == sorted(want), k


def test_reference_constant_terms_limb_shifts():
    """The terms kernel's k = (1, 7, 13) path (ops.hip perm_terms_at<SK>): from
    t with balanced digits (|t_i| <= 2^28, small signed top), 7 t by limb shifts,
    carried to balanced digits (v9_carry_balanced), then 7 t + 6 t carried
    again.  Every int32 step stays in range, the integers are exactly 7 T and
    13 T, and the digits come out balanced.  Emulates the 32-bit limb arithmetic."""
    import random
    rng = random.Random(7)

    def i32(v):
        assert -(1 << 31) <= v < (1 << 31), v
        return v

    def sext29(v):
        v &= (1 << 29) - 1
        return v - (1 << 29) if v >= (1 << 28) else v

    def carry(v):
        v = list(v)
        for i in range(8):
            d = sext29(v[i])
            c = i32(v[i] - d) >> 29
            v[i + 1] = i32(v[i + 1] + c)
            v[i] = d
        return v

    def value(v):
        return sum(l << (29 * i) for i, l in enumerate(v))

    cases = [[-(1 << 28)] * 8 + [-(1 << 23)], [(1 << 28) - 1] * 8 + [(1 << 23)],
             [(1 << 28) - 1, -(1 << 28)] * 4 + [0]]
    cases += [[rng.randrange(-(1 << 28), 1 << 28) for _ in range(8)] + [rng.randrange(-(1 << 23), 1 << 23)]
              for _ in range(20000)]
    for t in cases:
        T = value(t)
        t7 = carry([i32((x << 3) - x) for x in t])
        assert value(t7) == 7 * T and all(-(1 << 28) <= d < (1 << 28) for d in t7[:8])
        t13 = carry([i32(a + (x << 2) + (x << 1)) for a, x in zip(t7, t)])
        assert value(t13) == 13 * T and all(-(1 << 28) <= d < (1 << 28) for d in t13[:8])
        assert abs(t13[8]) < (1 << 28)
