"""The single-element inversion used by batch inversion (csrc/fr_inv.hpp),
built for the host with g++ and checked against Fermat's little theorem
(y^(r-2) mod r) and the defining identity y * y^-1 = 1 mod r.

The device path runs the same header; its results are checked on the GPU
through the element-wise inverse and the permutation grand product.
"""
import ctypes
import os
import random
import subprocess

import pytest

R = 0x73EDA753299D7D483339D80809A1D80553BDA402FFFE5BFEFFFFFFFF00000001
HDR = os.path.join(os.path.dirname(__file__), "..", "ppopp26-artifact_amd", "csrc", "fr_inv.hpp")


@pytest.fixture(scope="module")
def inv_host(tmp_path_factory):
    d = tmp_path_factory.mktemp("inv")
    src = d / "inv_host.cpp"
    src.write_text(
        '#include "%s"\n'
        'extern "C" void inv_host(const uint32_t* y, uint32_t* out) { pl::inv::inverse(y, out); }\n'
        % os.path.abspath(HDR)
    )
    lib = d / "libinvhost.so"
    subprocess.check_call(["g++", "-O2", "-std=c++17", "-shared", "-fPIC", str(src), "-o", str(lib)])
    so = ctypes.CDLL(str(lib))
    so.inv_host.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
    so.inv_host.restype = None

    def call(y):
        a = (ctypes.c_uint32 * 8)(*[(y >> (32 * i)) & 0xFFFFFFFF for i in range(8)])
        o = (ctypes.c_uint32 * 8)()
        so.inv_host(a, o)
        return sum(int(o[i]) << (32 * i) for i in range(8))

    return call


def _cases():
    rng = random.Random(20260)
    cs = [1, 2, 3, R - 1, R - 2, (R - 1) // 2, (R + 1) // 2, 7, 5 ** 100 % R]
    cs += [1 << k for k in range(255)]
    cs += [R - (1 << k) for k in range(1, 254)]
    cs += [(1 << k) - 1 for k in range(2, 255)]
    cs += [rng.randrange(1, R) for _ in range(1500)]
    cs += [rng.randrange(1, 1 << rng.randrange(1, 255)) for _ in range(500)]
    return cs


def test_zero_maps_to_zero(inv_host):
    assert inv_host(0) == 0


def test_inverse_matches_fermat(inv_host):
    bad = []
    for y in _cases():
        v = inv_host(y)
        if v != pow(y, R - 2, R) or (v * y) % R != 1:
            bad.append(y)
    assert not bad, "first failing inputs: %s" % [hex(b) for b in bad[:3]]


@pytest.fixture(scope="module")
def mod_update_host(tmp_path_factory):
    d = tmp_path_factory.mktemp("modupd")
    src = d / "mod_update_host.cpp"
    src.write_text(
        '#include "%s"\n'
        'extern "C" void mod_update_host(const uint32_t* u, int64_t f, const uint32_t* v, int64_t g,\n'
        '                                uint32_t* out) { pl::inv::mod_update(out, u, f, v, g); }\n'
        % os.path.abspath(HDR)
    )
    lib = d / "libmodupd.so"
    subprocess.check_call(["g++", "-O2", "-std=c++17", "-shared", "-fPIC", str(src), "-o", str(lib)])
    so = ctypes.CDLL(str(lib))
    so.mod_update_host.argtypes = [ctypes.c_void_p, ctypes.c_int64, ctypes.c_void_p, ctypes.c_int64, ctypes.c_void_p]
    so.mod_update_host.restype = None

    def call(u, f, v, g):
        a = (ctypes.c_uint32 * 8)(*[(u >> (32 * i)) & 0xFFFFFFFF for i in range(8)])
        b = (ctypes.c_uint32 * 8)(*[(v >> (32 * i)) & 0xFFFFFFFF for i in range(8)])
        o = (ctypes.c_uint32 * 8)()
        so.mod_update_host(a, f, b, g, o)
        return sum(int(o[i]) << (32 * i) for i in range(8))

    return call


def test_mod_update_is_the_exact_reduced_combination(mod_update_host):
    # (u f + v g) / 2^31 mod r, fully reduced, over u, v < r and |f| + |g| <= 2^31
    # (the bound the 31-step rounds keep); the extremes drive the quotient
    # estimate of the final reduction to both of its outcomes
    inv31 = pow(1 << 31, R - 2, R)
    rng = random.Random(4711)
    vals = [0, 1, 2, R - 1, R - 2, (R - 1) // 2, 1 << 254, R - (1 << 200)]
    coef = [(1 << 31, 0), (-(1 << 31), 0), (0, 1 << 31), (0, -(1 << 31)), ((1 << 30), (1 << 30)),
            (-(1 << 30), -(1 << 30)), ((1 << 30), -(1 << 30)), (1, -1), (0, 0), ((1 << 31) - 1, 1)]
    cases = [(u, f, v, g) for u in vals for v in vals for f, g in coef]
    for _ in range(3000):
        t = rng.randrange(0, (1 << 31) + 1)
        f = rng.randrange(-t, t + 1)
        g = (t - abs(f)) * rng.choice([-1, 1])
        cases.append((rng.randrange(R), f, rng.randrange(R), g))
    bad = []
    for u, f, v, g in cases:
        got = mod_update_host(u, f, v, g)
        if got != (u * f + v * g) * inv31 % R:
            bad.append((hex(u), f, hex(v), g))
    assert not bad, "first failing cases: %s" % bad[:3]


def _steps_if_form(a, b):
    """One round's 31 steps as the algorithm states them (branches, unpacked
    factors), on the same 64-bit approximations: low 31 bits and the 33 bits
    below the top of max(a, b) (at least 64 bits wide)."""
    n = max((a | b).bit_length(), 64)
    xa = (a & 0x7FFFFFFF) | (((a >> (n - 33)) & ((1 << 33) - 1)) << 31)
    xb = (b & 0x7FFFFFFF) | (((b >> (n - 33)) & ((1 << 33) - 1)) << 31)
    f0, g0, f1, g1 = 1, 0, 0, 1
    for _ in range(31):
        if xa & 1:
            if xa < xb:
                xa, xb, f0, f1, g0, g1 = xb, xa, f1, f0, g1, g0
            xa, f0, g0 = xa - xb, f0 - f1, g0 - g1
        xa >>= 1
        f1, g1 = 2 * f1, 2 * g1
    return f0, g0, f1, g1


def test_inner_steps_match_the_stated_steps(tmp_path):
    # inner_steps (its 64-bit approximations taken limb by limb) against the
    # steps on Python integers, on inputs that drive the factors to their
    # extremes: row 1 reaches +-2^31, the factor bound the round updates accept,
    # and row 0 passes 2^30 in magnitude; and on random pairs
    src = tmp_path / "steps_host.cpp"
    src.write_text(
        '#include "%s"\n'
        'extern "C" void steps_host(const uint32_t* a, const uint32_t* b, int64_t* m) '
        '{ pl::inv::inner_steps(a, b, m); }\n' % os.path.abspath(HDR)
    )
    lib = tmp_path / "libsteps.so"
    subprocess.check_call(["g++", "-O2", "-std=c++17", "-shared", "-fPIC", str(src), "-o", str(lib)])
    so = ctypes.CDLL(str(lib))
    so.steps_host.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
    so.steps_host.restype = None

    def call(a, b):
        A = (ctypes.c_uint32 * 8)(*[(a >> (32 * i)) & 0xFFFFFFFF for i in range(8)])
        B = (ctypes.c_uint32 * 8)(*[(b >> (32 * i)) & 0xFFFFFFFF for i in range(8)])
        m = (ctypes.c_int64 * 4)()
        so.steps_host(A, B, m)
        return tuple(m)

    rng = random.Random(3131)
    cases = [(1, 1), (1, 2), (2, 1), (1, R), (R, 1), (R - 2, R), (3, (1 << 64) - 1), ((1 << 64) - 1, 3),
             (1, (1 << 255) - 1), ((1 << 31) + 1, 1 << 62), (1 << 40, 1)]
    cases += [(rng.randrange(1, 1 << 256), rng.randrange(1, 1 << 256)) for _ in range(3000)]
    cases += [(rng.randrange(1, 1 << 64) | 1, rng.randrange(1, 1 << 64)) for _ in range(3000)]
    cases += [(rng.randrange(1, 1 << k) | 1, rng.randrange(1, R)) for k in range(1, 256, 5) for _ in range(10)]
    seen = [0, 0, 0, 0]
    bad = []
    for a, b in cases:
        want = _steps_if_form(a, b)
        got = call(a, b)
        if got != want:
            bad.append((hex(a), hex(b), got, want))
        seen = [max(s, abs(v)) for s, v in zip(seen, want)]
    assert not bad, "first failing cases: %s" % bad[:3]
    # the extremes were reached: row 1 at 2^31, row 0 past 2^30
    assert seen[2] == 1 << 31 and seen[3] == 1 << 31
    assert seen[0] > 1 << 30 and seen[1] > 1 << 30


def test_reduce_below_5r_at_the_boundaries(tmp_path):
    # x = k r + e for k < 5: e = 0, 1, r - 1 and e just below the width where
    # the top-bits estimate of k falls one short (its one correction step)
    src = tmp_path / "red_host.cpp"
    src.write_text(
        '#include "%s"\n'
        'extern "C" void red_host(uint32_t* x) { pl::inv::reduce_below_5r(x); }\n' % os.path.abspath(HDR)
    )
    lib = tmp_path / "libred.so"
    subprocess.check_call(["g++", "-O2", "-std=c++17", "-shared", "-fPIC", str(src), "-o", str(lib)])
    so = ctypes.CDLL(str(lib))
    so.red_host.argtypes = [ctypes.c_void_p]
    so.red_host.restype = None
    rng = random.Random(99)
    es = [0, 1, 2, R - 1, R - 2, 1 << 223, (1 << 224) - 1, 1 << 224, 3 << 223, 1 << 230]
    es += [rng.randrange(1 << 226) for _ in range(50)] + [rng.randrange(R) for _ in range(50)]
    for k in range(5):
        for e in es:
            x = k * R + e
            if x >= 5 * R:
                continue
            a = (ctypes.c_uint32 * 9)(*[(x >> (32 * i)) & 0xFFFFFFFF for i in range(9)])
            so.red_host(a)
            got = sum(int(a[i]) << (32 * i) for i in range(8))
            assert got == x % R, (k, hex(e))
