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
