"""ctypes view of the CPU oracle (oracle/_build/liboracle.so).

TEST INFRASTRUCTURE: imported only by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, as the checker.  Arrays are numpy uint64 of
shape (n, 4): canonical Montgomery limbs, byte-identical to the device layout.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "liboracle.so")
REF_PATH = os.path.join(HERE, "_ref", "libblstref.so")

_lib = None
_ref = None
_u64p = ctypes.POINTER(ctypes.c_uint64)


def build():
    subprocess.check_call(["make", "-s", "-C", HERE])


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        _lib = ctypes.CDLL(LIB_PATH)
        L = _lib
        for name in ["or_add", "or_sub", "or_mul"]:
            getattr(L, name).argtypes = [_u64p, _u64p, _u64p]
        for name in ["or_neg", "or_inv", "or_to_mont", "or_from_mont"]:
            getattr(L, name).argtypes = [_u64p, _u64p]
        L.or_pow.argtypes = [_u64p, _u64p, _u64p]
        L.or_one.argtypes = [_u64p]
        L.or_from_u64.argtypes = [_u64p, ctypes.c_uint64]
        L.or_is_canonical.argtypes = [_u64p]
        for name in ["or_vec_add", "or_vec_sub", "or_vec_mul"]:
            getattr(L, name).argtypes = [_u64p, _u64p, _u64p, ctypes.c_uint64]
        L.or_vec_inv.argtypes = [_u64p, _u64p, ctypes.c_uint64]
        L.or_vec_exp.argtypes = [_u64p, _u64p, ctypes.c_uint64, ctypes.c_uint32]
        L.or_sequence.argtypes = [_u64p, _u64p, ctypes.c_uint64]
        L.or_root_of_unity.argtypes = [_u64p, ctypes.c_int]
        L.or_domain_size_inverse.argtypes = [_u64p, ctypes.c_int]
        L.or_group_gen.argtypes = [_u64p]
        L.or_bitrev.argtypes = [ctypes.c_uint64, ctypes.c_int]
        L.or_bitrev.restype = ctypes.c_uint64
        L.or_fill_uniform.argtypes = [_u64p, ctypes.c_uint64, ctypes.c_uint64]
        L.or_seed.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int]
        L.or_seed.restype = ctypes.c_uint64
        L.or_set_threads.argtypes = [ctypes.c_int]
        L.or_get_threads.restype = ctypes.c_int
        L.or_ntt.argtypes = [_u64p, ctypes.c_int, ctypes.c_int]
        L.or_coset_ntt.argtypes = [_u64p, ctypes.c_int, _u64p]
        L.or_coset_intt.argtypes = [_u64p, ctypes.c_int, _u64p]
        L.or_root_of_unity_3n.argtypes = [_u64p, ctypes.c_int]
        L.or_lde3.argtypes = [_u64p, _u64p, ctypes.c_int, _u64p]
        L.or_lde.argtypes = [_u64p, _u64p, ctypes.c_int, ctypes.c_int, _u64p]
        L.or_naive_lde.argtypes = [_u64p, _u64p, ctypes.c_int, ctypes.c_int]
        L.or_grand_product.argtypes = [_u64p, ctypes.c_uint64]
        L.or_naive_gp.argtypes = [_u64p, _u64p, ctypes.c_uint64]
        L.or_eval.argtypes = [_u64p, ctypes.c_uint64, _u64p, _u64p]
        L.or_div.argtypes = [_u64p, ctypes.c_uint64, _u64p, _u64p]
        L.or_naive_div.argtypes = [_u64p, ctypes.c_uint64, _u64p, _u64p, _u64p]
        L.or_perm_ratio.argtypes = [_u64p] * 11 + [ctypes.c_int, _u64p]
        L.or_stage_no_rotate.argtypes = [_u64p, ctypes.c_uint64, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                         ctypes.c_uint64]
        L.or_stage_no_rotate.restype = ctypes.c_int
        L.or_ct_stage_indices.argtypes = [ctypes.c_uint64, ctypes.c_int, ctypes.c_int, _u64p, _u64p]
        L.or_rotate_stage_output_index.argtypes = [ctypes.c_uint64, ctypes.c_int, ctypes.c_int]
        L.or_rotate_stage_output_index.restype = ctypes.c_uint64
    return _lib


def ref_lib():
    """The reference's own blst field, or None when it was not built."""
    global _ref
    if _ref is None and os.path.exists(REF_PATH):
        _ref = ctypes.CDLL(REF_PATH)
        for name in ["ref_mul", "ref_add", "ref_sub"]:
            getattr(_ref, name).argtypes = [_u64p, _u64p, _u64p]
        _ref.ref_inv.argtypes = [_u64p, _u64p]
        _ref.ref_naive_gp.argtypes = [_u64p, _u64p, ctypes.c_uint64]
        _ref.ref_naive_div.argtypes = [_u64p, ctypes.c_uint64, _u64p, _u64p, _u64p]
        _ref.ref_eval.argtypes = [_u64p, ctypes.c_uint64, _u64p, _u64p]
        _ref.ref_perm_grand_product.argtypes = [ctypes.POINTER(_u64p), ctypes.POINTER(_u64p), _u64p, _u64p, _u64p,
                                                _u64p, ctypes.c_uint64, _u64p]
    return _ref


def P(a):
    assert a.dtype == np.uint64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_u64p)


def fr_array(n):
    return np.zeros((n, 4), dtype=np.uint64)


def one():
    o = fr_array(1)
    lib().or_one(P(o))
    return o


def from_u64(v):
    o = fr_array(1)
    lib().or_from_u64(P(o), v)
    return o


def uniform(n, seed):
    o = fr_array(n)
    lib().or_fill_uniform(P(o), n, seed)
    return o


def seed(op, lg, aux=0):
    return lib().or_seed(op, lg, aux)


OP_NTT, OP_INTT, OP_LDE, OP_GP, OP_EVAL, OP_DIV, OP_BATCH, OP_SERVE = range(1, 9)


def mul(a, b):
    o = fr_array(len(a))
    lib().or_vec_mul(P(o), P(np.ascontiguousarray(a)), P(np.ascontiguousarray(b)), len(a))
    return o


def add(a, b):
    o = fr_array(len(a))
    lib().or_vec_add(P(o), P(np.ascontiguousarray(a)), P(np.ascontiguousarray(b)), len(a))
    return o


def sub(a, b):
    o = fr_array(len(a))
    lib().or_vec_sub(P(o), P(np.ascontiguousarray(a)), P(np.ascontiguousarray(b)), len(a))
    return o


def inv(a):
    o = fr_array(len(a))
    lib().or_vec_inv(P(o), P(np.ascontiguousarray(a)), len(a))
    return o


def vec_exp(a, e):
    o = fr_array(len(a))
    lib().or_vec_exp(P(o), P(np.ascontiguousarray(a)), len(a), e)
    return o


def sequence(x, n):
    o = fr_array(n)
    lib().or_sequence(P(o), P(np.ascontiguousarray(x)), n)
    return o


def neg(a):
    o = fr_array(len(a))
    for i in range(len(a)):
        lib().or_neg(P(o[i:i + 1]), P(np.ascontiguousarray(a[i:i + 1])))
    return o


def pow_(a, e):
    o = fr_array(1)
    ee = np.array([(e >> (64 * i)) & 0xFFFFFFFFFFFFFFFF for i in range(4)], dtype=np.uint64)
    lib().or_pow(P(o), P(np.ascontiguousarray(a)), P(ee))
    return o


def root_of_unity(lg):
    o = fr_array(1)
    lib().or_root_of_unity(P(o), lg)
    return o


def domain_size_inverse(lg):
    o = fr_array(1)
    lib().or_domain_size_inverse(P(o), lg)
    return o


def set_threads(t):
    lib().or_set_threads(int(t))


def get_threads():
    return lib().or_get_threads()


def ntt(x, lg, inverse=False):
    y = np.array(x, dtype=np.uint64, copy=True, order="C")
    lib().or_ntt(P(y), lg, 1 if inverse else 0)
    return y


def coset_ntt(x, lg, shift):
    y = np.array(x, dtype=np.uint64, copy=True, order="C")
    lib().or_coset_ntt(P(y), lg, P(np.ascontiguousarray(shift)))
    return y


def coset_intt(x, lg, shift):
    y = np.array(x, dtype=np.uint64, copy=True, order="C")
    lib().or_coset_intt(P(y), lg, P(np.ascontiguousarray(shift)))
    return y


def lde(c, lg_n, lg_blowup, shift):
    out = fr_array(len(c) << lg_blowup)
    lib().or_lde(P(np.ascontiguousarray(c)), P(out), lg_n, lg_blowup, P(np.ascontiguousarray(shift)))
    return out


def root_of_unity_3n(lg):
    o = fr_array(1)
    lib().or_root_of_unity_3n(P(o), lg)
    return o


def lde3(c, lg_n, shift):
    out = fr_array(3 * len(c))
    lib().or_lde3(P(np.ascontiguousarray(c)), P(out), lg_n, P(np.ascontiguousarray(shift)))
    return out


def ref_lde_shift(lg_blowup):
    """our_lde's coset shift 7^(2^lg_blowup) (lde.cu:642-645, 659)."""
    return pow_(from_u64(7), 1 << lg_blowup)


def naive_lde(c, lg_n, lg_blowup):
    out = fr_array(len(c) << lg_blowup)
    lib().or_naive_lde(P(np.ascontiguousarray(c)), P(out), lg_n, lg_blowup)
    return out


def grand_product(p):
    y = np.array(p, dtype=np.uint64, copy=True, order="C")
    lib().or_grand_product(P(y), len(y))
    return y


def naive_gp(p):
    out = fr_array(len(p))
    lib().or_naive_gp(P(np.ascontiguousarray(p)), P(out), len(p))
    return out


def eval_(c, x):
    y = fr_array(1)
    lib().or_eval(P(np.ascontiguousarray(c)), len(c), P(np.ascontiguousarray(x)), P(y))
    return y


def div(a, c):
    y = np.array(a, dtype=np.uint64, copy=True, order="C")
    rem = fr_array(1)
    lib().or_div(P(y), len(y), P(np.ascontiguousarray(c)), P(rem))
    return y, rem


def naive_div(a, c, d1):
    q = fr_array(max(len(a) - 1, 0))
    lib().or_naive_div(P(np.ascontiguousarray(a)), len(a), P(np.ascontiguousarray(c)),
                       P(np.ascontiguousarray(d1)), P(q) if len(q) else None)
    return q


def perm_ratio(cols, beta, gamma, ks, lg):
    out = fr_array(1 << lg)
    args = [P(np.ascontiguousarray(c)) for c in cols] + [P(np.ascontiguousarray(beta)),
                                                         P(np.ascontiguousarray(gamma))]
    args += [P(np.ascontiguousarray(k)) for k in ks]
    lib().or_perm_ratio(*args, lg, P(out))
    return out


def stage_no_rotate(chunk, lg_domain, iterations, stage, chunk_id):
    """One Ntt::stage_no_rotate call on a copy of `chunk` (n x 4 uint64)."""
    out = np.ascontiguousarray(chunk, dtype=np.uint64).copy()
    rc = lib().or_stage_no_rotate(P(out), out.shape[0], lg_domain, iterations, stage, chunk_id)
    if rc != 0:
        raise ValueError("configuration rejected")
    return out


def ct_stage_indices(tids, stage, iterations):
    """ntt.cu:168-178, vectorised: global element pair of each gathered thread."""
    tids = np.asarray(tids, dtype=np.uint64)
    inp = np.uint64(1) << np.uint64(stage)
    outs = np.uint64(1) << np.uint64(stage + iterations - 1)
    pos = (tids & (outs - np.uint64(1))) >> np.uint64(iterations - 1)
    i0 = tids & ~(outs - np.uint64(1))
    i0 = i0 + ((tids << np.uint64(stage)) & (outs - np.uint64(1)))
    i0 = i0 * np.uint64(2) + pos
    return i0, i0 + inp


def rotate_stage_output_index(g, stage, iterations):
    """ntt.cu:190-195, vectorised."""
    g = np.asarray(g, dtype=np.uint64)
    mask = (np.uint64(1) << np.uint64(stage + iterations)) - (np.uint64(1) << np.uint64(stage))
    rot = g & mask
    rot = (rot >> np.uint64(1)) | (rot << np.uint64(iterations - 1))
    return (g & ~mask) | (rot & mask)


def ref_perm_grand_product(cols, beta, gamma, ks, omega):
    """The permutation grand product over the reference's own field (oracle/_ref);
    cols = [a, b, c, s1, s2, s3], ks = (3, 4) Montgomery constants."""
    ref = ref_lib()
    n = len(cols[0])
    cols = [np.ascontiguousarray(c) for c in cols]
    arr = _u64p * 3
    w = arr(*[P(cols[j]) for j in range(3)])
    s = arr(*[P(cols[3 + j]) for j in range(3)])
    z = fr_array(n)
    kk = np.ascontiguousarray(ks, dtype=np.uint64).reshape(3, 4)
    ref.ref_perm_grand_product(w, s, P(np.ascontiguousarray(beta)), P(np.ascontiguousarray(gamma)), P(kk),
                               P(np.ascontiguousarray(omega)), n, P(z))
    return z
