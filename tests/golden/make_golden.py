#!/usr/bin/env python3
"""Writes the golden fixtures under tests/golden/ from Python big integers.

Run once in a container that has the reference tree (it reads constant
tables and the reference's input file as DATA, never its code):

    python tests/golden/make_golden.py [--ref /root/reference]

Every expected value here comes from the closed-form definition of the
operator, evaluated with Python ints, independently of the C oracle and of
the GPU code:

  ref_constants.json  constant tables of the reference, parsed from
                      PLONK/utils/zkp/cuda/zksnark_ntt/parameters/bls12_381.h
                      (group_gen, roots of unity, 2^-k), Fr constants from
                      PLONK/utils/mont/cuda/ff/bls12-381.hpp:7-30, and the
                      miss_ops.cu literals (miss_ops.cu:29, 134, 164, 181).
  field_kat.npz       add/sub/mul/inv known answers incl. 0, 1, r-1.
  ntt_small.npz       brute-force DFT / inverse DFT, lg 0..8.
  lde_small.npz       brute-force coset LDE sum c_i (shift*w_M^k)^i and the
                      periodic naive_lde, lg_n 1..5, lg_blowup 1..3.
  scan_small.npz      grand product (exclusive, and naive_gp's inclusive
                      zero-tail layout), evaluation, division by (X + c),
                      naive_div, permutation ratio.
  input_prefix.bin    first 4096 elements of the reference's input.bin
                      (unzipped input.bin.zip), raw 32-byte limbs.
"""
import argparse
import io
import json
import os
import re
import zipfile

import numpy as np

R_MOD = 0x73EDA753299D7D483339D80809A1D80553BDA402FFFE5BFEFFFFFFFF00000001
RBITS = 1 << 256
RINV = pow(RBITS, -1, R_MOD)
HERE = os.path.dirname(os.path.abspath(__file__))


def mont(v):
    return (v % R_MOD) * RBITS % R_MOD


def unmont(m):
    return m * RINV % R_MOD


def limbs(m):
    return [(m >> (64 * i)) & 0xFFFFFFFFFFFFFFFF for i in range(4)]


def arr(ms):
    return np.array([limbs(m) for m in ms], dtype=np.uint64).reshape(-1, 4)


def from_u32(words):
    v = 0
    for i, w in enumerate(words):
        v |= (w & 0xFFFFFFFF) << (32 * i)
    return v


def from_u64(words):
    v = 0
    for i, w in enumerate(words):
        v |= (w & 0xFFFFFFFFFFFFFFFF) << (64 * i)
    return v


class SplitMix64:
    """SplitMix64 (Steele, Lea, Flood 2014), one 64-bit draw per call."""

    def __init__(self, seed):
        self.s = seed & 0xFFFFFFFFFFFFFFFF

    def next(self):
        self.s = (self.s + 0x9E3779B97F4A7C15) & 0xFFFFFFFFFFFFFFFF
        z = self.s
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & 0xFFFFFFFFFFFFFFFF
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & 0xFFFFFFFFFFFFFFFF
        return z ^ (z >> 31)


def uniform(n, seed):
    """Montgomery representations uniform in [0, r): 4 draws per element,
    top bit cleared, rejection of values >= r."""
    g = SplitMix64(seed)
    out = []
    while len(out) < n:
        l = [g.next() for _ in range(4)]
        l[3] &= 0x7FFFFFFFFFFFFFFF
        v = from_u64(l)
        if v < R_MOD:
            out.append(v)
    return out


def seed_of(op, lg, aux):
    return 0x504C4E4B00000000 ^ (op << 40) ^ (lg << 8) ^ aux


# all arithmetic below is on plain values; convert at the boundary
def omega(lg):
    g = 7
    w32 = pow(g, (R_MOD - 1) >> 32, R_MOD)
    return pow(w32, 1 << (32 - lg), R_MOD)


def dft(vals, w):
    n = len(vals)
    return [sum(vals[j] * pow(w, j * k, R_MOD) for j in range(n)) % R_MOD for k in range(n)]


def parse_u32_array(text, name):
    m = re.search(name + r"\s*=\s*\{(.*?)\}", text, re.S)
    return [int(x, 16) for x in re.findall(r"0x[0-9a-fA-F]+", m.group(1))]


def reference_constants(ref):
    p = os.path.join(ref, "PLONK/utils/zkp/cuda/zksnark_ntt/parameters/bls12_381.h")
    t = open(p).read()
    out = {}
    for name in ["group_gen", "group_gen_inverse", "forward_roots_of_unity",
                 "inverse_roots_of_unity", "domain_size_inverse"]:
        words = parse_u32_array(t, name)
        vals = [from_u32(words[i:i + 8]) for i in range(0, len(words), 8)]
        out[name] = [hex(v) for v in vals]
    h = open(os.path.join(ref, "PLONK/utils/mont/cuda/ff/bls12-381.hpp")).read()
    for name in ["BLS12_381_r", "BLS12_381_rRR", "BLS12_381_rone"]:
        m = re.search(name + r"\[8\]\s*=\s*\{(.*?)\}", h, re.S)
        words = [int(x, 16) for x in re.findall(r"0x[0-9a-fA-F]+", m.group(1))]
        out[name] = hex(from_u64(words))
    mo = open(os.path.join(ref, "miss_ops.cu")).read()
    m = re.search(r"fr_one\[4\]\s*=\s*\{(.*?)\}", mo, re.S)
    out["miss_ops_one"] = hex(from_u64([int(x.rstrip("UL")) for x in re.findall(r"\d+UL", m.group(1))]))
    m = re.search(r"pointt\[4\]\s*=\s*\{(.*?)\}", mo, re.S)
    out["miss_ops_point"] = hex(from_u64([int(x.rstrip("UL")) for x in re.findall(r"\d+UL", m.group(1))]))
    m = re.search(r"divisorr_gpu\[4\]\s*=\s*\{(.*?)\}", mo, re.S)
    out["miss_ops_divisor_gpu"] = hex(from_u64([int(x.rstrip("UL")) for x in re.findall(r"\d+UL", m.group(1))]))
    m = re.search(r"divisorr\[8\]\s*=\s*\{(.*?)\}", mo, re.S)
    w = [int(x.rstrip("UL")) for x in re.findall(r"\d+UL", m.group(1))]
    out["miss_ops_divisor_cpu"] = [hex(from_u64(w[0:4])), hex(from_u64(w[4:8]))]
    return out


def field_kat():
    special = [0, 1, R_MOD - 1, 2, R_MOD - 2, (R_MOD - 1) // 2, RBITS % R_MOD]
    rnd = [unmont(m) for m in uniform(64, seed_of(9, 0, 1))]
    vals = special + rnd
    a, b = [], []
    for x in vals:
        for y in special:
            a.append(x)
            b.append(y)
    for i in range(len(rnd)):
        a.append(rnd[i])
        b.append(rnd[(i * 7 + 3) % len(rnd)])
    add = [(x + y) % R_MOD for x, y in zip(a, b)]
    sub = [(x - y) % R_MOD for x, y in zip(a, b)]
    mul = [(x * y) % R_MOD for x, y in zip(a, b)]
    inv = [pow(x, -1, R_MOD) if x else 0 for x in a]
    np.savez(os.path.join(HERE, "field_kat.npz"),
             a=arr([mont(x) for x in a]), b=arr([mont(x) for x in b]),
             add=arr([mont(x) for x in add]), sub=arr([mont(x) for x in sub]),
             mul=arr([mont(x) for x in mul]), inv=arr([mont(x) for x in inv]))


def ntt_small():
    d = {}
    for lg in range(0, 9):
        n = 1 << lg
        x = [unmont(m) for m in uniform(n, seed_of(1, lg, 0))]
        w = omega(lg)
        y = dft(x, w)
        winv = pow(w, -1, R_MOD)
        ninv = pow(n, -1, R_MOD)
        z = [v * ninv % R_MOD for v in dft(x, winv)]
        d["in_%d" % lg] = arr([mont(v) for v in x])
        d["ntt_%d" % lg] = arr([mont(v) for v in y])
        d["intt_%d" % lg] = arr([mont(v) for v in z])
    np.savez(os.path.join(HERE, "ntt_small.npz"), **d)


def lde_small():
    d = {}
    for lg_n in range(1, 6):
        n = 1 << lg_n
        c = [unmont(m) for m in uniform(n, seed_of(3, lg_n, 0))]
        d["in_%d" % lg_n] = arr([mont(v) for v in c])
        for lb in range(1, 4):
            M = n << lb
            wM = omega(lg_n + lb)
            for tag, shift in (("ref", pow(7, 1 << lb, R_MOD)), ("g", 7)):
                out = []
                for k in range(M):
                    pt = shift * pow(wM, k, R_MOD) % R_MOD
                    out.append(sum(c[i] * pow(pt, i, R_MOD) for i in range(n)) % R_MOD)
                d["lde_%s_%d_%d" % (tag, lg_n, lb)] = arr([mont(v) for v in out])
            wn = omega(lg_n)
            y = dft(c, wn)
            d["naive_%d_%d" % (lg_n, lb)] = arr([mont(y[k % n]) for k in range(M)])
    np.savez(os.path.join(HERE, "lde_small.npz"), **d)


def scan_small(consts):
    d = {}
    for n in (1, 2, 3, 64, 257):
        p = [unmont(m) for m in uniform(n, seed_of(4, n, 0))]
        z, acc = [], 1
        for v in p:
            z.append(acc)
            acc = acc * v % R_MOD
        ng, acc = [], 1
        for i in range(n):
            if i + 1 < n:
                acc = acc * p[i] % R_MOD
                ng.append(acc)
            else:
                ng.append(0)
        if n == 1:
            ng = [1]  # naive_gp writes Mont(1) to out[0] and runs no step
        d["gp_in_%d" % n] = arr([mont(v) for v in p])
        d["gp_out_%d" % n] = arr([mont(v) for v in z])
        d["naive_gp_out_%d" % n] = arr([mont(v) for v in ng])
    for n in (1, 2, 3, 64, 100):
        c = [unmont(m) for m in uniform(n, seed_of(5, n, 0))]
        (x,) = [unmont(m) for m in uniform(1, seed_of(5, n, 1))]
        d["eval_in_%d" % n] = arr([mont(v) for v in c])
        for tag, pt in (("r", x), ("0", 0), ("1", 1)):
            d["eval_x_%s_%d" % (tag, n)] = arr([mont(pt)])
            d["eval_y_%s_%d" % (tag, n)] = arr([mont(sum(c[i] * pow(pt, i, R_MOD) for i in range(n)) % R_MOD)])
    for n in (1, 2, 3, 64, 100):
        a = [unmont(m) for m in uniform(n, seed_of(6, n, 0))]
        (cc,) = [unmont(m) for m in uniform(1, seed_of(6, n, 1))]
        z = (-cc) % R_MOD
        q = [sum(a[j] * pow(z, j - i - 1, R_MOD) for j in range(i + 1, n)) % R_MOD for i in range(n)]
        rem = sum(a[j] * pow(z, j, R_MOD) for j in range(n)) % R_MOD
        d["div_in_%d" % n] = arr([mont(v) for v in a])
        d["div_c_%d" % n] = arr([mont(cc)])
        d["div_q_%d" % n] = arr([mont(v) for v in q])
        d["div_rem_%d" % n] = arr([mont(rem)])
    # miss_ops literals: eval point and GPU divisor used as raw Montgomery limbs
    pt = unmont(int(consts["miss_ops_point"], 16))
    cgpu = unmont(int(consts["miss_ops_divisor_gpu"], 16))
    a = [unmont(m) for m in uniform(256, seed_of(6, 256, 7))]
    d["mo_poly"] = arr([mont(v) for v in a])
    d["mo_eval"] = arr([mont(sum(a[i] * pow(pt, i, R_MOD) for i in range(256)) % R_MOD)])
    z = (-cgpu) % R_MOD
    d["mo_div_q"] = arr([mont(sum(a[j] * pow(z, j - i - 1, R_MOD) for j in range(i + 1, 256)) % R_MOD)
                         for i in range(256)])
    # permutation ratio, k = (1, 7, 13), omega = omega_{2^lg}
    for lg in (3, 6):
        n = 1 << lg
        cols = [[unmont(m) for m in uniform(n, seed_of(4, lg, 10 + j))] for j in range(6)]
        beta, gamma = [unmont(m) for m in uniform(2, seed_of(4, lg, 20))]
        w = omega(lg)
        ks = (1, 7, 13)
        ratio = []
        for i in range(n):
            num = den = 1
            for j in range(3):
                num = num * (cols[j][i] + beta * ks[j] * pow(w, i, R_MOD) + gamma) % R_MOD
                den = den * (cols[j][i] + beta * cols[3 + j][i] + gamma) % R_MOD
            ratio.append(num * pow(den, -1, R_MOD) % R_MOD if den else 0)
        for j, nm in enumerate(["a", "b", "c", "s1", "s2", "s3"]):
            d["perm_%s_%d" % (nm, lg)] = arr([mont(v) for v in cols[j]])
        d["perm_beta_%d" % lg] = arr([mont(beta)])
        d["perm_gamma_%d" % lg] = arr([mont(gamma)])
        d["perm_ratio_%d" % lg] = arr([mont(v) for v in ratio])
    np.savez(os.path.join(HERE, "scan_small.npz"), **d)


def input_prefix(ref):
    with zipfile.ZipFile(os.path.join(ref, "input.bin.zip")) as z:
        name = [n for n in z.namelist() if n.endswith("input.bin")][0]
        with z.open(name) as f:
            data = f.read(4096 * 32)
    with open(os.path.join(HERE, "input_prefix.bin"), "wb") as f:
        f.write(data)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--ref", default="/root/reference")
    args = ap.parse_args()
    consts = reference_constants(args.ref)
    with open(os.path.join(HERE, "ref_constants.json"), "w") as f:
        json.dump(consts, f, indent=1)
    field_kat()
    ntt_small()
    lde_small()
    scan_small(consts)
    input_prefix(args.ref)


if __name__ == "__main__":
    main()
