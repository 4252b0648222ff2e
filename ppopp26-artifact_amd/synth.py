"""Synthetic inputs and workload shapes for the benchmarks and tests.

uniform(n, seed) draws n field elements uniform in [0, r) in Montgomery
limb form (the raw limbs are the Montgomery representation): each candidate
is four SplitMix64 outputs (little-endian limbs), the top bit of the highest
limb cleared, rejected while >= r.  The draw sequence is the one the CLIs'
seeded fill uses (cli/common.hpp), so a CLI run and a Python run on the same
seed see the same polynomial.  SplitMix64's k-th output depends only on
seed + k * golden, so the stream is generated in vectorised chunks.

ragged_batch(seed) is the ragged proof batch: 64 polynomials of lg drawn
uniformly from 12..20, 40 forward NTTs, 16 coset LDEs (lambda = 2) and 8
evaluation + division pairs, in a seeded order.
"""
import numpy as np

R_LIMBS = (0xffffffff00000001, 0x53bda402fffe5bfe, 0x3339d80809a1d805, 0x73eda753299d7d48)
_GOLDEN = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)

OP_NTT, OP_INTT, OP_LDE, OP_GP, OP_EVAL, OP_DIV, OP_BATCH, OP_SERVE = 1, 2, 3, 4, 5, 6, 7, 8


def seed(op, lg, aux=0):
    """Per-workload seed: 0x504C4E4B00000000 ^ op << 40 ^ lg << 8 ^ aux."""
    return (0x504C4E4B00000000 ^ (op << 40) ^ (lg << 8) ^ aux) & 0xFFFFFFFFFFFFFFFF


def _mix(z):
    z = (z ^ (z >> np.uint64(30))) * _M1
    z = (z ^ (z >> np.uint64(27))) * _M2
    return z ^ (z >> np.uint64(31))


def _below_r(c):
    """c: (m, 4) uint64 limbs, little-endian -> bool mask c < r"""
    lt = np.zeros(len(c), dtype=bool)
    eq = np.ones(len(c), dtype=bool)
    for j in (3, 2, 1, 0):
        rj = np.uint64(R_LIMBS[j])
        lt |= eq & (c[:, j] < rj)
        eq &= c[:, j] == rj
    return lt


def uniform(n, s, chunk=1 << 20):
    out = np.empty((n, 4), dtype=np.uint64)
    base = np.uint64(s & 0xFFFFFFFFFFFFFFFF)
    k = 0  # draws consumed
    filled = 0
    with np.errstate(over="ignore"):
        while filled < n:
            need = min(n - filled, chunk)
            m = need + need // 8 + 64
            t = np.arange(k + 1, k + 4 * m + 1, dtype=np.uint64)
            cand = _mix(base + t * _GOLDEN).reshape(m, 4)
            cand[:, 3] &= np.uint64(0x7FFFFFFFFFFFFFFF)
            idx = np.flatnonzero(_below_r(cand))
            if len(idx) >= need:
                idx = idx[:need]
                k += 4 * (int(idx[-1]) + 1)
            else:
                k += 4 * m
            out[filled:filled + len(idx)] = cand[idx]
            filled += len(idx)
    return out


def digest(a):
    """The device digest (pl_digest) on the host: sum_k (w_k ^ k*golden) *
    (2k+1) mod 2^64 over the 64-bit words of a (n, 4) uint64 array."""
    w = np.ascontiguousarray(a, dtype=np.uint64).ravel()
    k = np.arange(len(w), dtype=np.uint64)
    with np.errstate(over="ignore"):
        t = (w ^ (k * _GOLDEN)) * (np.uint64(2) * k + np.uint64(1))
        return int(np.sum(t, dtype=np.uint64))


def served_jobs(rate_index, n_jobs, lg_lo=14, lg_hi=24):
    """[(lg, lg_blowup, t_arrival_unit)] for the served workload: lg uniform
    in lg_lo..lg_hi, lg_blowup 2 or 4, Poisson arrivals with unit rate (scale
    the times by 1/rate)."""
    rng = np.random.default_rng(seed(OP_SERVE, 0, rate_index))
    lgs = rng.integers(lg_lo, lg_hi + 1, size=n_jobs)
    lbs = np.where(rng.random(n_jobs) < 0.5, 2, 4)
    t = np.cumsum(rng.exponential(1.0, size=n_jobs))
    t -= t[0]
    return [(int(a), int(b), float(c)) for a, b, c in zip(lgs, lbs, t)]


def ragged_batch(s=None):
    """[(op, lg)] for the ragged proof batch, op in ntt / lde / eval_div."""
    rng = np.random.default_rng(seed(OP_BATCH, 0) if s is None else s)
    lgs = rng.integers(12, 21, size=64)
    ops = np.array(["ntt"] * 40 + ["lde"] * 16 + ["eval_div"] * 8)
    ops = ops[rng.permutation(64)]
    return [(str(o), int(l)) for o, l in zip(ops, lgs)]
