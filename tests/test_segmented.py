"""The segmented NTT step (Ntt::stage_no_rotate) in the oracle, pinned by the
reference driver's own composition (CPU only).

The oracle restates _CT_NTT_without_rotate (ct_mixed_radix_wide.cuh:610-751)
as its thread program.  The driver (ntt.cu:801-977) runs a transform as a
sequence of such calls over chunks gathered with ct_stage_indices and
scattered with rotate_stage_output_index on bit-reversed input; that sequence
must equal the NTT.  The schedules below cover the three factor rules (stage
0; the windowed w_{2^28} powers; the w_{2^L} row-column tables at (6,6),
(6,12), (7,7), (8,8)), iteration counts below 6, and chunking with chunk ids.
"""
import sys
import os

import numpy as np
import pytest

sys.path.insert(0, os.path.join(os.path.dirname(__file__), "..", "oracle"))
import pyoracle as po  # noqa: E402


def composed_ntt(x, lg, schedule, lg_chunk, step=None):
    """The driver's flow: bit-reverse, then per (stage, it): gather, one call
    per chunk, rotate-scatter.  `step(chunk, lg, it, stage, chunk_id)` runs
    one call (default: the oracle)."""
    step = step or po.stage_no_rotate
    n = 1 << lg
    rev = np.array([po.lib().or_bitrev(i, lg) for i in range(n)], dtype=np.int64)
    data = np.empty_like(x)
    data[rev] = x
    for stage, it in schedule:
        tids = np.arange(n // 2, dtype=np.uint64)
        g0, g1 = po.ct_stage_indices(tids, stage, it)
        c = 1 << lg_chunk
        per = c // 2
        t = tids.astype(np.int64)
        pos0 = (t // per) * c + 2 * (t % per)
        pos1 = pos0 + 1
        gat = np.empty_like(data)
        gat[pos0] = data[g0.astype(np.int64)]
        gat[pos1] = data[g1.astype(np.int64)]
        for k in range(n // c):
            gat[k * c:(k + 1) * c] = step(gat[k * c:(k + 1) * c], lg, it, stage, k)
        out = np.empty_like(data)
        out[po.rotate_stage_output_index(g0, stage, it).astype(np.int64)] = gat[pos0]
        out[po.rotate_stage_output_index(g1, stage, it).astype(np.int64)] = gat[pos1]
        data = out
    return data


SCHEDULES = [
    (11, [(0, 5), (5, 6)], 11),
    (12, [(0, 6), (6, 6)], 12),
    (12, [(0, 6), (6, 6)], 8),
    (13, [(0, 6), (6, 7)], 9),
    (14, [(0, 7), (7, 7)], 10),
    (16, [(0, 8), (8, 8)], 12),
    (16, [(0, 6), (6, 10)], 11),
    (12, [(0, 3), (3, 4), (7, 5)], 6),
]


@pytest.mark.parametrize("lg,schedule,lg_chunk", SCHEDULES)
def test_driver_composition_is_the_ntt(lg, schedule, lg_chunk):
    x = po.uniform(1 << lg, po.seed(1, lg, 0))
    assert np.array_equal(composed_ntt(x, lg, schedule, lg_chunk), po.ntt(x, lg))


def test_three_step_schedule_with_the_stage_12_table():
    lg = 18
    x = po.uniform(1 << lg, po.seed(1, lg, 1))
    assert np.array_equal(composed_ntt(x, lg, [(0, 6), (6, 6), (12, 6)], 14), po.ntt(x, lg))


def test_index_helpers_match_the_restatement():
    import ctypes
    for stage, it in ((0, 6), (6, 6), (3, 5), (12, 10)):
        t = np.arange(0, 5000, 7, dtype=np.uint64)
        a, b = po.ct_stage_indices(t, stage, it)
        for k in range(0, len(t), 40):
            x, y = ctypes.c_uint64(), ctypes.c_uint64()
            po.lib().or_ct_stage_indices(int(t[k]), stage, it, ctypes.cast(ctypes.pointer(x), po._u64p),
                                         ctypes.cast(ctypes.pointer(y), po._u64p))
            assert (x.value, y.value) == (int(a[k]), int(b[k]))
            r = po.lib().or_rotate_stage_output_index(int(a[k]), stage, it)
            assert r == int(po.rotate_stage_output_index(a[k:k + 1], stage, it)[0])


def test_rejected_configurations():
    c = po.uniform(1 << 10, 5)
    for lg, it, stage in ((20, 9, 0), (10, 6, 6), (30, 6, 0)):
        with pytest.raises(ValueError):
            po.stage_no_rotate(c, lg, it, stage, 0)


def test_library_index_algebra_matches_the_oracle():
    """pl_ct_stage_indices / pl_rotate_stage_output_index (host functions of
    libpl, no device needed) against the oracle's restatement of
    ntt.cu:168-195."""
    import plffi
    if not os.path.exists(plffi.LIB_PATH):
        pytest.skip("libpl.so not built")
    for stage, it in ((0, 6), (6, 6), (3, 5), (12, 10), (14, 10), (0, 1), (9, 1), (16, 8)):
        t = np.arange(0, 20000, 13, dtype=np.uint64)
        a, b = po.ct_stage_indices(t, stage, it)
        ra = po.rotate_stage_output_index(a, stage, it)
        for k in range(0, len(t), 17):
            assert plffi.ct_stage_indices(int(t[k]), stage, it) == (int(a[k]), int(b[k]))
            assert plffi.rotate_stage_output_index(int(a[k]), stage, it) == int(ra[k])


def reference_flow_host(x, lg, lg_i, lg_k):
    """our_ntt's flow with the oracle's stage step and numpy permutations:
    bit reversal; pass 1 per chunk: step (0, lg_i/2) on the chunk as it
    lies, transition to the (lg_i/2, rest) gathered order, step, chunk-local
    rotate-scatter; pass 2: gather for (lg_i, lg - lg_i), step, scatter."""
    n = 1 << lg
    C = 1 << (lg_i + lg_k)
    first, second, it2 = lg_i // 2, lg_i - lg_i // 2, lg - lg_i
    rev = np.array([po.lib().or_bitrev(i, lg) for i in range(n)], dtype=np.int64)
    W = np.empty_like(x)
    W[rev] = x
    half = C // 2
    for c in range(n // C):
        chunk = po.stage_no_rotate(W[c * C:(c + 1) * C], lg, first, 0, c)
        tids = np.arange(c * half, (c + 1) * half, dtype=np.uint64)
        g0, g1 = po.ct_stage_indices(tids, first, second)
        # transition: source = the previous step's rotation undone
        def back(g):
            mask = np.uint64(((1 << (first)) - 1) << 0)
            r = g & mask
            r = ((r << np.uint64(1)) | (r >> np.uint64(first - 1))) & mask
            return ((g & ~mask) | r).astype(np.int64) - c * C
        S = np.empty_like(chunk)
        S[0::2] = chunk[back(g0)]
        S[1::2] = chunk[back(g1)]
        S = po.stage_no_rotate(S, lg, second, first, c)
        out = np.empty_like(S)
        out[po.rotate_stage_output_index(g0, first, second).astype(np.int64) - c * C] = S[0::2]
        out[po.rotate_stage_output_index(g1, first, second).astype(np.int64) - c * C] = S[1::2]
        W[c * C:(c + 1) * C] = out
    y = np.empty_like(x)
    for c in range(n // C):
        tids = np.arange(c * half, (c + 1) * half, dtype=np.uint64)
        g0, g1 = po.ct_stage_indices(tids, lg_i, it2)
        S = np.empty((C, 4), dtype=np.uint64)
        S[0::2] = W[g0.astype(np.int64)]
        S[1::2] = W[g1.astype(np.int64)]
        S = po.stage_no_rotate(S, lg, it2, lg_i, c)
        y[po.rotate_stage_output_index(g0, lg_i, it2).astype(np.int64)] = S[0::2]
        y[po.rotate_stage_output_index(g1, lg_i, it2).astype(np.int64)] = S[1::2]
    return y


@pytest.mark.parametrize("lg,lg_i,lg_k", [(12, 6, 1), (14, 8, 2), (16, 8, 0), (12, 4, 4), (13, 6, 1)])
def test_reference_flow_with_transition_is_the_ntt(lg, lg_i, lg_k):
    """The driver's pass-1 transition (regathering a chunk for the next step
    without the whole vector) composes to the NTT: the CPU model that the
    device flow pl_ntt_segmented follows."""
    x = po.uniform(1 << lg, po.seed(1, lg, 7))
    assert np.array_equal(reference_flow_host(x, lg, lg_i, lg_k), po.ntt(x, lg))
