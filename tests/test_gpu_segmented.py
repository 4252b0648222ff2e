"""GPU parity of the segmented NTT step (pl_ntt_stage_no_rotate, the
reference's Ntt::stage_no_rotate) against the oracle's restatement, through
the C-ABI, bit-exact; and the driver's composition of such calls on the GPU
equal to the NTT."""
import numpy as np
import pytest

import pyoracle as po
from test_segmented import composed_ntt

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def pl():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a HIP device"
    import plffi
    plffi.lib()
    return plffi


@pytest.fixture(scope="module")
def params(pl):
    return pl.params()


# (lg_domain, iterations, stage, lg_chunk, chunk_id): the three factor rules,
# blocks smaller than a radix block, iterations 1..8 and 10, chunk ids > 0
CASES = [
    (12, 6, 0, 12, 0), (12, 6, 0, 9, 3), (12, 6, 6, 10, 2), (20, 6, 6, 12, 5),
    (12, 5, 6, 8, 1), (18, 6, 12, 10, 7), (14, 7, 7, 12, 1), (16, 8, 8, 12, 9),
    (16, 10, 6, 11, 0), (22, 10, 12, 12, 100), (13, 7, 3, 8, 4), (12, 3, 7, 4, 17),
    (10, 1, 9, 1, 300), (10, 2, 0, 2, 5), (24, 8, 16, 14, 777), (28, 10, 18, 12, 2000),
]


@pytest.mark.parametrize("lg,it,stage,lgc,cid", CASES)
def test_stage_vs_oracle(pl, params, lg, it, stage, lgc, cid):
    x = po.uniform(1 << lgc, po.seed(po.OP_NTT, lg, stage * 16 + it))
    print("stage_no_rotate lg=%d it=%d stage=%d chunk 2^%d id %d" % (lg, it, stage, lgc, cid), flush=True)
    d = pl.to_dev(x)
    pl.stage_no_rotate(d, lg, it, stage, cid, params)
    np.testing.assert_array_equal(pl.to_host(d), po.stage_no_rotate(x, lg, it, stage, cid))


@pytest.mark.parametrize("lg,schedule,lgc", [(12, [(0, 6), (6, 6)], 9), (16, [(0, 8), (8, 8)], 12),
                                             (18, [(0, 6), (6, 6), (12, 6)], 14), (20, [(0, 10), (10, 10)], 16)])
def test_driver_composition_on_gpu(pl, params, lg, schedule, lgc):
    x = po.uniform(1 << lg, po.seed(po.OP_NTT, lg, 99))
    print("composed lg=%d %s" % (lg, schedule), flush=True)

    def step(chunk, lg_, it, stage, cid):
        d = pl.to_dev(chunk)
        pl.stage_no_rotate(d, lg_, it, stage, cid, params)
        return pl.to_host(d)

    np.testing.assert_array_equal(composed_ntt(x, lg, schedule, lgc, step), po.ntt(x, lg))


def test_argument_errors(pl, params):
    import ctypes
    L = pl.lib()
    d = pl.to_dev(po.uniform(64, 1))
    ptr = ctypes.c_void_p(d.data_ptr())
    pp = ctypes.c_void_p(params.data_ptr())
    assert L.pl_ntt_stage_no_rotate(ptr, 64, 20, 9, 0, 0, pp, None) == -3   # radix 9
    assert L.pl_ntt_stage_no_rotate(ptr, 64, 10, 6, 6, 0, pp, None) == -1   # stage + it > lg
    assert L.pl_ntt_stage_no_rotate(ptr, 32, 10, 6, 0, 0, pp, None) == -1   # chunk < 2^it
    assert L.pl_ntt_stage_no_rotate(ptr, 48, 10, 4, 0, 0, pp, None) == -1   # not a power of two
    assert L.pl_ntt_stage_no_rotate(ptr, 64, 10, 6, 0, 0, None, None) == -1


def _rot(g, stage, it, left=False):
    mask = np.uint64(((1 << it) - 1) << stage)
    r = g & mask
    if left:
        r = ((r << np.uint64(1)) | (r >> np.uint64(it - 1))) & mask
    else:
        r = ((r >> np.uint64(1)) | (r << np.uint64(it - 1))) & mask
    return (g & ~mask) | r


@pytest.mark.parametrize("lg,lgc,cid,stage,it", [(12, 8, 3, 0, 6), (12, 8, 5, 4, 4), (14, 10, 7, 8, 6),
                                                 (16, 12, 2, 6, 10)])
def test_chunk_permutations_vs_index_algebra(pl, lg, lgc, cid, stage, it):
    """pl_stage_gather / pl_stage_scatter / pl_stage_transition against
    numpy indexing by the oracle's ct_stage_indices and rotation."""
    import torch
    n, C = 1 << lg, 1 << lgc
    x = po.uniform(n, po.seed(po.OP_NTT, lg, 300 + stage))
    xd = pl.to_dev(x)
    half = C // 2
    tids = np.arange(cid * half, (cid + 1) * half, dtype=np.uint64)
    g0, g1 = po.ct_stage_indices(tids, stage, it)
    print("chunk permutations lg=%d chunk 2^%d id %d step (%d,%d)" % (lg, lgc, cid, stage, it), flush=True)
    # gather
    ch = torch.empty((C, 4), dtype=torch.int64, device="cuda")
    pl.stage_gather(xd, ch, cid, stage, it)
    want = np.empty((C, 4), dtype=np.uint64)
    want[0::2] = x[g0.astype(np.int64)]
    want[1::2] = x[g1.astype(np.int64)]
    np.testing.assert_array_equal(pl.to_host(ch), want)
    # scatter, global and chunk-local
    chunk = po.uniform(C, 9)
    out = torch.zeros_like(xd)
    pl.stage_scatter(pl.to_dev(chunk), out, cid, stage, it)
    want = np.zeros_like(x)
    want[_rot(g0, stage, it).astype(np.int64)] = chunk[0::2]
    want[_rot(g1, stage, it).astype(np.int64)] = chunk[1::2]
    np.testing.assert_array_equal(pl.to_host(out), want)
    if stage + it <= lgc:
        loc = torch.zeros((C, 4), dtype=torch.int64, device="cuda")
        pl.stage_scatter(pl.to_dev(chunk), loc, cid, stage, it, chunk_local=True)
        np.testing.assert_array_equal(pl.to_host(loc), want[cid * C:(cid + 1) * C])
        # transition from the step (0, stage) to (stage, it) within the chunk
        if stage >= 1:
            h0, h1 = po.ct_stage_indices(tids, stage, it)
            tr = torch.empty((C, 4), dtype=torch.int64, device="cuda")
            pl.stage_transition(pl.to_dev(chunk), tr, cid, 0, stage, stage, it)
            want_t = np.empty((C, 4), dtype=np.uint64)
            want_t[0::2] = chunk[_rot(h0, 0, stage, left=True).astype(np.int64) - cid * C]
            want_t[1::2] = chunk[_rot(h1, 0, stage, left=True).astype(np.int64) - cid * C]
            np.testing.assert_array_equal(pl.to_host(tr), want_t)


def test_bit_reverse(pl):
    lg = 13
    x = po.uniform(1 << lg, 11)
    rev = np.array([po.lib().or_bitrev(i, lg) for i in range(1 << lg)], dtype=np.int64)
    want = np.empty_like(x)
    want[rev] = x
    np.testing.assert_array_equal(pl.to_host(pl.bit_reverse(pl.to_dev(x))), want)


@pytest.mark.parametrize("lg,lg_i,lg_k", [(12, 6, 1), (16, 8, 2), (20, 10, 2), (22, 12, 4), (24, 14, 4)])
def test_segmented_flow_vs_oracle(pl, params, lg, lg_i, lg_k):
    """The reference's chunked flow (bit reversal, pass 1 with the
    transition, pass 2 with gather/scatter) run on the device equals the
    NTT, over the whole vector, at the reference's splits (22: 12+10, lg_k 4;
    24: 14+10, lg_k 4)."""
    from conftest import assert_fr_equal
    x = po.uniform(1 << lg, po.seed(po.OP_NTT, lg, 400))
    print("segmented flow lg=%d lg_i=%d lg_k=%d" % (lg, lg_i, lg_k), flush=True)
    y = pl.to_host(pl.ntt_segmented(pl.to_dev(x), lg_i, lg_k, params))
    assert_fr_equal(y, po.ntt(x, lg), "segmented %d" % lg)


def test_segmented_flow_rejects_reference_invalid_splits(pl, params):
    import torch
    t = torch.zeros((1 << 12, 4), dtype=torch.int64, device="cuda")
    L = pl.lib()
    import ctypes
    p = ctypes.c_void_p(t.data_ptr())
    q = ctypes.c_void_p(params.data_ptr())
    assert L.pl_ntt_segmented(p, 12, 2, 1, q, None, None) == -3   # pass 2 of 10 stages > chunk of 2^3
    assert L.pl_ntt_segmented(p, 12, 3, 9, q, None, None) == -3   # 12 - 3 = 9 stages: rejected as by the reference
    assert L.pl_ntt_segmented(p, 12, 6, 7, q, None, None) == -3   # chunk larger than the vector


@pytest.mark.parametrize("lg,npoly,lg_k,inverse", [(6, 1, 0, False), (10, 3, 2, False), (16, 1, 4, True),
                                                   (20, 2, 4, False), (20, 3, 8, True), (24, 1, 4, False)])
def test_host_to_host_ntt_vs_oracle(pl, lg, npoly, lg_k, inverse):
    """pl_ntt_host (pinned host in and out, grouped copies overlapping the
    first and last passes, two device slots) against the oracle."""
    import torch
    from conftest import assert_fr_equal
    n = 1 << lg
    x = po.uniform(npoly * n, po.seed(po.OP_NTT, lg, 500 + npoly))
    h = torch.from_numpy(x.view(np.int64)).pin_memory()
    print("host ntt lg=%d npoly=%d lg_k=%d inv=%s" % (lg, npoly, lg_k, inverse), flush=True)
    dom = pl.Domain(lg)
    out = pl.ntt_host(h, dom, lg_k=lg_k, npoly=npoly, inverse=inverse)
    got = out.numpy().view(np.uint64)
    for p in range(npoly):
        assert_fr_equal(got[p * n:(p + 1) * n], po.ntt(x[p * n:(p + 1) * n], lg, inverse=inverse),
                        "host ntt poly %d" % p)
    # in place in the host buffer, as our_ntt leaves its result in `input`
    pl.ntt_host(h, dom, lg_k=lg_k, npoly=npoly, inverse=inverse, out=h)
    assert_fr_equal(h.numpy().view(np.uint64), got, "host ntt in place")


def test_host_copy_groups_tuned(pl, tmp_path):
    """the autotuner's host pass (PL_TUNE_HOST) times pl_ntt_host over the
    copy-group counts, caches "h" lines and applies them; lg_k = -1 then runs
    with the tuned count, against the oracle"""
    import torch
    cache = tmp_path / "plans.txt"
    # splits already cached, so only the host-to-host copy groups are timed
    splits = ["%d %d %s 0.0" % (lg, len(pl.plan_get(lg)), " ".join(map(str, pl.plan_get(lg)))) for lg in (14, 15)]
    cache.write_text("\n".join(splits) + "\n")
    print("autotune host 14..15", flush=True)
    pl.autotune(14, 15, str(cache), reps=2, host=True)
    lines = [l.split() for l in cache.read_text().splitlines() if l.startswith("h ")]
    got = {int(l[1]): int(l[2]) for l in lines}
    assert sorted(got) == [14, 15] and all(0 <= k <= 6 for k in got.values())
    assert pl.host_groups(14) == got[14] and pl.host_groups(15) == got[15]
    lg = 15
    x = po.uniform(1 << lg, po.seed(po.OP_NTT, lg, 900))
    h = torch.from_numpy(x.view(np.int64)).pin_memory()
    out = pl.ntt_host(h, pl.Domain(lg))  # lg_k = -1: the tuned count
    np.testing.assert_array_equal(out.numpy().view(np.uint64), po.ntt(x, lg))
    for g in (14, 15):
        pl.lib().pl_ntt_host_groups_set(g, -1)
