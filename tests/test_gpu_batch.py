"""GPU parity of the batched form (pl_batch_*): every item of a batch equals
the per-call entry points and the oracle, bit for bit, on graph replays and
without a graph."""
import numpy as np
import pytest

import pyoracle as po
import synth

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def pl():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a HIP device"
    import plffi
    plffi.lib()
    return plffi


def make_items(pl, spec, seed_base):
    """spec: [(op, lg, lg_blowup)] -> batch items with seeded inputs."""
    import torch
    items = []
    for i, (op, lg, lb) in enumerate(spec):
        n = 1 << lg
        x = synth.uniform(n, synth.seed(synth.OP_BATCH, lg, seed_base + i))
        it = {"op": op, "lg": lg, "lg_blowup": lb, "host_in": x, "inp": pl.to_dev(x)}
        n_out = n << lb if op == "lde" else n
        it["out"] = torch.zeros((n_out, 4), dtype=torch.int64, device="cuda")
        if op == "eval_div":
            z = synth.uniform(1, synth.seed(synth.OP_BATCH, lg, seed_base + 1000 + i))
            it["host_z"] = z
            it["z"] = pl.to_dev(z)
            it["y"] = torch.zeros((1, 4), dtype=torch.int64, device="cuda")
        items.append(it)
    return items


def per_call(pl, it):
    """The same item through the per-call entry points."""
    if it["op"] in ("ntt", "intt"):
        return pl.to_host(pl.ntt(it["inp"], pl.Domain(it["lg"]), inverse=it["op"] == "intt")), None
    if it["op"] == "lde":
        return pl.to_host(pl.lde(it["inp"], pl.LdeDomain(it["lg"], it["lg_blowup"]))), None
    y = pl.evaluate(it["inp"], it["z"])
    q = it["inp"].clone()
    negz = pl.vec_op("neg", it["z"])
    pl.div_linear(q, negz)
    return pl.to_host(q), pl.to_host(y)


def oracle(it):
    x, lg = it["host_in"], it["lg"]
    if it["op"] in ("ntt", "intt"):
        return po.ntt(x, lg, inverse=it["op"] == "intt"), None
    if it["op"] == "lde":
        return po.lde(x, lg, it["lg_blowup"], po.ref_lde_shift(it["lg_blowup"])), None
    q, rem = po.div(x, po.neg(it["host_z"]))
    return q, po.eval_(x, it["host_z"])


def snapshot(pl, items):
    return [(pl.to_host(it["out"]), pl.to_host(it["y"]) if "y" in it else None) for it in items]


def test_small_batch_vs_oracle(pl):
    # every launch form: tiny (lg < 4), one pass, premultiplied one pass
    # (LDE lg <= 8), two and three passes; eval+div down to n = 1
    spec = [("ntt", lg, 0) for lg in (1, 2, 3, 4, 5, 8, 9, 12, 17)]
    spec += [("intt", lg, 0) for lg in (1, 4, 7, 10, 13)]
    spec += [("lde", 1, 1), ("lde", 2, 3), ("lde", 3, 4), ("lde", 6, 2), ("lde", 8, 2), ("lde", 9, 2),
             ("lde", 12, 4), ("lde", 13, 1), ("lde", 16, 2)]
    spec += [("eval_div", lg, 0) for lg in (0, 1, 5, 12, 13)]
    items = make_items(pl, spec, 0)
    b = pl.Batch(items, graph=True)
    b.run()
    got = snapshot(pl, items)
    for it, (out, y) in zip(items, got):
        print("batch item", it["op"], it["lg"], it["lg_blowup"], flush=True)
        ref, ry = oracle(it)
        np.testing.assert_array_equal(out, ref)
        if ry is not None:
            np.testing.assert_array_equal(y, ry)
    n, nbytes = b.info()
    assert n > 0 and nbytes > 0
    b.close()


def test_ragged_batch_matches_per_call(pl):
    import torch
    spec = [(op, lg, 2 if op == "lde" else 0) for op, lg in synth.ragged_batch()]
    items = make_items(pl, spec, 0)
    b = pl.Batch(items, graph=True)
    b.run()
    torch.cuda.synchronize()
    first = snapshot(pl, items)
    for it, (out, y) in zip(items, first):
        print("ragged item", it["op"], it["lg"], flush=True)
        ref, ry = per_call(pl, it)
        np.testing.assert_array_equal(out, ref)
        if ry is not None:
            np.testing.assert_array_equal(y, ry)
    # graph replays with eager work in between give the same results
    for rep in range(3):
        for it in items:
            it["out"].zero_()
            if "y" in it:
                it["y"].zero_()
        torch.cuda.synchronize()
        b.run()
        torch.cuda.synchronize()
        again = snapshot(pl, items)
        for (o1, y1), (o2, y2) in zip(first, again):
            np.testing.assert_array_equal(o1, o2)
            if y1 is not None:
                np.testing.assert_array_equal(y1, y2)
    # the plain sequence (one launch sequence per polynomial) agrees too
    for it in items:
        it["out"].zero_()
    b.run_plain()
    torch.cuda.synchronize()
    for (o1, y1), (o2, y2) in zip(first, snapshot(pl, items)):
        np.testing.assert_array_equal(o1, o2)
        if y1 is not None:
            np.testing.assert_array_equal(y1, y2)
    b.close()


def test_ragged_batch_vs_oracle(pl):
    # oracle check on every item of the ragged workload (lg 12..20; the
    # oracle's transforms run on the host's threads)
    spec = [(op, lg, 2 if op == "lde" else 0) for op, lg in synth.ragged_batch()]
    items = make_items(pl, spec, 500)
    b = pl.Batch(items, graph=False)
    b.run()
    for it, (out, y) in zip(items, snapshot(pl, items)):
        print("ragged oracle item", it["op"], it["lg"], flush=True)
        ref, ry = oracle(it)
        np.testing.assert_array_equal(out, ref)
        if ry is not None:
            np.testing.assert_array_equal(y, ry)
    b.close()


def test_batch_in_place_ntt(pl):
    import torch
    x = synth.uniform(1 << 12, 5)
    t = pl.to_dev(x)
    items = [{"op": "ntt", "lg": 12, "lg_blowup": 0, "inp": t, "out": t},
             {"op": "intt", "lg": 6, "lg_blowup": 0, "inp": t[:64], "out": torch.zeros((64, 4), dtype=torch.int64,
                                                                                     device="cuda")}]
    # items are independent: the second reads a region the first overwrites,
    # which the batch does not allow, so use separate buffers instead
    items[1]["inp"] = pl.to_dev(x[:64])
    b = pl.Batch(items)
    b.run()
    np.testing.assert_array_equal(pl.to_host(t), po.ntt(x, 12))
    np.testing.assert_array_equal(pl.to_host(items[1]["out"]), po.ntt(x[:64], 6, inverse=True))
    b.close()


def test_batch_argument_errors(pl):
    import ctypes
    import torch
    t = torch.zeros((16, 4), dtype=torch.int64, device="cuda")
    h = ctypes.c_void_p()
    L = pl.lib()
    bad = (pl.BatchItem * 1)()
    bad[0].op, bad[0].lg, bad[0].d_in, bad[0].d_out = 0, 29, t.data_ptr(), t.data_ptr()
    assert L.pl_batch_create(ctypes.cast(bad, ctypes.c_void_p), 1, 1, ctypes.byref(h)) == -3
    bad[0].op, bad[0].lg = 3, 4  # eval_div without a point
    assert L.pl_batch_create(ctypes.cast(bad, ctypes.c_void_p), 1, 1, ctypes.byref(h)) == -1
    bad[0].op, bad[0].lg, bad[0].lg_blowup = 2, 4, 5
    assert L.pl_batch_create(ctypes.cast(bad, ctypes.c_void_p), 1, 1, ctypes.byref(h)) == -3
    assert L.pl_batch_create(None, 1, 1, ctypes.byref(h)) == -1
    assert not h.value


def test_batch_point_rewritten_between_runs(pl):
    """Both forms read the evaluation point when they run: a new challenge
    written after pl_batch_create gives p(z_new) and the quotient by
    (X - z_new), against the oracle."""
    import torch
    lg = 10
    a = synth.uniform(1 << lg, 321)
    z0 = synth.uniform(1, 322)
    z1 = synth.uniform(1, 323)
    inp = pl.to_dev(a)
    zt = pl.to_dev(z0)
    items = [{"op": "eval_div", "lg": lg, "lg_blowup": 0, "inp": inp,
              "out": torch.zeros_like(inp), "z": zt, "y": torch.zeros((1, 4), dtype=torch.int64, device="cuda")}]
    for graph in (True, False):
        b = pl.Batch(items, graph=graph)
        for z in (z0, z1, z0):
            zt.copy_(pl.to_dev(z))
            y_ref = po.eval_(a, z)
            q_ref, _ = po.div(a, po.neg(z))
            for plain in (False, True):
                print("rewritten point graph=%s plain=%s" % (graph, plain), flush=True)
                items[0]["out"].zero_()
                items[0]["y"].zero_()
                torch.cuda.synchronize()
                (b.run_plain if plain else b.run)()
                torch.cuda.synchronize()
                np.testing.assert_array_equal(pl.to_host(items[0]["y"]), y_ref)
                np.testing.assert_array_equal(pl.to_host(items[0]["out"]), q_ref)
        b.close()
