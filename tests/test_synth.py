"""The package's synthetic input generator draws the oracle's stream, and the
ragged batch workload has the stated shape."""
from collections import Counter

import numpy as np

import pyoracle as po
import synth


def test_uniform_is_the_oracle_stream():
    for n, s in [(1, 5), (7, 123), (5000, po.seed(po.OP_NTT, 12)), (300000, 99)]:
        np.testing.assert_array_equal(synth.uniform(n, s), po.uniform(n, s))
    # chunk boundaries do not change the stream
    np.testing.assert_array_equal(synth.uniform(10000, 3, chunk=777), po.uniform(10000, 3))


def test_seed_matches_oracle():
    for op, lg, aux in [(1, 24, 0), (3, 20, 2), (7, 12, 1033)]:
        assert synth.seed(op, lg, aux) == po.seed(op, lg, aux)


def test_uniform_range():
    x = synth.uniform(20000, 11)
    r = synth.R_LIMBS
    # every element < r, top bit clear
    assert (x[:, 3] <= np.uint64(r[3])).all()
    assert (x[:, 3] >> np.uint64(63) == 0).all()


def test_ragged_batch_shape():
    b = synth.ragged_batch()
    assert len(b) == 64
    assert Counter(o for o, _ in b) == {"ntt": 40, "lde": 16, "eval_div": 8}
    assert all(12 <= lg <= 20 for _, lg in b)
    assert synth.ragged_batch() == b  # seeded
    assert len(set(lg for _, lg in b)) >= 7
