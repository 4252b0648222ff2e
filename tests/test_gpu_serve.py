"""GPU checks of the served form (pl_server_*) and of the device input
generator / digest it is driven with: every job's outputs equal the oracle's
composition and are the same run alone and in a mix of concurrent jobs."""
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


def test_fill_uniform_is_the_host_stream(pl):
    import torch
    for n, s in [(1, 5), (7, 123), (4096, synth.seed(1, 12)), (100003, 99), (1 << 20, 7)]:
        print("fill_uniform", n, s, flush=True)
        t = torch.empty((n, 4), dtype=torch.int64, device="cuda")
        pl.fill_uniform(t, s)
        np.testing.assert_array_equal(pl.to_host(t), synth.uniform(n, s))


def test_digest_matches_host(pl):
    for n in (1, 3, 1000, 65537):
        x = synth.uniform(n, n)
        assert pl.digest(pl.to_dev(x)) == synth.digest(x)
    x = synth.uniform(64, 1)
    y = x.copy()
    y[[3, 4]] = y[[4, 3]]
    assert synth.digest(x) != synth.digest(y)


def make_job(pl, lg, lb, seed_aux):
    """8 seeded columns, beta, gamma, zeta as device tensors + host copies."""
    n = 1 << lg
    cols = [synth.uniform(n, synth.seed(synth.OP_SERVE, lg, 100 * seed_aux + c)) for c in range(8)]
    sc = synth.uniform(3, synth.seed(synth.OP_SERVE, lg, 100 * seed_aux + 50))
    dev = [pl.to_dev(c) for c in cols]
    return {"lg": lg, "lb": lb, "cols": cols, "sc": sc, "dev": dev, "beta": pl.to_dev(sc[0:1]),
            "gamma": pl.to_dev(sc[1:2]), "zeta": pl.to_dev(sc[2:3])}


def oracle_job(job):
    lg, lb, cols, sc = job["lg"], job["lb"], job["cols"], job["sc"]
    coeffs = [po.ntt(c, lg, inverse=True) for c in cols]
    lde = [po.lde(c, lg, lb, po.ref_lde_shift(lb)) for c in coeffs]
    ks = [po.from_u64(1), po.from_u64(7), po.from_u64(13)]
    z = po.grand_product(po.perm_ratio(cols[:6], sc[0:1], sc[1:2], ks, lg))
    y = np.concatenate([po.eval_(c, sc[2:3]) for c in coeffs])
    q = [po.div(c, po.neg(sc[2:3]))[0] for c in coeffs]
    return {"coeffs": np.concatenate(coeffs), "lde": np.concatenate(lde), "z": z, "y": y, "q": np.concatenate(q)}


def digests(pl, r, lg, lb):
    n, m = 1 << lg, 1 << (lg + lb)
    return {"coeffs": pl.digest_ptr(r.d_coeffs, 8 * n), "lde": pl.digest_ptr(r.d_lde, 8 * m),
            "z": pl.digest_ptr(r.d_z, n), "y": pl.digest_ptr(r.d_y, 8), "q": pl.digest_ptr(r.d_q, 8 * n)}


@pytest.mark.parametrize("mode", [0, 1])
def test_job_vs_oracle(pl, mode):
    srv = pl.Server(hbm_bytes=1 << 30, mode=mode)
    for lg, lb in [(1, 1), (3, 2), (6, 4), (10, 2), (13, 1)]:
        print("served job vs oracle", mode, lg, lb, flush=True)
        job = make_job(pl, lg, lb, lg)
        jid = srv.submit(lg, lb, job["dev"], job["beta"], job["gamma"], job["zeta"])
        r = srv.wait(jid)
        assert r.done_s >= r.admit_s >= r.arrival_s >= 0
        ref = oracle_job(job)
        got = digests(pl, r, lg, lb)
        if lg <= 6:  # element-level view of the smallest ones
            np.testing.assert_array_equal(pl.to_host(pl.copy_from_ptr(r.d_lde, 8 << (lg + lb))), ref["lde"])
        for k in ref:
            assert got[k] == synth.digest(ref[k]), (lg, lb, k)
        srv.release(jid)
    st = srv.stats()
    assert st["completed"] == 5 and st["arena_used"] == 0
    srv.close()


def test_same_alone_and_in_mix(pl):
    # 24 jobs of mixed sizes: each run alone (plain server, one at a time),
    # then all submitted at once to a batched server that merges and
    # overlaps them; every output must be identical
    import torch
    rng = np.random.default_rng(11)
    spec = [(int(rng.integers(6, 17)), int(rng.choice([1, 2, 4]))) for _ in range(24)]
    spec[7] = (18, 4)  # eight LDEs of 2^22: run one after another in the batched form
    jobs = [make_job(pl, lg, lb, 1000 + i) for i, (lg, lb) in enumerate(spec)]
    alone = []
    srv = pl.Server(hbm_bytes=4 << 30, mode=pl.SERVE_PLAIN)
    for j in jobs:
        jid = srv.submit(j["lg"], j["lb"], j["dev"], j["beta"], j["gamma"], j["zeta"])
        alone.append(digests(pl, srv.wait(jid), j["lg"], j["lb"]))
        srv.release(jid)
    srv.close()
    print("alone done", flush=True)
    srv = pl.Server(hbm_bytes=4 << 30, slots=3, mode=pl.SERVE_BATCHED, group_elems=1 << 18)
    ids = [srv.submit(j["lg"], j["lb"], j["dev"], j["beta"], j["gamma"], j["zeta"]) for j in jobs]
    for i, (jid, j) in enumerate(zip(ids, jobs)):
        r = srv.wait(jid)
        assert digests(pl, r, j["lg"], j["lb"]) == alone[i], (i, j["lg"], j["lb"])
        srv.release(jid)
    st = srv.stats()
    print("mix stats", st, flush=True)
    assert st["completed"] == 24 and st["groups"] < 24  # some jobs were merged
    srv.close()
    # spot-check two against the oracle
    for i in (0, 5):
        ref = oracle_job(jobs[i])
        assert alone[i]["lde"] == synth.digest(ref["lde"])
        assert alone[i]["z"] == synth.digest(ref["z"])
    torch.cuda.synchronize()


def per_call_digests(pl, job):
    """the job's five outputs from the per-call entry points (pl_ntt inverse,
    pl_lde, pl_perm_grand_product, pl_eval, pl_div_linear), as digests"""
    import torch
    lg, lb = job["lg"], job["lb"]
    n, m = 1 << lg, 1 << (lg + lb)
    dom, ldom = pl.Domain(lg), pl.LdeDomain(lg, lb)
    coeffs = torch.empty((8 * n, 4), dtype=torch.int64, device="cuda")
    for j, d in enumerate(job["dev"]):
        pl.ntt(d, dom, inverse=True, out=coeffs[j * n:(j + 1) * n])
    out = {"coeffs": pl.digest(coeffs)}
    lde = torch.empty((8 * m, 4), dtype=torch.int64, device="cuda")
    for j in range(8):
        pl.lde(coeffs[j * n:(j + 1) * n], ldom, out=lde[j * m:(j + 1) * m])
    out["lde"] = pl.digest(lde)
    del lde
    ks = np.concatenate([po.from_u64(1), po.from_u64(7), po.from_u64(13)])
    z = pl.perm_ratio(job["dev"][0:3], job["dev"][3:6], job["beta"], job["gamma"], ks, dom, grand_product=True)
    out["z"] = pl.digest(z)
    y = torch.cat([pl.evaluate(coeffs[j * n:(j + 1) * n], job["zeta"]) for j in range(8)])
    out["y"] = pl.digest(y)
    negz = pl.vec_op("neg", job["zeta"])
    for j in range(8):
        pl.div_linear(coeffs[j * n:(j + 1) * n], negz)
    out["q"] = pl.digest(coeffs)
    return out


@pytest.mark.parametrize("lg,lb", [(20, 4), (22, 2), (22, 4), (24, 2)])
def test_large_jobs_vs_per_call(pl, lg, lb):
    # the sizes of the served workload (lg 14..24, lambda 2 and 4), in both
    # server modes, against the per-call entry points (themselves pinned to
    # the oracle) by digest
    import torch
    job = make_job(pl, lg, lb, 7)
    want = per_call_digests(pl, job)
    torch.cuda.empty_cache()
    for mode in (pl.SERVE_BATCHED, pl.SERVE_PLAIN):
        print("large served job", lg, lb, mode, flush=True)
        srv = pl.Server(hbm_bytes=pl.job_bytes(lg, lb) * 2 + (8 << 30), mode=mode)
        jid = srv.submit(lg, lb, job["dev"], job["beta"], job["gamma"], job["zeta"])
        got = digests(pl, srv.wait(jid), lg, lb)
        srv.release(jid)
        srv.close()
        assert got == want, (lg, lb, mode)


def test_admission_by_footprint(pl):
    # an arena that holds one lg-12 job at a time: jobs still all complete,
    # one after another; a job larger than the arena fails cleanly
    lg, lb = 12, 2
    one = pl.job_bytes(lg, lb)
    srv = pl.Server(hbm_bytes=int(one * 1.6) + (8 << 20), slots=3)
    jobs = [make_job(pl, lg, lb, 7000 + i) for i in range(4)]
    ids = [srv.submit(lg, lb, j["dev"], j["beta"], j["gamma"], j["zeta"]) for j in jobs]
    rs = []
    for jid in ids:
        r = srv.wait(jid)
        rs.append((r.admit_s, r.done_s, digests(pl, r, lg, lb)))
        srv.release(jid)
    # no two jobs were resident at once: each admitted after the previous finished
    for a, b in zip(rs, rs[1:]):
        assert b[0] >= a[1]
    st = srv.stats()
    assert st["arena_peak"] <= st["arena_bytes"]
    big = make_job(pl, 16, 4, 9)
    jid = srv.submit(16, 4, big["dev"], big["beta"], big["gamma"], big["zeta"])
    with pytest.raises(pl.PlError) as ei:
        srv.wait(jid)
    assert ei.value.status == -4
    srv.close()
    ref = oracle_job(jobs[0])
    assert rs[0][2]["q"] == synth.digest(ref["q"])


def test_waits_for_release(pl):
    # the arena holds one job's outputs: the next job must wait until the
    # client releases the first (not fail), also when no group is running
    import threading
    import time
    lg, lb = 12, 2
    one = pl.job_bytes(lg, lb)
    # one job's outputs + its work block (~0.75 of the outputs) fit, two
    # jobs' outputs do not
    srv = pl.Server(hbm_bytes=int(one * 1.9), slots=1)
    jobs = [make_job(pl, lg, lb, 8000 + i) for i in range(2)]
    ids = [srv.submit(lg, lb, j["dev"], j["beta"], j["gamma"], j["zeta"]) for j in jobs]
    r0 = srv.wait(ids[0])
    got = {}

    def waiter():
        try:
            got["r"] = srv.wait(ids[1])
        except Exception as e:  # reported below
            got["e"] = e

    th = threading.Thread(target=waiter)
    th.start()
    time.sleep(0.3)            # the server is idle, the arena full of job 0
    assert th.is_alive() and "e" not in got
    d0 = digests(pl, r0, lg, lb)
    srv.release(ids[0])
    th.join(timeout=60)
    assert "r" in got, got.get("e")
    assert got["r"].admit_s >= r0.done_s
    srv.release(ids[1])
    srv.close()
    ref = oracle_job(jobs[0])
    assert d0["lde"] == synth.digest(ref["lde"])
