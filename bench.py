#!/usr/bin/env python3
"""Benchmark of the hot path on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--lg 24]
                    [--no-cpu] [--no-lde] [--no-batch] [--no-serve] [--serve-jobs J]

Prints ONE JSON line.  `value` = Fr elements per second of the device-resident
NTT of 2^24 elements (the reference's `NTT --bench-split 24` workload with the
input already in HBM): K back-to-back NTTs bracketed by HIP events on the
launch stream.  Beside it, in the same line:

  host_to_host  NTT 2^24 from pinned host memory back to pinned host memory
           (the reference's `our-ntt execution time` scope), 1 and 4 polynomials
  lde      coset LDE 2^24 -> 2^26 (lambda = 2, shift 7^4), device-resident
  lde_l3   coset LDE 2^24 -> 2^27 (lambda = 3, shift 7^8)
  batched  the ragged proof batch (64 polynomials, lg 12..20: 40 NTT, 16 LDE
           lambda = 2, 8 evaluation + division), one hipGraph launch versus one
           launch sequence per polynomial; per-run latency p50 / p99
  served   proof jobs (8 columns of 2^lg, lg 14..24, lambda 2 or 4: iNTT ->
           LDE -> permutation grand product -> evaluation -> division) through
           the request loop, batched versus plain form: throughput with every
           job queued at once, and p50 / p99 latency under Poisson arrivals at
           three rates; outputs compared by digest between the two forms
  roofline the dominant kernel (NTT pass 0), per-launch HIP events on its
           stream, against the measured v_mad_u64_u32 peak
Each timed section first runs its own step, untimed, for at least 300 ms of
device time, so that the clocks have settled under that load.
  cpu_baseline  the oracle's radix-2 NTT and LDE of 2^24 on all host cores (at
           most 16 threads), and the miss_ops workload at 2^20 over the
           reference's own blst field on one pinned core
"""
import argparse
import json
import os
import platform
import queue
import sys
import threading
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "ppopp26-artifact_amd"))

# MI355X: 256 CUs x 64 lanes x 2.4 GHz peak engine clock; v_mad_u64_u32 is a
# full-rate VALU op on gfx950 (one 32x32->64 multiply-add per lane-cycle)
PEAK_MAD_SPEC = 256 * 64 * 2.4e9
PEAK_HBM = 8.0e12
MAD_PER_MULMOD = 128  # 8-limb Montgomery product: 2 * 8^2 limb multiply-adds
# HBM bytes per launch of the dominant kernel (NTT pass 0 at 2^24), from
# rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE runs scaled by a plain copy of
# known size (profiles/ntt_pass_2p24_hbm_traffic.txt): the data (2^24 x 32 B
# read and written) plus the pass's 2^24-entry twiddle table read once
TRAFFIC_PASS0_2P24 = 1611643008


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--lg", type=int, default=24)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--no-lde", action="store_true")
    ap.add_argument("--no-h2h", action="store_true", help="skip the host-to-host NTT (our_ntt's scope)")
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--no-miss", action="store_true", help="skip grand product / evaluation / division")
    ap.add_argument("--serve-jobs", type=int, default=200, help="jobs in the served mix (every run uses the same mix)")
    ap.add_argument("--serve-slots", type=int, default=2, help="job groups running at once (batched form)")
    ap.add_argument("--serve-large", type=int, default=25, help="log2 of the work that makes a job large")
    ap.add_argument("--plans", default=os.path.join(ROOT, "ppopp26-artifact_amd", "tune", "mi355x_plans.txt"),
                    help="autotuned pass-split cache (measured and written when absent)")
    ap.add_argument("--no-tune", action="store_true", help="default pass splits")
    ap.add_argument("--settle-ms", type=float, default=300.0,
                    help="untimed device time each section runs before timing (0: none, e.g. under counters)")
    ap.add_argument("--pcie-probe", action="store_true",
                    help="only time pinned host <-> device copies of 2^lg * 32 bytes (1-D, 2-D, both ways at once)")
    ap.add_argument("--copy-probe", action="store_true",
                    help="only run a plain device copy of 2^lg * 32 bytes (HBM counter calibration)")
    return ap.parse_args()


def time_events(fn, reps, stream):
    import torch
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for _ in range(reps):
        fn()
    e1.record(stream)
    e1.synchronize()
    return e0.elapsed_time(e1) / reps  # ms


SETTLE_MS = 300.0


def settle(fn, stream, min_ms=None, chunk=4):
    """Untimed: run fn until at least min_ms of device time has passed.  The
    clocks drop for some tens of milliseconds when a heavy VALU load starts
    and then recover; timing starts once they have settled under this load."""
    min_ms = SETTLE_MS if min_ms is None else min_ms
    spent = 0.0
    while spent < min_ms:
        spent += time_events(fn, chunk, stream) * chunk


def time_each(fn, reps, stream):
    """per-run milliseconds (event pair around each run)"""
    import torch
    out = []
    for _ in range(reps):
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        fn()
        e1.record(stream)
        e1.synchronize()
        out.append(e0.elapsed_time(e1))
    return out


def progress(msg):
    print("[bench] " + msg, file=sys.stderr, flush=True)


def pct(v, q):
    import numpy as np
    return float(np.percentile(np.asarray(v), q)) if len(v) else None


def dev_uniform(pl, n, seed):
    import torch
    t = torch.empty((n, 4), dtype=torch.int64, device="cuda")
    return pl.fill_uniform(t, seed)


# ---------------------------------------------------------------------------
def bench_ntt(pl, synth, lg, steps, warmup, stream, peak):
    import torch
    n = 1 << lg
    dom = pl.Domain(lg)
    x = dev_uniform(pl, n, synth.seed(synth.OP_NTT, lg))
    y = torch.empty_like(x)
    work = torch.empty_like(x)
    # correctness gate on the measured buffers: iNTT(NTT(x)) == x
    pl.ntt(x, dom, out=y, work=work)
    z = pl.ntt(y, dom, inverse=True, work=work)
    torch.cuda.synchronize()
    verified = bool(torch.equal(z, x))
    del z
    step = lambda: pl.ntt(x, dom, out=y, work=work)  # noqa: E731
    for _ in range(warmup):
        step()
    settle(step, stream)
    torch.cuda.synchronize()
    ms = time_events(step, steps, stream)
    npass = pl.ntt_num_passes(lg)
    pass_ms = []
    for p in range(npass):
        src = x if p == 0 else work
        dst = work if p < npass - 1 else y
        for _ in range(2):
            pl.ntt_pass(src, dst, dom, p)
        pass_ms.append(time_events(lambda: pl.ntt_pass(src, dst, dom, p), steps, stream))
    # the pass split in use (tuned or default); pass kernels hold 2^3 elements
    # per lane, so a 2^m-point column runs ntt_pass_kernel<m - 3, ...>
    m_pass = pl.plan_get(lg)
    assert len(m_pass) == npass
    dom_ms = pass_ms[0]
    mad = (n // 2) * m_pass[0] * MAD_PER_MULMOD
    achieved = mad / (dom_ms * 1e-3)
    byts = 64 * n
    roof = {
        "bound": "valu",
        "achieved": round(achieved / 1e12, 3),
        "peak": round(peak / 1e12, 3),
        "unit": "Tmad/s",
        "frac": round(achieved / peak, 4),
        "traffic": TRAFFIC_PASS0_2P24 if lg == 24 else None,
        "peak_source": "measured v_mad_u64_u32 rate (pl_probe_mad_rate); spec full rate %.2f" % (
            PEAK_MAD_SPEC / 1e12),
        "kernel": "ntt_pass_kernel<%d,false,false,false> (pass 0 of %d, 2^%d-point columns)" % (m_pass[0] - 3, npass,
                                                                                          m_pass[0]),
        "alg_per_launch": "%d mulmod = (N/2)*%d, x128 u32 multiply-adds; %d bytes = 64 N" % (
            (n // 2) * m_pass[0], m_pass[0], byts),
        "avg_launch_ms": round(dom_ms, 4),
        "pass_ms": [round(v, 4) for v in pass_ms],
        "hbm_gbps_alg": round(byts / (dom_ms * 1e-3) / 1e9, 1),
        "hbm_frac_alg": round(byts / (dom_ms * 1e-3) / PEAK_HBM, 4),
        "whole_ntt_frac": round(((n // 2) * lg * MAD_PER_MULMOD / peak) / (ms * 1e-3), 4),
    }
    del x, y, work
    return {"value": n / (ms * 1e-3), "ms": ms, "verified": verified, "npass": npass}, roof


def bench_host_to_host(pl, synth, lg, stream, reps=5):
    """our_ntt's scope (ntt.cu:979-1037, the reference's `our-ntt execution
    time`): pinned host buffer in, NTT, pinned host buffer out, through
    pl_ntt_host (grouped 2-D copies overlapping the first and last passes;
    with several polynomials the next one's copy-in overlaps the current
    one's transform and copy-out).  Wall time of the blocking call."""
    import torch
    n = 1 << lg
    dom = pl.Domain(lg)
    tuned = pl.host_groups(lg)  # the autotuner's copy-group count (plan file), else the default
    out = {"tuned_lg_k": tuned}
    for npoly in (1, 4):
        h = torch.empty((npoly * n, 4), dtype=torch.int64).pin_memory()
        h.copy_(dev_uniform(pl, npoly * n, synth.seed(synth.OP_NTT, lg, 7)).cpu())
        o = torch.empty_like(h).pin_memory()
        # the device slots, allocated once as a repeated caller would
        work = torch.empty((pl.lib().pl_ntt_host_work_bytes(lg, npoly) // 32, 4), dtype=torch.int64, device="cuda")
        best = {}
        for lg_k in sorted({0, 1, 2, 3, 4, 6, tuned}):
            pl.ntt_host(h, dom, lg_k=lg_k, npoly=npoly, out=o, work=work)
            ts = []
            for _ in range(reps):
                t0 = time.perf_counter()
                pl.ntt_host(h, dom, lg_k=lg_k, npoly=npoly, out=o, work=work)
                ts.append((time.perf_counter() - t0) * 1e3)
            best[lg_k] = pct(ts, 50)
        k = min(best, key=best.get)
        ms = best[k]
        out["npoly_%d" % npoly] = {"ms": round(ms, 3), "ms_per_poly": round(ms / npoly, 3),
                                   "value": round(npoly * n / (ms * 1e-3), 1), "unit": "Fr elems/s",
                                   "pcie_gbps_each_way": round(npoly * n * 32 / (ms * 1e-3) / 1e9, 2),
                                   "best_lg_k": k, "tuned_ms": round(best[tuned], 3),
                                   "ms_by_lg_k": {str(a): round(v, 3) for a, v in best.items()}}
        del h, o, work
    out["workload"] = ("NTT 2^%d host-to-host (pinned in, pinned out, device slots allocated once), median of %d "
                       "calls, best copy-group count" % (lg, reps))
    return out


def bench_lde(pl, synth, lg, lb, steps, warmup, stream, peak):
    """lb >= 1: blowup 2^lb; lb = -1: blowup 3 (cosets 7 w_3N^t)."""
    import torch
    n = 1 << lg
    m = 3 * n if lb < 0 else 1 << (lg + lb)
    dom = pl.LdeDomain(lg, 0, blowup3=True) if lb < 0 else pl.LdeDomain(lg, lb)
    c = dev_uniform(pl, n, synth.seed(synth.OP_LDE, lg, max(lb, 0)))
    out = torch.empty((m, 4), dtype=torch.int64, device="cuda")
    work = torch.empty((max(dom.work_bytes // 32, 1), 4), dtype=torch.int64, device="cuda")
    step = lambda: pl.lde(c, dom, out=out, work=work)  # noqa: E731
    for _ in range(warmup):
        step()
    settle(step, stream)
    torch.cuda.synchronize()
    ms = time_events(step, steps, stream)
    mulmod = (m // 2) * lg + m
    if lb < 0:
        what = "coset LDE 2^%d -> 3 x 2^%d (blowup 3, cosets 7 w_3N^t), device-resident" % (lg, lg)
    else:
        what = "coset LDE 2^%d -> 2^%d (lambda=%d, shift 7^%d), device-resident" % (lg, lg + lb, lb, 1 << lb)
    res = {
        "workload": what,
        "ms": round(ms, 4),
        "value": round(n / (ms * 1e-3), 1),
        "unit": "Fr elems/s (input N)",
        "output_elems_per_s": round(m / (ms * 1e-3), 1),
        "valu_frac": round((mulmod * MAD_PER_MULMOD / peak) / (ms * 1e-3), 4),
        "alg": "%d mulmod = (M/2)*%d + M" % (mulmod, lg),
    }
    del c, out, work
    return res


def bench_batch(pl, synth, steps, stream):
    import torch
    spec = [(op, lg, 2 if op == "lde" else 0) for op, lg in synth.ragged_batch()]
    items, n_in, n_out = [], 0, 0
    for i, (op, lg, lb) in enumerate(spec):
        n = 1 << lg
        it = {"op": op, "lg": lg, "lg_blowup": lb, "inp": dev_uniform(pl, n, synth.seed(synth.OP_BATCH, lg, i))}
        it["out"] = torch.empty(((n << lb), 4), dtype=torch.int64, device="cuda")
        n_in += n
        n_out += n << lb
        if op == "eval_div":
            it["z"] = dev_uniform(pl, 1, synth.seed(synth.OP_BATCH, lg, 1000 + i))
            it["y"] = torch.empty((1, 4), dtype=torch.int64, device="cuda")
        items.append(it)
    b = pl.Batch(items, graph=True)
    nk, nbytes = b.info()
    # bit-exactness gate: batched == per-polynomial form, by digest
    b.run()
    torch.cuda.synchronize()
    d_b = [pl.digest(it["out"]) for it in items]
    b.run_plain()
    torch.cuda.synchronize()
    d_p = [pl.digest(it["out"]) for it in items]
    for _ in range(3):
        b.run()
        b.run_plain()
    settle(b.run, stream, chunk=2)
    torch.cuda.synchronize()
    t_b = time_each(b.run, steps, stream)
    settle(b.run_plain, stream, chunk=2)
    torch.cuda.synchronize()
    t_p = time_each(b.run_plain, steps, stream)
    mb, mp = pct(t_b, 50), pct(t_p, 50)
    b.close()
    return {
        "workload": "ragged proof batch: 64 polys lg 12..20 (40 NTT, 16 LDE lambda=2, 8 eval+div)",
        "value": round(n_in / (mb * 1e-3), 1),
        "unit": "Fr elems/s (inputs)",
        "plain_value": round(n_in / (mp * 1e-3), 1),
        "speedup_vs_plain": round(mp / mb, 3),
        "p50_ms": round(mb, 4), "p99_ms": round(pct(t_b, 99), 4),
        "plain_p50_ms": round(mp, 4), "plain_p99_ms": round(pct(t_p, 99), 4),
        "input_elems": n_in, "output_elems": n_out,
        "graph_kernel_nodes": nk, "plan_bytes": nbytes,
        "bit_exact_vs_plain": d_b == d_p,
    }


def pcie_probe(lg):
    """Pinned host <-> device copy rates for 2^lg elements (the ceiling of the
    host-to-host NTT): plain copies each way, both ways at once on two
    streams, and 2-D copies of 2^6-row column groups as the host-to-host
    NTT issues them (and the same groups as one 1-D copy per row)."""
    import ctypes
    import torch
    n = 1 << lg
    nbytes = n * 32
    h_in = torch.empty((n, 4), dtype=torch.int64).pin_memory()
    h_out = torch.empty((n, 4), dtype=torch.int64).pin_memory()
    d_a = torch.empty((n, 4), dtype=torch.int64, device="cuda")
    d_b = torch.empty((n, 4), dtype=torch.int64, device="cuda")
    s1, s2 = torch.cuda.Stream(), torch.cuda.Stream()

    def rate(fn, reps=5):
        fn()
        torch.cuda.synchronize()
        t = time.perf_counter()
        for _ in range(reps):
            fn()
        torch.cuda.synchronize()
        return round(nbytes / ((time.perf_counter() - t) / reps) / 1e9, 2)

    def both():
        with torch.cuda.stream(s1):
            d_a.copy_(h_in, non_blocking=True)
        with torch.cuda.stream(s2):
            h_out.copy_(d_b, non_blocking=True)

    out = {"bytes": nbytes, "h2d_gbps": rate(lambda: d_a.copy_(h_in, non_blocking=True)),
           "d2h_gbps": rate(lambda: h_out.copy_(d_a, non_blocking=True)), "both_ways_gbps_each": rate(both)}
    hip = ctypes.CDLL("libamdhip64.so")
    hip.hipMemcpy2DAsync.argtypes = [ctypes.c_void_p, ctypes.c_size_t, ctypes.c_void_p, ctypes.c_size_t,
                                     ctypes.c_size_t, ctypes.c_size_t, ctypes.c_int, ctypes.c_void_p]
    hip.hipMemcpyAsync.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int, ctypes.c_void_p]
    cols = n >> 6  # 2^6 rows of 2^(lg-6) elements
    for groups in (1, 4, 64):
        w = cols // groups * 32

        def g2d():
            st = torch.cuda.current_stream().cuda_stream
            for g in range(groups):
                r = hip.hipMemcpy2DAsync(ctypes.c_void_p(d_a.data_ptr() + g * w), cols * 32,
                                         ctypes.c_void_p(h_in.data_ptr() + g * w), cols * 32, w, 64, 1,
                                         ctypes.c_void_p(st))
                if r != 0:
                    raise RuntimeError("hipMemcpy2DAsync: %d" % r)
        out["h2d_2d_groups_%d_gbps" % groups] = rate(g2d)

        def d2d():
            st = torch.cuda.current_stream().cuda_stream
            for g in range(groups):
                r = hip.hipMemcpy2DAsync(ctypes.c_void_p(h_out.data_ptr() + g * w), cols * 32,
                                         ctypes.c_void_p(d_a.data_ptr() + g * w), cols * 32, w, 64, 2,
                                         ctypes.c_void_p(st))
                if r != 0:
                    raise RuntimeError("hipMemcpy2DAsync: %d" % r)
        out["d2h_2d_groups_%d_gbps" % groups] = rate(d2d)

        def d2d_side():
            with torch.cuda.stream(s2):
                d2d()
        out["d2h_2d_groups_%d_side_stream_gbps" % groups] = rate(d2d_side)

        def d2h_rows():
            st = torch.cuda.current_stream().cuda_stream
            for g in range(groups):
                for row in range(64):
                    off = row * cols * 32 + g * w
                    r = hip.hipMemcpyAsync(ctypes.c_void_p(h_out.data_ptr() + off),
                                           ctypes.c_void_p(d_a.data_ptr() + off), w, 2, ctypes.c_void_p(st))
                    if r != 0:
                        raise RuntimeError("hipMemcpyAsync: %d" % r)
        out["d2h_rows_groups_%d_gbps" % groups] = rate(d2h_rows)
    return out


def bench_miss_ops(pl, synth, steps, stream, peak):
    """The miss_ops workload (grand product with the permutation ratio,
    evaluation, division by X - z) at 2^20 and 2^24, device-resident."""
    import torch
    out = {}
    # Montgomery forms of the coset constants 1, 7, 13, computed on the device
    kraw = torch.zeros((3, 4), dtype=torch.int64, device="cuda")
    kraw[0, 0], kraw[1, 0], kraw[2, 0] = 1, 7, 13
    ks = pl.to_host(pl.vec_op("to_mont", kraw))
    for lg in (20, 24):
        n = 1 << lg
        dom = pl.Domain(lg)
        cols = [dev_uniform(pl, n, synth.seed(synth.OP_GP, lg, 10 + j)) for j in range(6)]
        bg = dev_uniform(pl, 2, synth.seed(synth.OP_GP, lg, 20))
        beta, gamma = bg[0:1], bg[1:2]
        z = torch.empty_like(cols[0])
        wv = torch.empty((pl.lib().pl_vec_work_bytes(n) + 31) // 32, 4, dtype=torch.int64, device="cuda")
        gp = lambda: pl.perm_ratio(cols[:3], cols[3:], beta, gamma, ks, dom, out=z, grand_product=True,  # noqa: E731
                                   work=wv)
        poly = dev_uniform(pl, n, synth.seed(synth.OP_EVAL, lg))
        pt = dev_uniform(pl, 1, synth.seed(synth.OP_EVAL, lg, 1))
        y = torch.empty((1, 4), dtype=torch.int64, device="cuda")
        sw = torch.empty((pl.lib().pl_scan_work_bytes(n) + 31) // 32, 4, dtype=torch.int64, device="cuda")
        ev = lambda: pl.evaluate(poly, pt, y=y, work=sw)  # noqa: E731
        q = poly.clone()
        dv = lambda: pl.div_linear(q, pt, work=sw)  # noqa: E731 (in place; timing only)
        res = {}
        for name, fn, mm, byts in [("perm_grand_product", gp, 17 * n, 224 * n), ("eval", ev, n, 32 * n),
                                   ("div", dv, n, 64 * n)]:
            for _ in range(3):
                fn()
            torch.cuda.synchronize()
            ms = time_events(fn, steps, stream)
            t_roof = max(byts / PEAK_HBM, mm * MAD_PER_MULMOD / peak)
            res[name] = {"ms": round(ms, 4), "elems_per_s": round(n / (ms * 1e-3), 1),
                         "roofline_frac": round(t_roof / (ms * 1e-3), 4)}
        out["2^%d" % lg] = res
        del cols, z, wv, poly, q, sw
        torch.cuda.empty_cache()
    out["roofline"] = ("max(bytes / 8 TB/s, 128 MACs x mulmods / measured P_mac); mulmods 17 N (perm GP), "
                       "N (eval), N (div); bytes 224 N, 32 N, 64 N")
    return out


def run_server(pl, mode, jobs, times, pool, scal, verify, slots=0, large_lg=0):
    """Submit jobs at their arrival times (seconds, None = all at once);
    returns per-job (arrival, admit, done) and digests."""
    import torch
    srv = pl.Server(mode=mode, slots=slots, large_work=(1 << large_lg) if large_lg else 0)
    res = [None] * len(jobs)
    q = queue.Queue()
    err = []

    def releaser():
        s = torch.cuda.Stream()
        try:
            for i in range(len(jobs)):
                jid = q.get()
                r = srv.wait(jid)
                d = None
                if verify:
                    lg, lb = jobs[i]
                    n, m = 1 << lg, 1 << (lg + lb)
                    d = (pl.digest_ptr(r.d_coeffs, 8 * n, s), pl.digest_ptr(r.d_lde, 8 * m, s),
                         pl.digest_ptr(r.d_z, n, s), pl.digest_ptr(r.d_y, 8, s), pl.digest_ptr(r.d_q, 8 * n, s))
                res[i] = (r.arrival_s, r.admit_s, r.done_s, d)
                srv.release(jid)
        except Exception as e:  # surfaced by the caller
            err.append(e)

    th = threading.Thread(target=releaser)
    th.start()
    t0 = srv.now()
    for i, (lg, lb) in enumerate(jobs):
        if times is not None:
            while True:
                d = t0 + times[i] - srv.now()
                if d <= 0:
                    break
                time.sleep(min(d, 0.001))
        beta, gamma, zeta = scal[i]
        q.put(srv.submit(lg, lb, pool[lg], beta, gamma, zeta))
    th.join()
    st = srv.stats()
    srv.close()
    if err:
        raise err[0]
    return res, st


def summarize(jobs, res):
    lat = [(r[2] - r[0]) * 1e3 for r in res]
    span = max(r[2] for r in res) - min(r[0] for r in res)
    elems = sum(8 << lg for lg, _ in jobs)
    return {"p50_ms": round(pct(lat, 50), 2), "p99_ms": round(pct(lat, 99), 2),
            "elems_per_s": round(elems / span, 1), "span_s": round(span, 3)}


def bench_served(pl, synth, n_jobs, slots, large_lg):
    """One seeded job mix (n_jobs proof jobs, unit-rate Poisson arrival
    times) run through the request loop in both forms: first with every job
    queued at once (each form's capacity on this mix, outputs compared by
    digest), then with the mix's arrival times scaled to 0.5, 0.8 and 1.1
    times the plain form's capacity on the same mix."""
    import torch
    # resident inputs: 8 columns per size (jobs of one size share columns,
    # each job has its own beta, gamma, zeta)
    pool = {lg: [dev_uniform(pl, 1 << lg, synth.seed(synth.OP_SERVE, lg, c)) for c in range(8)]
            for lg in range(14, 25)}
    rates_rel = [0.5, 0.8, 1.1]
    wl = synth.served_jobs(0, n_jobs)
    jobs = [(lg, lb) for lg, lb, _ in wl]
    scal = [tuple(dev_uniform(pl, 1, synth.seed(synth.OP_SERVE, 0, 10000 + 3 * i + k)) for k in range(3))
            for i in range(n_jobs)]
    # warm-up: one job of every size and blowup (domains and coset tables are
    # built on first use and stay resident)
    warm = sorted(set(jobs))
    run_server(pl, pl.SERVE_BATCHED, warm, None, pool, [scal[0]] * len(warm), False, slots, large_lg)
    torch.cuda.synchronize()
    progress("served: warm-up done (%d sizes)" % len(warm))
    # capacity of each form on this mix: every job queued at once; outputs
    # verified between the forms
    sat_p, _ = run_server(pl, pl.SERVE_PLAIN, jobs, None, pool, scal, True)
    sat_b, st_b = run_server(pl, pl.SERVE_BATCHED, jobs, None, pool, scal, True, slots, large_lg)
    exact = all(a[3] == b[3] for a, b in zip(sat_p, sat_b))
    cap_p = summarize(jobs, sat_p)
    cap_b = summarize(jobs, sat_b)
    progress("served: capacity %.3g vs %.3g elems/s (batched / plain), bit-exact %s" % (
        cap_b["elems_per_s"], cap_p["elems_per_s"], exact))
    plain_jobs_per_s = len(jobs) / cap_p["span_s"]
    rates = []
    for rel in rates_rel:
        rate = rel * plain_jobs_per_s
        times = [t / rate for _, _, t in wl]
        rp, _ = run_server(pl, pl.SERVE_PLAIN, jobs, times, pool, scal, False)
        rb, _ = run_server(pl, pl.SERVE_BATCHED, jobs, times, pool, scal, False, slots, large_lg)
        sp, sb = summarize(jobs, rp), summarize(jobs, rb)
        progress("served: %.1f x plain capacity (%.1f jobs/s): p99 batched %.1f ms, plain %.1f ms" % (
            rel, rate, sb["p99_ms"], sp["p99_ms"]))
        rates.append({"rate_jobs_per_s": round(rate, 2), "load_vs_plain_capacity": rel,
                      "batched_p50_ms": sb["p50_ms"], "batched_p99_ms": sb["p99_ms"],
                      "plain_p50_ms": sp["p50_ms"], "plain_p99_ms": sp["p99_ms"],
                      "batched_elems_per_s": sb["elems_per_s"], "plain_elems_per_s": sp["elems_per_s"]})
    del pool
    torch.cuda.synchronize()
    return {
        "workload": "proof jobs: 8 cols x 2^lg (lg 14..24), iNTT -> LDE (lambda 2|4) -> perm GP -> eval -> div; "
                    "one seeded mix of %d jobs for every run" % n_jobs,
        "value": cap_b["elems_per_s"], "unit": "Fr elems/s (job inputs, all jobs queued at once)",
        "plain_value": cap_p["elems_per_s"],
        "speedup_vs_plain": round(cap_b["elems_per_s"] / cap_p["elems_per_s"], 3),
        "saturated_p50_ms": cap_b["p50_ms"], "saturated_p99_ms": cap_b["p99_ms"],
        "plain_saturated_p50_ms": cap_p["p50_ms"], "plain_saturated_p99_ms": cap_p["p99_ms"],
        "plain_capacity_jobs_per_s": round(plain_jobs_per_s, 2),
        "poisson": rates,
        "bit_exact_vs_plain": exact,
        "slots": slots, "large_work_lg": large_lg, "groups": st_b["groups"],
        "arena_peak_gib": round(st_b["arena_peak"] / 2**30, 2),
        "arena_gib": round(st_b["arena_bytes"] / 2**30, 1), "domain_cache_gib": round(st_b["domain_bytes"] / 2**30, 2),
    }


def _cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or platform.machine()


def cpu_baseline(lg):
    """The CPU baselines of the plan, timed on this host (reported only):
      * NTT 2^lg and LDE 2^lg -> 2^(lg+2) by the oracle's radix-2 transform
        on all host cores available to this process (at most 16 threads);
      * the miss_ops workload at 2^20 (the reference's CPU path: its serial
        naive_gp and naive_div loops, plus the permutation ratio composed
        element by element and a Horner evaluation) over the reference's own
        blst field build (oracle/_ref), on one pinned core."""
    import ctypes
    import numpy as np
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle as po  # the CPU baseline leg only
    cores = min(16, len(os.sched_getaffinity(0)))
    po.set_threads(cores)
    n = 1 << lg
    xh = po.uniform(n, po.seed(po.OP_NTT, lg))
    t0 = time.perf_counter()
    po.ntt(xh, lg)
    dt = time.perf_counter() - t0
    del xh
    c = po.uniform(n, po.seed(po.OP_LDE, lg, 2))
    t0 = time.perf_counter()
    po.lde(c, lg, 2, po.ref_lde_shift(2))
    dt_lde = time.perf_counter() - t0
    del c
    out = {"value": round(n / dt, 1), "unit": "Fr elems/s", "cores": cores, "kind": "port",
           "sample": "one NTT 2^%d by the oracle's radix-2 restatement on %d threads (%.2f s)" % (lg, cores, dt),
           "cpu_model": _cpu_model(), "nproc": os.cpu_count(), "threads_used": cores,
           "lde": {"workload": "LDE 2^%d -> 2^%d (lambda=2, shift 7^4), oracle on %d threads" % (lg, lg + 2, cores),
                   "s": round(dt_lde, 3), "value": round(n / dt_lde, 1), "unit": "Fr elems/s (input N)",
                   "kind": "port", "cores": cores}}
    ref = po.ref_lib()
    if ref is None:
        out["miss_ops"] = "reference blst build (oracle/_ref) absent"
        return out
    lgm = 20
    nm = 1 << lgm
    cols = [po.uniform(nm, po.seed(po.OP_GP, lgm, 10 + j)) for j in range(6)]
    bg = po.uniform(2, po.seed(po.OP_GP, lgm, 20))
    ks = np.concatenate([po.from_u64(1), po.from_u64(7), po.from_u64(13)])
    poly = po.uniform(nm, po.seed(po.OP_EVAL, lgm))
    pt = po.uniform(1, po.seed(po.OP_EVAL, lgm, 1))
    cdiv = po.neg(pt)
    gp_out = po.fr_array(nm)
    q = po.fr_array(nm - 1)
    y = po.fr_array(1)
    prev = os.sched_getaffinity(0)
    core = min(prev)
    os.sched_setaffinity(0, {core})
    try:
        t0 = time.perf_counter()
        z = po.ref_perm_grand_product(cols, bg[0:1], bg[1:2], ks, po.root_of_unity(lgm))
        t_perm = time.perf_counter() - t0
        t0 = time.perf_counter()
        ref.ref_naive_gp(po.P(cols[0]), po.P(gp_out), nm)
        t_gp = time.perf_counter() - t0
        t0 = time.perf_counter()
        ref.ref_naive_div(po.P(poly), nm, po.P(cdiv), po.P(po.one()), po.P(q))
        t_div = time.perf_counter() - t0
        t0 = time.perf_counter()
        ref.ref_eval(po.P(poly), nm, po.P(pt), po.P(y))
        t_ev = time.perf_counter() - t0
    finally:
        os.sched_setaffinity(0, prev)
    del z, cols, poly, gp_out, q
    ms = lambda t: round(t * 1e3, 3)  # noqa: E731
    out["miss_ops"] = {
        "workload": "miss_ops 2^20: permutation grand product (ratio over a,b,c / s1,s2,s3, k = 1, 7, 13, "
                    "then z), naive_gp scan, naive_div by (X - z), Horner evaluation",
        "kind": "reference", "cores": 1, "pinned_core": core,
        "perm_grand_product_ms": ms(t_perm), "naive_gp_ms": ms(t_gp), "naive_div_ms": ms(t_div),
        "eval_ms": ms(t_ev),
        "perm_grand_product_elems_per_s": round(nm / t_perm, 1), "eval_elems_per_s": round(nm / t_ev, 1),
        "div_elems_per_s": round(nm / t_div, 1),
        "field": "the reference's blst (PLONK/blst/src/assembly.S + cpu/ff/bls12-381.hpp) built from source",
    }
    return out


def main():
    args = parse()
    global SETTLE_MS
    SETTLE_MS = args.settle_ms
    if args.gpus != 1:
        print("bench.py: this build measures one GPU only (no multi-GPU path); run with --gpus 1",
              file=sys.stderr)
        sys.exit(2)
    import torch

    import plffi as pl
    import synth

    torch.cuda.init()
    stream = torch.cuda.current_stream()
    lg = args.lg
    if args.copy_probe:
        n = 1 << lg
        a = torch.ones((n, 4), dtype=torch.int64, device="cuda")
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_probe_bytes_read": n * 32, "copy_probe_bytes_written": n * 32, "launches": 3}))
        return
    if args.pcie_probe:
        print(json.dumps(pcie_probe(lg)))
        return
    peak = pl.probe_mad_rate() * 1e12
    progress("measured v_mad_u64_u32 peak %.2f T/s" % (peak / 1e12))
    plans = None
    if not args.no_tune:
        t0 = time.time()
        os.makedirs(os.path.dirname(args.plans), exist_ok=True)
        pl.autotune(4, 24, args.plans, reps=5)
        plans = {lg: pl.plan_get(lg) for lg in range(12, 25)}
        progress("pass splits from %s (%.1f s)" % (args.plans, time.time() - t0))
    ntt, roof = bench_ntt(pl, synth, lg, args.steps, args.warmup, stream, peak)
    progress("ntt 2^%d: %.3f ms" % (lg, ntt["ms"]))
    torch.cuda.empty_cache()
    result = {
        "metric": "Fr elems/s for 2^24 NTT & LDE; ragged-batch p50/p99 vs one-call-per-launch",
        "value": round(ntt["value"], 1),
        "unit": "Fr elems/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ntt["ms"], 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u32",
        "data": "synthetic (SplitMix64 uniform Fr, Montgomery form, generated on the device)",
        "config": {"workload": "NTT 2^%d, natural order in/out, device-resident" % lg, "lg_n": lg,
                   "passes": ntt["npass"], "verified_intt_roundtrip": ntt["verified"]},
        "roofline": roof,
        "measured_mad_peak_tmad_s": round(peak / 1e12, 3),
        "pass_splits": {str(k): v for k, v in plans.items()} if plans else "default",
    }
    if not args.no_h2h:
        result["host_to_host"] = bench_host_to_host(pl, synth, lg, stream)
        progress("host-to-host ntt: %s" % json.dumps(result["host_to_host"]["npoly_1"]))
        torch.cuda.empty_cache()
    if not args.no_lde:
        result["lde"] = bench_lde(pl, synth, lg, 2, args.steps, args.warmup, stream, peak)
        progress("lde: %.3f ms" % result["lde"]["ms"])
        torch.cuda.empty_cache()
        result["lde_l3"] = bench_lde(pl, synth, lg, 3, max(1, args.steps // 2), args.warmup, stream, peak)
        progress("lde lambda 3: %.3f ms" % result["lde_l3"]["ms"])
        torch.cuda.empty_cache()
        result["lde_b3"] = bench_lde(pl, synth, lg, -1, max(1, args.steps // 2), args.warmup, stream, peak)
        progress("lde blowup 3: %.3f ms" % result["lde_b3"]["ms"])
        torch.cuda.empty_cache()
    if not args.no_miss:
        result["miss_ops"] = bench_miss_ops(pl, synth, args.steps, stream, peak)
        progress("miss_ops: %s" % json.dumps(result["miss_ops"]["2^20"]))
        torch.cuda.empty_cache()
    if not args.no_batch:
        result["batched"] = bench_batch(pl, synth, args.steps, stream)
        progress("batched: %.3f ms vs plain %.3f ms" % (result["batched"]["p50_ms"], result["batched"]["plain_p50_ms"]))
        torch.cuda.empty_cache()
    if not args.no_serve:
        result["served"] = bench_served(pl, synth, args.serve_jobs, args.serve_slots, args.serve_large)
        torch.cuda.empty_cache()
    if not args.no_cpu:
        progress("cpu baseline (NTT/LDE on all cores, miss_ops loops on one core, ~20 s)")
        result["cpu_baseline"] = cpu_baseline(lg)
        if "miss_ops" in result and isinstance(result["cpu_baseline"].get("miss_ops"), dict):
            g, c = result["miss_ops"]["2^20"], result["cpu_baseline"]["miss_ops"]
            result["cpu_baseline"]["miss_ops"]["gpu_speedup"] = {
                "perm_grand_product": round(c["perm_grand_product_ms"] / g["perm_grand_product"]["ms"], 1),
                "eval": round(c["eval_ms"] / g["eval"]["ms"], 1), "div": round(c["naive_div_ms"] / g["div"]["ms"], 1)}
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
