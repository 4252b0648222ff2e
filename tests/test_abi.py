"""CPU-side checks of the C-ABI library: it loads without a GPU, exports every
symbol include/pl.h declares, and its pure host functions behave."""
import ctypes
import os
import re
import subprocess

import pytest

from conftest import ROOT

HDR = os.path.join(ROOT, "include", "pl.h")


def declared():
    txt = open(HDR).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:const\s+)?[a-z_0-9]+\**\s+\**(pl_[a-z_0-9]+)\s*\(", txt, re.M)))


def test_header_declares_entry_points():
    names = declared()
    for n in ("pl_ntt", "pl_lde", "pl_grand_product", "pl_eval", "pl_div_linear", "pl_domain_create"):
        assert n in names


def test_library_exports_every_declared_symbol():
    import plffi
    if not os.path.exists(plffi.LIB_PATH):
        pytest.skip("libpl.so not built")
    out = subprocess.check_output(["nm", "-D", "--defined-only", plffi.LIB_PATH]).decode()
    exported = set(l.split()[-1] for l in out.splitlines() if " T " in l)
    missing = [n for n in declared() if n not in exported]
    assert not missing, missing
    # and the ctypes binding covers them all with signatures
    assert set(declared()) <= set(plffi.SIGNATURES), set(declared()) - set(plffi.SIGNATURES)


def test_library_loads_and_host_functions():
    import plffi
    if not os.path.exists(plffi.LIB_PATH):
        pytest.skip("libpl.so not built")
    L = plffi.lib()
    assert L.pl_version() >= 100
    assert L.pl_status_string(0) == b"ok"
    assert L.pl_status_string(-3) == b"unsupported size"
    # plan: ceil(lg/6) passes of 3..6 points per column; tiny sizes one pass
    assert [L.pl_ntt_num_passes(lg) for lg in (1, 3, 4, 6, 7, 12, 13, 18, 19, 24, 25, 28)] == \
        [1, 1, 1, 1, 2, 2, 3, 3, 4, 4, 5, 5]
    assert L.pl_ntt_num_passes(29) == -3 and L.pl_ntt_num_passes(0) == -3
    assert L.pl_ntt_work_bytes(5, 1) == 0          # every split of 2^5 is one pass, in place
    assert L.pl_ntt_work_bytes(6, 1) == 32 << 6    # 2^6 may be split 3+3 (size fixed by lg alone)
    assert L.pl_ntt_work_bytes(24, 1) == 32 << 24  # one intermediate buffer
    assert L.pl_ntt_work_bytes(12, 3) == 3 * (32 << 12)
    assert L.pl_lde_work_bytes(24, 2) == 32 << 26
    assert L.pl_scan_work_bytes(1 << 20) >= 32 * 2 * 256


def test_no_device_is_an_error_not_a_crash():
    import plffi
    if not os.path.exists(plffi.LIB_PATH):
        pytest.skip("libpl.so not built")
    try:
        import torch
        if torch.cuda.is_available():
            pytest.skip("a device is present")
    except ImportError:
        pass
    h = ctypes.c_void_p()
    st = plffi.lib().pl_domain_create(12, ctypes.byref(h))
    assert st == -5 and not h.value


def test_argument_errors_without_device():
    import plffi
    if not os.path.exists(plffi.LIB_PATH):
        pytest.skip("libpl.so not built")
    L = plffi.lib()
    h = ctypes.c_void_p()
    assert L.pl_domain_create(0, ctypes.byref(h)) == -1
    assert L.pl_domain_create(29, ctypes.byref(h)) == -3
    assert L.pl_lde_domain_create(24, 5, None, ctypes.byref(h)) == -3
    assert L.pl_ntt(None, None, 12, 0, None, None, None) == -1
    assert L.pl_grand_product(None, 10, None, None, None) == -1
    assert L.pl_div_linear(None, 0, None, None, None, None) == -1
    # element-wise entry points: the checks run before any device call
    big = (1 << 36) + 1
    assert L.pl_exp_mod(4096, 8192, 4, -1, None) == -1  # negative exponent
    assert L.pl_exp_mod(4096, 8192, big, 3, None) == -1  # beyond the grid bound
    assert L.pl_exp_mod(4100, 8192, 4, 3, None) == -1  # misaligned
    assert L.pl_vec_op(0, 4096, 8192, 12288, big, None, None) == -1
    assert L.pl_vec_op(99, 4096, 8192, 12288, 4, None, None) == -1  # no such op
    assert L.pl_gen_sequence((1 << 28) + 1, 4096, 8192, None) == -1
    assert L.pl_gen_sequence(4, 4096 + 3 * 32, 4096, None) == -1  # x inside the output
    assert L.pl_repeat_to_poly(4096 + 32, 4096, 4, None) == -1  # x inside the output
    assert L.pl_repeat_to_poly(4096, 8192, big, None) == -1


def test_batch_argument_errors_without_device():
    import plffi
    if not os.path.exists(plffi.LIB_PATH):
        pytest.skip("libpl.so not built")
    L = plffi.lib()
    h = ctypes.c_void_p()
    assert L.pl_batch_create(None, 1, 1, ctypes.byref(h)) == -1
    items = (plffi.BatchItem * 2)()
    for it in items:
        it.op, it.lg, it.d_in, it.d_out = 0, 12, 256, 512
    assert L.pl_batch_create(ctypes.cast(items, ctypes.c_void_p), 2, 2, ctypes.byref(h)) == -1  # flags
    assert L.pl_batch_create(ctypes.cast(items, ctypes.c_void_p), 0, 1, ctypes.byref(h)) == -1
    items[1].lg = 29
    assert L.pl_batch_create(ctypes.cast(items, ctypes.c_void_p), 2, 1, ctypes.byref(h)) == -3
    items[1].lg, items[1].op = 12, 7
    assert L.pl_batch_create(ctypes.cast(items, ctypes.c_void_p), 2, 1, ctypes.byref(h)) == -1
    items[1].op, items[1].d_out = 3, 256  # eval_div: out aliases in, no point
    assert L.pl_batch_create(ctypes.cast(items, ctypes.c_void_p), 2, 1, ctypes.byref(h)) == -1
    items[1].op, items[1].lg_blowup, items[1].d_out = 2, 0, 512
    assert L.pl_batch_create(ctypes.cast(items, ctypes.c_void_p), 2, 1, ctypes.byref(h)) == -1
    items[1].d_in = 260  # misaligned
    items[1].lg_blowup = 2
    assert L.pl_batch_create(ctypes.cast(items, ctypes.c_void_p), 2, 1, ctypes.byref(h)) == -1
    assert not h.value
    assert L.pl_batch_info(None, None, None) == -1
    assert L.pl_batch_run(None, None) == -1
    L.pl_batch_destroy(None)


def test_generated_montgomery_columns_current():
    # mont_columns.inc is generated; it must match its generator
    gen = os.path.join(ROOT, "ppopp26-artifact_amd", "csrc", "gen_mont_columns.py")
    inc = os.path.join(ROOT, "ppopp26-artifact_amd", "csrc", "mont_columns.inc")
    out = subprocess.check_output(["python3", gen]).decode()
    assert out == open(inc).read()
    # every fold reads its carry two or more instructions after the mad wrote it
    import re
    for blk in re.findall(r'asm\("([^"]*)"', out):
        ins = blk.split("\\n\\t")
        written = {}
        for pos, t in enumerate(ins):
            if t.startswith("s_nop"):
                continue
            m = re.match(r"v_mad_u64_u32 %0, (%\d+),", t)
            if m:
                written[m.group(1)] = pos
                continue
            c = t.split(", ")[-1]
            assert c in written, t
            nops = sum(int(x.split()[1]) + 1 for x in ins[written[c] + 1:pos] if x.startswith("s_nop"))
            valu = sum(1 for x in ins[written[c] + 1:pos] if not x.startswith("s_nop"))
            assert valu + nops >= 2, blk
