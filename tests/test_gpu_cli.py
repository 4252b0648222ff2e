"""The drop-in layers on the GPU: the C++ operator API (SyncedMemory,
Ntt/Intt, accumulate_mul_poly, evaluate, poly_div_poly, element-wise ops) and
the NTT / LDE / miss_ops command lines with the reference's flags."""
import os
import subprocess

import numpy as np
import pytest

import pyoracle as po
from conftest import ROOT

pytestmark = pytest.mark.gpu
BIN = os.path.join(ROOT, "ppopp26-artifact_amd", "bin")


def run(args, cwd=None, timeout=300):
    print("run", args, flush=True)
    p = subprocess.run([os.path.join(BIN, args[0])] + args[1:], cwd=cwd, capture_output=True, text=True,
                       timeout=timeout)
    print(p.stdout[-3000:], p.stderr[-2000:], flush=True)
    return p


def load(path):
    return np.fromfile(path, dtype=np.uint64).reshape(-1, 4)


def test_cpp_api(tmp_path):
    lg = 14
    p = run(["api_check", str(tmp_path), str(lg)])
    assert p.returncode == 0
    f = lambda n: load(os.path.join(tmp_path, n))  # noqa: E731
    x, y, pt, c = f("x.bin"), f("y.bin"), f("pt.bin"), f("c.bin")
    np.testing.assert_array_equal(f("ntt.bin"), po.ntt(x, lg))
    np.testing.assert_array_equal(f("intt.bin"), x)
    gp = po.grand_product(x)
    np.testing.assert_array_equal(f("gp.bin"), gp)
    np.testing.assert_array_equal(f("gp_incl.bin")[:-1], gp[1:])
    np.testing.assert_array_equal(f("eval.bin"), po.eval_(x, pt))
    np.testing.assert_array_equal(f("div.bin"), po.div(x, c)[0])
    np.testing.assert_array_equal(f("add.bin"), po.add(x, y))
    np.testing.assert_array_equal(f("mul.bin"), po.mul(x, y))
    np.testing.assert_array_equal(f("div_ew.bin"), po.mul(x, po.inv(y)))
    np.testing.assert_array_equal(f("add_cpu.bin"), po.add(x, y))
    np.testing.assert_array_equal(f("mul_cpu.bin"), po.mul(x, y))
    np.testing.assert_array_equal(f("exp5.bin"), po.vec_exp(x, 5))
    np.testing.assert_array_equal(f("exp5_cpu.bin"), po.vec_exp(x, 5))
    np.testing.assert_array_equal(f("seq.bin"), po.sequence(pt, len(x) + 5))
    np.testing.assert_array_equal(f("rep.bin"), np.repeat(pt, 33, axis=0))
    np.testing.assert_array_equal(f("lde.bin"), po.lde(x, lg, 2, po.ref_lde_shift(2)))
    np.testing.assert_array_equal(f("stage0.bin"), po.stage_no_rotate(x, lg, 6, 0, 0))
    np.testing.assert_array_equal(f("stage6.bin"), po.stage_no_rotate(x[:1024], lg, 6, 6, 1))
    g = po.from_u64(7)
    np.testing.assert_array_equal(f("coset.bin"), po.coset_ntt(x, lg, g))
    np.testing.assert_array_equal(f("icoset.bin"), po.coset_intt(x, lg, g))
    np.testing.assert_array_equal(f("coset_steps.bin"), po.coset_ntt(x, lg, g))
    np.testing.assert_array_equal(f("icoset_steps.bin"), po.coset_intt(x, lg, g))
    q = len(x) // 4
    np.testing.assert_array_equal(f("coset_init0.bin"), po.mul(x[:q], np.concatenate([po.pow_(g, i) for i in range(q)])))
    np.testing.assert_array_equal(f("coset_lde3.bin"), po.lde(x, lg, 3, g))


@pytest.mark.parametrize("mode", ["--check-split", "--check-baseline"])
def test_ntt_cli_check(mode):
    p = run(["NTT", mode, "22"])
    assert p.returncode == 0
    assert "outputs match for 4194304 elements" in p.stdout


def test_ntt_cli_bench_split():
    p = run(["NTT", "--bench-split", "24"])
    assert p.returncode == 0
    assert p.stdout.count("our-ntt execution time:") == 3


def test_ntt_cli_errors():
    assert run(["NTT"]).returncode == 1
    p = run(["NTT", "--bench-split", "21"])
    assert p.returncode == 1 and "Unsupported lg_N = 21" in p.stderr


def test_lde_cli():
    p = run(["LDE", "--bench", "20", "2"])
    assert p.returncode == 0
    assert p.stdout.count("our-lde execution time:") == 3
    assert "[our_lde] avg lg_N=20, lambda=2" in p.stdout
    p = run(["LDE", "--bench", "21"])
    assert p.returncode == 1 and "error: unsupported LDE test config" in p.stderr


def test_miss_ops_cli():
    p = run(["miss_ops", "--lg", "16", "--perm"])
    assert p.returncode == 0
    for tag in ("[naive_gp]", "[ours_gp]", "[ours_perm_gp]", "[ours_eval]", "[naive_div]", "[ours_div]"):
        assert p.stdout.count(tag + " execution time:") == 3, tag


def test_ntt_cli_autotune(tmp_path):
    cache = str(tmp_path / "plans.txt")
    p = run(["NTT", "--autotune", "10", "12", cache])
    assert p.returncode == 0
    lines = [l for l in p.stdout.splitlines() if l.startswith("[autotune]")]
    assert len(lines) == 3
    rows = [l.split() for l in open(cache) if not l.startswith("#")]
    assert [int(r[0]) for r in rows] == [10, 11, 12]
    for r in rows:
        m = [int(v) for v in r[2:2 + int(r[1])]]
        assert sum(m) == int(r[0]) and all(3 <= v <= 6 for v in m)


def test_ntt_check_split_uses_the_reference_flow():
    """--check-split's expected output comes from the reference's chunked
    flow on the device (pl_ntt_segmented), a different set of kernels from
    our_ntt's pipelined passes, so a match says something."""
    p = run(["NTT", "--check-split", "24"])
    assert p.returncode == 0
    assert "[check] expected: the reference's chunked flow on the device, split 14+10, 2^4 segments" in p.stdout
    assert "outputs match for 16777216 elements" in p.stdout


def test_miss_ops_cli_outputs_vs_oracle(tmp_path):
    """Every miss_ops output, the CPU loops included (naive_gp, naive_div),
    against the oracle, on the reference's own input file (the first 4096
    elements of input.bin, read through the CLI's loader)."""
    from conftest import GOLDEN
    src = os.path.join(GOLDEN, "input_prefix.bin")
    p = run(["miss_ops", "--lg", "12", "--perm", "--input", src, "--dump", str(tmp_path)])
    assert p.returncode == 0
    f = lambda n: load(os.path.join(tmp_path, n))  # noqa: E731
    x = f("input.bin")
    np.testing.assert_array_equal(x, load(src))
    np.testing.assert_array_equal(f("naive_gp.bin"), po.naive_gp(x))
    np.testing.assert_array_equal(f("ours_gp.bin"), po.grand_product(x))
    point = np.array([[123456789, 987654321, 1122334455, 6677889900]], dtype=np.uint64)
    np.testing.assert_array_equal(f("ours_eval.bin"), po.eval_(x, point))
    c = np.array([[123456789123456789, 987654321987654321, 1122334455667788, 1122334455667788]], dtype=np.uint64)
    np.testing.assert_array_equal(f("naive_div.bin"), po.naive_div(x, c, po.one()))
    np.testing.assert_array_equal(f("ours_div.bin"), po.div(x, c)[0])
    cols = f("perm_cols.bin").reshape(6, -1, 4)
    bg = f("perm_bg.bin")
    ks = [po.from_u64(1), po.from_u64(7), po.from_u64(13)]
    ratio = po.perm_ratio([np.ascontiguousarray(cols[j]) for j in range(6)], bg[0:1], bg[1:2], ks, 12)
    np.testing.assert_array_equal(f("perm_z.bin"), po.grand_product(ratio))
    # the reference's own blst loops agree with the CLI's CPU loops too
    ref = po.ref_lib()
    if ref is not None:
        out = po.fr_array(len(x))
        ref.ref_naive_gp(po.P(x), po.P(out), len(x))
        np.testing.assert_array_equal(f("naive_gp.bin"), out)


def test_input_loader_short_file(tmp_path):
    """A file shorter than the buffer is read and the rest zero-filled, with
    a message (the reference reads whatever the file holds)."""
    from conftest import GOLDEN
    src = os.path.join(GOLDEN, "input_prefix.bin")
    p = run(["miss_ops", "--lg", "13", "--input", src, "--dump", str(tmp_path)])
    assert p.returncode == 0
    assert "rest zero-filled" in p.stderr
    x = load(os.path.join(tmp_path, "input.bin"))
    np.testing.assert_array_equal(x[:4096], load(src))
    assert not x[4096:].any()
    np.testing.assert_array_equal(load(os.path.join(tmp_path, "naive_gp.bin")), po.naive_gp(x))


def test_input_loader_missing_file(tmp_path):
    """A missing file is reported and replaced by the seeded generator."""
    p = run(["miss_ops", "--lg", "10", "--input", str(tmp_path / "absent.bin"), "--dump", str(tmp_path)])
    assert p.returncode == 0
    assert "Error opening file" in p.stderr and "seeded synthetic data" in p.stderr
    x = load(os.path.join(tmp_path, "input.bin"))
    np.testing.assert_array_equal(x, po.uniform(1 << 10, po.seed(po.OP_GP, 10)))


def test_ntt_cli_bench_sweep():
    # the reference's segment-count sweep: lg_k = 1..8 at one size
    p = run(["NTT", "--bench-sweep", "22", "1"])
    assert p.returncode == 0
    assert p.stdout.count("=== lg_N = 22") == 8
    assert p.stdout.count("our-ntt execution time:") == 8


def test_lde_cli_naive_and_sweep():
    p = run(["LDE", "--bench-naive", "20", "2"])
    assert p.returncode == 0 and "[naive-lde] lg_N=20, lambda=2" in p.stdout
    p = run(["LDE", "--bench-sweep", "20", "2", "1"])
    assert p.returncode == 0 and p.stdout.count("=== lde lg_N = 20, lambda = 2") == 8


def test_miss_ops_cli_no_arguments(tmp_path):
    # the reference's default: 2^22 from ../../input.bin (absent here: the
    # seeded fallback), every timing line three times
    d = tmp_path / "a" / "b"
    d.mkdir(parents=True)
    p = run(["miss_ops"], cwd=str(d), timeout=600)
    assert p.returncode == 0
    for tag in ("[naive_gp]", "[ours_gp]", "[ours_eval]", "[naive_div]", "[ours_div]"):
        assert p.stdout.count(tag + " execution time:") == 3, tag
