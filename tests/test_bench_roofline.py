"""bench.py on the CPU: the roofline of the dominant kernel (the Hadamard's
bound is the longer of its two roofline times, with the figures of the N=2M,
p=10 Laplace tree — 104,368 M2L pairs, 4,680 source and 4,680 target nodes,
4,400 frequencies — worked out by hand) and the refusal of multi-GPU runs."""
import os
import subprocess
import sys

import pytest

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
import bench  # noqa: E402

PAIRS, NFREQ, NSRC, NTRG = 104368, 4400, 4680, 4680
FLOPS = 512 * PAIRS * NFREQ  # 2.351e11


def test_float_hadamard_is_hbm_bound():
    r = bench.roofline("HAD", 1.4211, True, True, 0, PAIRS, NFREQ, NSRC, NTRG)
    # spectra in and out: 9,360 nodes x 4,400 freq x 8 children x (re, im) x 4 B;
    # operators: 26 x 4,400 x 128 reals x 4 B
    assert r["algorithmic_bytes"] == 9360 * 4400 * 64 + 26 * 4400 * 512 == 2694348800
    assert r["bound"] == "hbm" and r["unit"] == "GB/s" and r["peak"] == 8000.0
    assert r["achieved"] == pytest.approx(2694348800 / 1.4211e-3 / 1e9)
    assert r["frac"] == pytest.approx(r["achieved"] / 8000.0)
    # the other view: FP32-equivalent flops against the f16 rate / 3 products
    assert r["mfma"]["peak"] == pytest.approx(2500.0 / 3)
    assert r["mfma"]["achieved"] == pytest.approx(FLOPS / 1.4211e-3 / 1e12)
    assert r["ms"] == 1.4211


def test_fp64_hadamard_is_mfma_bound():
    # 2.35e11 flop / 78.6 TFLOP/s = 3.0 ms > 5.39 GB / 8 TB/s = 0.67 ms
    r = bench.roofline("HAD", 4.8, False, False, 0, PAIRS, NFREQ, NSRC, NTRG)
    assert r["bound"] == "mfma" and r["peak"] == pytest.approx(78.6)
    assert r["algorithmic_flops"] == FLOPS
    assert r["frac"] == pytest.approx(FLOPS / 4.8e-3 / 78.6e12)
    assert r["hbm"]["achieved"] == pytest.approx((9360 * 4400 * 128 + 26 * 4400 * 1024) / 4.8e-3 / 1e9)


def test_p2p_is_valu_bound():
    r = bench.roofline("P2P", 1.3, True, True, 3097030512, PAIRS, NFREQ, NSRC, NTRG)
    assert r["bound"] == "valu" and r["peak"] == pytest.approx(157.3)
    assert r["achieved"] == pytest.approx(20 * 3097030512 / 1.3e-3 / 1e12)


def test_transform_rows_are_hbm_bound():
    r = bench.roofline("IFFT", 0.6, True, True, 0, PAIRS, NFREQ, NSRC, 4000)
    assert r["bound"] == "hbm"
    assert r["achieved"] == pytest.approx(4000 * 4400 * 64 / 0.6e-3 / 1e9)


@pytest.mark.parametrize("argv,env", [(["--gpus", "2"], {}), (["--gpus", "1"], {"WORLD_SIZE": "8"})])
def test_more_than_one_gpu_is_refused(argv, env):
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    r = subprocess.run([sys.executable, os.path.join(root, "bench.py")] + argv, capture_output=True, text=True,
                       env=dict(os.environ, **env), timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr and r.stdout == ""


def test_committed_traffic_matches_the_kernel_source():
    """The Hadamard traffic that bench.py reports comes from a committed FETCH_SIZE / WRITE_SIZE
    measurement; it is keyed by the sha256 of the kernel's source text, so it is reported only while the
    kernel is the one that was measured (a changed kernel reports null until it is measured again)."""
    import argparse
    a = argparse.Namespace(n=2_000_000, p=10, ncrit=128)
    t = bench.measured_traffic("HAD", True, a)
    assert t is not None and 2.0e9 < t < 6.0e9
    h = bench.kernel_source_sha256("k_hadamard_h")
    assert h is not None and len(h) == 64
    src = os.path.join(os.path.dirname(bench.__file__), "exafmm-t_amd", "csrc", "m2l.hip")
    text = open(src).read()
    import tempfile
    with tempfile.NamedTemporaryFile("w", suffix=".hip", delete=False) as f:
        f.write(text.replace("void k_hadamard_h(", "void k_hadamard_h(\n", 1))
    try:
        assert bench.kernel_source_sha256("k_hadamard_h", f.name) != h
    finally:
        os.unlink(f.name)
