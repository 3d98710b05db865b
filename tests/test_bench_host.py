"""Host-side parts of bench.py: the one-GPU gate and the roofline record."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_more_than_one_gpu_exits_with_a_message():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr
    assert r.stdout == ""


def test_gemm_roofline_record():
    sys.path.insert(0, ROOT)
    import bench
    m, n, k, ms = 820, 23104, 23104, 4.0
    roof = bench.gemm_roofline(m, n, k, ms)
    assert {"bound", "achieved", "peak", "unit", "frac", "traffic"} <= set(roof)
    assert roof["bound"] == "mfma" and roof["unit"] == "TFLOP/s" and roof["peak"] == 2500.0
    # algorithmic work: 2 M N K per launch; the kernel's three bf16 products reported apart
    assert abs(roof["achieved"] - 2.0 * m * n * k / 4.0e-3 / 1e12) < 1e-9
    assert abs(roof["hardware_bf16_tflops"] - 3 * roof["achieved"]) < 1e-9
    assert abs(roof["frac"] - roof["achieved"] / 2500.0) < 1e-15
