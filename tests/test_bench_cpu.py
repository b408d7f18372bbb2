"""bench.py's command-line contract that needs no GPU: more than one GPU is
refused with a message and exit status 2, the option list is stable, and the
roofline finds measured traffic for both FFT passes."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _run(*args):
    return subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), *args],
                          capture_output=True, text=True, timeout=120)


def test_more_than_one_gpu_is_refused():
    r = _run("--gpus", "2")
    assert r.returncode == 2
    assert "one GPU" in r.stderr
    assert r.stdout == ""


def test_options():
    r = _run("--help")
    assert r.returncode == 0
    for opt in ("--gpus", "--steps", "--warmup", "--tiles", "--only", "--sweep-steps",
                "--plain-steps", "--copy-calibration", "--no-cpu"):
        assert opt in r.stdout


def test_traffic_for_either_dominant_pass():
    """the roofline's traffic comes from the committed counter summary for
    whichever pass measures longer (the two are within a few percent)"""
    sys.path.insert(0, ROOT)
    import bench

    for k in ("k_fft_rows", "k_wave_cols_split"):
        t = bench.measured_traffic(k)
        assert isinstance(t, int) and 10_000_000 < t < 20_000_000
    assert bench.measured_traffic("k_unknown") is None
