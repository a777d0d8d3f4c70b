"""The CPU oracle under AddressSanitizer and UndefinedBehaviorSanitizer
(oracle/asan_check.cpp, `make -C oracle asan`, g++): float32, float64 and
mixed runs, screened and unscreened, base depth 0 with V-cycles, every stage
array exported, fields evaluated, and the stand-alone sliced relaxations
(serial and tiled, compared bit for bit).  A sanitizer finding aborts the
driver with a non-zero status."""
import subprocess
from pathlib import Path

ROOT = Path(__file__).resolve().parents[1]


def test_oracle_clean_under_sanitizers():
    subprocess.run(["make", "-s", "-C", str(ROOT / "oracle"), "asan"], check=True, timeout=600)
    out = subprocess.run([str(ROOT / "oracle" / "_build" / "asan_check")], capture_output=True, text=True, timeout=600)
    assert out.returncode == 0, out.stdout + out.stderr
    assert "asan ok" in out.stdout
