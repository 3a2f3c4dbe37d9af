"""AddressSanitizer + UndefinedBehaviorSanitizer run of the host layer's text
parsers (camera, image and grid files, valid and corrupt: truncated,
non-numeric, counts beyond the data, indices outside the cage), the camera
and shuffle code, and the oracle's C restatement on a small caged grid
(tests/san: driver.cpp, Makefile).  CPU only: the host sources are built by
hipcc with the sanitizers on the host side, the oracle and the driver by the
ROCm clang; the run must end with "san: ok" and no sanitizer report."""
import os
import shutil
import subprocess

import pytest

SAN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "san")


@pytest.mark.timeout(600)
def test_host_layer_and_oracle_under_sanitizers(tmp_path):
    if shutil.which("hipcc") is None and not os.path.exists("/opt/rocm/bin/hipcc"):
        pytest.skip("no hipcc")
    subprocess.run(["make", "-s", "-C", SAN, "-j4"], check=True, capture_output=True)
    env = dict(os.environ)
    env["ASAN_OPTIONS"] = "detect_leaks=1:halt_on_error=1:abort_on_error=0"
    env["UBSAN_OPTIONS"] = "halt_on_error=1:print_stacktrace=1"
    r = subprocess.run([os.path.join(SAN, "_build", "san_driver"), str(tmp_path)], env=env,
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr[-4000:]
    assert "san: ok" in r.stdout
    assert "ERROR: AddressSanitizer" not in r.stderr and "runtime error" not in r.stderr
