"""The host-only C++ of the library (beam geometry, doseMatFolder reader and merger) built with
AddressSanitizer and UndefinedBehaviorSanitizer and driven with random, degenerate and corrupted
inputs (tests/host/sanitize_driver.cpp).  CPU only: no GPU code is built or run."""
import shutil
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
CSRC = ROOT / "fastdose_amd" / "csrc"


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
def test_host_code_under_asan_ubsan(tmp_path):
    exe = tmp_path / "sanitize_driver"
    cmd = ["g++", "-std=c++17", "-O1", "-g", "-fno-omit-frame-pointer", "-ffp-contract=off",
           "-fsanitize=address,undefined", "-fno-sanitize-recover=all", f"-I{ROOT / 'include'}", f"-I{CSRC}",
           str(ROOT / "tests" / "host" / "sanitize_driver.cpp"), str(CSRC / "geometry.cpp"),
           str(CSRC / "dosemat_files.cpp"), "-o", str(exe)]
    res = subprocess.run(cmd, capture_output=True, text=True)
    assert res.returncode == 0, res.stderr[-4000:]
    work = tmp_path / "folders"
    work.mkdir()
    env = {"ASAN_OPTIONS": "detect_leaks=1:abort_on_error=1", "UBSAN_OPTIONS": "print_stacktrace=1",
           "PATH": "/usr/bin:/bin"}
    run = subprocess.run([str(exe), str(work)], capture_output=True, text=True, env=env, timeout=300)
    assert run.returncode == 0, (run.stdout + run.stderr)[-6000:]
    assert "geometry:" in run.stdout and "folders:" in run.stdout
