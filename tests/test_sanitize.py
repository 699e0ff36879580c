"""The oracle and the host helpers of the drop-in executables built under
AddressSanitizer and UndefinedBehaviorSanitizer (gcc / g++, CPU only, no GPU
code): each driver in tests/sanitize/ runs every function of its library on
small inputs (log and linear grids, spot, facula and solar modes, the mask CCF,
the fits, the differentiable model; .bin / text / .npz I/O, polyfit) and must
exit cleanly with no sanitizer report.  Skipped where the compiler has no
sanitizer runtime."""
from __future__ import annotations

import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SAN = ["-fsanitize=address,undefined", "-fno-sanitize-recover=all", "-fno-omit-frame-pointer", "-g", "-O1"]
ENV = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0", UBSAN_OPTIONS="print_stacktrace=1")


def _have(cc, tmp_path):
    if not shutil.which(cc):
        return False
    src = tmp_path / ("probe.c" if cc == "gcc" else "probe.cpp")
    src.write_text("int main(void) { return 0; }\n")
    r = subprocess.run([cc, *SAN, str(src), "-o", str(tmp_path / "probe")], capture_output=True)
    return r.returncode == 0


def _run(exe, args, tmp_path):
    r = subprocess.run([exe, *args], capture_output=True, text=True, timeout=600, env=ENV, cwd=str(tmp_path))
    report = r.stderr
    assert r.returncode == 0, report[-4000:]
    assert "runtime error" not in report and "ERROR: AddressSanitizer" not in report, report[-4000:]
    assert "ERROR: LeakSanitizer" not in report, report[-4000:]
    return r.stdout


def test_oracle_under_sanitizers(tmp_path):
    if not _have("gcc", tmp_path):
        pytest.skip("gcc without the sanitizer runtime")
    exe = str(tmp_path / "oracle_asan")
    subprocess.run(["gcc", *SAN, "-std=c11", "-ffp-contract=off", "-Wno-unknown-pragmas", "-I",
                    os.path.join(ROOT, "oracle"), os.path.join(ROOT, "tests", "sanitize", "oracle_asan.c"),
                    os.path.join(ROOT, "oracle", "soap_oracle.c"), "-lm", "-o", exe], check=True)
    assert "ok" in _run(exe, [], tmp_path)


def test_host_helpers_under_sanitizers(tmp_path):
    if not _have("g++", tmp_path):
        pytest.skip("g++ without the sanitizer runtime")
    exe = str(tmp_path / "host_asan")
    subprocess.run(["g++", *SAN, "-std=c++17", "-I", os.path.join(ROOT, "include"), "-I",
                    os.path.join(ROOT, "soap_gpu_amd", "driver"), os.path.join(ROOT, "tests", "sanitize", "host_asan.cpp"),
                    "-o", exe], check=True)
    assert "ok" in _run(exe, [str(tmp_path)], tmp_path)
