"""The oracle's C code under AddressSanitizer and UndefinedBehaviorSanitizer
(host code only): tests/sanitize/oracle_main.c drives every entry point on
boxes of odd and tiny sizes, 2D and 3D, every solver and EMF average, and
ConsToPrim floors; an out-of-bounds access or undefined behaviour aborts
the run with a sanitizer report."""
import os
import shutil
import subprocess

import pytest

from conftest import ROOT


@pytest.mark.skipif(shutil.which("gcc") is None, reason="gcc not available")
def test_oracle_under_asan_ubsan(tmp_path):
    exe = tmp_path / "oracle_asan"
    cmd = ["gcc", "-O1", "-g", "-fno-omit-frame-pointer", "-ffp-contract=off",
           "-fsanitize=address,undefined", "-fno-sanitize-recover=undefined",
           os.path.join(ROOT, "tests", "sanitize", "oracle_main.c"),
           os.path.join(ROOT, "oracle", "mhd_oracle.c"), "-lm", "-o", str(exe)]
    subprocess.run(cmd, check=True, capture_output=True, text=True)
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0",
               UBSAN_OPTIONS="print_stacktrace=1")
    env.pop("LD_PRELOAD", None)
    p = subprocess.run([str(exe)], capture_output=True, text=True, env=env, timeout=600)
    assert p.returncode == 0, p.stdout[-3000:] + p.stderr[-6000:]
    assert "sanitized oracle run: ok" in p.stdout
    assert "runtime error" not in p.stderr and "AddressSanitizer" not in p.stderr


HIPCC = "/opt/rocm/bin/hipcc"
CLANG = "/opt/rocm/llvm/bin/clang++"


@pytest.mark.skipif(not (os.path.exists(HIPCC) and os.path.exists(CLANG)),
                    reason="ROCm compilers not available")
def test_dbl_host_code_under_asan_ubsan(tmp_path):
    """The library's .dbl writer and reader (host code of dbl_io.hip) with the
    host compilation pass sanitized (-Xarch_host: the device pass is not)."""
    obj = tmp_path / "dbl_io.o"
    exe = tmp_path / "dbl_asan"
    san = ["-Xarch_host", "-fsanitize=address", "-Xarch_host", "-fsanitize=undefined"]
    subprocess.run([HIPCC, "-x", "hip", "--offload-arch=gfx950", "-O1", "-g", "-std=c++17"] + san +
                   ["-I" + os.path.join(ROOT, "include"), "-c",
                    os.path.join(ROOT, "pluto_amd", "csrc", "dbl_io.hip"), "-o", str(obj)],
                   check=True, capture_output=True, text=True)
    subprocess.run([CLANG, "-O1", "-g", "-fsanitize=address,undefined",
                    os.path.join(ROOT, "tests", "sanitize", "dbl_main.cpp"), str(obj),
                    "-L/opt/rocm/lib", "-lamdhip64", "-Wl,-rpath,/opt/rocm/lib", "-o", str(exe)],
                   check=True, capture_output=True, text=True)
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=0", UBSAN_OPTIONS="print_stacktrace=1")
    env.pop("LD_PRELOAD", None)
    out = tmp_path / "out"
    out.mkdir()
    p = subprocess.run([str(exe), str(out)], capture_output=True, text=True, env=env, timeout=300)
    assert p.returncode == 0, p.stdout[-3000:] + p.stderr[-6000:]
    assert "sanitized dbl run: ok" in p.stdout
    assert "runtime error" not in p.stderr and "AddressSanitizer" not in p.stderr
