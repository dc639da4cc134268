"""AddressSanitizer / UndefinedBehaviorSanitizer runs of the host code (CPU only).

The oracle's C restatement is rebuilt with both sanitizers (oracle/Makefile,
target `sanitize`, no recovery: the first report aborts) and its pin tests run
against that build in a child process with libasan preloaded; the facade's host
numerics (bb_precompute, pca_compute, the Kronecker PCA, the eigensolver and the
closed forms of cpp/utils.cpp) run under both sanitizers through
tests/cpp/utils_check.cpp and must print what the plain build prints.
"""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "financial-derivatives-montecarlo_amd")
SAN = ["-fsanitize=address,undefined", "-fno-sanitize-recover=undefined", "-fno-omit-frame-pointer", "-g"]


def _libasan():
    p = subprocess.run(["gcc", "-print-file-name=libasan.so"], capture_output=True, text=True)
    path = p.stdout.strip()
    if p.returncode != 0 or not os.path.isabs(path) or not os.path.exists(path):
        pytest.skip("libasan is not available to gcc")
    return path


def test_oracle_pins_under_asan_ubsan():
    asan = _libasan()
    subprocess.run(["make", "-C", os.path.join(ROOT, "oracle"), "sanitize"], check=True, capture_output=True)
    env = dict(os.environ, LD_PRELOAD=asan, ASAN_OPTIONS="detect_leaks=0:abort_on_error=1",
               UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1",
               FDMC_ORACLE_LIB=os.path.join(ROOT, "oracle", "_build", "liboracle_san.so"))
    tests = [os.path.join(ROOT, "tests", f) for f in ("test_oracle_pins.py", "test_book_oracle.py",
                                                       "test_hh_seeding.py")]
    p = subprocess.run([sys.executable, "-m", "pytest", "-x", "-q", "-p", "no:cacheprovider"] + tests,
                       env=env, capture_output=True, text=True, cwd=ROOT, timeout=1200)
    tail = (p.stdout + p.stderr)[-3000:]
    assert p.returncode == 0, tail
    assert "ERROR: AddressSanitizer" not in tail and "runtime error:" not in tail, tail


def test_host_numerics_under_asan_ubsan(tmp_path):
    _libasan()
    src = [os.path.join(ROOT, "tests", "cpp", "utils_check.cpp"), os.path.join(PKG, "cpp", "utils.cpp")]
    outs = []
    for flags, name in (([], "plain"), (SAN, "san")):
        exe = str(tmp_path / name)
        subprocess.run(["g++", "-O1", "-std=c++17", "-I" + os.path.join(PKG, "cpp")] + flags + src + ["-o", exe],
                       check=True)
        p = subprocess.run([exe], capture_output=True, text=True,
                           env=dict(os.environ, ASAN_OPTIONS="detect_leaks=1", UBSAN_OPTIONS="halt_on_error=1"))
        assert p.returncode == 0 and "runtime error" not in p.stderr, p.stderr[-2000:]
        outs.append(p.stdout)
    assert outs[0] == outs[1]
