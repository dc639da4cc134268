"""The reference's own example programs (examples/ejemplo01..11*.cpp) against the
C++ facade: compiled unchanged (CPU), and run on the GPU at the first precision of
each example's section in the reference's report.

The compile test reads the reference checkout and is skipped where it is absent;
the run test uses the executables that build() leaves in lib/reference_examples/
and is skipped where they were not built.
"""
import math
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "financial-derivatives-montecarlo_amd")
sys.path.insert(0, PKG)
sys.path.insert(0, os.path.join(PKG, "cpp", "examples"))

import build  # noqa: E402
from reference_sweep import BIN, SWEEP, parse_table  # noqa: E402

NAMES = sorted(SWEEP)


@pytest.mark.skipif(not os.path.isdir(build.REF_EXAMPLES), reason="reference checkout not present")
def test_reference_examples_compile_unchanged(tmp_path):
    if not os.path.exists(build.LIB):
        pytest.skip("libfdmc.so not built")
    built = build.build_reference_examples(out_dir=str(tmp_path))
    assert sorted(os.path.basename(b) for b in built) == NAMES
    for b in built:
        assert os.access(b, os.X_OK)


@pytest.mark.gpu
@pytest.mark.parametrize("name", NAMES)
def test_reference_example_runs(name):
    exe = os.path.join(BIN, name)
    if not os.path.exists(exe):
        pytest.skip("reference example programs not built (no reference checkout at build time)")
    eps = SWEEP[name][0]
    print("running", name, "eps", eps, flush=True)
    p = subprocess.run([exe, repr(eps)], capture_output=True, text=True, timeout=90)
    assert p.returncode == 0, p.stderr[-2000:]
    ref, rows = parse_table(p.stdout)
    assert ref is not None and len(rows) >= 2, p.stdout
    for r in rows:
        assert math.isfinite(r["price"]) and r["std_error"] >= 0 and r["n"] > 0, r
        # the examples' own criterion: |price - their reference| < 2 eps
        assert r["ok"], (name, r)
