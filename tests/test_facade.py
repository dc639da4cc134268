"""C++ facade (methods_hip.hpp / utils.hpp): host numerics on the CPU against the
oracle and closed forms; the drop-in example program on the GPU."""
import json
import os
import subprocess

import numpy as np
import pytest

import pyoracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "financial-derivatives-montecarlo_amd")


def test_facade_utils_match_oracle(tmp_path):
    exe = str(tmp_path / "utils_check")
    subprocess.run(["g++", "-O2", "-std=c++17", "-I" + os.path.join(PKG, "cpp"),
                    os.path.join(ROOT, "tests", "cpp", "utils_check.cpp"), os.path.join(PKG, "cpp", "utils.cpp"),
                    "-o", exe], check=True)
    out = json.loads(subprocess.run([exe], check=True, capture_output=True, text=True).stdout)
    assert abs(out["bs"] - O.bs_call(100, 100, 1, 0.05, 0.2)) < 1e-12
    assert abs(out["geo"] - O.geom_asian_analytic(100, 95, 1, 0.05, 0.3, 64)) < 1e-12
    M = np.array(out["pca8"]).reshape(8, 8).T          # column-major -> M[i, k]
    Mo = O.pca_compute(8, 1.0)
    assert np.allclose(M @ M.T, np.eye(8) / 8, atol=1e-14)
    assert np.allclose(np.abs(M), np.abs(Mo), atol=1e-12)   # eigenvector signs are free
    K = np.array(out["kron"]).reshape(12, 12).T
    corr = np.array([[1.0, 0.3, 0.2], [0.3, 1.0, 0.4], [0.2, 0.4, 1.0]])
    assert np.allclose(K @ K.T, np.kron(np.eye(4), corr) / 4, atol=1e-14)
    bb = O.OracleBB(16, 2.0).arrays()
    assert out["bb_map"] == list(bb["map_idx"])
    assert np.allclose(out["bb_sd"], bb["sd"], atol=1e-15)
    assert np.allclose(out["eig"], np.sort(np.linalg.eigvalsh(np.array([[4, 1, 2], [1, 3, 0], [2, 0, 5.0]])))[::-1],
                       atol=1e-13)


@pytest.mark.gpu
def test_facade_example_runs_reference_call_sequence():
    exe = os.path.join(PKG, "lib", "asian_table")
    p = subprocess.run([exe, "0.01"], capture_output=True, text=True, timeout=300)
    print(p.stdout)
    assert p.returncode == 0, p.stdout + p.stderr
    assert "MLQMC BB" in p.stdout and "facade rows outside tolerance: 0" in p.stdout


@pytest.mark.gpu
def test_facade_book_example_batched_equals_plain():
    exe = os.path.join(PKG, "lib", "book_table")
    p = subprocess.run([exe, "0.01", "48"], capture_output=True, text=True, timeout=300)
    print(p.stdout)
    assert p.returncode == 0, p.stdout + p.stderr
    assert "batched == plain: yes" in p.stdout


@pytest.mark.gpu
def test_facade_heston_example_and_long_bridges():
    """The reference's Heston example call sequence (Richardson c1, MC, MLMC, QMC
    Raw, MLQMC Raw) plus MLMC with the Milstein variance term, and a 2048-date
    GBM Asian with bridge and PCA noise, all within tolerance of fixed-step runs."""
    exe = os.path.join(PKG, "lib", "heston_table")
    p = subprocess.run([exe, "0.05"], capture_output=True, text=True, timeout=300)
    print(p.stdout)
    assert p.returncode == 0, p.stdout + p.stderr
    assert "MLMC Milstein" in p.stdout and "QMC PCA 2048" in p.stdout
    assert "facade rows outside tolerance: 0" in p.stdout
