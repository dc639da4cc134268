"""C ABI surface (CPU only): the library loads, exports every symbol that
include/fdmc.h declares, and rejects bad arguments on the host before any
device work (no GPU compute is issued by these tests)."""
import ctypes as C
import os
import re

import pytest

import fdmc

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "fdmc.h")


def declared_symbols():
    txt = open(HEADER).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(fdmc_[a-z0-9_]+)\s*\(", txt)))


def test_header_declares_reference_entry_points():
    syms = declared_symbols()
    for s in ("fdmc_run_mc", "fdmc_run_qmc", "fdmc_run_mlmc", "fdmc_run_mlqmc", "fdmc_bb_upload",
              "fdmc_pca_upload", "fdmc_cv_pilot", "fdmc_run_mc_cv", "fdmc_run_qmc_cv", "fdmc_run_mc_fixed",
              "fdmc_run_is"):
        assert s in syms


def test_library_exports_every_declared_symbol():
    lib = fdmc.load_library()
    missing = [s for s in declared_symbols() if not hasattr(lib, s)]
    assert not missing, missing
    assert lib.fdmc_abi_version() == 1


def test_ctypes_struct_layout_matches_header():
    # fdmc_model: 2 ints + 15 doubles + L[4] + 4 doubles + 2 ints + 3 pointers
    assert C.sizeof(fdmc.CModel) == 8 + 8 * (4 + 5 + 4 + 3 + 4) + 8 + 3 * 8
    assert C.sizeof(fdmc.CPayoff) == 8 + 5 * 8 + 8
    assert C.sizeof(fdmc.CResult) == 8 * 5 + 8
    assert C.sizeof(fdmc.CQMCConfig) == 32
    # contract book: kind, n_steps + 5 doubles; 8 ints; 8 doubles + 2 int64 + 4 ints; 2 doubles + int64 + 6 ints
    assert C.sizeof(fdmc.CBookContract) == 8 + 5 * 8
    assert C.sizeof(fdmc.CBookConfig) == 8 * 4
    assert C.sizeof(fdmc.CBookResult) == 8 * 8 + 2 * 8 + 4 * 4
    assert C.sizeof(fdmc.CBookStats) == 3 * 8 + 6 * 4


def test_null_context_is_rejected_on_host():
    lib = fdmc.load_library()
    m, p, c, r = fdmc.GBMParams().to_c(), fdmc.European(100, 0.05, 1).to_c(), fdmc.QMCConfig().to_c(), fdmc.CResult()
    rc = lib.fdmc_run_qmc(None, C.byref(m), C.byref(p), 1e-3, 64, C.byref(c), 0, None, None, C.byref(r))
    assert rc == 1
    assert b"null" in lib.fdmc_last_error()


def test_null_context_is_rejected_by_every_estimator():
    lib = fdmc.load_library()
    m, p, r = fdmc.GBMParams().to_c(), fdmc.European(180, 0.05, 1).to_c(), fdmc.CResult()
    mc, ml, qc = fdmc.MCConfig().to_c(), fdmc.MLMCConfig().to_c(), fdmc.QMCConfig().to_c()
    calls = [
        lambda: lib.fdmc_run_mc(None, C.byref(m), C.byref(p), 1e-2, 64, C.byref(mc), C.byref(r)),
        lambda: lib.fdmc_run_is(None, C.byref(m), C.byref(p), 2.8, 2e-3, C.byref(mc), C.byref(r)),
        lambda: lib.fdmc_run_mlmc(None, C.byref(m), C.byref(p), 1e-2, C.byref(ml), C.byref(r)),
        lambda: lib.fdmc_run_mlqmc(None, C.byref(m), C.byref(p), 1e-2, C.byref(ml), C.byref(qc), 2, None, 0,
                                   None, 0, C.byref(r)),
    ]
    for call in calls:
        lib.fdmc_last_error()
        assert call() == 1
        assert b"null" in lib.fdmc_last_error()
    lv, rp = (C.c_int32 * 1)(0), (C.c_int32 * 1)(0)
    fi, co = (C.c_int64 * 1)(0), (C.c_int64 * 1)(256)
    out = (C.c_double * 4)()
    assert lib.fdmc_levels_accumulate(None, C.byref(m), C.byref(p), 2, 1, 2, 1, 42, 1, lv, rp, fi, co, None, 0,
                                      None, 0, out) == 1


def test_context_create_without_device_reports_device_error():
    try:
        import torch
        if torch.cuda.is_available():
            pytest.skip("a GPU is present")
    except ImportError:
        pass
    lib = fdmc.load_library()
    h = C.c_void_p()
    rc = lib.fdmc_context_create(0, C.byref(h))
    assert rc == 4 and not h.value
    assert lib.fdmc_last_error()


def test_bb_upload_rejects_non_dyadic_tables():
    lib = fdmc.load_library()
    # null context is checked first; a bad table must also be rejected without a device
    bb = fdmc.bb_precompute(8, 1.0)
    N = 8
    ia = lambda v: (C.c_int32 * N)(*v)
    da = lambda v: (C.c_double * N)(*v)
    h = C.c_void_p()
    rc = lib.fdmc_bb_upload(None, N, 1.0, ia(bb.map_idx), ia(bb.left_idx), ia(bb.right_idx),
                            da(bb.weight_left), da(bb.weight_right), da(bb.std_dev), C.byref(h))
    assert rc == 1


def test_python_bb_precompute_matches_oracle():
    import numpy as np
    import pyoracle as O
    for N in (1, 2, 16, 256):
        a = fdmc.bb_precompute(N, 1.3)
        b = O.OracleBB(N, 1.3).arrays()
        assert list(b["map_idx"]) == a.map_idx and list(b["left_idx"]) == a.left_idx
        assert np.allclose(b["sd"], a.std_dev, rtol=0, atol=1e-15)
        assert np.allclose(b["wl"], a.weight_left, rtol=0, atol=1e-15)
