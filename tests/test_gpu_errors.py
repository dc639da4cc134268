"""Argument validation of the C ABI: every rejected input comes back as an
FdmcError with the documented status (1 invalid, 2 unsupported, 3 range) and
the reason, before any device work, and the context stays usable (the
reference throws std::runtime_error from the same checks, e.g.
methods_cuda.cu:834-837, 1695-1704, 1772-1789).
"""
import numpy as np
import pytest

import fdmc
import pyoracle as O

pytestmark = pytest.mark.gpu

GBM = fdmc.GBMParams(100.0, 0.05, 0.2, 1.0)
HES = fdmc.HestonParams(100.0, 0.05, 2.0, 0.04, 0.5, -0.7, 0.04, 1.0).compute_cholesky()
EURO = fdmc.European(100.0, 0.05, 1.0)
ASIAN = fdmc.Asian(100.0)
INVALID, UNSUPPORTED, RANGE = 1, 2, 3


def _raises(code, fn, *args, **kw):
    with pytest.raises(fdmc.FdmcError) as e:
        fn(*args, **kw)
    assert e.value.code == code, str(e.value)
    return str(e.value)


def test_estimator_arguments(engine):
    _raises(INVALID, engine.run_mc_cuda, GBM, EURO, 0.0, 64, fdmc.MCConfig())
    _raises(INVALID, engine.run_mc_cuda, GBM, EURO, 1e-2, 64, fdmc.MCConfig(pilot_n=1))
    _raises(INVALID, engine.run_mc_cuda, GBM, EURO, 1e-2, 0, fdmc.MCConfig())
    _raises(INVALID, engine.run_qmc_cuda, GBM, EURO, 1e-3, 64, fdmc.QMCConfig(N0=1000))
    _raises(INVALID, engine.run_qmc_cuda, GBM, EURO, 1e-3, 64, fdmc.QMCConfig(R=0))
    _raises(INVALID, engine.run_mlmc_cuda, GBM, EURO, 1e-2, fdmc.MLMCConfig(M=1))
    _raises(INVALID, engine.run_mlmc_cuda, GBM, EURO, 1e-2, fdmc.MLMCConfig(max_L=40))
    _raises(UNSUPPORTED, engine.run_is_cuda, HES, EURO, 1.0, 1e-2)
    bad = fdmc.HestonParams(scheme=2).compute_cholesky()
    assert "variance scheme" in _raises(INVALID, engine.run_mlmc_cuda, bad, EURO, 1e-2, fdmc.MLMCConfig(M=4))
    _raises(UNSUPPORTED, engine.run_is_cuda, GBM, ASIAN, 1.0, 1e-2)


def test_pass_arguments(engine):
    _raises(INVALID, engine.qmc_accumulate, GBM, EURO, 64, 2, 42, 0, 0, fdmc.NOISE_RAW)      # no points
    _raises(INVALID, engine.qmc_accumulate, GBM, EURO, 64, 0, 42, 0, 256, fdmc.NOISE_RAW)    # no replicas
    _raises(INVALID, engine.qmc_accumulate, GBM, EURO, 64, 2, 42, -1, 256, fdmc.NOISE_RAW)
    msg = _raises(RANGE, engine.qmc_accumulate, GBM, EURO, 64, 2, 42, (1 << 32) - 100, 256, fdmc.NOISE_RAW)
    assert "32 bits" in msg
    basket = fdmc.MultiGBMParams(4, [100.0] * 4, [0.2] * 4, 0.05, 1.0, uncorrelated=True)
    _raises(RANGE, engine.qmc_accumulate, basket, fdmc.Basket(100.0, 0.05, 1.0, 4), 16, 1, 42, (1 << 32) - 10,
            256, fdmc.NOISE_RAW)
    wide = fdmc.MultiDupireParams(n=100, uncorrelated=True)
    _raises(RANGE, engine.qmc_accumulate, wide, fdmc.Basket(100.0, 0.05, 1.0, 100), 8, 1, 42, (1 << 32) - 10,
            256, fdmc.NOISE_RAW)
    # Brownian bridge / PCA need a one-factor model and matching sizes
    bb = engine.bb_upload(fdmc.bb_precompute(32, 1.0))
    _raises(UNSUPPORTED, engine.qmc_accumulate, HES, EURO, 32, 1, 42, 0, 256, fdmc.NOISE_BB, dev_bb=bb)
    _raises(INVALID, engine.qmc_accumulate, GBM, EURO, 64, 1, 42, 0, 256, fdmc.NOISE_BB, dev_bb=bb)
    bb.free()
    bb = engine.bb_upload(fdmc.bb_precompute(2048, 1.0))
    msg = _raises(UNSUPPORTED, engine.greeks_accumulate, GBM, EURO, 2048, 1, 42, 0, 256, fdmc.NOISE_BB, dev_bb=bb)
    assert "1024" in msg
    bb.free()
    pca = engine.pca_upload(16, O.pca_compute(16, 1.0).T.ravel())
    _raises(INVALID, engine.qmc_accumulate, GBM, EURO, 32, 1, 42, 0, 256, fdmc.NOISE_PCA, dev_pca=pca)
    pca.free()


def test_multilevel_lists(engine):
    ml, qc = fdmc.MLMCConfig(M=2, max_L=4, pilot_n=256), fdmc.QMCConfig(R=4)
    bbl = [engine.bb_upload(fdmc.bb_precompute(2 ** l, 1.0)) for l in range(2)]       # levels 0, 1 only
    msg = _raises(INVALID, engine.run_mlqmc_cuda, GBM, ASIAN, 1e-2, ml, qc, fdmc.NOISE_BB, bbl)
    assert "bb_list[2]" in msg
    for b in bbl:
        b.free()
    pcl = [engine.pca_upload(2 ** l, O.pca_compute(2 ** l, 1.0).T.ravel()) for l in range(3)]
    pcl[1], wrong = engine.pca_upload(4, O.pca_compute(4, 1.0).T.ravel()), pcl[1]    # level 1 with m = 4
    msg = _raises(INVALID, engine.run_mlqmc_cuda, GBM, ASIAN, 1e-2, ml, qc, fdmc.NOISE_PCA, pca_list=pcl)
    assert "pca_list[1]" in msg
    for x in pcl + [wrong]:
        x.free()
    _raises(UNSUPPORTED, engine.run_mlqmc_cuda, HES, EURO, 1e-2, ml, qc, fdmc.NOISE_BB, [])


def test_context_usable_after_errors(engine):
    for _ in range(3):
        _raises(INVALID, engine.qmc_accumulate, GBM, EURO, 64, 2, 42, 0, 0, fdmc.NOISE_RAW)
    got = np.array(engine.qmc_accumulate(GBM, EURO, 64, 2, 42, 0, 1000, fdmc.NOISE_RAW))
    ref = O.qmc_accumulate(GBM, EURO, 64, 2, 42, 0, 1000, fdmc.NOISE_RAW)
    assert np.allclose(got[:, 0] / 1000, ref[:, 0] / 1000, rtol=2e-5)
