"""Baskets of more than 32 assets (reference examples 07 and 08): the
uncorrelated 1000-asset local-vol basket at the Sobol dimension limit
(20 dates x 1000 assets = 20000) and a 100-asset Cholesky-correlated basket,
against the oracle's restatement of kernel_multi_dupire (methods_cuda.cu:526-551)
at the same Sobol points / Philox paths, and the reference's report price.
"""
import json
import math
import os

import numpy as np
import pytest

import fdmc
import pyoracle as O

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))


def _corr_100(n=100, seed=0):
    # ejemplo08_basket_corr.cpp:29-47: C = A A^T / n + 0.01 I, scaled to unit diagonal
    A = np.random.default_rng(seed).standard_normal((n, n))
    Cm = A @ A.T / n + 0.01 * np.eye(n)
    d = np.sqrt(np.diag(Cm))
    return Cm / np.outer(d, d)


UNCORR = fdmc.MultiDupireParams(n=1000, uncorrelated=True)
PAY1000 = fdmc.Basket(100.0, 0.05, 1.0, 1000)
CORR100 = fdmc.MultiDupireParams(n=100, L=np.linalg.cholesky(_corr_100()).ravel())
PAY100 = fdmc.Basket(100.0, 0.05, 1.0, 100)
GBM64 = fdmc.MultiGBMParams(64, [90.0 + a * 0.3 for a in range(64)], [0.15 + 0.003 * a for a in range(64)], 0.05, 1.0,
                            L=np.linalg.cholesky(_corr_100(64, 3)).ravel())
PAY64 = fdmc.Basket(100.0, 0.05, 1.0, 64, weights=[1.0 / 64 + (a - 31.5) * 1e-4 for a in range(64)])


@pytest.mark.parametrize("model,payoff,n_steps,npts", [
    (UNCORR, PAY1000, 20, 300),
    (CORR100, PAY100, 16, 512),
    (GBM64, PAY64, 12, 512),
], ids=["uncorr1000-D20000", "corr100", "gbm64-corr-weights"])
def test_wide_basket_qmc_matches_oracle(engine, model, payoff, n_steps, npts):
    R = 2
    got = np.array(engine.qmc_accumulate(model, payoff, n_steps, R, 42, 100, npts, fdmc.NOISE_RAW))
    ref = O.qmc_accumulate(model, payoff, n_steps, R, 42, 100, npts, fdmc.NOISE_RAW)
    assert np.allclose(got[:, 0] / npts, ref[:, 0] / npts, rtol=2e-5, atol=1e-6), (got[:, 0], ref[:, 0])
    assert np.allclose(got[:, 1] / npts, ref[:, 1] / npts, rtol=5e-5, atol=1e-6)


UNCORR37 = fdmc.MultiDupireParams(n=37, uncorrelated=True, S0=[100.0 + 0.5 * (a % 5) for a in range(37)])
PAYU37 = fdmc.Basket(101.0, 0.05, 1.0, 37, weights=[1.0 / 37 + (a - 18) * 2e-4 for a in range(37)])


@pytest.mark.parametrize("model,payoff,n_steps", [(UNCORR, PAY1000, 10), (CORR100, PAY100, 24),
                                                  (UNCORR37, PAYU37, 7)],
                         ids=["uncorr1000", "corr100", "uncorr37"])
def test_wide_basket_prng_matches_oracle(engine, model, payoff, n_steps):
    gm, gv = engine.run_mc_fixed(model, payoff, n_steps, 1000, 99)
    rm, rv = O.run_mc_fixed(model, payoff, n_steps, 1000, 99)
    assert abs(gm - rm) <= 2e-5 * abs(rm) + 1e-6
    assert abs(gv - rv) <= 1e-3 * rv


# shapes of the MFMA correlation kernel's edges: asset counts that are not a
# multiple of 4 (Philox groups of four normals straddle asset blocks and dates)
# or of 16 (a partial last block), offsets inside a 64-path tile, and the
# kernel's compilations for ceil(n / 16) blocks: 3 (37), 8 (113), 10 (150),
# 13 (201) and the largest, 16 (256 assets, full blocks)
GBM37 = fdmc.MultiGBMParams(37, [95.0 + a * 0.25 for a in range(37)], [0.12 + 0.005 * a for a in range(37)], 0.04, 1.0,
                            L=np.linalg.cholesky(_corr_100(37, 5)).ravel())
PAY37 = fdmc.Basket(101.0, 0.04, 1.0, 37, weights=[1.0 / 37 + (a - 18) * 2e-4 for a in range(37)])
DUP150 = fdmc.MultiDupireParams(n=150, L=np.linalg.cholesky(_corr_100(150, 7)).ravel(),
                                S0=[100.0 + 0.1 * (a % 7) for a in range(150)])
PAY150 = fdmc.Basket(100.0, 0.05, 1.0, 150)
GBM113 = fdmc.MultiGBMParams(113, [100.0] * 113, [0.2 + 0.001 * a for a in range(113)], 0.05, 1.0,
                             L=np.linalg.cholesky(_corr_100(113, 9)).ravel())
PAY113 = fdmc.Basket(100.0, 0.05, 1.0, 113)
GBM201 = fdmc.MultiGBMParams(201, [100.0 + 0.02 * a for a in range(201)], [0.15 + 0.0005 * a for a in range(201)],
                             0.05, 1.0, L=np.linalg.cholesky(_corr_100(201, 11)).ravel())
PAY201 = fdmc.Basket(100.0, 0.05, 1.0, 201)
GBM256 = fdmc.MultiGBMParams(256, [100.0] * 256, [0.18 + 0.0004 * a for a in range(256)], 0.03, 1.0,
                             L=np.linalg.cholesky(_corr_100(256, 13)).ravel())
PAY256 = fdmc.Basket(99.0, 0.03, 1.0, 256)


@pytest.mark.parametrize("model,payoff,n_steps,first,npts", [
    (GBM37, PAY37, 7, 37, 300),
    (GBM113, PAY113, 9, 64, 128),
    (DUP150, PAY150, 5, 5, 200),
    (GBM201, PAY201, 3, 17, 150),
    (GBM256, PAY256, 2, 0, 128),
], ids=["gbm37", "gbm113", "dupire150", "gbm201", "gbm256"])
def test_corr_basket_edge_shapes_qmc(engine, model, payoff, n_steps, first, npts):
    R = 2
    print(model.n, "assets,", n_steps, "dates, points", first, "+", npts, flush=True)
    got = np.array(engine.qmc_accumulate(model, payoff, n_steps, R, 42, first, npts, fdmc.NOISE_RAW))
    ref = O.qmc_accumulate(model, payoff, n_steps, R, 42, first, npts, fdmc.NOISE_RAW)
    assert np.allclose(got[:, 0] / npts, ref[:, 0] / npts, rtol=2e-5, atol=1e-6), (got[:, 0], ref[:, 0])
    assert np.allclose(got[:, 1] / npts, ref[:, 1] / npts, rtol=5e-5, atol=1e-6)


@pytest.mark.parametrize("model,payoff,n_steps", [(GBM37, PAY37, 6), (GBM113, PAY113, 3), (DUP150, PAY150, 4),
                                                   (GBM201, PAY201, 3), (GBM256, PAY256, 2)],
                         ids=["gbm37", "gbm113", "dupire150", "gbm201", "gbm256"])
def test_corr_basket_edge_shapes_prng(engine, model, payoff, n_steps):
    print(model.n, "assets,", n_steps, "dates, Philox", flush=True)
    gm, gv = engine.run_mc_fixed(model, payoff, n_steps, 1000, 7)
    rm, rv = O.run_mc_fixed(model, payoff, n_steps, 1000, 7)
    assert abs(gm - rm) <= 2e-5 * abs(rm) + 1e-6, (gm, rm)
    assert abs(gv - rv) <= 1e-3 * rv


def test_ejemplo07_qmc_and_mc_vs_report(engine):
    """Config of examples/ejemplo07_basket_uncorr.cpp: run_qmc_cuda Raw and
    run_mc_cuda at eps = 0.01 with 20 dates against the report's 5000-path
    reference price (its own standard error is ~0.007)."""
    rep = [e for e in json.load(open(os.path.join(HERE, "golden", "report_prices.json")))
           if e["example"] == "ejemplo07_basket_uncorr"][0]
    q = engine.run_qmc_cuda(UNCORR, PAY1000, 0.01, 20, fdmc.QMCConfig(R=16, N0=1 << 10))
    assert q.flags == 0
    assert abs(q.price - rep["price"]) < 3 * math.hypot(q.std_error, 0.0072) + 0.005, (q.price, q.std_error)
    m = engine.run_mc_cuda(UNCORR, PAY1000, 0.01, 20, fdmc.MCConfig())
    assert abs(m.price - rep["price"]) < 3 * math.hypot(m.std_error, 0.0072) + 0.005, (m.price, m.std_error)
    assert abs(q.price - m.price) < 3 * math.hypot(q.std_error, m.std_error) + 1e-3


def test_wide_basket_rejections(engine):
    with pytest.raises(fdmc.FdmcError):
        engine.greeks_accumulate(UNCORR, PAY1000, 4, 1, 42, 0, 256, fdmc.NOISE_RAW)
    big = fdmc.MultiDupireParams(n=300, L=np.eye(300).ravel())
    with pytest.raises(fdmc.FdmcError):
        engine.qmc_accumulate(big, fdmc.Basket(100.0, 0.05, 1.0, 300), 4, 1, 42, 0, 256, fdmc.NOISE_RAW)
