"""GPU parity of the fused path kernels against the oracle at the same noise.

Each case runs one pass through fdmc_qmc_accumulate (QMC) or fdmc_run_mc_fixed
(Philox) on the GPU and the oracle's restatement of kernel_mc / the CV kernels
(methods_cuda.cu:414-437, 603-677) on the CPU, with identical scrambled Sobol
points or Philox streams.  The GPU evaluates its own FP32 inverse CDF and
hardware exp/log; tolerances below are written per payoff:
 - Lipschitz payoffs: replica means agree to 2e-5 relative (+1e-6 absolute);
 - barrier: a knock-out indicator can flip for paths within FP32 rounding of
   the barrier, so the mean may differ by a few max-payoff/N.
"""
import math

import numpy as np
import pytest

import fdmc
import pyoracle as O

pytestmark = pytest.mark.gpu

GBM = fdmc.GBMParams(100.0, 0.05, 0.2, 1.0)
GBM3 = fdmc.GBMParams(100.0, 0.05, 0.3, 1.0)
HES = fdmc.HestonParams(100.0, 0.05, 2.0, 0.04, 0.5, -0.7, 0.04, 1.0).compute_cholesky()
HES_MIL = fdmc.HestonParams(100.0, 0.05, 2.0, 0.04, 0.5, -0.7, 0.04, 1.0,
                            scheme=fdmc.HESTON_MILSTEIN).compute_cholesky()
DUP = fdmc.DupireLocalParams(100.0, 0.05, 0.2, 0.5, 0.7, 1.0)
LVQ = fdmc.LocalVolQuadParams(100.0, 0.05, 0.2, 0.1, 1.0, 100.0, 1.0)
# vol level 1: about 9 % of the four-date Euler paths step to S <= 0, where the
# local vol's log is NaN; such a path pays 0 under every payoff
LVQ_STEEP = fdmc.LocalVolQuadParams(100.0, 0.05, 1.0, 0.05, 1.0, 100.0, 1.0)

EURO = fdmc.European(100.0, 0.05, 1.0)
ASIAN = fdmc.Asian(100.0)
GEO = fdmc.GeomAsian(100.0)
LOOK = fdmc.Lookback(0.2)
BAR = fdmc.Barrier(100.0, 130.0, 0.2, 0.05, 1.0)


def _close(got, ref, n, payoff, what):
    got, ref = np.asarray(got, float), np.asarray(ref, float)
    tol = 2e-5 * np.maximum(1.0, np.abs(ref)) + 1e-6
    if isinstance(payoff, fdmc.Barrier):
        tol = tol + 4.0 * 40.0 / n
    bad = np.abs(got - ref) > tol
    assert not bad.any(), (what, got[bad][:4], ref[bad][:4])


QMC_CASES = [
    ("gbm-euro-raw", GBM, EURO, 64, fdmc.NOISE_RAW),
    ("gbm-asian-bb", GBM3, ASIAN, 256, fdmc.NOISE_BB),
    ("gbm-geo-bb8", GBM, GEO, 8, fdmc.NOISE_BB),
    ("gbm-look-bb", GBM, LOOK, 32, fdmc.NOISE_BB),
    ("gbm-bar-raw", GBM, BAR, 64, fdmc.NOISE_RAW),
    ("heston-euro-raw", HES, EURO, 32, fdmc.NOISE_RAW),
    ("heston-bar-raw", HES, BAR, 32, fdmc.NOISE_RAW),
    ("heston-milstein-asian-raw", HES_MIL, ASIAN, 64, fdmc.NOISE_RAW),
    ("dupire-euro-bb", DUP, EURO, 64, fdmc.NOISE_BB),
    ("dupire-look-raw", DUP, LOOK, 16, fdmc.NOISE_RAW),
    ("lvq-asian-bb", LVQ, ASIAN, 128, fdmc.NOISE_BB),
    ("lvq-bar-bb", LVQ, BAR, 64, fdmc.NOISE_BB),
    ("lvq-steep-look-raw", LVQ_STEEP, LOOK, 4, fdmc.NOISE_RAW),
    ("lvq-steep-look-bb", LVQ_STEEP, LOOK, 8, fdmc.NOISE_BB),
    ("lvq-steep-euro-bb", LVQ_STEEP, EURO, 4, fdmc.NOISE_BB),
]


@pytest.mark.parametrize("name,model,payoff,n_steps,mode", QMC_CASES, ids=[c[0] for c in QMC_CASES])
def test_qmc_pass_parity(engine, name, model, payoff, n_steps, mode):
    R, off, n = 3, 300, 1500    # unaligned first point and count
    bb = engine.bb_upload(fdmc.bb_precompute(n_steps, model.T)) if mode == fdmc.NOISE_BB else None
    got = np.array(engine.qmc_accumulate(model, payoff, n_steps, R, 42, off, n, mode, dev_bb=bb))
    if bb:
        bb.free()
    obb = O.OracleBB(n_steps, model.T) if mode == fdmc.NOISE_BB else None
    ref = O.qmc_accumulate(model, payoff, n_steps, R, 42, off, n, mode, bb=obb)
    _close(got[:, 0] / n, ref[:, 0] / n, n, payoff, name + " mean")
    # second moments: relative check
    assert np.allclose(got[:, 1] / n, ref[:, 1] / n, rtol=1e-4, atol=1e-4 + (1e4 * 4 / n if isinstance(payoff, fdmc.Barrier) else 0))


@pytest.mark.parametrize("model,payoff,n_steps", [(GBM3, ASIAN, 2048), (LVQ, LOOK, 2048), (GBM, BAR, 4096)],
                         ids=["gbm-asian-2048", "lvq-look-2048", "gbm-bar-4096"])
def test_qmc_bridge_above_1024_dates(engine, model, payoff, n_steps):
    """Bridges of more than 1024 dates (the reference's examples ask for up to
    2^11 with BB) read their direction records from global memory instead of
    LDS: the same points and paths as the oracle."""
    R, off, n = 2, 300, 700
    bb = engine.bb_upload(fdmc.bb_precompute(n_steps, model.T))
    got = np.array(engine.qmc_accumulate(model, payoff, n_steps, R, 42, off, n, fdmc.NOISE_BB, dev_bb=bb))
    bb.free()
    ref = O.qmc_accumulate(model, payoff, n_steps, R, 42, off, n, fdmc.NOISE_BB, bb=O.OracleBB(n_steps, model.T))
    _close(got[:, 0] / n, ref[:, 0] / n, n, payoff, "bridge %d mean" % n_steps)
    assert np.allclose(got[:, 1] / n, ref[:, 1] / n, rtol=1e-4,
                       atol=1e-4 + (1e4 * 4 / n if isinstance(payoff, fdmc.Barrier) else 0))


def test_run_qmc_bridge_2048_matches_oracle_loop(engine):
    cfg = fdmc.QMCConfig(R=8, N0=1 << 10, max_doublings=2)
    bb = engine.bb_upload(fdmc.bb_precompute(2048, 1.0))
    g = engine.run_qmc_cuda(GBM, ASIAN, 1e-5, 2048, cfg, fdmc.NOISE_BB, dev_bb=bb)
    bb.free()
    r = O.run_qmc(GBM, ASIAN, 1e-5, 2048, cfg, fdmc.NOISE_BB, bb=O.OracleBB(2048, 1.0))
    assert g.n_samples == r.n_samples == 8 * (1 << 11)
    assert abs(g.price - r.price) < 1e-4
    assert abs(g.std_error - r.std_error) < 0.05 * r.std_error


def test_qmc_cv_pass_parity(engine):
    """C2 kernel: arithmetic Asian + geometric Asian control on the same bridge path."""
    n_steps, R, n = 256, 4, 2048
    E = fdmc.geom_asian_analytic(100.0, 100.0, 1.0, 0.05, 0.3, n_steps)
    bb = engine.bb_upload(fdmc.bb_precompute(n_steps, 1.0))
    got = np.array(engine.qmc_accumulate(GBM3, ASIAN, n_steps, R, 42, 0, n, fdmc.NOISE_BB, dev_bb=bb,
                                         ctrl_model=GBM3, ctrl_payoff=GEO, E_ctrl=E, beta=0.9))
    bb.free()
    ref = O.qmc_accumulate(GBM3, ASIAN, n_steps, R, 42, 0, n, fdmc.NOISE_BB, bb=O.OracleBB(n_steps, 1.0),
                           ctrl_model=GBM3, ctrl_payoff=GEO, E_ctrl=E, beta=0.9)
    for i in (0, 2, 4):
        _close(got[:, i] / n, ref[:, i] / n, n, ASIAN, "acc%d" % i)
    assert np.allclose(got[:, 6] / n, ref[:, 6] / n, rtol=1e-4)


def test_dupire_gbm_cv_parity(engine):
    """kernel_dupire_gbm_cv analogue: Dupire European with a GBM(sigma0) control leg."""
    ctrl = fdmc.GBMParams(100.0, 0.05, 0.2, 1.0)
    E = fdmc.bs_call(100.0, 100.0, 1.0, 0.05, 0.2)
    n_steps, R, n = 32, 2, 1024
    got = np.array(engine.qmc_accumulate(DUP, EURO, n_steps, R, 7, 0, n, fdmc.NOISE_RAW,
                                         ctrl_model=ctrl, ctrl_payoff=EURO, E_ctrl=E, beta=1.0))
    ref = O.qmc_accumulate(DUP, EURO, n_steps, R, 7, 0, n, fdmc.NOISE_RAW, ctrl_model=ctrl, ctrl_payoff=EURO,
                           E_ctrl=E, beta=1.0)
    for i in (0, 2, 4):
        _close(got[:, i] / n, ref[:, i] / n, n, EURO, "acc%d" % i)


def test_prng_fallback_above_20000_dims(engine):
    """D > 20000 uses Philox points, as the reference falls back to XORWOW (methods_cuda.cu:1521)."""
    n_steps, R, n = 10001, 2, 256   # Heston: D = 20002
    got = np.array(engine.qmc_accumulate(HES, EURO, n_steps, R, 42, 0, n, fdmc.NOISE_RAW))
    ref = O.qmc_accumulate(HES, EURO, n_steps, R, 42, 0, n, fdmc.NOISE_RAW)
    _close(got[:, 0] / n, ref[:, 0] / n, n, EURO, "fallback")


@pytest.mark.parametrize("model,payoff,n_steps", [(GBM, EURO, 64), (HES, BAR, 32), (HES_MIL, EURO, 32),
                                                  (DUP, ASIAN, 16), (LVQ, LOOK, 24)])
def test_mc_fixed_parity(engine, model, payoff, n_steps):
    n = 5000
    gm, gv = engine.run_mc_fixed(model, payoff, n_steps, n, 99)
    rm, rv = O.run_mc_fixed(model, payoff, n_steps, n, 99)
    _close([gm], [rm], n, payoff, "mc_fixed mean")
    assert abs(gv - rv) <= 1e-3 * rv + (1e-3 if isinstance(payoff, fdmc.Barrier) else 0.0)


def test_run_qmc_matches_oracle_loop(engine):
    cfg = fdmc.QMCConfig(R=8, N0=1 << 10, max_doublings=4)
    bb = engine.bb_upload(fdmc.bb_precompute(64, 1.0))
    g = engine.run_qmc_cuda(GBM, ASIAN, 1e-4, 64, cfg, fdmc.NOISE_BB, dev_bb=bb)
    bb.free()
    r = O.run_qmc(GBM, ASIAN, 1e-4, 64, cfg, fdmc.NOISE_BB, bb=O.OracleBB(64, 1.0))
    assert g.n_samples == r.n_samples == 8 * (1 << 13)
    assert g.levels == r.levels == 4
    assert abs(g.price - r.price) < 1e-4
    assert abs(g.std_error - r.std_error) < 0.05 * r.std_error


@pytest.mark.parametrize("eps", [2e-3, 3e-4])
def test_run_is_matches_oracle_and_black_scholes(engine, eps):
    """run_is_cuda (methods_cuda.cu:2317-2365) on ejemplo11's deep-OTM call
    (K = 180, z* = 2.79): the same Philox paths as the oracle's restatement of
    kernel_is_gbm, and Black-Scholes within 3 standard errors."""
    g, e = fdmc.GBMParams(100.0, 0.05, 0.2, 1.0), fdmc.European(180.0, 0.05, 1.0)
    zs = (math.log(1.8) - (0.05 - 0.02)) / 0.2
    r = engine.run_is_cuda(g, e, zs, eps, fdmc.MCConfig())
    o = O.run_is(g, e, zs, eps, fdmc.MCConfig())
    # N = ceil(2 Var / eps^2) of the pilot rounds up to even: one step of rounding apart at most
    assert abs(r.n_samples - o.n_samples) <= 2
    if r.n_samples == o.n_samples:
        assert abs(r.price - o.price) <= 2e-5 * o.price, (r.price, o.price)
        assert abs(r.std_error - o.std_error) <= 1e-4 * o.std_error
    bs = fdmc.bs_call(100, 180, 1, 0.05, 0.2)
    assert abs(r.price - bs) < 3 * r.std_error + 1e-4
    assert r.std_error < eps / math.sqrt(2) * 1.08


def test_c1_mc_european_vs_black_scholes(engine):
    """Config 1 at eps = 1e-2: within 3 standard errors of BS after the Euler bias."""
    r = engine.run_mc_cuda(GBM, EURO, 1e-2, 64, fdmc.MCConfig())
    bs = fdmc.bs_call(100, 100, 1, 0.05, 0.2)
    assert r.n_samples > 3_000_000
    assert abs(r.price - (bs - 0.002)) < 3 * r.std_error + 2e-4, (r.price, r.std_error)
    assert r.std_error < 1e-2 / math.sqrt(2) * 1.05


def test_c2_rqmc_asian_cv_within_eps(engine):
    """Config 2 price: RQMC (R=32, N0=2^14) Asian with BB and geometric CV reaches eps=1e-3."""
    n = 256
    E = fdmc.geom_asian_analytic(100, 100, 1, 0.05, 0.3, n)
    cfg = fdmc.QMCConfig(R=32, N0=1 << 14, N_max=1 << 20)
    bb = engine.bb_upload(fdmc.bb_precompute(n, 1.0))
    r = engine.run_qmc_cv_cuda(GBM3, GBM3, ASIAN, GEO, E, 1.0, 1e-3, n, cfg, fdmc.NOISE_BB, dev_bb=bb)
    bb.free()
    assert r.flags == 0 and r.std_error < 1e-3 / math.sqrt(2)
    # the oracle's doubling loop on the same points, capped at the doublings the
    # device run used (so the CPU work stays bounded): same stop, same price
    ref = O.run_qmc_cv(GBM3, GBM3, ASIAN, GEO, E, 1.0, 1e-3, n,
                       fdmc.QMCConfig(R=32, N0=1 << 14, N_max=1 << 20, max_doublings=r.levels), fdmc.NOISE_BB,
                       bb=O.OracleBB(n, 1.0))
    assert ref.flags == 0 and ref.levels == r.levels and ref.n_samples == r.n_samples
    assert abs(r.price - ref.price) < 1e-4
    assert abs(r.std_error - ref.std_error) < 0.05 * ref.std_error
