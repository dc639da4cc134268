"""GPU checks of the fused price + pathwise-adjoint Greeks kernel.

Oracles:
 - common-random-number central differences of the GPU price (same Sobol points
   or Philox streams, parameters bumped up and down);
 - bump-and-reprice with the CPU oracle (the reference's own approach to
   sensitivities, run through its path restatement);
 - Black-Scholes delta/vega/rho and the geometric-Asian closed form's
   derivatives (Euler bias at 256 steps is below the stated tolerance).
The pathwise estimator equals the derivative of the CRN estimator except for
paths whose payoff kink lies between the bumped runs, so they agree to O(bump).
"""
import math

import numpy as np
import pytest
from scipy.stats import norm

import fdmc
import pyoracle as O

pytestmark = pytest.mark.gpu

EURO = fdmc.European(100.0, 0.05, 1.0)
ASIAN, GEO, LOOK = fdmc.Asian(100.0), fdmc.GeomAsian(100.0), fdmc.Lookback(0.2)


def gbm(S0=100.0, mu=0.05, sig=0.2):
    return fdmc.GBMParams(S0, mu, sig, 1.0)


def dup(S0=100.0, mu=0.05, s0=0.2):
    return fdmc.DupireLocalParams(S0, mu, s0, 0.5, 0.7, 1.0)


def lvq(S0=100.0, mu=0.05, a=0.2):
    return fdmc.LocalVolQuadParams(S0, mu, a, 0.1, 1.0, 100.0, 1.0)


def _bumped(make, payoff, which, h):
    """(model, payoff) with parameter `which` moved by h; rho moves mu and r together."""
    if which == "delta":
        return make(S0=100.0 + h), payoff
    if which == "vega":
        key = {gbm: "sig", dup: "s0", lvq: "a"}[make]
        return make(**{key: {gbm: 0.2, dup: 0.2, lvq: 0.2}[make] + h}), payoff
    p = payoff
    if isinstance(payoff, fdmc.European):
        p = fdmc.European(payoff.K, payoff.r + h, payoff.T)
    return make(mu=0.05 + h), p


def _price_means(engine, model, payoff, n_steps, R, n, mode, bb, prng=False):
    if prng:
        rows = [engine.run_mc_fixed(model, payoff, n_steps, n, 42 + r)[0] for r in range(R)]
        return np.array(rows)
    return np.array(engine.qmc_accumulate(model, payoff, n_steps, R, 42, 0, n, mode, dev_bb=bb))[:, 0] / n


CASES = [
    ("gbm-euro-bb", gbm, EURO, 64, fdmc.NOISE_BB),
    ("gbm-asian-bb", gbm, ASIAN, 64, fdmc.NOISE_BB),
    ("gbm-geo-raw", gbm, GEO, 32, fdmc.NOISE_RAW),
    ("gbm-look-bb", gbm, LOOK, 32, fdmc.NOISE_BB),
    ("dupire-euro-bb", dup, EURO, 32, fdmc.NOISE_BB),
    ("lvq-asian-raw", lvq, ASIAN, 32, fdmc.NOISE_RAW),
    ("gbm-euro-prng", gbm, EURO, 32, fdmc.NOISE_PRNG),
]
BUMP = {"delta": 0.05, "vega": 0.002, "rho": 0.002}


@pytest.mark.parametrize("name,make,payoff,n_steps,mode", CASES, ids=[c[0] for c in CASES])
def test_adjoint_matches_crn_central_differences(engine, name, make, payoff, n_steps, mode):
    R, n = 2, 1 << 14
    bb = engine.bb_upload(fdmc.bb_precompute(n_steps, 1.0)) if mode == fdmc.NOISE_BB else None
    prng = mode == fdmc.NOISE_PRNG
    rows = np.array(engine.greeks_accumulate(make(), payoff, n_steps, R, 42, 0, n, mode, dev_bb=bb))
    # price part equals the plain kernel's
    base = _price_means(engine, make(), payoff, n_steps, R, n, mode, bb, prng)
    assert np.allclose(rows[:, 0] / n, base, rtol=1e-6, atol=1e-7)
    for k, which in enumerate(("delta", "vega", "rho")):
        adj = rows[:, 2 + 2 * k] / n
        h = BUMP[which]
        up = _price_means(engine, *_bumped(make, payoff, which, +h), n_steps, R, n, mode, bb, prng)
        dn = _price_means(engine, *_bumped(make, payoff, which, -h), n_steps, R, n, mode, bb, prng)
        fd = (up - dn) / (2 * h)
        assert np.all(np.abs(adj - fd) <= 0.01 * np.abs(fd) + 2e-3), (which, adj, fd)
    if bb:
        bb.free()


def test_adjoint_of_paths_leaving_the_domain_is_finite(engine):
    """Vol level 1 over four dates: about 9 % of the Euler paths step to S <= 0,
    where the local vol's log is NaN.  They pay 0 (the lookback too) and add 0
    to the Greeks, so every sum is finite and the price part equals the plain
    kernel's."""
    m = fdmc.LocalVolQuadParams(100.0, 0.05, 1.0, 0.05, 1.0, 100.0, 1.0)
    for payoff in (LOOK, EURO, ASIAN):
        R, n = 2, 1 << 14
        rows = np.array(engine.greeks_accumulate(m, payoff, 4, R, 42, 0, n, fdmc.NOISE_RAW))
        assert np.all(np.isfinite(rows)), (payoff, rows)
        base = _price_means(engine, m, payoff, 4, R, n, fdmc.NOISE_RAW, None)
        assert np.allclose(rows[:, 0] / n, base, rtol=1e-6, atol=1e-7)
        assert np.all(rows[:, 0] > 0)


def test_adjoint_matches_oracle_bump_and_reprice(engine):
    """Bump-and-reprice through the CPU restatement (same Sobol points, FP64 ICDF)."""
    R, n, n_steps = 2, 4096, 32
    bb = engine.bb_upload(fdmc.bb_precompute(n_steps, 1.0))
    rows = np.array(engine.greeks_accumulate(gbm(), ASIAN, n_steps, R, 42, 0, n, fdmc.NOISE_BB, dev_bb=bb))
    bb.free()
    obb = O.OracleBB(n_steps, 1.0)
    for k, which in enumerate(("delta", "vega", "rho")):
        h = BUMP[which]
        mu_ = O.qmc_accumulate(*_bumped(gbm, ASIAN, which, +h), n_steps, R, 42, 0, n, fdmc.NOISE_BB, bb=obb)[:, 0] / n
        md_ = O.qmc_accumulate(*_bumped(gbm, ASIAN, which, -h), n_steps, R, 42, 0, n, fdmc.NOISE_BB, bb=obb)[:, 0] / n
        fd = (mu_ - md_) / (2 * h)
        adj = rows[:, 2 + 2 * k] / n
        assert np.all(np.abs(adj - fd) <= 0.01 * np.abs(fd) + 2e-3), (which, adj, fd)


def test_european_greeks_vs_black_scholes(engine):
    n_steps = 256
    bb = engine.bb_upload(fdmc.bb_precompute(n_steps, 1.0))
    r, g = engine.run_qmc_greeks(gbm(), EURO, 2e-3, n_steps, fdmc.QMCConfig(R=16, N0=1 << 14), fdmc.NOISE_BB, bb)
    bb.free()
    d1 = (math.log(1.0) + (0.05 + 0.02) * 1.0) / 0.2
    delta, vega = norm.cdf(d1), 100 * norm.pdf(d1)
    rho = 100 * math.exp(-0.05) * norm.cdf(d1 - 0.2)
    assert abs(g.delta - delta) < 3 * g.delta_se + 2e-3 * delta, (g, delta)
    assert abs(g.vega - vega) < 3 * g.vega_se + 5e-3 * vega, (g, vega)
    assert abs(g.rho - rho) < 3 * g.rho_se + 5e-3 * rho, (g, rho)


def test_c2_price_and_greeks_with_control_variate(engine):
    """Config 2: RQMC Asian (sigma = 0.3, D = 256, R = 32, N0 = 2^14, BB) with the
    geometric control on the same path; eps = 1e-3 price, adjoint delta/vega/rho."""
    m = gbm(sig=0.3)
    n_steps = 256
    E = fdmc.geom_asian_analytic(100, 100, 1, 0.05, 0.3, n_steps)
    dE = fdmc.geom_asian_greeks(100, 100, 1, 0.05, 0.3, n_steps)
    cfg = fdmc.QMCConfig(R=32, N0=1 << 14, N_max=1 << 20)
    bb = engine.bb_upload(fdmc.bb_precompute(n_steps, 1.0))
    r, g = engine.run_qmc_greeks(m, ASIAN, 1e-3, n_steps, cfg, fdmc.NOISE_BB, bb, ctrl_model=m, ctrl_payoff=GEO,
                                 E_ctrl=E, dE_ctrl=dE, beta=1.0)
    rp = engine.run_qmc_cv_cuda(m, m, ASIAN, GEO, E, 1.0, 1e-3, n_steps, cfg, fdmc.NOISE_BB, dev_bb=bb)
    r2, g2 = engine.run_qmc_greeks(m, ASIAN, 1e-3, n_steps, cfg, fdmc.NOISE_BB, bb)   # no control
    bb.free()
    assert r.flags == 0 and abs(r.price - rp.price) < 1e-9
    for a_, b_, sa, sb in ((g.delta, g2.delta, g.delta_se, g2.delta_se), (g.vega, g2.vega, g.vega_se, g2.vega_se),
                           (g.rho, g2.rho, g.rho_se, g2.rho_se)):
        assert abs(a_ - b_) < 4 * math.hypot(sa, sb) + 1e-3, (g, g2)
    assert g.delta_se < g2.delta_se and g.vega_se < g2.vega_se      # the control reduces Greek variance too


def test_c2_adjoint_with_control_matches_crn_differences(engine):
    """Config 2's own setting: sigma = 0.3, D = 256 bridge, R = 32 replicas, the
    geometric-Asian control with beta = 1.  The adjoint Greeks of the controlled
    estimator Y_a - beta (Y_g - E(theta)) (kernel sums of d(Y_a - beta Y_g) plus
    beta dE/dtheta of the closed form) against central differences of the same
    controlled estimator repriced at theta +- h on the same points, with E(theta)
    re-evaluated at each bump."""
    n_steps, R, n, beta = 256, 32, 1 << 12, 1.0

    def E_at(S0=100.0, mu=0.05, sig=0.3):
        return fdmc.geom_asian_analytic(S0, 100.0, 1.0, mu, sig, n_steps)

    bb = engine.bb_upload(fdmc.bb_precompute(n_steps, 1.0))
    m = gbm(sig=0.3)
    rows = np.array(engine.greeks_accumulate(m, ASIAN, n_steps, R, 42, 0, n, fdmc.NOISE_BB, dev_bb=bb,
                                             ctrl_model=m, ctrl_payoff=GEO, E_ctrl=E_at(), beta=beta))
    dE = fdmc.geom_asian_greeks(100.0, 100.0, 1.0, 0.05, 0.3, n_steps)
    npr = rows.shape[1] - 6
    bumps = {"delta": ("S0", 0.05), "vega": ("sig", 0.002), "rho": ("mu", 0.002)}
    for k, which in enumerate(("delta", "vega", "rho")):
        key, h = bumps[which]
        cv = []
        for sgn in (+1, -1):
            kw = {"S0": 100.0, "mu": 0.05, "sig": 0.3}
            kw[key] += sgn * h
            mb = gbm(**kw)
            got = np.array(engine.qmc_accumulate(mb, ASIAN, n_steps, R, 42, 0, n, fdmc.NOISE_BB, dev_bb=bb,
                                                 ctrl_model=mb, ctrl_payoff=GEO, E_ctrl=E_at(**kw), beta=beta))
            cv.append(got[:, 2] / n)
        fd = (cv[0] - cv[1]) / (2 * h)
        adj = rows[:, npr + 2 * k] / n + beta * dE[k]
        print(which, "adjoint", adj[:3], "crn", fd[:3], flush=True)
        assert np.all(np.abs(adj - fd) <= 0.01 * np.abs(fd) + 2e-3), (which, adj, fd)
    bb.free()


def test_autograd_binding_returns_adjoint_greeks(engine, torch_cuda):
    torch = torch_cuda
    import autograd as ag
    bb = engine.bb_upload(fdmc.bb_precompute(64, 1.0))
    S0 = torch.tensor(100.0, dtype=torch.float64, requires_grad=True)
    sg = torch.tensor(0.2, dtype=torch.float64, requires_grad=True)
    mu = torch.tensor(0.05, dtype=torch.float64, requires_grad=True)
    price = ag.QMCPrice.apply(S0, sg, mu, engine, EURO, 64, 2, 4096, fdmc.NOISE_BB, bb)
    price.backward()
    rows = np.array(engine.greeks_accumulate(gbm(), EURO, 64, 2, 42, 0, 4096, fdmc.NOISE_BB, dev_bb=bb))
    bb.free()
    n = 2 * 4096
    assert abs(S0.grad.item() - rows[:, 2].sum() / n) < 1e-12
    assert abs(sg.grad.item() - rows[:, 4].sum() / n) < 1e-12
    assert abs(mu.grad.item() - rows[:, 6].sum() / n) < 1e-12
