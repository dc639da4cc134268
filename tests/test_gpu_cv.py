"""GPU parity of the control-variate drivers cv_pilot and run_mc_cv_cuda
(methods_cuda.cu:2026-2076, 2078-2163) against the oracle's restatements on the
same Philox streams, at the parameters of the reference's callers:
examples/ejemplo09_asian_cv.cpp:38,55-57 (GBM arithmetic Asian with the
geometric-Asian control, E_ctrl from the discrete closed form, 50,000-path
pilot) and examples/ejemplo10_dupire_cv.cpp:66 (Dupire European with a GBM(sigma0)
European control, E_ctrl from Black-Scholes).

The device evaluates FP32 paths with its own Box-Muller and hardware exp/log; the
oracle runs the same FP32 recursion from FP64 normals.  Per-path payoffs agree to
~1e-6 relative, so the pilot's moments (beta, var_plain, var_cv) agree to 2e-5
relative (var_cv, a difference, on var_plain's scale), the sample count N = ceil(2 v / eps^2) to one path, and price and
standard error to 2e-5 relative when N agrees.
"""
import math

import pytest

import fdmc
import pyoracle as O

pytestmark = pytest.mark.gpu

GBM = fdmc.GBMParams(100.0, 0.05, 0.2, 1.0)
DUP = fdmc.DupireLocalParams(100.0, 0.05, 0.2, 0.5, 0.7, 1.0)
ASIAN, GEO = fdmc.Asian(100.0), fdmc.GeomAsian(100.0)
EURO = fdmc.European(100.0, 0.05, 1.0)


def _rel(a, b):
    return abs(a - b) / max(abs(b), 1e-300)


def _ej09(n_steps):
    return GBM, GBM, ASIAN, GEO, fdmc.geom_asian_analytic(100.0, 100.0, 1.0, 0.05, 0.2, n_steps)


def _ej10():
    return DUP, GBM, EURO, EURO, fdmc.bs_call(100.0, 100.0, 1.0, 0.05, 0.2)


@pytest.mark.parametrize("case,n_steps,seed", [("ej09", 64, 0), ("ej09", 32, 7), ("ej10", 64, 0),
                                                ("ej10", 128, 3)])
def test_cv_pilot_matches_oracle(engine, case, n_steps, seed):
    mm, cm, mp, cp, E = _ej09(n_steps) if case == "ej09" else _ej10()
    g = engine.cv_pilot(mm, cm, mp, cp, E, n_steps, 50000, seed)
    o = O.cv_pilot(mm, cm, mp, cp, E, n_steps, 50000, seed)
    print(case, n_steps, seed, g, o, flush=True)
    assert _rel(g.beta, o.beta) < 2e-5, (g.beta, o.beta)
    assert _rel(g.var_plain, o.var_plain) < 2e-5, (g.var_plain, o.var_plain)
    # var_cv = var_plain - cov^2 / var_ctrl cancels most of var_plain: compare on var_plain's scale
    assert abs(g.var_cv - o.var_cv) < 2e-5 * o.var_plain, (g.var_cv, o.var_cv)
    # the control helps: strongly for the geometric Asian, moderately for the GBM leg
    assert 0.0 < g.beta <= 5.0 and g.var_cv < (0.05 if case == "ej09" else 0.5) * g.var_plain


@pytest.mark.parametrize("case,n_steps,eps,beta", [("ej09", 64, 1e-2, None), ("ej09", 64, 2e-2, 0.0),
                                                    ("ej10", 64, 1e-2, None), ("ej10", 32, 5e-3, 1.0)])
def test_run_mc_cv_matches_oracle(engine, case, n_steps, eps, beta):
    mm, cm, mp, cp, E = _ej09(n_steps) if case == "ej09" else _ej10()
    if beta is None:     # the examples' flow: beta from a 50,000-path pilot
        beta = O.cv_pilot(mm, cm, mp, cp, E, n_steps, 50000).beta
    cfg = fdmc.MCConfig()
    g = engine.run_mc_cv_cuda(mm, cm, mp, cp, E, beta, eps, n_steps, cfg)
    o = O.run_mc_cv(mm, cm, mp, cp, E, beta, eps, n_steps, cfg)
    print(case, n_steps, eps, beta, g.price, g.std_error, g.n_samples, o.price, o.std_error, o.n_samples,
          flush=True)
    # N = max(ceil(2 v / eps^2), pilot_n) from the pilot on stream seed, main run on seed + 100
    assert abs(g.n_samples - o.n_samples) <= 1
    assert g.n_samples >= cfg.pilot_n
    if g.n_samples == o.n_samples:
        assert _rel(g.price, o.price) < 2e-5
        assert _rel(g.std_error, o.std_error) < 2e-5
    assert g.std_error < eps / math.sqrt(2) * 1.1
    # the estimator is unbiased for the Euler price: agree with the plain MC on the same grid
    plain = engine.run_mc_cuda(mm, mp, eps, n_steps, fdmc.MCConfig())
    assert abs(g.price - plain.price) < 4 * math.hypot(g.std_error, plain.std_error)
