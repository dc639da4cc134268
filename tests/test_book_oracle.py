"""Pins of the contract-book oracle (CPU, no GPU).

The book prices each contract by multilevel QMC with M = 2 and a fixed finest
level (its own dates), with pathwise Greeks.  The oracle (oracle_drivers.c,
orc_book_*) is pinned here against what does not come from its own code:

- the reference-equivalent MLQMC level sums (orc_mlqmc_level_sums, itself pinned
  by test_oracle_pins) on the same salts: contract c, level l, replica r uses
  salt 42 + (10 c + l) R + r, i.e. run_mlqmc_cuda's 42 + l R + r for contract 0
  (methods_cuda.cu:1845);
- closed forms of the one-step Euler payoff at level 0 (price and its three
  pathwise derivatives, normal integrals);
- the telescoping identity sum_l E[P_l - P_{l-1}] = E[P_L] against a plain QMC
  run on the finest grid;
- central finite differences, per path, of the same Euler recursion in FP64 for
  every (model, payoff) with Greeks;
- Black-Scholes price, delta, vega and rho for a GBM European contract;
- the uniform-grid bridge the device kernel assumes (weights 1/2, one sd per depth).
"""
import math

import numpy as np
import pytest

import fdmc
import pyoracle as po
from scipy.stats import norm

R = 8


def lvq(T=1.0, S0=100.0, a=0.2):
    return fdmc.LocalVolQuadParams(S0=S0, mu=0.05, a=a, b=0.1, c=1.0, s_ref=100.0, T=T)


def contract(kind, n, K=100.0, T=1.0, r=0.05, B=None):
    return fdmc.BookContract(kind, n, K, 1.3 * K if B is None else B, T, r, 0.2)


def cfg(**kw):
    base = dict(R=R, pilot_n=400, salt_base=42, greeks=True, max_iter=50)
    base.update(kw)
    return fdmc.BookConfig(**base)


# ------------------------------------------------------------------ bridge
@pytest.mark.parametrize("T", [0.25, 0.5, 1.0, 2.0])
def test_uniform_bridge_is_halves_with_one_sd_per_depth(T):
    for l in range(10):
        N = 1 << l
        a = po.OracleBB(N, T).arrays()
        wl, wr, sd = (a[k].astype(np.float32) for k in ("wl", "wr", "sd"))
        assert sd[0] == np.float32(math.sqrt(T))
        assert np.all(wl[1:] == np.float32(0.5)) and np.all(wr[1:] == np.float32(0.5))
        for lev in range(l):
            seg = sd[1 << lev:2 << lev]
            assert np.all(seg == seg[0]), (T, l, lev)
            h = T / N
            assert abs(float(seg[0]) - math.sqrt((N >> lev) * h / 4.0)) < 1e-6


# ------------------------------------------------------------------ salts / coupling
@pytest.mark.parametrize("c", [0, 3])
@pytest.mark.parametrize("kind,payoff", [
    (fdmc.PAYOFF_EUROPEAN, lambda T: fdmc.European(100.0, 0.05, T)),
    (fdmc.PAYOFF_ASIAN, lambda T: fdmc.Asian(100.0)),
    (fdmc.PAYOFF_LOOKBACK, lambda T: fdmc.Lookback(0.2)),
    (fdmc.PAYOFF_BARRIER, lambda T: fdmc.Barrier(100.0, 130.0, 0.2, 0.05, T)),
])
def test_level_sums_equal_reference_mlqmc_on_the_same_salts(c, kind, payoff):
    T = 0.5
    m = lvq(T)
    k = contract(kind, 64, T=T, B=130.0)
    for l in (0, 3, 6):
        book = po.book_level_sums(m, k, c, l, R, 42, 1000, 512)
        ref = po.mlqmc_level_sums(m, payoff(T), 2, l, R, 42 + 10 * R * c, 1000, 512, mode=fdmc.NOISE_BB,
                                  bb=po.OracleBB(1 << l, T))
        # FP32 path arithmetic: the lookback's S_T - min * bgk contracts to one FMA in
        # the book and rounds min * bgk first in the reference form (1 ulp per path);
        # a salt, offset or coupling mistake moves the sums by O(1) per path
        np.testing.assert_allclose(book[:, 0], ref[:, 0], rtol=1e-6, atol=1e-5 * 512)


# ------------------------------------------------------------------ level 0 closed forms
def test_level0_european_matches_one_step_euler_closed_forms():
    S0, mu, sig, T, K = 100.0, 0.05, 0.3, 1.0, 105.0
    m = fdmc.GBMParams(S0, mu, sig, T)
    k = contract(fdmc.PAYOFF_EUROPEAN, 1, K=K, T=T, r=mu)
    n = 1 << 16
    s = po.book_level_sums(m, k, 0, 0, R, 42, 0, n).sum(axis=0) / (R * n)
    disc = math.exp(-mu * T)
    mm, ss = S0 * (1 + mu * T) - K, S0 * sig * math.sqrt(T)
    x = mm / ss
    price = disc * (mm * norm.cdf(x) + ss * norm.pdf(x))
    delta = disc * ((1 + mu * T) * norm.cdf(x) + sig * math.sqrt(T) * norm.pdf(x))
    vega = disc * S0 * math.sqrt(T) * norm.pdf(x)
    rho = disc * S0 * T * norm.cdf(x) - T * price
    np.testing.assert_allclose(s, [price, delta, vega, rho], rtol=2e-4, atol=2e-5)


# ------------------------------------------------------------------ telescoping
@pytest.mark.parametrize("kind,payoff", [
    (fdmc.PAYOFF_EUROPEAN, lambda T: fdmc.European(95.0, 0.05, T)),
    (fdmc.PAYOFF_ASIAN, lambda T: fdmc.Asian(95.0)),
    (fdmc.PAYOFF_LOOKBACK, lambda T: fdmc.Lookback(0.2)),
    (fdmc.PAYOFF_BARRIER, lambda T: fdmc.Barrier(95.0, 120.0, 0.2, 0.05, T)),
])
def test_book_contract_telescopes_to_single_level_qmc(kind, payoff):
    T, n = 1.0, 32
    m = lvq(T)
    k = contract(kind, n, K=95.0, T=T, B=120.0)
    eps = 1e-2
    b = po.book_contract_run(m, k, 0, eps, cfg(R=16))
    q = po.run_qmc(m, payoff(T), 2e-3, n, fdmc.QMCConfig(R=16, N0=1 << 12, salt_base=7),
                   mode=fdmc.NOISE_BB, bb=po.OracleBB(n, T))
    assert b.flags == 0 and b.std_error <= eps / math.sqrt(2) * (1 + 1e-12)
    assert abs(b.price - q.price) < 3 * math.hypot(b.std_error, q.std_error) + 1e-4, (b.price, q.price)


# ------------------------------------------------------------------ Greeks vs FP64 finite differences
# The oracle's tangents are FP32.  Finite differences of the FP32 path are
# dominated by rounding (a level difference and its derivative are small), so
# the differences are taken on the same Euler recursion evaluated in FP64 below;
# that recursion is itself checked against the oracle's forward value on every
# path.  No derivative formula appears here: only the forward model.
def _leg64(mkind, par, kind, K, B, T, r, inc):
    P, n = inc.shape
    h = T / n
    S0, mu = par["S0"], par["mu"]
    S = np.full(P, S0)
    acc, mn, mx = np.zeros(P), np.full(P, S0), np.zeros(P)
    for j in range(n):
        t = j * h
        if mkind == "gbm":
            vol = par["v"]
        elif mkind == "dupire":        # sigma0 e^{-alpha t} (S/S0)^(beta-1), alpha 0.5, beta 0.7
            vol = par["v"] * np.exp(-0.5 * t) * (S / S0) ** (0.7 - 1.0)
        else:                          # a + b e^{-c t} ln(S/100)^2, b 0.1, c 1
            vol = par["v"] + 0.1 * np.exp(-t) * np.log(S / 100.0) ** 2
        S = S + mu * S * h + vol * S * inc[:, j]
        acc += S
        mn = np.minimum(mn, S)
        mx = np.maximum(mx, S)
    bgk = np.exp(-0.5826 * 0.2 * np.sqrt(h))
    if kind == fdmc.PAYOFF_EUROPEAN:
        return np.maximum(S - K, 0.0) * np.exp(-r * T)
    if kind == fdmc.PAYOFF_ASIAN:
        return np.maximum(acc / n - K, 0.0)
    return np.where(S > 0, S - mn * bgk, 0.0)     # lookback; 0 for a path that left the domain


def _diff64(mkind, par, kind, K, T, r, dW):
    f = _leg64(mkind, par, kind, K, 0.0, T, r, dW)
    if dW.shape[1] >= 2:
        f = f - _leg64(mkind, par, kind, K, 0.0, T, r, dW[:, 0::2] + dW[:, 1::2])
    return f


MODELS = {
    "gbm": lambda S0, v, mu: fdmc.GBMParams(S0, mu, v, 1.0),
    "dupire": lambda S0, v, mu: fdmc.DupireLocalParams(S0, mu, v, 0.5, 0.7, 1.0),
    "lvq": lambda S0, v, mu: fdmc.LocalVolQuadParams(S0, mu, v, 0.1, 1.0, 100.0, 1.0),
}


@pytest.mark.parametrize("model", sorted(MODELS))
@pytest.mark.parametrize("kind", [fdmc.PAYOFF_EUROPEAN, fdmc.PAYOFF_ASIAN, fdmc.PAYOFF_LOOKBACK])
@pytest.mark.parametrize("nf", [1, 4, 32])
def test_path_greeks_match_fp64_central_differences(model, kind, nf):
    rng = np.random.default_rng(11 * nf + kind)
    T, K, r = 1.0, 98.0, 0.05
    dW32 = (rng.standard_normal((64, nf)) * math.sqrt(T / nf)).astype(np.float32)
    dW = dW32.astype(np.float64)
    par = dict(S0=100.0, v=0.22, mu=0.05)
    k = fdmc.BookContract(kind, 32, K, 1.3 * K, T, r, 0.2)
    orc = np.array([po.book_path(MODELS[model](**par), k, dW32[i]) for i in range(len(dW))], np.float64)
    np.testing.assert_allclose(orc[:, 0], _diff64(model, par, kind, K, T, r, dW), atol=3e-4)
    for col, key, h in ((1, "S0", 1e-4), (2, "v", 1e-7), (3, "mu", 1e-7)):
        up, dn = dict(par), dict(par)
        up[key] += h
        dn[key] -= h
        dr = h if key == "mu" else 0.0     # rho moves mu and the discount rate together
        fd = (_diff64(model, up, kind, K, T, r + dr, dW) - _diff64(model, dn, kind, K, T, r - dr, dW)) / (2 * h)
        np.testing.assert_allclose(orc[:, col], fd, rtol=1e-5, atol=5e-4, err_msg=key)


# ------------------------------------------------------------------ closed forms of the limit
def test_gbm_european_contract_matches_black_scholes():
    S0, mu, sig, T, K = 100.0, 0.05, 0.25, 1.0, 100.0
    m = fdmc.GBMParams(S0, mu, sig, T)
    k = contract(fdmc.PAYOFF_EUROPEAN, 256, K=K, T=T, r=mu)
    b = po.book_contract_run(m, k, 0, 4e-3, cfg(R=16))
    d1 = (math.log(S0 / K) + (mu + 0.5 * sig * sig) * T) / (sig * math.sqrt(T))
    d2 = d1 - sig * math.sqrt(T)
    bs = S0 * norm.cdf(d1) - K * math.exp(-mu * T) * norm.cdf(d2)
    # Euler bias at 256 steps is below 1e-3 on this call (about -0.002 at 64 steps)
    assert abs(b.price - bs) < 3 * b.std_error + 1e-3, (b.price, bs)
    assert abs(b.delta - norm.cdf(d1)) < 3 * b.delta_se + 2e-4
    assert abs(b.vega - S0 * norm.pdf(d1) * math.sqrt(T)) < 3 * b.vega_se + 0.05
    assert abs(b.rho - K * T * math.exp(-mu * T) * norm.cdf(d2)) < 3 * b.rho_se + 0.05


def test_stopping_rule_and_barrier_greeks():
    m = lvq(0.5)
    for kind in (fdmc.PAYOFF_ASIAN, fdmc.PAYOFF_BARRIER):
        k = contract(kind, 16, K=100.0, T=0.5, B=125.0)
        b = po.book_contract_run(m, k, 2, 1e-2, cfg(R=16))
        assert b.flags == 0 and b.iterations >= 1
        assert b.std_error <= 1e-2 / math.sqrt(2) * (1 + 1e-12)
        assert b.levels == 5 and b.n_samples >= 5 * 16 * 400
        assert math.isnan(b.delta) == (kind == fdmc.PAYOFF_BARRIER)
    b = po.book_contract_run(m, contract(fdmc.PAYOFF_ASIAN, 16, T=0.5), 2, 1e-2, cfg(R=16, max_iter=1))
    assert b.iterations == 1 and b.flags == 2      # FDMC_FLAG_NOT_CONVERGED: out of passes


# ------------------------------------------------------------------ paths that leave the domain
# An Euler path under a local vol can step to S <= 0, where ln S is NaN and so is
# every later state.  Every payoff of such a path is 0: the calls through fmaxf
# (as in the reference's kernels), the lookback by the explicit S_T > 0 test
# (the reference's S_T - min would be NaN), and its Greeks are 0.
@pytest.mark.parametrize("kind", [fdmc.PAYOFF_EUROPEAN, fdmc.PAYOFF_ASIAN, fdmc.PAYOFF_LOOKBACK,
                                  fdmc.PAYOFF_BARRIER])
def test_book_paths_that_leave_the_domain_pay_zero(kind):
    k = contract(kind, 4, K=90.0)
    # h = 1/4: 1 + mu h + 0.2 (-6) < 0, so S_1 < 0 and S_2.. are NaN; the coarse
    # leg's first step (-5.9) leaves too
    out = po.book_path(lvq(), k, np.array([-6.0, 0.1, 0.2, -0.1], np.float32))
    assert list(out) == [0.0, 0.0, 0.0, 0.0]
    # one date: S_T < 0 and finite; the lookback's S_T - min would be 0 here but
    # with the BGK factor negative
    out = po.book_path(lvq(), contract(kind, 1, K=90.0), np.array([-6.0], np.float32))
    assert list(out) == [0.0, 0.0, 0.0, 0.0]
    # inside the domain the lookback pays S_T - min bgk > 0
    if kind == fdmc.PAYOFF_LOOKBACK:
        out = po.book_path(lvq(), k, np.array([0.3, -0.2, 0.1, 0.05], np.float32))
        assert out[0] != 0.0 and all(math.isfinite(x) for x in out)


@pytest.mark.parametrize("kind", [fdmc.PAYOFF_EUROPEAN, fdmc.PAYOFF_ASIAN, fdmc.PAYOFF_LOOKBACK])
def test_book_pair_with_one_leg_outside_the_domain(kind):
    """Fine leg out (its first step -6 takes S below 0), coarse leg in (its one
    step is the sum -0.05): the pair is 0 - (coarse leg), price and Greeks, and
    the coarse leg of a two-date contract is the one-date contract's only leg."""
    k2, k1 = contract(kind, 2, K=90.0), contract(kind, 1, K=90.0)
    pair = po.book_path(lvq(), k2, np.array([-6.0, 5.95], np.float32))
    leg = po.book_path(lvq(), k1, np.array([np.float32(-6.0) + np.float32(5.95)], np.float32))
    assert leg[0] > 0.0 and all(math.isfinite(x) and x != 0.0 for x in leg)
    assert list(pair) == [-x for x in leg]


@pytest.mark.parametrize("payoff", [fdmc.European(90.0, 0.05, 1.0), fdmc.Asian(90.0), fdmc.GeomAsian(90.0),
                                    fdmc.Lookback(0.2), fdmc.Barrier(90.0, 130.0, 0.2, 0.05, 1.0)],
                         ids=["euro", "asian", "geo", "look", "barrier"])
def test_reference_paths_that_leave_the_domain_pay_zero(payoff):
    P = po.make_params(lvq(), payoff, 4)
    assert po.path_single(P, np.array([-6.0, 0.1, 0.2, -0.1], np.float32), 4) == 0.0
    P1 = po.make_params(lvq(), payoff, 1)
    assert po.path_single(P1, np.array([-6.0], np.float32), 1) == 0.0


def _exit_fraction(a, b, T, n, P=100000, seed=5):
    """Fraction of FP32 Euler paths of the quadratic local vol a + b e^{-t}
    ln(S/100)^2 that reach S <= 0 (or NaN) within n steps."""
    rng = np.random.default_rng(seed)
    h = np.float32(T / n)
    S = np.full(P, 100.0, np.float32)
    out = np.zeros(P, bool)
    with np.errstate(all="ignore"):
        for j in range(n):
            l = np.log(S / np.float32(100.0))
            vol = np.float32(a) + np.float32(b) * np.float32(math.exp(-j * float(h))) * l * l
            S = S + np.float32(0.05) * S * h + vol * S * (rng.standard_normal(P).astype(np.float32) * np.sqrt(h))
            out |= ~(S > 0)
    return out.mean()


# a = 1 over four dates of T = 1: about 9 % of the fine paths (and more of the
# coarser levels') leave the domain
STEEP = dict(a=1.0, b=0.05, T=1.0, n=4)


def test_steep_local_vol_book_is_finite():
    """A book contract whose paths leave the domain often: the estimate and its
    Greeks are finite and the stopping rule is met, in both differentiation
    modes (the lookback's price was NaN before paths outside the domain paid 0)."""
    assert _exit_fraction(STEEP["a"], STEEP["b"], STEEP["T"], STEEP["n"]) > 0.05
    m = fdmc.LocalVolQuadParams(S0=100.0, mu=0.05, a=STEEP["a"], b=STEEP["b"], c=1.0, s_ref=100.0, T=1.0)
    for kind in (fdmc.PAYOFF_LOOKBACK, fdmc.PAYOFF_EUROPEAN, fdmc.PAYOFF_ASIAN):
        for greeks in (1, 2):
            b = po.book_contract_run(m, contract(kind, STEEP["n"], T=STEEP["T"]), 0, 5e-2, cfg(greeks=greeks))
            assert b.flags == 0 and b.std_error <= 5e-2 / math.sqrt(2) * (1 + 1e-9)
            vals = (b.price, b.delta, b.vega, b.rho, b.delta_se, b.vega_se, b.rho_se)
            assert all(math.isfinite(x) for x in vals), (kind, greeks, b)
