"""Pins of the CPU oracle (no GPU).

Each group checks the oracle against something it was not derived from:
scipy's independent Sobol generator and ndtri, the Random123 known-answer
vectors, closed-form eigenpairs and covariances, numerical integration of the
closed forms, hand-computed paths, exact special-case reductions of the models,
net properties that an ordering or masking mistake in the scramble would break,
and the reference's own report prices.
"""
import json
import math
import os

import numpy as np
import pytest
from scipy import integrate, stats
from scipy.special import ndtri

import fdmc
import pyoracle as O

HERE = os.path.dirname(os.path.abspath(__file__))


# ---------------------------------------------------------------- Sobol
def test_sobol_raw_matches_scipy_first_dims():
    D, m = 64, 10
    X = (stats.qmc.Sobol(D, scramble=False, bits=32).random_base2(m) * 2.0 ** 32).astype(np.uint64)
    for d in range(D):
        for i in range(0, 1 << m, 7):
            assert O.sobol_raw(d, i) == X[i, d], (d, i)


def test_sobol_raw_matches_scipy_high_dims():
    D = 20000
    X = (stats.qmc.Sobol(D, scramble=False, bits=32).random(33) * 2.0 ** 32).astype(np.uint64)
    for d in (100, 1023, 5000, 12345, 19990, 19999):
        for i in range(33):
            assert O.sobol_raw(d, i) == X[i, d], (d, i)


def test_dimension_zero_is_van_der_corput():
    v = O.dirvectors(0)
    assert [int(x) for x in v] == [1 << (31 - j) for j in range(32)]
    # Gray-code order: 0, 1/2, 3/4, 1/4, ...
    assert [O.sobol_raw(0, i) for i in range(4)] == [0, 1 << 31, 3 << 30, 1 << 30]
    assert [O.sobol_raw(1, i) for i in range(4)] == [0, 1 << 31, 1 << 30, 3 << 30]


def test_scramble_is_lower_triangular_with_unit_diagonal():
    rng = np.random.default_rng(1)
    for _ in range(200):
        base = int(rng.integers(0, 2 ** 32))
        raw = int(rng.integers(0, 2 ** 32))
        b = int(rng.integers(0, 32))
        y0 = O.lib().orc_hh_scramble(raw, base)
        y1 = O.lib().orc_hh_scramble(raw ^ (1 << b), base)
        diff = y0 ^ y1
        # flipping raw bit b flips output bit b and possibly lower bits only
        assert (diff >> b) == 1, (base, raw, b, hex(diff))


def test_scramble_is_f2_affine():
    rng = np.random.default_rng(2)
    for _ in range(200):
        base = int(rng.integers(0, 2 ** 32))
        a, b = (int(x) for x in rng.integers(0, 2 ** 32, 2))
        f = lambda x: O.lib().orc_hh_scramble(x, base)
        assert f(a ^ b) == f(a) ^ f(b) ^ f(0)


def test_scrambled_points_keep_the_net_property():
    """Dims 0 and 1 form a (0, m, 2)-net in base 2; the lower-triangular scramble
    with a digital shift keeps every elementary box of area 2^-m holding one point."""
    m = 8
    for salt in (42, 43, 1234567):
        pts = O.sobol_scrambled_block(salt, 2, 0, 1 << m)
        for a in range(m + 1):
            bx = (pts[0] >> np.uint32(32 - a)) if a else np.zeros(1 << m, np.uint32)
            by = (pts[1] >> np.uint32(32 - (m - a))) if m - a else np.zeros(1 << m, np.uint32)
            cells = bx.astype(np.uint64) * (1 << (m - a)) + by
            assert len(np.unique(cells)) == 1 << m, (salt, a)


def test_linear_table_generator_equals_per_point_scramble():
    rng = np.random.default_rng(3)
    for salt, D, off, n in ((42, 16, 0, 300), (99, 40, 123456, 200), (7, 3, 0xFFFFFFFF - 100, 101)):
        blk = O.sobol_scrambled_block(salt, D, off, n)
        for _ in range(150):
            d, p = int(rng.integers(D)), int(rng.integers(n))
            assert blk[d, p] == O.sobol_scrambled(salt, d, off + p)


def test_golden_scrambled_vectors():
    for v in json.load(open(os.path.join(HERE, "golden", "sobol_scrambled.json"))):
        assert O.sobol_scrambled(v["salt"], v["dim"], v["index"]) == v["value"], v


# ---------------------------------------------------------------- uniform map / ICDF
def test_reference_uniform_map_hits_one():
    # methods_cuda.cu:823: ((float)s + 1) * 2^-32 rounds to 1.0f near the top
    assert O.lib().orc_u_reference(0xFFFFFFFF) == 1.0
    assert O.lib().orc_u_reference(0xFFFFFF80) == 1.0
    assert O.lib().orc_u_open(0xFFFFFFFF) < 1.0 and O.lib().orc_u_open(0) > 0.0


def test_ndtri_matches_scipy():
    us = np.concatenate([np.logspace(-300, -1, 300), np.linspace(0.01, 0.99, 199), 1 - np.logspace(-16, -1, 80)])
    ref = ndtri(us)
    got = np.array([O.lib().orc_ndtri(float(u)) for u in us])
    assert np.max(np.abs(got - ref) / np.maximum(1, np.abs(ref))) < 1e-13


def test_normal_from_u32_is_odd_and_exact():
    for s in (0, 1, 12345, 2 ** 31 - 1, 2 ** 24 + 3):
        x = O.lib().orc_normal_u32(s)
        assert x == -O.lib().orc_normal_u32(2 ** 32 - 1 - s)
        assert abs(x - ndtri((s + 0.5) / 2 ** 32)) < 1e-12 * max(1, abs(x))


# ---------------------------------------------------------------- Philox / Box-Muller
def test_philox_known_answers():
    # Random123 kat_vectors, philox4x32_10
    assert O.philox([0, 0, 0, 0], [0, 0]) == [0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8]
    assert O.philox([0xffffffff] * 4, [0xffffffff] * 2) == [0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd]
    assert O.philox([0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344], [0xa4093822, 0x299f31d0]) == \
        [0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1]


def test_box_muller_normals_are_standard():
    z = np.concatenate([O.prng_normals(123, 0, p, 64) for p in range(4000)]).astype(np.float64)
    assert abs(z.mean()) < 4 / math.sqrt(z.size)
    assert abs(z.var() - 1) < 4 * math.sqrt(2 / z.size)
    assert stats.kstest(z, "norm").pvalue > 1e-3


# ---------------------------------------------------------------- Brownian bridge
@pytest.mark.parametrize("N,T", [(1, 1.0), (2, 0.5), (8, 1.0), (64, 2.0)])
def test_bridge_covariance_is_min_t(N, T):
    bb = O.OracleBB(N, T)
    B = np.stack([bb.apply_f64(np.eye(N)[j]) for j in range(N)], axis=1)   # dW = B z
    W = np.cumsum(B, axis=0)
    t = np.arange(1, N + 1) * T / N
    assert np.allclose(W @ W.T, np.minimum.outer(t, t), atol=1e-12)


def test_bridge_f32_matches_f64_and_raw_for_one_step():
    bb = O.OracleBB(256, 1.0)
    z = np.random.default_rng(5).standard_normal(256)
    assert np.allclose(bb.apply_f32(z.astype(np.float32)), bb.apply_f64(z), atol=2e-6)
    one = O.OracleBB(1, 0.7)
    assert abs(one.apply_f64(np.array([1.3]))[0] - 1.3 * math.sqrt(0.7)) < 1e-15


# ---------------------------------------------------------------- PCA
@pytest.mark.parametrize("m", [4, 33, 128])
def test_pca_matches_closed_form_eigenpairs(m):
    T = 1.0
    h = T / m
    C = h * np.minimum.outer(np.arange(1, m + 1), np.arange(1, m + 1)).astype(float)
    w, V = O.eigh_jacobi(C)
    k = np.arange(1, m + 1)
    lam = h / (4 * np.sin((2 * k - 1) * np.pi / (2 * (2 * m + 1))) ** 2)
    assert np.allclose(w, lam, rtol=1e-12)
    i = np.arange(1, m + 1)[:, None]
    Vc = 2 / np.sqrt(2 * m + 1) * np.sin((2 * k[None, :] - 1) * i * np.pi / (2 * m + 1))
    assert np.allclose(np.abs(V), np.abs(Vc), atol=1e-10)
    M = O.pca_compute(m, T)
    assert np.allclose(M @ M.T, h * np.eye(m), atol=1e-13)


def test_jacobi_matches_numpy_on_random_symmetric():
    A = np.random.default_rng(7).standard_normal((40, 40))
    A = A + A.T
    w, V = O.eigh_jacobi(A)
    assert np.allclose(w, np.sort(np.linalg.eigvalsh(A))[::-1], atol=1e-11)
    assert np.allclose(V @ np.diag(w) @ V.T, A, atol=1e-10)


def test_kronecker_pca_reproduces_increment_covariance():
    n, steps, T = 4, 8, 1.0
    B = np.random.default_rng(9).uniform(0.3, 0.6, (n, 2))
    corr = B @ B.T
    np.fill_diagonal(corr, 1.0)
    M = O.pca_kron_compute(steps, T, corr)
    h = T / steps
    assert np.allclose(M @ M.T, h * np.kron(np.eye(steps), corr), atol=1e-13)


# ---------------------------------------------------------------- closed forms
def test_black_scholes_report_value_and_integral():
    bs = O.bs_call(100, 100, 1, 0.05, 0.2)
    assert round(bs, 6) == 10.450584          # informes_finales/informe_completo.txt:8
    f = lambda z: max(100 * math.exp((0.05 - 0.02) + 0.2 * z) - 100, 0) * stats.norm.pdf(z)
    assert abs(bs - math.exp(-0.05) * integrate.quad(f, -10, 10, points=[0.0])[0]) < 1e-8


@pytest.mark.parametrize("n", [1, 12, 64, 256])
def test_geometric_asian_closed_form_vs_integral(n):
    S0, K, T, mu, sig = 100.0, 95.0, 1.0, 0.05, 0.3
    t = np.arange(1, n + 1) * T / n
    m = math.log(S0) + (mu - 0.5 * sig * sig) * t.mean()
    v = sig * sig * np.minimum.outer(t, t).sum() / n ** 2
    f = lambda x: max(math.exp(m + math.sqrt(v) * x) - K, 0) * stats.norm.pdf(x)
    ref = integrate.quad(f, -12, 12, points=[(math.log(K) - m) / math.sqrt(v)])[0]
    assert abs(O.geom_asian_analytic(S0, K, T, mu, sig, n) - ref) < 1e-9


# ---------------------------------------------------------------- path kernels
def test_gbm_two_steps_by_hand():
    P = O.make_params(fdmc.GBMParams(100, 0.05, 0.2, 1.0), fdmc.European(90, 0.05, 1.0), 2)
    dW = np.array([0.1, -0.05], np.float32)
    S = 100.0
    for d in dW:
        S = S + 0.05 * S * 0.5 + 0.2 * S * float(d)
    assert abs(O.path_single(P, dW, 2) - (S - 90) * math.exp(-0.05)) < 1e-4


def test_asian_is_undiscounted_and_excludes_S0():
    P = O.make_params(fdmc.GBMParams(100, 0.0, 0.0, 1.0), fdmc.Asian(90), 4)
    assert abs(O.path_single(P, np.zeros(4, np.float32), 4) - 10.0) < 1e-5
    P2 = O.make_params(fdmc.GBMParams(100, 0.1, 0.0, 1.0), fdmc.Asian(100), 2)
    S1, S2 = 100 * 1.05, 100 * 1.05 * 1.05
    assert abs(O.path_single(P2, np.zeros(2, np.float32), 2) - ((S1 + S2) / 2 - 100)) < 1e-4


def test_lookback_and_barrier_bgk_shift():
    h = 1.0 / 4
    corr = math.exp(-0.5826 * 0.2 * math.sqrt(h))
    dW = np.array([-0.1, 0.05, 0.2, -0.02], np.float32)
    S, path = 100.0, []
    for d in dW:
        S = S + 0.05 * S * h + 0.2 * S * float(d)
        path.append(S)
    P = O.make_params(fdmc.GBMParams(100, 0.05, 0.2, 1.0), fdmc.Lookback(0.2), 4)
    assert abs(O.path_single(P, dW, 4) - (path[-1] - min(100.0, min(path)) * corr)) < 1e-4
    hi = max(path)
    Pk = O.make_params(fdmc.GBMParams(100, 0.05, 0.2, 1.0), fdmc.Barrier(90, hi / corr * 0.999, 0.2, 0.05, 1.0), 4)
    assert O.path_single(Pk, dW, 4) == 0.0           # max / corr >= B: knocked out
    Pa = O.make_params(fdmc.GBMParams(100, 0.05, 0.2, 1.0), fdmc.Barrier(90, hi / corr * 1.001, 0.2, 0.05, 1.0), 4)
    assert abs(O.path_single(Pa, dW, 4) - (path[-1] - 90) * math.exp(-0.05)) < 1e-4


def test_model_reductions_are_exact():
    rng = np.random.default_rng(11)
    dW = (rng.standard_normal(64) * math.sqrt(1 / 64)).astype(np.float32)
    euro = fdmc.European(100, 0.05, 1.0)
    g = O.path_single(O.make_params(fdmc.GBMParams(100, 0.05, 0.2, 1.0), euro, 64), dW, 64)
    d = O.path_single(O.make_params(fdmc.DupireLocalParams(100, 0.05, 0.2, 0.0, 1.0, 1.0), euro, 64), dW, 64)
    assert g == d                                    # alpha = 0, beta = 1 -> GBM(sigma0), bit-exact
    lv = O.path_single(O.make_params(fdmc.LocalVolQuadParams(100, 0.05, 0.2, 0.0, 1.0, 100.0, 1.0), euro, 64), dW, 64)
    assert abs(g - lv) < 1e-4                        # b = 0 -> GBM(a)
    hz = np.zeros(128, np.float32)
    hz[0::2] = dW
    hz[1::2] = rng.standard_normal(64).astype(np.float32)
    hes = fdmc.HestonParams(100, 0.05, 2.0, 0.04, 0.0, -0.5, 0.04, 1.0).compute_cholesky()
    hv = O.path_single(O.make_params(hes, euro, 64), hz, 64)
    g2 = O.path_single(O.make_params(fdmc.GBMParams(100, 0.05, float(np.sqrt(np.float32(0.04))), 1.0), euro, 64), dW, 64)
    assert abs(hv - g2) < 1e-5                       # xi = 0, v0 = theta -> GBM(sqrt(theta))
    bk = fdmc.MultiGBMParams(1, [100.0], [0.2], 0.05, 1.0)
    b = O.path_basket(O.make_params(bk, fdmc.Basket(100, 0.05, 1.0, 1), 64), dW, 64)
    assert abs(b - g) < 1e-5                         # one-asset basket -> European


def test_coupled_path_consistency():
    rng = np.random.default_rng(13)
    M, nf = 4, 16
    dW = (rng.standard_normal(nf) * math.sqrt(1 / nf)).astype(np.float32)
    P = O.make_params(fdmc.GBMParams(100, 0.05, 0.2, 1.0), fdmc.Asian(100), 1)
    pf, pc = O.path_coupled(P, dW, nf, M)
    Pf = O.make_params(fdmc.GBMParams(100, 0.05, 0.2, 1.0), fdmc.Asian(100), nf)
    Pc = O.make_params(fdmc.GBMParams(100, 0.05, 0.2, 1.0), fdmc.Asian(100), nf // M)
    assert abs(pf - O.path_single(Pf, dW, nf)) < 1e-5
    dWc = dW.reshape(-1, M).sum(axis=1).astype(np.float32)
    assert abs(pc - O.path_single(Pc, dWc, nf // M)) < 1e-4
    pf0, pc0 = O.path_coupled(P, dW[:1], 1, M)
    assert pc0 == 0.0


# ---------------------------------------------------------------- estimators
def test_mc_european_within_three_se_of_black_scholes():
    r = O.run_mc(fdmc.GBMParams(), fdmc.European(100, 0.05, 1), 5e-2, 64, fdmc.MCConfig())
    assert abs(r.price - (10.450584 - 0.002)) < 3 * r.std_error


def test_qmc_bb_european_within_eps():
    bb = O.OracleBB(64, 1.0)
    r = O.run_qmc(fdmc.GBMParams(), fdmc.European(100, 0.05, 1), 5e-3, 64, fdmc.QMCConfig(R=16, N0=512), 1, bb=bb)
    assert r.flags == 0
    assert abs(r.price - 10.450584) < 3 * r.std_error + 0.004


def test_mlmc_gbm_european():
    r = O.run_mlmc(fdmc.GBMParams(), fdmc.European(100, 0.05, 1), 5e-2, fdmc.MLMCConfig(M=2, max_L=8, pilot_n=1000))
    assert abs(r.price - 10.450584) < 3 * r.std_error + 5e-2


def _bridge_matrix(n, T):
    """The Brownian bridge as a dense increment matrix: column j = bridge of e_j (FP64)."""
    bb = O.OracleBB(n, T)
    return bb, np.stack([bb.apply_f64(e) for e in np.eye(n)], axis=1)


def test_mlqmc_pca_level_sums_equal_bridge_through_its_matrix():
    # PCA mode with the bridge written as a matrix must give the bridge's level
    # sums up to FP32 rounding of the increments; the transposed matrix must not
    g, A = fdmc.GBMParams(100, 0.05, 0.2, 1.0), fdmc.Asian(100.0)
    for M, l in ((2, 4), (4, 2)):
        bb, Mb = _bridge_matrix(M ** l, 1.0)
        ref = O.mlqmc_level_sums(g, A, M, l, 2, 42, 7, 300, fdmc.NOISE_BB, bb=bb)
        got = O.mlqmc_level_sums(g, A, M, l, 2, 42, 7, 300, fdmc.NOISE_PCA, pca_M=Mb)
        assert np.allclose(got, ref, rtol=2e-5, atol=1e-3), (got, ref)
        bad = O.mlqmc_level_sums(g, A, M, l, 2, 42, 7, 300, fdmc.NOISE_PCA, pca_M=Mb.T)
        assert not np.allclose(bad[:, 2], ref[:, 2], rtol=1e-3)


def test_mlqmc_pca_loop_matches_bridge_loop():
    g, A = fdmc.GBMParams(100, 0.05, 0.2, 1.0), fdmc.Asian(100.0)
    ml, qc = fdmc.MLMCConfig(M=2, max_L=6, pilot_n=256), fdmc.QMCConfig(R=8)
    rp = O.run_mlqmc(g, A, 1e-2, ml, qc, fdmc.NOISE_PCA, pcas=[O.pca_compute(2 ** l, 1.0) for l in range(7)])
    rb = O.run_mlqmc(g, A, 1e-2, ml, qc, fdmc.NOISE_BB, [O.OracleBB(2 ** l, 1.0) for l in range(7)])
    assert abs(rp.price - rb.price) < 3 * math.hypot(rp.std_error, rb.std_error)
    assert rp.std_error < 1e-2 / math.sqrt(2) * 1.01


# ---------------------------------------------------------------- importance sampling
def _z_star(S0, K, mu, sigma, T):
    # ejemplo11_otm_is.cpp:25-27: the shift that centres the terminal log-price on the strike
    return (math.log(K / S0) - (mu - 0.5 * sigma ** 2) * T) / (sigma * math.sqrt(T))


def test_is_with_zero_shift_is_plain_mc():
    g, e = fdmc.GBMParams(100, 0.05, 0.2, 1.0), fdmc.European(100, 0.05, 1.0)
    s = O.is_sums(g, e, 16, 0.0, 7, 0, 3000)
    m, v = O.run_mc_fixed(g, e, 16, 3000, 7)
    assert s[0] / 3000 == pytest.approx(m, rel=1e-12)
    assert (s[1] / 3000 - (s[0] / 3000) ** 2) / 3000 == pytest.approx(v, rel=1e-9)


@pytest.mark.parametrize("zs", [0.1, 0.5])
def test_is_likelihood_ratio_keeps_the_euler_mean(zs):
    # K = 0: the payoff is S_T e^{-rT}; under any shift the weighted mean is
    # E[S_T] e^{-rT} = S0 (1 + mu h)^n e^{-rT} of the Euler scheme
    g, e = fdmc.GBMParams(100, 0.05, 0.2, 1.0), fdmc.European(0.0, 0.05, 1.0)
    n, N = 8, 200000
    s = O.is_sums(g, e, n, zs, 11, 0x300, N)
    m = s[0] / N
    se = math.sqrt(max(s[1] / N - m * m, 0) / N)
    exact = 100 * (1 + 0.05 / n) ** n * math.exp(-0.05)
    assert abs(m - exact) < 4 * se, (m, exact, se)
    bad = O.is_sums(g, e, n, -zs, 11, 0x300, N)[0] / N   # the shift's sign matters for the draws only
    assert abs(bad - exact) < 6 * se


def _is_exact_moments(S0, K, mu, sigma, T, zs):
    """Mean and variance of the importance-sampled payoff in the continuous-time
    limit, by quadrature over the terminal normal: E_q[Y^2] = E_p[f^2 LR]."""
    f = lambda x: max(S0 * math.exp((mu - 0.5 * sigma ** 2) * T + sigma * math.sqrt(T) * x) - K, 0.0) \
        * math.exp(-mu * T)
    m = integrate.quad(lambda x: f(x) * stats.norm.pdf(x), -10, 12, points=[zs])[0]
    m2 = integrate.quad(lambda x: f(x) ** 2 * math.exp(-zs * x + zs * zs / 2) * stats.norm.pdf(x), -10, 12,
                        points=[zs])[0]
    return m, m2 - m * m


def test_is_deep_otm_matches_black_scholes_and_report():
    g, e = fdmc.GBMParams(100, 0.05, 0.2, 1.0), fdmc.European(180, 0.05, 1.0)
    zs = _z_star(100, 180, 0.05, 0.2, 1.0)
    assert ("%f" % zs)[:4] == "2.78"                 # printed truncated: informe_completo.txt:611
    bs = O.bs_call(100, 180, 1.0, 0.05, 0.2)
    rows = [x for x in REPORT if x["example"] == "ejemplo11_otm_is"]
    assert abs(bs - [x for x in rows if x["kind"] == "black_scholes"][0]["price"]) < 5e-7
    m_ex, v_ex = _is_exact_moments(100, 180, 0.05, 0.2, 1.0, zs)
    assert abs(m_ex - bs) < 1e-9
    # the oracle's per-path variance (10000-path pilot of the default MCConfig) is the exact one
    n = 500
    s = O.is_sums(g, e, n, zs / math.sqrt(n), 123, 0x300, 10000)
    assert abs((s[1] / 1e4 - (s[0] / 1e4) ** 2) / v_ex - 1) < 0.05
    for row in (x for x in rows if x["kind"] == "is_run"):
        r = O.run_is(g, e, zs, row["eps"], fdmc.MCConfig())
        # N = max(pilot_n, 2 Var / eps^2); at eps = 2e-4 the reference printed about
        # half of that (its pilot variance is half the exact 1.28e-3), so its
        # std error there is ~eps rather than eps / sqrt(2)
        want = 2 * v_ex / row["eps"] ** 2
        if want < 10000:
            assert r.n_samples == 10000
        else:
            assert abs(r.n_samples / want - 1) < 0.08, (r.n_samples, want)
        assert r.std_error < row["eps"] / math.sqrt(2) * 1.08
        assert abs(r.price - bs) < 3 * r.std_error + 1e-4, (r.price, bs, r.std_error)
        assert abs(r.price - row["price"]) < 3 * math.hypot(r.std_error, row["std_error"]) + 1e-4


def test_cv_pilot_beta_for_asian_geometric():
    g = fdmc.GBMParams(100, 0.05, 0.2, 1.0)
    E = fdmc.geom_asian_analytic(100, 100, 1, 0.05, 0.2, 64)
    p = O.cv_pilot(g, g, fdmc.Asian(100), fdmc.GeomAsian(100), E, 64, 20000, 0)
    assert 0.9 < p.beta < 1.2 and p.var_cv < 0.02 * p.var_plain


REPORT = json.load(open(os.path.join(HERE, "golden", "report_prices.json")))


def _build(entry):
    md, pd = entry["model"], entry["payoff"]
    if md["kind"] == "gbm":
        m = fdmc.GBMParams(md["S0"], md["mu"], md["sigma"], md["T"])
    elif md["kind"] == "heston":
        m = fdmc.HestonParams(md["S0"], md["mu"], md["kappa"], md["theta"], md["xi"], md["rho"], md["v0"],
                              md["T"]).compute_cholesky()
    else:
        m = fdmc.DupireLocalParams(md["S0"], md["mu"], md["sigma0"], md["alpha"], md["beta_d"], md["T"])
    k = pd["kind"]
    p = {"european": lambda: fdmc.European(pd["K"], pd["r"], pd["T"]), "asian": lambda: fdmc.Asian(pd["K"]),
         "lookback": lambda: fdmc.Lookback(pd["sigma"]),
         "barrier": lambda: fdmc.Barrier(pd["K"], pd["B"], pd["sigma"], pd["r"], pd["T"])}[k]()
    return m, p


@pytest.mark.parametrize("entry", [e for e in REPORT if e["kind"] == "mc500k"], ids=lambda e: e["example"])
def test_report_reference_prices(entry):
    """The reference's own printed prices (fixed-seed 500k-path runs, +-0.01-0.02)."""
    m, p = _build(entry)
    mean, var = O.run_mc_fixed(m, p, entry["n_steps"], 200000, 2024)
    assert abs(mean - entry["price"]) < 4 * math.sqrt(var + var * 200000 / 500000), (mean, entry["price"])


# ---- Heston variance schemes (methods_cuda.cu:276-284 and the optional Milstein term)
def _heston(scheme, xi=0.3, rho=-0.5):
    return fdmc.HestonParams(100.0, 0.05, 2.0, 0.09, xi, rho, 0.09, 1.0, scheme=scheme).compute_cholesky()


def _coarsen(dW, m):
    n_paths, n, _ = dW.shape
    return dW.reshape(n_paths, n // m, m, 2).sum(axis=2)


def test_heston_milstein_has_strong_order_one():
    """Strong convergence of V_T on shared Brownian paths (Feller ratio
    2 kappa theta / xi^2 = 4, so truncation is idle): RMS error against a
    2048-step Milstein solution falls about 4x from 16 to 64 steps with the
    Milstein term (order 1) and about 2x without it (order 1/2).  A wrong
    coefficient or sign of the term, or a dropped h, leaves order 1/2 or a bias."""
    rng = np.random.default_rng(11)
    n_paths, nf, T, rho = 3000, 2048, 1.0, -0.5
    z = rng.standard_normal((n_paths, nf, 2)) * math.sqrt(T / nf)
    dW = np.stack([z[..., 0], rho * z[..., 0] + math.sqrt(1 - rho * rho) * z[..., 1]], axis=-1)
    mil, eul = _heston(fdmc.HESTON_MILSTEIN), _heston(fdmc.HESTON_EXP_EULER)
    eur = fdmc.European(100.0, 0.05, 1.0)
    _, v_ref = O.heston_terminal(O.make_params(mil, eur, nf), dW, nf)

    def rms(model, n):
        _, v = O.heston_terminal(O.make_params(model, eur, n), _coarsen(dW, nf // n), n)
        return math.sqrt(np.mean((v.astype(np.float64) - v_ref) ** 2))

    em16, em64 = rms(mil, 16), rms(mil, 64)
    ee16, ee64 = rms(eul, 16), rms(eul, 64)
    assert em16 / em64 > 3.2, (em16, em64)
    assert 1.5 < ee16 / ee64 < 2.6, (ee16, ee64)
    assert em64 < ee64 / 2


def test_heston_milstein_is_the_reference_step_without_vol_of_vol():
    """xi = 0 makes the Milstein coefficient xi^2/4 zero: both schemes give the
    same paths bit for bit, and the deterministic variance theta + e^{-kappa t}(v0 - theta)."""
    rng = np.random.default_rng(3)
    n = 32
    dW = (rng.standard_normal((64, n, 2)) * math.sqrt(1.0 / n)).astype(np.float32)
    eur = fdmc.European(100.0, 0.05, 1.0)
    m0 = fdmc.HestonParams(100.0, 0.05, 2.0, 0.09, 0.0, -0.5, 0.04, 1.0, scheme=fdmc.HESTON_EXP_EULER)
    m1 = fdmc.HestonParams(100.0, 0.05, 2.0, 0.09, 0.0, -0.5, 0.04, 1.0, scheme=fdmc.HESTON_MILSTEIN)
    s0, v0 = O.heston_terminal(O.make_params(m0.compute_cholesky(), eur, n), dW, n)
    s1, v1 = O.heston_terminal(O.make_params(m1.compute_cholesky(), eur, n), dW, n)
    assert np.array_equal(s0, s1) and np.array_equal(v0, v1)
    assert np.allclose(v1, 0.09 + math.exp(-2.0) * (0.04 - 0.09), rtol=1e-5)


def test_heston_milstein_coupled_coarse_leg_uses_the_coarse_step():
    """The coupled MLMC path's coarse leg (M = 4) equals a 16-step path on the
    summed fine increments: the Milstein term of the coarse step takes the
    coarse increment and h_c, not h_f."""
    rng = np.random.default_rng(8)
    M, nf = 4, 64
    model = _heston(fdmc.HESTON_MILSTEIN, xi=0.6)
    eur = fdmc.European(0.0, 0.05, 1.0)
    Pf = O.make_params(model, eur, nf)
    Pc = O.make_params(model, eur, nf // M)
    sq = np.float32(math.sqrt(1.0 / nf))
    L10, L11 = np.float32(model.L[2]), np.float32(model.L[3])
    for _ in range(20):
        zz = rng.standard_normal((nf, 2)).astype(np.float32)
        _, pc = O.path_coupled(Pf, zz, nf, M)
        d1 = zz[:, 0] * sq
        d2 = (L10 * zz[:, 0] + L11 * zz[:, 1]) * sq
        acc = np.zeros((1, nf // M, 2), np.float32)
        for j in range(nf // M):       # sequential FP32 sums, as the coupled path accumulates
            a1 = a2 = np.float32(0.0)
            for k in range(M):
                a1 = np.float32(a1 + d1[j * M + k])
                a2 = np.float32(a2 + d2[j * M + k])
            acc[0, j] = (a1, a2)
        s_c, _ = O.heston_terminal(Pc, acc, nf // M)
        assert abs(pc - math.exp(-0.05) * float(s_c[0])) <= 2e-6 * abs(pc), (pc, s_c)


def test_closed_form_pca_matches_oracle_jacobi():
    """The test helper's closed-form increment matrix equals the oracle's Jacobi
    PCA up to eigenvector signs, and M M^T = h I."""
    from closed_forms import pca_uniform

    for m in (1, 2, 8, 64):
        M = pca_uniform(m, 1.0)
        Mo = O.pca_compute(m, 1.0)
        assert np.allclose(M @ M.T, np.eye(m) / m, atol=1e-13)
        assert np.allclose(np.abs(M), np.abs(Mo), atol=1e-10), m
