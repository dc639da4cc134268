"""GPU parity of the persistent multilevel kernel (MLMC / MLQMC).

The per-level sums are compared with the oracle's restatement of kernel_mlmc
(methods_cuda.cu:445-518, with the Heston running-payoff fix) on the same Philox
paths or scrambled Sobol points; the estimator loops (run_mlmc_cuda, :1305-1487;
run_mlqmc_cuda, :1680-2019) are compared with the oracle's loops and with
closed forms (Black-Scholes, Heston's semi-closed form).
"""
import math

import numpy as np
import pytest

import fdmc
import pyoracle as O
from closed_forms import heston_call, pca_uniform

pytestmark = pytest.mark.gpu

GBM = fdmc.GBMParams(100.0, 0.05, 0.2, 1.0)
HES = fdmc.HestonParams(100.0, 0.05, 2.0, 0.04, 0.5, -0.7, 0.04, 1.0).compute_cholesky()
HES_MIL = fdmc.HestonParams(100.0, 0.05, 2.0, 0.04, 0.5, -0.7, 0.04, 1.0,
                            scheme=fdmc.HESTON_MILSTEIN).compute_cholesky()
DUP = fdmc.DupireLocalParams(100.0, 0.05, 0.2, 0.5, 0.7, 1.0)
LVQ = fdmc.LocalVolQuadParams(100.0, 0.05, 0.2, 0.1, 1.0, 100.0, 1.0)
LVQ_STEEP = fdmc.LocalVolQuadParams(100.0, 0.05, 1.0, 0.05, 1.0, 100.0, 1.0)   # many paths reach S <= 0
EURO = fdmc.European(100.0, 0.05, 1.0)
ASIAN = fdmc.Asian(100.0)
LOOK = fdmc.Lookback(0.2)
BAR = fdmc.Barrier(100.0, 130.0, 0.2, 0.05, 1.0)


def _check_sums(got, ref, n, barrier):
    got, ref = np.asarray(got) / n, np.asarray(ref) / n
    tol0 = 2e-5 * np.maximum(1.0, np.abs(ref[2])) + 1e-6 + (4 * 40.0 / n if barrier else 0.0)
    assert abs(got[0] - ref[0]) <= tol0, (got, ref)
    assert abs(got[2] - ref[2]) <= tol0, (got, ref)
    assert abs(got[3] - ref[3]) <= 1e-4 * abs(ref[3]) + (4 * 1600.0 / n if barrier else 1e-6), (got, ref)


@pytest.mark.parametrize("model,payoff,M,levels", [
    (HES, EURO, 4, (0, 1, 2, 3, 4)),
    (HES, BAR, 4, (0, 1, 3)),
    (HES, BAR, 4, (4, 5, 6, 7, 8)),
    (HES_MIL, EURO, 4, (0, 1, 2, 4)),
    (GBM, ASIAN, 2, (0, 1, 2, 5, 7)),
    (DUP, LOOK, 2, (1, 4)),
    (LVQ, EURO, 4, (2, 3)),
    (LVQ_STEEP, LOOK, 2, (0, 1, 2)),
], ids=["heston-euro-M4", "heston-barrier-M4", "heston-barrier-M4-deep", "heston-milstein-euro-M4", "gbm-asian-M2", "dupire-look-M2",
        "lvq-euro-M4", "lvq-steep-look-M2"])
def test_mlmc_level_sums_parity(engine, model, payoff, M, levels):
    reqs = [(l, 0, 1000 * l, 700 + 37 * l) for l in levels]   # unaligned ranges
    got = engine.levels_accumulate(model, payoff, M, reqs, qmc=False, salt_or_seed=42)
    for (l, _, first, n), g in zip(reqs, got):
        print("level %d: %d paths of %d fine steps" % (l, n, M ** l), flush=True)
        ref = O.mlmc_level_sums(model, payoff, M, l, 42, first, n)
        _check_sums(g, ref, n, isinstance(payoff, fdmc.Barrier))


@pytest.mark.parametrize("model,payoff,M,levels", [
    (GBM, ASIAN, 2, (3, 11)),
    (LVQ, LOOK, 4, (6,)),
], ids=["gbm-asian-M2-8+2048", "lvq-look-M4-4096"])
def test_mlqmc_bridge_levels_above_1024_dates(engine, model, payoff, M, levels):
    """Bridge levels of 2048 and 4096 fine dates (the skeleton is walked on a
    per-lane stack, direction records come from global memory above 1024
    dimensions), alone and in one launch with a short level: level sums equal
    the oracle's at the same points."""
    R, n = 2, 513
    bbl = {l: engine.bb_upload(fdmc.bb_precompute(M ** l, 1.0)) for l in levels}
    lst = [bbl.get(l) for l in range(max(levels) + 1)]
    reqs = [(l, r, 300 * r, n) for l in levels for r in range(R)]
    got = engine.levels_accumulate(model, payoff, M, reqs, qmc=True, mode=fdmc.NOISE_BB, R=R, salt_or_seed=42,
                                   bb_list=lst)
    for b in bbl.values():
        b.free()
    for (l, r, first, cnt), g in zip(reqs, got):
        ref = O.mlqmc_level_sums(model, payoff, M, l, R, 42, first, cnt, fdmc.NOISE_BB, bb=O.OracleBB(M ** l, 1.0))
        print("level %d replica %d" % (l, r), flush=True)
        _check_sums(g, ref[r], cnt, False)


@pytest.mark.parametrize("model,payoff,M,mode", [
    (GBM, ASIAN, 2, fdmc.NOISE_BB),
    (DUP, EURO, 2, fdmc.NOISE_RAW),
    (HES, EURO, 4, fdmc.NOISE_RAW),
    (HES_MIL, ASIAN, 4, fdmc.NOISE_RAW),
    (LVQ, LOOK, 2, fdmc.NOISE_BB),
], ids=["gbm-asian-bb", "dupire-euro-raw", "heston-euro-raw", "heston-milstein-asian-raw", "lvq-look-bb"])
def test_mlqmc_level_sums_parity(engine, model, payoff, M, mode):
    R, L = 3, 5
    bbl = [engine.bb_upload(fdmc.bb_precompute(M ** l, 1.0)) for l in range(L + 1)] if mode == fdmc.NOISE_BB else []
    reqs = [(l, r, 100 * r, 513) for l in (0, 2, L) for r in range(R)]
    got = engine.levels_accumulate(model, payoff, M, reqs, qmc=True, mode=mode, R=R, salt_or_seed=42, bb_list=bbl)
    for b in bbl:
        b.free()
    obbs = {l: O.OracleBB(M ** l, 1.0) for l in range(L + 1)} if mode == fdmc.NOISE_BB else {}
    for l in (0, 2, L):
        ref = O.mlqmc_level_sums(model, payoff, M, l, R, 42, 0, 513, mode, bb=obbs.get(l))
        ref2 = O.mlqmc_level_sums(model, payoff, M, l, R, 42, 100, 513, mode, bb=obbs.get(l))
        ref3 = O.mlqmc_level_sums(model, payoff, M, l, R, 42, 200, 513, mode, bb=obbs.get(l))
        rows = {0: ref[0], 1: ref2[1], 2: ref3[2]}
        for r in range(R):
            g = got[[i for i, q in enumerate(reqs) if q[0] == l and q[1] == r][0]]
            _check_sums(g, rows[r], 513, False)


@pytest.mark.parametrize("model,payoff,M,bf16", [
    (GBM, ASIAN, 2, False),
    (LVQ, EURO, 4, False),
    (DUP, LOOK, 2, True),
], ids=["gbm-asian-fp16", "lvq-euro-fp16", "dupire-look-bf16"])
def test_mlqmc_pca_level_sums_parity(engine, model, payoff, M, bf16):
    """PCA noise per level (methods_cuda.cu:1907-1913): the fine increments are
    dW = M_l z on matrix cores with FP16/BF16 operands; the oracle applies M_l in
    FP64.  Replicas with their own ranges run one GEMM each, replicas sharing a
    range one GEMM together."""
    R, L = 3, 4
    levels = (0, 2, L)
    mats = {l: O.pca_compute(M ** l, 1.0) for l in range(L + 1)}
    pcl = [engine.pca_upload(M ** l, mats[l].T.ravel()) for l in range(L + 1)]
    reqs = [(l, r, 100 * r, 513) for l in levels for r in range(R)] + \
           [(l, r, 1000, 600) for l in levels for r in range(R)]
    mode = fdmc.NOISE_PCA | (fdmc.PCA_BF16 if bf16 else 0)
    got = engine.levels_accumulate(model, payoff, M, reqs, qmc=True, mode=mode, R=R, salt_or_seed=42,
                                   pca_list=pcl)
    for x in pcl:
        x.free()
    tol = 1e-2 if bf16 else 2e-3     # operand quantisation of the increments
    for i, (l, r, first, n) in enumerate(reqs):
        ref = O.mlqmc_level_sums(model, payoff, M, l, R, 42, first, n, fdmc.NOISE_PCA, pca_M=mats[l])[r] / n
        g = np.asarray(got[i]) / n
        scale = max(1.0, abs(ref[2]))
        assert abs(g[0] - ref[0]) <= tol * scale, (l, r, g, ref)
        assert abs(g[2] - ref[2]) <= tol * scale, (l, r, g, ref)
        assert abs(g[3] - ref[3]) <= 2 * tol * max(1.0, abs(ref[3])), (l, r, g, ref)


def test_run_mlqmc_pca_matches_oracle_loop_and_bridge(engine):
    """A run that converges with a decisive bias test: the Asian's level
    corrections halve per level (E_7 ~ 0.038, E_8 ~ 0.019 against eps / sqrt(2) =
    0.028 at eps = 0.04), so the loop stops at L = 9 of max_L = 10 whatever the
    FP16 quantisation of the increments does to the estimates.  The loop with
    per-level PCA noise on matrix-core FP16 increments against the oracle's loop
    with the FP64 transform, and against the bridge run."""
    ml = fdmc.MLMCConfig(M=2, max_L=10, pilot_n=256)
    qc = fdmc.QMCConfig(R=8)
    mats = [pca_uniform(2 ** l, 1.0) for l in range(11)]
    pcl = [engine.pca_upload(2 ** l, m.T.ravel()) for l, m in enumerate(mats)]
    print("running the device loop", flush=True)
    g = engine.run_mlqmc_cuda(GBM, ASIAN, 0.04, ml, qc, fdmc.NOISE_PCA, pca_list=pcl)
    for x in pcl:
        x.free()
    bbl = [engine.bb_upload(fdmc.bb_precompute(2 ** l, 1.0)) for l in range(11)]
    gb = engine.run_mlqmc_cuda(GBM, ASIAN, 0.04, ml, qc, fdmc.NOISE_BB, bbl)
    for b in bbl:
        b.free()
    r = O.run_mlqmc(GBM, ASIAN, 0.04, ml, qc, fdmc.NOISE_PCA, pcas=mats)
    print("gpu", g, "oracle", r, "gpu bridge", gb, flush=True)
    assert g.flags == 0 and r.flags == 0 and g.levels == r.levels == 10
    assert abs(g.n_samples - r.n_samples) <= 0.02 * r.n_samples
    assert abs(g.path_steps - r.path_steps) <= 0.02 * r.path_steps
    assert abs(g.price - r.price) < 0.5 * r.std_error
    assert abs(g.price - gb.price) < 3 * math.hypot(g.std_error, gb.std_error)
    assert g.std_error < 0.04 / math.sqrt(2) * 1.01


def test_run_mlmc_matches_oracle_loop(engine):
    cfg = fdmc.MLMCConfig(M=2, max_L=6, pilot_n=2000, seed=7)
    g = engine.run_mlmc_cuda(GBM, EURO, 2e-2, cfg)
    r = O.run_mlmc(GBM, EURO, 2e-2, cfg)
    assert g.levels == r.levels
    assert abs(g.n_samples - r.n_samples) <= 0.01 * r.n_samples
    assert abs(g.price - r.price) < 0.2 * r.std_error + 1e-4


def test_run_mlqmc_matches_oracle_loop(engine):
    """A bridge run that converges (L = 8 of max_L = 10, every level allocated to
    the variance target): levels, sample counts and work agree with the oracle's
    loop to 2 % (allocations ceil(N_opt) of variance estimates whose last digits
    differ between the FP32 paths of the two implementations), prices within half
    a standard error (the counts select slightly different point sets)."""
    ml = fdmc.MLMCConfig(M=2, max_L=10, pilot_n=256)
    qc = fdmc.QMCConfig(R=8)
    bbl = [engine.bb_upload(fdmc.bb_precompute(2 ** l, 1.0)) for l in range(11)]
    g = engine.run_mlqmc_cuda(GBM, EURO, 5e-3, ml, qc, fdmc.NOISE_BB, bbl)
    for b in bbl:
        b.free()
    r = O.run_mlqmc(GBM, EURO, 5e-3, ml, qc, fdmc.NOISE_BB, [O.OracleBB(2 ** l, 1.0) for l in range(11)])
    print("gpu", g, "oracle", r, flush=True)
    assert g.flags == 0 and r.flags == 0 and g.levels == r.levels
    assert abs(g.n_samples - r.n_samples) <= 0.02 * r.n_samples
    assert abs(g.path_steps - r.path_steps) <= 0.02 * r.path_steps
    assert abs(g.price - r.price) < 0.5 * r.std_error
    bs = fdmc.bs_call(100, 100, 1, 0.05, 0.2)
    assert abs(g.price - bs) < 3 * g.std_error + 5e-3 / math.sqrt(2)


def test_run_mlqmc_asian_at_level_cap_matches_oracle_loop(engine):
    """A run that ends at its iteration cap (Asian, eps = 5e-3, max_L = 6): the
    allocation compounds last-digit differences of the variance estimates over
    the loop, so counts agree to 10 % here (the per-point level sums are exact
    above, and the converging run agrees to 2 %)."""
    ml = fdmc.MLMCConfig(M=2, max_L=6, pilot_n=256)
    qc = fdmc.QMCConfig(R=8)
    bbl = [engine.bb_upload(fdmc.bb_precompute(2 ** l, 1.0)) for l in range(7)]
    g = engine.run_mlqmc_cuda(GBM, ASIAN, 5e-3, ml, qc, fdmc.NOISE_BB, bbl)
    for b in bbl:
        b.free()
    r = O.run_mlqmc(GBM, ASIAN, 5e-3, ml, qc, fdmc.NOISE_BB, [O.OracleBB(2 ** l, 1.0) for l in range(7)])
    assert g.levels == r.levels
    assert abs(g.n_samples - r.n_samples) <= 0.1 * r.n_samples
    assert abs(g.price - r.price) < 1.0 * r.std_error


def test_mlmc_gbm_european_vs_black_scholes(engine):
    r = engine.run_mlmc_cuda(GBM, EURO, 5e-3, fdmc.MLMCConfig(M=2, max_L=10, pilot_n=2000))
    bs = fdmc.bs_call(100, 100, 1, 0.05, 0.2)
    assert abs(r.price - bs) < 3 * r.std_error + 5e-3 / math.sqrt(2)


def test_c4_heston_mlmc_european_vs_closed_form(engine):
    """Config 4: Heston MLMC, M=4, L <= 8, eps = 5e-4, pilot 1e4 per level."""
    cfg = fdmc.MLMCConfig(M=4, max_L=8, pilot_n=10000, seed=42)
    r = engine.run_mlmc_cuda(HES, EURO, 5e-4, cfg)
    ref = heston_call(100, 100, 1, 0.05, 2.0, 0.04, 0.5, -0.7, 0.04)
    assert r.std_error < 5e-4 / math.sqrt(2) * 1.01
    assert abs(r.price - ref) < 3 * r.std_error + 5e-4 / math.sqrt(2), (r.price, ref, r.std_error, r.levels)


def test_heston_milstein_mlmc_vs_closed_form(engine):
    """The optional Milstein variance step: run_mlmc reaches the semi-closed-form
    Heston price (Milstein keeps Euler's weak order 1 in S) at eps = 1e-3, and
    its level variances of the coupled V-dependent payoff are no larger than the
    reference scheme's."""
    cfg = fdmc.MLMCConfig(M=4, max_L=8, pilot_n=10000, seed=42)
    r = engine.run_mlmc_cuda(HES_MIL, EURO, 1e-3, cfg)
    ref = heston_call(100, 100, 1, 0.05, 2.0, 0.04, 0.5, -0.7, 0.04)
    assert r.flags == 0 and r.std_error < 1e-3 / math.sqrt(2) * 1.01
    assert abs(r.price - ref) < 3 * r.std_error + 1e-3 / math.sqrt(2), (r.price, ref, r.std_error, r.levels)
    n = 200_000
    for l in (3, 4):
        vm = engine.levels_accumulate(HES_MIL, EURO, 4, [(l, 0, 0, n)], qmc=False, salt_or_seed=7)[0]
        ve = engine.levels_accumulate(HES, EURO, 4, [(l, 0, 0, n)], qmc=False, salt_or_seed=7)[0]
        var = lambda s: s[1] / n - (s[0] / n) ** 2
        assert var(vm) <= 1.1 * var(ve), (l, vm, ve)


def test_c4_heston_mlmc_barrier_vs_single_level(engine):
    """Config 4 barrier: the coupled kernel applies the running maximum for Heston
    (the reference never updates it there, methods_cuda.cu:464-470, so its
    barrier never knocks out).  With the finest level capped at L = 4 (4^4 = 256
    steps) the telescoping sum estimates the 256-step BGK-corrected price, which
    single-level MC at 256 steps estimates independently."""
    euro = heston_call(100, 100, 1, 0.05, 2.0, 0.04, 0.5, -0.7, 0.04)
    r = engine.run_mlmc_cuda(HES, BAR, 2e-3, fdmc.MLMCConfig(M=4, max_L=4, pilot_n=10000, seed=42))
    assert r.levels == 5
    m, v = engine.run_mc_fixed(HES, BAR, 256, 4_000_000, 5)
    assert r.price < euro - 1.0          # knock-out is active
    assert abs(r.price - m) < 3 * math.sqrt(r.std_error ** 2 + v), (r.price, m)


def test_c4_heston_mlmc_barrier_at_config(engine):
    """Config 4's up-and-out leg at its own settings (M = 4, L <= 8, eps = 5e-4,
    pilot 1e4; run_mlmc_cuda, methods_cuda.cu:1305-1487).  The BGK-corrected
    barrier's level corrections decay slowly (E_5..E_8 ~ 0.04-0.07, 0.03, 0.007),
    so the bias test |E_L| / (M - 1) < eps / sqrt(2) cannot fire by L = 8: the run
    ends at the level cap with the not-converged flag, every level allocated to
    the variance target.  Its telescoping sum estimates the 65,536-step price,
    which equals the 256-step single-level price plus the oracle's E_5..E_8 on
    independent Philox streams."""
    cfg = fdmc.MLMCConfig(M=4, max_L=8, pilot_n=10000, seed=42)
    r = engine.run_mlmc_cuda(HES, BAR, 5e-4, cfg)
    print("barrier MLMC:", r.price, r.std_error, r.levels, r.flags, r.n_samples, flush=True)
    assert r.levels == 9
    assert r.flags & fdmc.FLAG_NOT_CONVERGED
    assert r.std_error < 5e-4 / math.sqrt(2) * 1.01
    m256, v256 = engine.run_mc_fixed(HES, BAR, 256, 4_000_000, 5)
    corr, cvar = 0.0, 0.0
    for l, n in ((5, 80000), (6, 40000), (7, 20000), (8, 10000)):
        s = O.mlmc_level_sums(HES, BAR, 4, l, 1234, 0, n)
        m = s[0] / n
        corr += m
        cvar += (s[1] / n - m * m) / n
    print("256-step price %.5f + oracle E5..E8 %.5f +- %.5f" % (m256, corr, math.sqrt(cvar)), flush=True)
    assert 0.05 < corr < 0.3                     # the bias the loop could not remove below L = 8
    assert abs(r.price - (m256 + corr)) < 3 * math.sqrt(r.std_error ** 2 + v256 + cvar), (r.price, m256, corr)


def test_pca_level_request_above_8_gib_runs_in_chunks(engine):
    """One (level, replica) PCA request of 1.5M points on a 2048-date level needs
    18 GB of normals and GEMM rows: it runs in chunks of a multiple of 256 points
    with at most 8 GiB of buffers each (the reference batches such a level by free
    memory, methods_cuda.cu:1791-1803), and its sums equal those of three requests
    of 500,000 points over the same range (each below 8 GiB: one pass each) up to
    FP64 summation order."""
    pca = engine.pca_upload(2048, pca_uniform(2048, 1.0).T.ravel())
    lst = [None] * 11 + [pca]
    try:
        print("one request of 1.5M points", flush=True)
        big = engine.levels_accumulate(GBM, ASIAN, 2, [(11, 0, 1000, 1_500_000)], qmc=True, mode=fdmc.NOISE_PCA,
                                       R=1, salt_or_seed=42, pca_list=lst)[0]
        print("three requests of 500k points", flush=True)
        parts = engine.levels_accumulate(GBM, ASIAN, 2, [(11, 0, 1000 + k * 500_000, 500_000) for k in range(3)],
                                         qmc=True, mode=fdmc.NOISE_PCA, R=1, salt_or_seed=42, pca_list=lst)
    finally:
        pca.free()
    tot = np.sum(np.array(parts), axis=0)
    assert np.allclose(big, tot, rtol=1e-10, atol=1e-9), (big, tot)
    assert abs(big[2] / 1.5e6 - 6.06) < 0.05          # the 2048-date Asian's mean payoff
