"""Seeded random parity sweep of the fused path kernels: models (GBM, Dupire,
quadratic local vol, Heston), payoffs (European, Asian, geometric Asian,
lookback, barrier), noise (raw Sobol, Brownian bridge), odd step counts,
replicas, unaligned point ranges and parameters drawn from a fixed numpy
generator, each pass against the oracle's restatement of kernel_mc
(methods_cuda.cu:414-437) on the same scrambled Sobol points, to the
tolerances of the fixed-case tests (replica means within 2e-5 relative for
Lipschitz payoffs, a few max-payoff/N more for a barrier's indicator flips).
The same generator draws multilevel requests, baskets, Greek cases and a
contract book, each checked against its oracle or CRN differences.
"""
import numpy as np
import pytest

import fdmc
import pyoracle as O
from book_parity import check_greek_sums

pytestmark = pytest.mark.gpu


def _cases(n=24, seed=20261204):
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n):
        T = float(rng.choice([0.25, 0.5, 1.0, 2.0]))
        mu = float(rng.uniform(0.0, 0.08))
        kind = int(rng.integers(4))
        if kind == 0:
            model = fdmc.GBMParams(100.0, mu, float(rng.uniform(0.1, 0.5)), T)
        elif kind == 1:
            model = fdmc.DupireLocalParams(100.0, mu, float(rng.uniform(0.1, 0.3)), float(rng.uniform(0.0, 1.0)),
                                           float(rng.uniform(0.5, 1.0)), T)
        elif kind == 2:
            model = fdmc.LocalVolQuadParams(100.0, mu, float(rng.uniform(0.1, 0.3)), float(rng.uniform(0.0, 0.2)),
                                            float(rng.uniform(0.0, 2.0)), 100.0, T)
        else:
            model = fdmc.HestonParams(100.0, mu, float(rng.uniform(0.5, 3.0)), float(rng.uniform(0.02, 0.09)),
                                      float(rng.uniform(0.1, 0.7)), float(rng.uniform(-0.9, 0.2)),
                                      float(rng.uniform(0.02, 0.09)), T).compute_cholesky()
        K = float(rng.uniform(70.0, 130.0))
        pk = int(rng.integers(5))
        payoff = [fdmc.European(K, mu, T), fdmc.Asian(K), fdmc.GeomAsian(K), fdmc.Lookback(0.2),
                  fdmc.Barrier(K, K * float(rng.uniform(1.1, 1.5)), 0.2, mu, T)][pk]
        bb = kind != 3 and rng.random() < 0.5
        n_steps = int(rng.choice([1, 2, 4, 8, 16, 32, 64, 128, 256])) if bb else \
            int(rng.choice([1, 2, 3, 5, 7, 16, 33, 64, 100, 255]))
        R = int(rng.integers(1, 4))
        first = int(rng.integers(0, 5000))
        count = int(rng.integers(64, 1500))
        out.append(("c%02d-%s-%s-%s-%d" % (i, ["gbm", "dup", "lvq", "hes"][kind],
                                           ["eur", "asi", "geo", "lkb", "bar"][pk], "bb" if bb else "raw", n_steps),
                    model, payoff, n_steps, fdmc.NOISE_BB if bb else fdmc.NOISE_RAW, R, first, count))
    return out


CASES = _cases()


def _long_cases(n=6, seed=20261205):
    """Long paths: bridges of 2048-8192 dates (direction records from global
    memory) on one-factor models, raw Heston up to 12000 dimensions."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n):
        T = float(rng.choice([0.5, 1.0, 2.0]))
        mu = float(rng.uniform(0.0, 0.08))
        kind = i % 4
        if kind == 0:
            model = fdmc.GBMParams(100.0, mu, float(rng.uniform(0.1, 0.5)), T)
        elif kind == 1:
            model = fdmc.DupireLocalParams(100.0, mu, float(rng.uniform(0.1, 0.3)), float(rng.uniform(0.0, 1.0)),
                                           float(rng.uniform(0.5, 1.0)), T)
        elif kind == 2:
            model = fdmc.LocalVolQuadParams(100.0, mu, float(rng.uniform(0.1, 0.3)), float(rng.uniform(0.0, 0.2)),
                                            float(rng.uniform(0.0, 2.0)), 100.0, T)
        else:
            model = fdmc.HestonParams(100.0, mu, float(rng.uniform(0.5, 3.0)), float(rng.uniform(0.02, 0.09)),
                                      float(rng.uniform(0.1, 0.7)), float(rng.uniform(-0.9, 0.2)),
                                      float(rng.uniform(0.02, 0.09)), T, scheme=int(rng.integers(2))).compute_cholesky()
        K = float(rng.uniform(80.0, 120.0))
        pk = int(rng.integers(5))
        payoff = [fdmc.European(K, mu, T), fdmc.Asian(K), fdmc.GeomAsian(K), fdmc.Lookback(0.2),
                  fdmc.Barrier(K, K * float(rng.uniform(1.2, 1.5)), 0.2, mu, T)][pk]
        bb = kind != 3
        n_steps = int(rng.choice([2048, 4096, 8192])) if bb else int(rng.choice([1500, 3000, 6000]))
        out.append(("l%02d-%s-%s-%s-%d" % (i, ["gbm", "dup", "lvq", "hes"][kind],
                                           ["eur", "asi", "geo", "lkb", "bar"][pk], "bb" if bb else "raw", n_steps),
                    model, payoff, n_steps, fdmc.NOISE_BB if bb else fdmc.NOISE_RAW, 2,
                    int(rng.integers(0, 5000)), int(rng.integers(200, 600))))
    return out


LONG_CASES = _long_cases()


@pytest.mark.parametrize("name,model,payoff,n_steps,mode,R,first,count", CASES, ids=[c[0] for c in CASES])
def test_random_pass_parity(engine, name, model, payoff, n_steps, mode, R, first, count):
    bb = engine.bb_upload(fdmc.bb_precompute(n_steps, model.T)) if mode == fdmc.NOISE_BB else None
    got = np.array(engine.qmc_accumulate(model, payoff, n_steps, R, 77, first, count, mode, dev_bb=bb))
    if bb is not None:
        bb.free()
    obb = O.OracleBB(n_steps, model.T) if mode == fdmc.NOISE_BB else None
    ref = O.qmc_accumulate(model, payoff, n_steps, R, 77, first, count, mode, bb=obb)
    g, r = got[:, 0] / count, ref[:, 0] / count
    tol = 2e-5 * np.maximum(1.0, np.abs(r)) + 1e-6
    if isinstance(payoff, fdmc.Barrier):
        tol = tol + 4.0 * 40.0 / count
    assert (np.abs(g - r) <= tol).all(), (name, g, r)
    g2, r2 = got[:, 1] / count, ref[:, 1] / count
    tol2 = 1e-4 * np.maximum(1.0, np.abs(r2)) + (4.0 * 1600.0 / count if isinstance(payoff, fdmc.Barrier) else 1e-6)
    assert (np.abs(g2 - r2) <= tol2).all(), (name, g2, r2)


@pytest.mark.parametrize("name,model,payoff,n_steps,mode,R,first,count", LONG_CASES, ids=[c[0] for c in LONG_CASES])
def test_random_long_path_parity(engine, name, model, payoff, n_steps, mode, R, first, count):
    test_random_pass_parity(engine, name, model, payoff, n_steps, mode, R, first, count)


def _level_cases(n=16, seed=1204):
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n):
        T = float(rng.choice([0.5, 1.0, 2.0]))
        mu = float(rng.uniform(0.0, 0.08))
        kind = int(rng.integers(4))
        if kind == 0:
            model = fdmc.GBMParams(100.0, mu, float(rng.uniform(0.1, 0.5)), T)
        elif kind == 1:
            model = fdmc.DupireLocalParams(100.0, mu, float(rng.uniform(0.1, 0.3)), float(rng.uniform(0.0, 1.0)),
                                           float(rng.uniform(0.5, 1.0)), T)
        elif kind == 2:
            model = fdmc.LocalVolQuadParams(100.0, mu, float(rng.uniform(0.1, 0.3)), float(rng.uniform(0.0, 0.2)),
                                            float(rng.uniform(0.0, 2.0)), 100.0, T)
        else:
            model = fdmc.HestonParams(100.0, mu, float(rng.uniform(0.5, 3.0)), float(rng.uniform(0.02, 0.09)),
                                      float(rng.uniform(0.1, 0.7)), float(rng.uniform(-0.9, 0.2)),
                                      float(rng.uniform(0.02, 0.09)), T).compute_cholesky()
        K = float(rng.uniform(80.0, 120.0))
        pk = int(rng.integers(4))   # Lipschitz payoffs: level sums to the fixed-case tolerances
        payoff = [fdmc.European(K, mu, T), fdmc.Asian(K), fdmc.GeomAsian(K), fdmc.Lookback(0.2)][pk]
        M = int(rng.choice([2, 4]))
        lmax = 5 if M == 2 else 3
        qmc = bool(rng.random() < 0.6)
        mode = fdmc.NOISE_BB if (qmc and kind != 3 and rng.random() < 0.5) else fdmc.NOISE_RAW
        levels = sorted(set(int(x) for x in rng.integers(0, lmax + 1, size=3)))
        R = int(rng.integers(1, 4)) if qmc else 1
        first = int(rng.integers(0, 3000))
        count = int(rng.integers(100, 900))
        out.append(("l%02d-%s-%s-M%d-%s" % (i, ["gbm", "dup", "lvq", "hes"][kind], ["eur", "asi", "geo", "lkb"][pk], M,
                                             ("bb" if mode == fdmc.NOISE_BB else "sobol") if qmc else "prng"),
                    model, payoff, M, qmc, mode, levels, R, first, count))
    return out


LEVEL_CASES = _level_cases()


@pytest.mark.parametrize("name,model,payoff,M,qmc,mode,levels,R,first,count", LEVEL_CASES,
                         ids=[c[0] for c in LEVEL_CASES])
def test_random_level_sums_parity(engine, name, model, payoff, M, qmc, mode, levels, R, first, count):
    """Coupled fine/coarse level sums of the persistent multilevel kernel
    (kernel_mlmc, methods_cuda.cu:445-518) against the oracle on the same
    Philox streams or scrambled Sobol points."""
    bbl = [engine.bb_upload(fdmc.bb_precompute(M ** l, model.T)) for l in range(max(levels) + 1)] \
        if mode == fdmc.NOISE_BB else []
    reqs = [(l, r, first, count) for l in levels for r in range(R)]
    got = engine.levels_accumulate(model, payoff, M, reqs, qmc=qmc, mode=mode, R=R, salt_or_seed=91, bb_list=bbl)
    for b in bbl:
        b.free()
    for (l, r, f, n), g in zip(reqs, got):
        if qmc:
            obb = O.OracleBB(M ** l, model.T) if mode == fdmc.NOISE_BB else None
            ref = O.mlqmc_level_sums(model, payoff, M, l, R, 91, f, n, mode, bb=obb)[r]
        else:
            ref = O.mlmc_level_sums(model, payoff, M, l, 91, f, n)
        g, ref = np.asarray(g) / n, np.asarray(ref) / n
        tol = 2e-5 * max(1.0, abs(ref[2])) + 1e-6
        assert abs(g[0] - ref[0]) <= tol and abs(g[2] - ref[2]) <= tol, (name, l, r, g, ref)
        assert abs(g[3] - ref[3]) <= 1e-4 * abs(ref[3]) + 1e-6, (name, l, r, g, ref)


def _greek_cases(n=8, seed=4242):
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n):
        kind = int(rng.integers(3))
        T = float(rng.choice([0.5, 1.0, 2.0]))
        mu0 = float(rng.uniform(0.01, 0.08))
        v0 = float(rng.uniform(0.15, 0.35))
        extra = (float(rng.uniform(0.0, 1.0)), float(rng.uniform(0.5, 1.0))) if kind == 1 else \
            (float(rng.uniform(0.0, 0.2)), float(rng.uniform(0.0, 2.0)))

        def make(S0=100.0, mu=mu0, vol=v0, kind=kind, T=T, extra=extra):
            if kind == 0:
                return fdmc.GBMParams(S0, mu, vol, T)
            if kind == 1:
                return fdmc.DupireLocalParams(S0, mu, vol, extra[0], extra[1], T)
            return fdmc.LocalVolQuadParams(S0, mu, vol, extra[0], extra[1], 100.0, T)

        K = float(rng.uniform(85.0, 115.0))
        pk = int(rng.integers(4))
        bb = bool(rng.random() < 0.5)
        n_steps = int(rng.choice([8, 16, 32, 64])) if bb else int(rng.choice([5, 12, 31, 64]))
        out.append(("g%02d-%s-%s-%s-%d" % (i, ["gbm", "dup", "lvq"][kind], ["eur", "asi", "geo", "lkb"][pk],
                                            "bb" if bb else "raw", n_steps),
                    make, pk, K, mu0, v0, T, n_steps, fdmc.NOISE_BB if bb else fdmc.NOISE_RAW))
    return out


GREEK_CASES = _greek_cases()


def _payoff(pk, K, r, T):
    return [fdmc.European(K, r, T), fdmc.Asian(K), fdmc.GeomAsian(K), fdmc.Lookback(0.2)][pk]


@pytest.mark.parametrize("name,make,pk,K,mu0,v0,T,n_steps,mode", GREEK_CASES, ids=[c[0] for c in GREEK_CASES])
def test_random_adjoint_vs_crn_differences(engine, name, make, pk, K, mu0, v0, T, n_steps, mode):
    """Pathwise-adjoint delta, vega and rho against common-random-number central
    differences of the price on the same Sobol points (rho moves mu and r)."""
    R, n = 2, 1 << 13
    bb = engine.bb_upload(fdmc.bb_precompute(n_steps, T)) if mode == fdmc.NOISE_BB else None

    def price(model, r):
        return np.array(engine.qmc_accumulate(model, _payoff(pk, K, r, T), n_steps, R, 42, 0, n, mode,
                                              dev_bb=bb))[:, 0] / n

    rows = np.array(engine.greeks_accumulate(make(), _payoff(pk, K, mu0, T), n_steps, R, 42, 0, n, mode, dev_bb=bb))
    bumps = {"delta": 0.05, "vega": 0.002, "rho": 0.002}
    for k, which in enumerate(("delta", "vega", "rho")):
        h = bumps[which]
        if which == "delta":
            up, dn = price(make(S0=100.0 + h), mu0), price(make(S0=100.0 - h), mu0)
        elif which == "vega":
            up, dn = price(make(vol=v0 + h), mu0), price(make(vol=v0 - h), mu0)
        else:
            up, dn = price(make(mu=mu0 + h), mu0 + h), price(make(mu=mu0 - h), mu0 - h)
        fd = (up - dn) / (2 * h)
        adj = rows[:, 2 + 2 * k] / n
        assert np.all(np.abs(adj - fd) <= 0.01 * np.abs(fd) + 2e-3), (name, which, adj, fd)
    if bb:
        bb.free()


def _basket_cases(n=12, seed=777):
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n):
        na = int(rng.choice([2, 3, 5, 8, 13, 16, 32, 40, 64]))
        steps = int(rng.choice([1, 4, 12, 32]))
        gbm = bool(rng.random() < 0.5)
        corr = bool(rng.random() < 0.6)
        L = None
        if corr:
            Bf = rng.uniform(0.2, 0.5, (na, 3))
            Cm = Bf @ Bf.T
            np.fill_diagonal(Cm, 1.0)
            L = np.linalg.cholesky(Cm).ravel()
        S0 = list(rng.uniform(80.0, 120.0, na))
        if gbm:
            model = fdmc.MultiGBMParams(na, S0, list(rng.uniform(0.1, 0.4, na)), 0.05, 1.0, L=L, uncorrelated=not corr)
        else:
            model = fdmc.MultiDupireParams(na, 0.05, 0.2, 0.5, 0.7, 1.0, uncorrelated=not corr, S0=S0, L=L)
        w = list(rng.dirichlet(np.ones(na)))
        payoff = fdmc.Basket(float(rng.uniform(90.0, 110.0)), 0.05, 1.0, na, weights=w)
        prng = bool(rng.random() < 0.3) or na % 4 == 0 and rng.random() < 0.2
        if prng and na <= 32 and na % 4:
            prng = False   # pseudo-random baskets up to 32 assets need a multiple of 4
        out.append(("b%02d-%s-n%d-s%d-%s-%s" % (i, "gbm" if gbm else "dup", na, steps, "corr" if corr else "ind",
                                                 "prng" if prng else "sobol"), model, payoff, steps, prng))
    return out


BASKET_CASES = _basket_cases()


@pytest.mark.parametrize("name,model,payoff,steps,prng", BASKET_CASES, ids=[c[0] for c in BASKET_CASES])
def test_random_basket_parity(engine, name, model, payoff, steps, prng):
    """Baskets of 2 to 64 assets (in registers up to 32, the wide kernel above),
    weights, per-asset S0 and sigma, Cholesky correlation or none, against the
    oracle's kernel_multi_dupire restatement on the same points or Philox paths."""
    if prng:
        gm, gv = engine.run_mc_fixed(model, payoff, steps, 1500, 5)
        rm, rv = O.run_mc_fixed(model, payoff, steps, 1500, 5)
        assert abs(gm - rm) <= 2e-5 * max(1.0, abs(rm)) + 1e-6, (name, gm, rm)
        assert abs(gv - rv) <= 1e-3 * rv + 1e-12, (name, gv, rv)
    else:
        R, cnt = 2, 700
        got = np.array(engine.qmc_accumulate(model, payoff, steps, R, 42, 33, cnt, fdmc.NOISE_RAW))
        ref = O.qmc_accumulate(model, payoff, steps, R, 42, 33, cnt, fdmc.NOISE_RAW)
        assert np.allclose(got[:, 0] / cnt, ref[:, 0] / cnt, rtol=2e-5, atol=1e-6), (name, got[:, 0], ref[:, 0])


def _random_book(seed=99, n=16):
    rng = np.random.default_rng(seed)
    model = fdmc.LocalVolQuadParams(S0=100.0, mu=float(rng.uniform(0.0, 0.08)), a=float(rng.uniform(0.15, 0.3)),
                                    b=float(rng.uniform(0.0, 0.2)), c=float(rng.uniform(0.0, 2.0)), s_ref=100.0,
                                    T=1.0)
    book = []
    for _ in range(n):
        kind = [fdmc.PAYOFF_EUROPEAN, fdmc.PAYOFF_ASIAN, fdmc.PAYOFF_LOOKBACK, fdmc.PAYOFF_BARRIER][rng.integers(4)]
        K = float(rng.uniform(70.0, 130.0))
        book.append(fdmc.BookContract(kind, int(2 ** rng.integers(0, 10)), K, float(rng.uniform(1.15, 1.4)) * K,
                                      float(rng.choice([0.25, 0.5, 1.0, 2.0])), model.mu, 0.2))
    return model, book


@pytest.mark.parametrize("gm", [1, 2])
def test_random_book_level_sums_match_oracle(engine, gm):
    """A seeded random book (local-vol parameters, payoff kinds, strikes,
    barriers, maturities, 1-512 dates): the adjoint (gm=1) and tangent (gm=2)
    kernels' per-(contract, level, replica) price and Greek sums against the
    book oracle on the same points, to the fixed-book tolerances (a barrier's
    knock-out and a payoff kink may flip for a point on the edge to FP32
    rounding)."""
    model, book = _random_book()
    R, n, off = 4, 512, 4097
    b = engine.book_create(model, book, fdmc.BookConfig(R=R, greeks=gm))
    try:
        for c, k in enumerate(book):
            L = int(np.log2(k.n_steps))
            for l in sorted({0, L // 2, L}):
                got = np.array(b.level_sums(c, l, off, n))
                ref = O.book_level_sums(model, k, c, l, R, 42, off, n)
                tol = 2e-4 * n + (2 * k.K if k.kind == fdmc.PAYOFF_BARRIER else 0.0)
                np.testing.assert_allclose(got[:, 0], ref[:, 0], atol=tol, rtol=1e-6,
                                           err_msg="contract %d kind %d n %d level %d" % (c, k.kind, k.n_steps, l))
                if k.kind != fdmc.PAYOFF_BARRIER:
                    print("contract %d kind %d n %d level %d" % (c, k.kind, k.n_steps, l), flush=True)
                    check_greek_sums(got[:, 1:], ref[:, 1:], n,
                                     lambda i: O.book_level_sums(model, k, c, l, R, 42, off + i, 1))
    finally:
        b.free()
