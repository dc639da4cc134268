"""GPU parity of the contract book (k_book.hip through fdmc_book_*).

- per-(contract, level, replica) sums of the price difference and of the three
  Greeks against the book oracle (orc_book_level_sums: the reference's MLQMC
  coupling with FP32 forward tangents) on the same scrambled-Sobol points, for
  the adjoint (greeks = 1) and the tangent (greeks = 2) kernels;
- contract 0 against the engine's multilevel kernel with run_mlqmc_cuda's
  salts 42 + l R + r (methods_cuda.cu:1845);
- whole runs: batched and plain modes bit-identical, each contract against the
  oracle's allocation loop, a GBM European contract against Black-Scholes;
- argument validation.
"""
import math

import numpy as np
import pytest
from scipy.stats import norm

import fdmc
import pyoracle as po
from book_parity import check_greek_sums

pytestmark = pytest.mark.gpu

R = 8
E, A, LB, BR = fdmc.PAYOFF_EUROPEAN, fdmc.PAYOFF_ASIAN, fdmc.PAYOFF_LOOKBACK, fdmc.PAYOFF_BARRIER
LVQ = fdmc.LocalVolQuadParams(S0=100.0, mu=0.05, a=0.2, b=0.1, c=1.0, s_ref=100.0, T=1.0)


def C(kind, n, K=100.0, T=1.0, B=None):
    return fdmc.BookContract(kind, n, K, 1.3 * K if B is None else B, T, 0.05, 0.2)


# every payoff, small levels (n < 16: skeleton-only path) and blocked levels, four maturities
BOOK = [C(E, 16, 95.0, 0.25), C(A, 32, 100.0, 0.5), C(LB, 64, 100.0, 1.0), C(BR, 16, 100.0, 0.5, B=120.0),
        C(E, 512, 105.0, 2.0), C(A, 128, 98.0, 1.0), C(LB, 8, 100.0, 0.25), C(BR, 256, 90.0, 1.0, B=117.0),
        C(E, 1, 100.0, 0.5), C(A, 2, 101.0, 2.0), C(LB, 512, 100.0, 0.5), C(A, 4, 99.0, 1.0)]


@pytest.fixture(scope="module")
def books(engine):
    out = {gm: engine.book_create(LVQ, BOOK, fdmc.BookConfig(R=R, greeks=gm)) for gm in (0, 1, 2)}
    yield out
    for b in out.values():
        b.free()


def _levels(k):
    L = int(math.log2(k.n_steps))
    return sorted({0, min(1, L), L // 2, L})


def _tol(k, n):
    # FP32 paths: |error| per point ~1e-5 of S0; a barrier may flip knock-out on a
    # point that sits on the barrier to rounding (payoff up to ~S - K)
    return 2e-4 * n + (2 * k.K if k.kind == BR else 0.0)




@pytest.mark.parametrize("gm", [1, 2])
def test_level_sums_match_oracle(books, gm):
    b = books[gm]
    n, off = 1024, 777
    for c, k in enumerate(BOOK):
        for l in _levels(k):
            got = np.array(b.level_sums(c, l, off, n))
            ref = po.book_level_sums(LVQ, k, c, l, R, 42, off, n)
            tol = _tol(k, n)
            print("contract %d kind %d n %d level %d" % (c, k.kind, k.n_steps, l), flush=True)
            np.testing.assert_allclose(got[:, 0], ref[:, 0], atol=tol, rtol=1e-6)
            if k.kind != BR:
                check_greek_sums(got[:, 1:], ref[:, 1:], n,
                                 lambda i: po.book_level_sums(LVQ, k, c, l, R, 42, off + i, 1))


def test_adjoint_and_tangent_kernels_agree(books):
    n = 2048
    for c, k in enumerate(BOOK):
        if k.kind == BR:
            continue
        L = int(math.log2(k.n_steps))
        adj = np.array(books[1].level_sums(c, L, 0, n))
        tan = np.array(books[2].level_sums(c, L, 0, n))
        np.testing.assert_allclose(adj[:, 0], tan[:, 0], rtol=1e-6, atol=1e-6 * n)
        np.testing.assert_allclose(adj[:, 1:], tan[:, 1:], rtol=1e-4, atol=2e-5 * n)


@pytest.mark.parametrize("kind,payoff", [(E, fdmc.European(100.0, 0.05, 0.5)), (A, fdmc.Asian(100.0)),
                                         (LB, fdmc.Lookback(0.2)), (BR, fdmc.Barrier(100.0, 120.0, 0.2, 0.05, 0.5))])
def test_contract0_matches_multilevel_kernel_salts(engine, kind, payoff):
    T, nst = 0.5, 64
    m = fdmc.LocalVolQuadParams(S0=100.0, mu=0.05, a=0.2, b=0.1, c=1.0, s_ref=100.0, T=T)
    k = C(kind, nst, 100.0, T, B=120.0)
    b = engine.book_create(m, [k], fdmc.BookConfig(R=R, greeks=1))
    bbs = [engine.bb_upload(fdmc.bb_precompute(1 << l, T)) for l in range(7)]
    try:
        n, off = 2048, 4096
        for l in (0, 2, 4, 6):
            got = np.array(b.level_sums(0, l, off, n))
            ref = engine.levels_accumulate(m, payoff, 2, [(l, r, off, n) for r in range(R)], qmc=True,
                                           mode=fdmc.NOISE_BB, R=R, salt_or_seed=42, bb_list=bbs)
            np.testing.assert_allclose(got[:, 0], np.array(ref)[:, 0], rtol=1e-6, atol=_tol(k, n))
    finally:
        for x in bbs:
            x.free()
        b.free()


def test_batched_and_plain_are_bit_identical_and_match_oracle(engine):
    book = [k for k in BOOK if k.kind != BR] + [C(BR, 16, 100.0, 0.5, B=120.0)]
    b = engine.book_create(LVQ, book, fdmc.BookConfig(R=16, greeks=1))
    try:
        eps = 2e-2
        bat, st_b = b.run(eps, fdmc.BOOK_BATCHED)
        pla, st_p = b.run(eps, fdmc.BOOK_PLAIN)
        assert st_b.graph_replays >= 1 and st_p.graph_replays == 0
        assert st_b.path_steps == st_p.path_steps > 0
        for x, y in zip(bat, pla):
            assert tuple(x) == tuple(y) or all((u == v) or (u != u and v != v) for u, v in zip(x, y)), (x, y)
        cfg = fdmc.BookConfig(R=16, greeks=1)
        for c, (k, g) in enumerate(zip(book, bat)):
            o = po.book_contract_run(LVQ, k, c, eps, cfg)
            print("contract", c, g.price, o.price, g.std_error, g.n_samples, o.n_samples, flush=True)
            assert g.flags == 0 and o.flags == 0
            assert g.std_error <= eps / math.sqrt(2) * (1 + 1e-9)
            assert abs(g.price - o.price) <= 0.25 * g.std_error + 1e-5
            assert abs(g.n_samples - o.n_samples) <= 0.02 * o.n_samples
            if k.kind != BR:
                for a_, b_, s_ in ((g.delta, o.delta, g.delta_se), (g.vega, o.vega, g.vega_se),
                                   (g.rho, o.rho, g.rho_se)):
                    assert abs(a_ - b_) <= 0.25 * s_ + 1e-4 * max(1.0, abs(b_)), (c, a_, b_, s_)
            else:
                assert math.isnan(g.delta) and math.isnan(g.vega) and math.isnan(g.rho)
    finally:
        b.free()


def test_steep_local_vol_book_matches_oracle(engine):
    """Vol level 1: on four dates over T = 1 about 9 % of the Euler paths step
    to S <= 0 (more on the coarser levels), on 32 dates over T = 2 about 0.2 %.
    Such a path pays 0 under every payoff and adds 0 to the Greeks, so the
    estimates are finite, converge, and match the oracle's contract loop, in
    the register (4 and 16 dates) and in the block (32 dates) kernels; the
    tangent-mode book agrees with the adjoint one."""
    m = fdmc.LocalVolQuadParams(S0=100.0, mu=0.05, a=1.0, b=0.05, c=1.0, s_ref=100.0, T=1.0)
    book = [C(LB, 4), C(E, 4), C(A, 4), C(LB, 16), C(A, 32, T=2.0), C(LB, 32, T=2.0)]
    eps = 0.1
    ba = engine.book_create(m, book, fdmc.BookConfig(R=16, greeks=1))
    bt = engine.book_create(m, book, fdmc.BookConfig(R=16, greeks=2))
    try:
        adj, _ = ba.run(eps, fdmc.BOOK_BATCHED)
        tan, _ = bt.run(eps, fdmc.BOOK_BATCHED)
    finally:
        ba.free()
        bt.free()
    cfg = fdmc.BookConfig(R=16, greeks=1)
    for c, (k, g, t) in enumerate(zip(book, adj, tan)):
        o = po.book_contract_run(m, k, c, eps, cfg)
        print("contract", c, g.price, o.price, g.std_error, g.n_samples, o.n_samples, flush=True)
        for x in (g, t):
            assert x.flags == 0 and x.std_error <= eps / math.sqrt(2) * (1 + 1e-9)
            assert all(math.isfinite(v) for v in (x.price, x.delta, x.vega, x.rho, x.delta_se, x.vega_se, x.rho_se))
        assert o.flags == 0
        assert abs(g.price - o.price) <= 0.25 * g.std_error + 1e-5
        assert abs(g.n_samples - o.n_samples) <= 0.02 * o.n_samples
        for a_, b_, s_ in ((g.delta, o.delta, g.delta_se), (g.vega, o.vega, g.vega_se), (g.rho, o.rho, g.rho_se)):
            assert abs(a_ - b_) <= 0.25 * s_ + 1e-4 * max(1.0, abs(b_)), (c, a_, b_, s_)
        assert abs(t.price - g.price) <= 0.25 * g.std_error + 1e-5
        assert abs(t.delta - g.delta) <= 0.25 * g.delta_se + 1e-4


def test_gbm_european_contract_matches_black_scholes(engine):
    S0, mu, sig, T, K = 100.0, 0.05, 0.25, 1.0, 100.0
    m = fdmc.GBMParams(S0, mu, sig, T)
    b = engine.book_create(m, [fdmc.BookContract(E, 256, K, 0.0, T, mu, 0.2)], fdmc.BookConfig(R=16, greeks=1))
    try:
        (g,), _ = b.run(1e-3, fdmc.BOOK_BATCHED)
    finally:
        b.free()
    d1 = (math.log(S0 / K) + (mu + 0.5 * sig * sig) * T) / (sig * math.sqrt(T))
    d2 = d1 - sig * math.sqrt(T)
    bs = S0 * norm.cdf(d1) - K * math.exp(-mu * T) * norm.cdf(d2)
    assert g.flags == 0 and g.std_error <= 1e-3 / math.sqrt(2)
    assert abs(g.price - bs) < 3 * g.std_error + 5e-4, (g.price, bs)      # Euler bias at 256 steps < 5e-4
    assert abs(g.delta - norm.cdf(d1)) < 3 * g.delta_se + 1e-4
    assert abs(g.vega - S0 * norm.pdf(d1) * math.sqrt(T)) < 3 * g.vega_se + 0.02
    assert abs(g.rho - K * T * math.exp(-mu * T) * norm.cdf(d2)) < 3 * g.rho_se + 0.02


def test_price_only_and_tangent_runs(books):
    eps = 3e-2
    p0, _ = books[0].run(eps, fdmc.BOOK_BATCHED)
    p1, _ = books[1].run(eps, fdmc.BOOK_BATCHED)
    p2, _ = books[2].run(eps, fdmc.BOOK_BATCHED)
    for x, y, z in zip(p0, p1, p2):
        assert math.isnan(x.delta) and math.isnan(x.rho)
        assert abs(x.price - y.price) <= 0.25 * y.std_error + 1e-5
        assert abs(z.price - y.price) <= 0.25 * y.std_error + 1e-5
        if not math.isnan(y.delta):
            assert abs(z.delta - y.delta) <= 0.25 * y.delta_se + 1e-4


def test_local_vol_book_generator_runs_batched(engine):
    model, book = fdmc.local_vol_book(96, 2026)
    b = engine.book_create(model, book, fdmc.BookConfig(R=16, greeks=1))
    try:
        res, st = b.run(2e-2, fdmc.BOOK_BATCHED)
    finally:
        b.free()
    assert st.path_steps == sum(r.path_steps for r in res) > 0
    assert st.wg_per_cu >= 1 and st.variant in (0, 1)
    for k, r in zip(book, res):
        assert r.flags == 0 and r.std_error <= 2e-2 / math.sqrt(2) * (1 + 1e-9)
        assert r.levels == int(math.log2(k.n_steps)) + 1
        assert math.isnan(r.delta) == (k.kind == BR)


@pytest.mark.parametrize("bad", ["steps3", "steps1024", "geom", "ninemat", "heston", "greeks3", "R1"])
def test_invalid_books_are_rejected(engine, bad):
    model, book, cfg = LVQ, [C(E, 16)], fdmc.BookConfig(R=R)
    if bad == "steps3":
        book = [C(E, 3)]
    elif bad == "steps1024":
        book = [C(E, 1024)]
    elif bad == "geom":
        book = [C(fdmc.PAYOFF_GEOM_ASIAN, 16)]
    elif bad == "ninemat":
        book = [C(E, 16, T=0.1 * (i + 1)) for i in range(9)]
    elif bad == "heston":
        model = fdmc.HestonParams(100.0, 0.05, 2.0, 0.04, 0.5, -0.7, 0.04, 1.0).compute_cholesky()
    elif bad == "greeks3":
        cfg = fdmc.BookConfig(R=R, greeks=3)
    elif bad == "R1":
        cfg = fdmc.BookConfig(R=1)
    with pytest.raises(fdmc.FdmcError) as ei:
        engine.book_create(model, book, cfg)
    assert ei.value.code in (1, 2)
