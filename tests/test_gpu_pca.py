"""PCA on matrix cores and the multi-asset basket (config 3).

The MFMA GEMM is checked against an FP64 GEMM of the same FP16/BF16-rounded
operands (only FP32 accumulation error remains, so any lane-layout mistake
fails), and against an FP64 GEMM of the unrounded normals (the quantisation
error the precision choice costs).  Basket prices are compared with the
oracle's restatement of kernel_multi_dupire (methods_cuda.cu:526-551) at the
same points; deltas with common-random-number differences.
"""
import numpy as np
import pytest

import fdmc
import pyoracle as O
from c3_setup import c3_parameters, cholesky_lower

pytestmark = pytest.mark.gpu


def _round_bf16(x):
    u = np.ascontiguousarray(x, np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16 << 16
    return u.astype(np.uint32).view(np.float32)


@pytest.fixture(scope="module")
def c3(engine):
    sig, C = c3_parameters()
    M = O.pca_kron_compute(64, 1.0, C)          # D x D, dW = M z, index k * 16 + a
    pca = engine.pca_upload(1024, M.T.ravel())   # column-major
    yield sig, C, M, pca
    pca.free()


@pytest.mark.parametrize("bf16", [False, True], ids=["fp16", "bf16"])
@pytest.mark.parametrize("n", [4096, 4224], ids=["tiles256", "tiles128"])
def test_gemm_matches_fp64_gemm(engine, torch_cuda, c3, bf16, n):
    # 4096 rows take the 256 x 256-tile kernel, 4224 (not a multiple of 256) the 128 x 128 one
    torch = torch_cuda
    sig, C, M, pca = c3
    salt, off = 42, 512
    out = torch.empty((n, 1024), dtype=torch.float32, device="cuda")
    engine.pca_transform(pca, salt, off, n, bf16, out.data_ptr())
    got = out.cpu().numpy().astype(np.float64)
    z = O.sobol_normal_block(salt, 1024, off, n).T.astype(np.float32)     # [n][D]
    q = (lambda a: _round_bf16(a)) if bf16 else (lambda a: a.astype(np.float16).astype(np.float32))
    exact_q = q(z).astype(np.float64) @ q(M.astype(np.float32)).astype(np.float64).T
    # The device rounds its own FP32 normals (within ~1e-7 of the FP64 oracle), so
    # a small fraction of operands round to the neighbouring half-precision value;
    # beyond that only FP32 accumulation error remains (a layout error is O(1)).
    err_acc = np.linalg.norm(got - exact_q) / np.linalg.norm(exact_q)
    assert err_acc < (1e-4 if bf16 else 3e-5), err_acc
    ref = z.astype(np.float64) @ M.T
    rel_fro = np.linalg.norm(got - ref) / np.linalg.norm(ref)
    print("GEMM %s: rel Frobenius error vs FP64 GEMM of the exact normals %.3e, vs rounded operands %.3e"
          % ("bf16" if bf16 else "fp16", rel_fro, err_acc))
    assert rel_fro < (6e-3 if bf16 else 8e-4), rel_fro    # operand quantisation


def _basket(sig, C, corr=True):
    L = cholesky_lower(C).ravel() if corr else None
    return fdmc.MultiGBMParams(16, [100.0] * 16, list(sig), 0.05, 1.0, L=L, uncorrelated=not corr)


PAY = fdmc.Basket(100.0, 0.05, 1.0, 16)


def test_basket_raw_cholesky_matches_oracle(engine, c3):
    sig, C, M, pca = c3
    m = _basket(sig, C)
    R, n = 2, 1500
    got = np.array(engine.qmc_accumulate(m, PAY, 64, R, 42, 100, n, fdmc.NOISE_RAW))
    ref = O.qmc_accumulate(m, PAY, 64, R, 42, 100, n, fdmc.NOISE_RAW)
    assert np.allclose(got[:, 0] / n, ref[:, 0] / n, rtol=2e-5, atol=1e-6)


def test_basket_pca_matches_oracle_fp64_gemm(engine, c3):
    sig, C, M, pca = c3
    m = _basket(sig, C)
    R, n = 2, 2048
    for bf16, tol in ((False, 2e-3), (True, 1e-2)):
        mode = fdmc.NOISE_PCA | (fdmc.PCA_BF16 if bf16 else 0)
        got = np.array(engine.qmc_accumulate(m, PAY, 64, R, 42, 0, n, mode, dev_pca=pca))
        ref = O.qmc_accumulate(m, PAY, 64, R, 42, 0, n, fdmc.NOISE_PCA, pca_M=M)
        assert np.allclose(got[:, 0] / n, ref[:, 0] / n, rtol=tol), (bf16, got[:, 0] / n, ref[:, 0] / n)


@pytest.mark.parametrize("n,steps", [(8, 32), (4, 48), (2, 96), (16, 24), (32, 8)],
                         ids=["8x32-fused", "4x48-fused-padded", "2x96-fused-padded", "16x24-rows", "32x8-fused"])
def test_basket_pca_product_epilogue_and_rows_match_oracle(engine, n, steps):
    """GBM baskets whose padded dimension is a multiple of 256 price through the
    GEMM's product epilogue (the Euler factors multiply per 128-column slab,
    padding columns count as 1); the others through stored dW rows.  Both
    against the oracle's FP64 PCA and sequential Euler steps."""
    sig, C = c3_parameters(n, seed=n)
    M = O.pca_kron_compute(steps, 1.0, C)
    D = n * steps
    pca = engine.pca_upload(D, M.T.ravel())
    m = fdmc.MultiGBMParams(n, [100.0 + a for a in range(n)], list(sig), 0.05, 1.0, L=cholesky_lower(C).ravel())
    pay = fdmc.Basket(100.0, 0.05, 1.0, n)
    R, cnt = 3, 1500
    for bf16, tol in ((False, 2e-3), (True, 1e-2)):
        mode = fdmc.NOISE_PCA | (fdmc.PCA_BF16 if bf16 else 0)
        got = np.array(engine.qmc_accumulate(m, pay, steps, R, 42, 300, cnt, mode, dev_pca=pca))
        ref = O.qmc_accumulate(m, pay, steps, R, 42, 300, cnt, fdmc.NOISE_PCA, pca_M=M)
        assert np.allclose(got[:, 0] / cnt, ref[:, 0] / cnt, rtol=tol, atol=1e-3), (bf16, got[:, 0], ref[:, 0])
        assert np.allclose(got[:, 1] / cnt, ref[:, 1] / cnt, rtol=2 * tol, atol=1e-2), (bf16, got[:, 1], ref[:, 1])
    pca.free()


def test_basket_pca_dupire_rows_match_oracle(engine, c3):
    """Local-vol assets step sequentially, so they read the stored dW rows."""
    sig, C, M, pca = c3
    m = fdmc.MultiDupireParams(16, 0.05, 0.2, 0.5, 0.7, 1.0, L=cholesky_lower(C).ravel())
    R, n = 2, 1024
    got = np.array(engine.qmc_accumulate(m, PAY, 64, R, 42, 0, n, fdmc.NOISE_PCA, dev_pca=pca))
    ref = O.qmc_accumulate(m, PAY, 64, R, 42, 0, n, fdmc.NOISE_PCA, pca_M=M)
    assert np.allclose(got[:, 0] / n, ref[:, 0] / n, rtol=2e-3, atol=1e-3), (got[:, 0] / n, ref[:, 0] / n)


def test_basket_prng_matches_oracle(engine, c3):
    sig, C, M, pca = c3
    m = _basket(sig, C)
    gm, gv = engine.run_mc_fixed(m, PAY, 64, 3000, 7)
    rm, rv = O.run_mc_fixed(m, PAY, 64, 3000, 7)
    assert abs(gm - rm) < 2e-5 * abs(rm) + 1e-6 and abs(gv - rv) < 1e-3 * rv


def test_single_asset_qmc_pca_mode(engine):
    """run_qmc_cuda PCA mode for a one-factor model (D = 64 padded to 128)."""
    g = fdmc.GBMParams(100, 0.05, 0.2, 1.0)
    A = fdmc.Asian(100.0)
    M = O.pca_compute(64, 1.0)
    pca = engine.pca_upload(64, M.T.ravel())
    R, n = 3, 2048
    got = np.array(engine.qmc_accumulate(g, A, 64, R, 42, 0, n, fdmc.NOISE_PCA, dev_pca=pca))
    ref = O.qmc_accumulate(g, A, 64, R, 42, 0, n, fdmc.NOISE_PCA, pca_M=M)
    assert np.allclose(got[:, 0] / n, ref[:, 0] / n, rtol=2e-3)
    r = engine.run_qmc_cuda(g, A, 2e-3, 64, fdmc.QMCConfig(R=16, N0=1 << 12), fdmc.NOISE_PCA, dev_pca=pca)
    bb = engine.bb_upload(fdmc.bb_precompute(64, 1.0))
    rb = engine.run_qmc_cuda(g, A, 2e-3, 64, fdmc.QMCConfig(R=16, N0=1 << 12), fdmc.NOISE_BB, dev_bb=bb)
    bb.free()
    pca.free()
    assert abs(r.price - rb.price) < 3 * np.hypot(r.std_error, rb.std_error) + 1e-3


def test_c3_price_and_per_asset_deltas(engine, c3):
    """Config 3: 2^18 points x R = 16 through the FP16 MFMA PCA; the price agrees
    with the Cholesky/raw estimator and the per-asset deltas with CRN differences."""
    sig, C, M, pca = c3
    m = _basket(sig, C)
    R, n = 16, 1 << 18
    s = np.array(engine.qmc_accumulate(m, PAY, 64, R, 42, 0, n, fdmc.NOISE_PCA, dev_pca=pca))
    means = s[:, 0] / n
    p_pca, se_pca = means.mean(), means.std(ddof=1) / np.sqrt(R)
    s2 = np.array(engine.qmc_accumulate(m, PAY, 64, R, 4242, 0, n, fdmc.NOISE_RAW))
    m2 = s2[:, 0] / n
    assert abs(p_pca - m2.mean()) < 4 * np.hypot(se_pca, m2.std(ddof=1) / np.sqrt(R)) + 2e-3
    assert se_pca < 1e-3
    # per-asset deltas (raw noise, one replica) vs CRN central differences
    g = np.array(engine.greeks_accumulate(m, PAY, 64, 1, 42, 0, 1 << 16, fdmc.NOISE_RAW))[0]
    for a in (0, 7, 15):
        h = 0.05
        up, dn = [100.0] * 16, [100.0] * 16
        up[a] += h
        dn[a] -= h
        L = cholesky_lower(C).ravel()
        pu = engine.qmc_accumulate(fdmc.MultiGBMParams(16, up, list(sig), 0.05, 1.0, L=L), PAY, 64, 1, 42, 0,
                                   1 << 16, fdmc.NOISE_RAW)[0][0]
        pd = engine.qmc_accumulate(fdmc.MultiGBMParams(16, dn, list(sig), 0.05, 1.0, L=L), PAY, 64, 1, 42, 0,
                                   1 << 16, fdmc.NOISE_RAW)[0][0]
        fd = (pu - pd) / (2 * h) / (1 << 16)
        adj = g[2 + 2 * a] / (1 << 16)
        assert abs(adj - fd) < 0.01 * abs(fd) + 1e-4, (a, adj, fd)


@pytest.mark.parametrize("bf16", [False, True], ids=["fp16", "bf16"])
def test_c3_pca_per_asset_deltas_vs_crn(engine, c3, bf16):
    """Config 3 through the PCA pipeline itself (2^18 points x R = 16, Kronecker
    PCA of min(t_i, t_j) x Corr on matrix cores, the Euler factors multiplied in
    the GEMM's epilogue): the pass with Greeks returns the same price sums as the
    price-only pass, and every asset's pathwise delta disc 1{B > K} w_a S_T^a /
    S0^a agrees per replica with central differences of the price over the same
    points with S0^a bumped by +-0.05 (the price is linear in S0^a on each path,
    so the difference moves only with paths crossing the strike)."""
    sig, C, M, pca = c3
    m = _basket(sig, C)
    R, n, h = 16, 1 << 18, 0.05
    mode = fdmc.NOISE_PCA | (fdmc.PCA_BF16 if bf16 else 0)
    g = np.array(engine.greeks_accumulate_pca(m, PAY, 64, R, 42, 0, n, pca, bf16))
    assert g.shape == (R, 2 + 2 * 16)
    p = np.array(engine.qmc_accumulate(m, PAY, 64, R, 42, 0, n, mode, dev_pca=pca))
    assert np.allclose(g[:, :2], p[:, :2], rtol=1e-12, atol=0.0)
    deltas = g[:, 2::2] / n                                  # [R][asset]
    for a in range(16):
        px = []
        for sgn in (+1, -1):
            s0 = [100.0] * 16
            s0[a] += sgn * h
            mb = fdmc.MultiGBMParams(16, s0, list(sig), 0.05, 1.0, L=cholesky_lower(C).ravel())
            px.append(np.array(engine.qmc_accumulate(mb, PAY, 64, R, 42, 0, n, mode, dev_pca=pca))[:, 0] / n)
        fd = (px[0] - px[1]) / (2 * h)
        assert np.all(np.abs(deltas[:, a] - fd) <= 2e-3 * np.abs(fd) + 2e-5), (a, deltas[:, a], fd)
    # the replicas' mean deltas: positive, weights 1/16, summing to about the basket's delta
    mean = deltas.mean(axis=0)
    assert np.all(mean > 0.0) and 0.3 < mean.sum() < 1.0, mean


def test_pca_greeks_entry_rejects_other_models(engine, c3):
    sig, C, M, pca = c3
    with pytest.raises(fdmc.FdmcError):
        engine.greeks_accumulate_pca(fdmc.GBMParams(100.0, 0.05, 0.2, 1.0), fdmc.European(100.0, 0.05, 1.0),
                                     1024, 1, 42, 0, 256, pca)


def test_run_qmc_pca_chunks_large_doublings(engine):
    """run_qmc_cuda with PCA noise over 2048 dates (the largest step count the
    reference's examples use) and 32 replicas: the doubling from 2^16 to 2^17
    points per replica needs 26 GB of normals and GEMM rows, so it runs in
    chunks of at most 8 GiB.  Its replica sums equal those of single passes over
    the same point ranges (FP64 summation order apart), so chunks neither skip
    nor repeat points."""
    n, R = 2048, 32
    t = np.arange(1, n + 1) / n
    w, V = np.linalg.eigh(np.minimum.outer(t, t))
    W = V[:, ::-1] * np.sqrt(np.maximum(w[::-1], 0.0))     # Brownian path W = W z, leading components first
    M = np.diff(W, axis=0, prepend=0.0)                    # increments dW = M z, as pca_compute returns
    pca = engine.pca_upload(n, M.T.ravel())
    gbm, asian = fdmc.GBMParams(100.0, 0.05, 0.3, 1.0), fdmc.Asian(100.0)
    try:
        r = engine.run_qmc_cuda(gbm, asian, 1e-9, n, fdmc.QMCConfig(R=R, N0=1 << 14, max_doublings=4),
                                fdmc.NOISE_PCA, dev_pca=pca)
        assert r.n_samples == R * (1 << 17) and r.flags & fdmc.FLAG_NOT_CONVERGED
        tot = np.zeros(R)
        for first, count in [(0, 1 << 14), (1 << 14, 1 << 14), (1 << 15, 1 << 15), (1 << 16, 1 << 15),
                             (3 << 15, 1 << 15)]:
            tot += np.array(engine.qmc_accumulate(gbm, asian, n, R, 42, first, count, fdmc.NOISE_PCA,
                                                  dev_pca=pca))[:, 0]
    finally:
        pca.free()
    means = tot / (1 << 17)
    assert abs(r.price - means.mean()) <= 1e-12 * abs(r.price)
    assert abs(r.std_error - means.std(ddof=1) / np.sqrt(R)) <= 1e-9 * r.std_error


def _slab_products_ref(dw, D, n, c0, sig):
    """The basket's slab products from FP32 dW rows [points][ld], in the order
    the GEMM's product epilogue forms them: lane l of a 128-column slab holds
    columns 32 j + l (asset l % n) and multiplies their factors c0 + sig dW
    (one FP32 fused multiply-add each; padding columns >= D count as 1) for
    j = 0..3, then lanes l and l ^ n, l ^ 2n, .. 16 combine pairwise."""
    npts, ld = dw.shape
    lanes = np.arange(32)
    # the kernel receives sig and c0 as FP32
    sg = np.asarray(sig, np.float32).astype(np.float64)[lanes % n]
    c0 = np.float64(np.float32(c0))
    out = np.empty((ld // 128, npts, n), np.float32)
    for s in range(ld // 128):
        p = np.ones((npts, 32), np.float32)
        for j in range(4):
            c = s * 128 + j * 32 + lanes
            f = (sg * dw[:, c].astype(np.float64) + c0).astype(np.float32)   # exact fma, one rounding
            p = p * np.where(c < D, f, np.float32(1.0))
        sft = n
        while sft < 32:
            p = p * p[:, lanes ^ sft]
            sft <<= 1
        out[s] = p[:, :n]
    return out


@pytest.mark.parametrize("n,steps,bf16,npts", [(16, 64, False, 1000), (16, 64, True, 1000), (8, 120, False, 1000),
                                               (32, 8, True, 1000), (2, 384, False, 1000), (16, 64, False, 70000)],
                         ids=["c3-16x64-fp16", "c3-16x64-bf16", "8x120-padded", "32x8", "2x384", "c3-multitile"])
def test_gemm_product_epilogue_rows_match_products_of_dw(engine, torch_cuda, n, steps, bf16, npts):
    """The fused product epilogue (k_gemm_prod) against the same products formed
    on the host from the plain GEMM's dW rows of the same points: both GEMMs
    accumulate each output in the same k order, so the products agree to the
    last bit or two (the host's fused multiply-add is emulated in FP64).  A
    mistake in the epilogue's row, lane or asset mapping moves whole factors:
    basket prices would not show it (paths assembled from other rows' slabs have
    the same law), this comparison does.  Point counts not a multiple of 256
    leave padding rows in the GEMM; 8 x 120 = 960 dates leave padding columns;
    70,000 points are 1,096 tiles of 256 x 256, about four per persistent
    workgroup, so later tiles (and the operand stream that runs from one tile
    into the next) are checked too."""
    torch = torch_cuda
    sig, C = c3_parameters(n, seed=3 + n)
    M = O.pca_kron_compute(steps, 1.0, C)
    D = n * steps
    pca = engine.pca_upload(D, M.T.ravel())
    ld = (D + 127) // 128 * 128
    off = 768
    c0 = 1.0 + 0.05 / steps
    dw = torch.empty((npts, D), dtype=torch.float32, device="cuda")
    engine.pca_transform(pca, 42, off, npts, bf16, dw.data_ptr())
    prod = torch.empty((ld // 128, npts, n), dtype=torch.float32, device="cuda")
    engine.pca_products(pca, 42, off, npts, bf16, c0, list(sig), prod.data_ptr())
    torch.cuda.synchronize()
    dw_full = np.zeros((npts, ld), np.float32)
    dw_full[:, :D] = dw.cpu().numpy()
    ref = _slab_products_ref(dw_full, D, n, c0, sig)
    got = prod.cpu().numpy()
    rel = np.abs(got.astype(np.float64) - ref) / np.abs(ref)
    assert rel.max() < 3e-7, (rel.max(), np.unravel_index(rel.argmax(), rel.shape))
    assert np.mean(got == ref) > 0.99
    pca.free()


def test_pca_products_rejects_bad_shapes(engine, torch_cuda, c3):
    """Shapes the product epilogue cannot take are rejected on the host before
    any device work (the output buffer is a real one all the same)."""
    sig, C, M, pca = c3
    buf = torch_cuda.empty((8, 256, 4), dtype=torch_cuda.float32, device="cuda")
    with pytest.raises(fdmc.FdmcError):
        engine.pca_products(pca, 42, 0, 256, False, 1.0, [0.2] * 3, buf.data_ptr())   # 3 assets
    small = engine.pca_upload(128, np.eye(128).ravel())                                # mp = 128
    with pytest.raises(fdmc.FdmcError):
        engine.pca_products(small, 42, 0, 256, False, 1.0, [0.2] * 2, buf.data_ptr())
    small.free()


@pytest.mark.parametrize("off", [256 * 7 + 100, 65525 * 256 + 37, ((1 << 23) - 11) * 256 + 37],
                         ids=["low", "tile-65536", "tile-2^23"])
@pytest.mark.parametrize("bf16", [False, True], ids=["fp16", "bf16"])
def test_pca_normals_are_the_oracle_points(engine, torch_cuda, bf16, off):
    """The GEMM's A operand itself: with M = I the transform returns the
    half-precision normals unchanged (products with 1, sums with 0), so they
    are compared with the oracle's scrambled-Sobol normals rounded the same
    way, row for row.  786,688 points are 3,073 tiles, more than k_gen_z
    launches workgroups (three per CU), so workgroups run several tiles and
    the rows checked include later tiles (Sobol words advanced in registers from the
    first tile by the XOR of two direction numbers per dimension, v_7 and
    v_{8 + ctz(n)} of tile n); D = 160 has a partial 64-dimension chunk and a
    padding chunk (mp = 256).  At the two high offsets tile 2^16 (points
    2^24..) and tile 2^23 (points 2^31..) are the 12th tile of the range, not
    the first of a workgroup for either workgroup size (two or five tiles), so
    their words come from the tile-to-tile update with ctz 16 and 23 (the
    update rows that are not precomputed); rows 2779.. are those tiles."""
    torch = torch_cuda
    D, npts, salt = 160, 3073 * 256, 77
    pca = engine.pca_upload(D, np.eye(D).ravel())
    out = torch.empty((npts, D), dtype=torch.float32, device="cuda")
    engine.pca_transform(pca, salt, off, npts, bf16, out.data_ptr())
    torch.cuda.synchronize()
    q = (lambda a: _round_bf16(a)) if bf16 else (lambda a: a.astype(np.float16).astype(np.float32))
    ulp = 2.0 ** (-7 if bf16 else -10)
    # rows in first and second tiles of several workgroups, and the last rows
    for r0 in (0, 156, 256 * 1 + 156, 2779, 256 * 1000 + 156, 256 * 2047 + 156, npts - 256):
        got = out[r0:r0 + 256].cpu().numpy()
        ref = q(O.sobol_normal_block(salt, D, off + r0, 256).T)
        diff = np.abs(got - ref)
        # the device's FP32 normal can sit across a rounding boundary from the
        # oracle's (within ~4e-7): at most one half-precision step apart
        assert np.all(diff <= ulp * np.maximum(1.0, np.abs(ref)) * 1.01), (r0, diff.max())
        assert np.mean(diff == 0) > 0.995, (r0, np.mean(diff == 0))
    pca.free()
