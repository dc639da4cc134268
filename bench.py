"""Benchmark of the path-simulation hot path on one MI355X.

Workload (the second configuration of the baseline): scrambled-Sobol RQMC
arithmetic-Asian call under GBM with a Brownian bridge and a geometric-Asian
control variate: S0 = K = 100, r = mu = 0.05, sigma = 0.3, T = 1, D = 256 dates,
R = 32 replicas (salts 42 + r).  One step = one fused pass over the largest
doubling of that configuration, 2^20 points of every replica
(32 x 2^20 x 256 = 8.6e9 path-steps), inputs (direction tables, bridge weights)
resident in HBM.

Printed line: path-steps/s of the timed steps (value), the dominant kernel's
FP32-VALU roofline from HIP-event timings on the engine stream (31 algorithmic
flops per path-step), the time to reach eps = 1e-3 with the full doubling loop,
and the oracle's CPU throughput on a bounded sample of the same workload.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "financial-derivatives-montecarlo_amd"))

FLOPS_PER_PATH_STEP = 31        # C2 forward: Euler 3 + BB 6 + Asian 1 + geo 1 + ICDF 20
VALU_PEAK_TFLOPS = 157.3        # MI355X FP32 vector peak (packed)
MFMA_PEAK_TFLOPS = 2500.0       # MI355X dense FP16/BF16 matrix peak
METRIC = "time-to-eps=1e-3 price+Greeks, batched vs plain (s); path-steps/s vs VALU peak"


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_baseline(target_s=10.0):
    """Oracle (FP32 paths, FP64 ICDF) on all host threads, same workload, fewer points."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import fdmc
    import pyoracle as O

    m = fdmc.GBMParams(100.0, 0.05, 0.3, 1.0)
    a, g = fdmc.Asian(100.0), fdmc.GeomAsian(100.0)
    E = fdmc.geom_asian_analytic(100, 100, 1, 0.05, 0.3, 256)
    bb = O.OracleBB(256, 1.0)
    R = 32
    dts = []
    for n in (256, 1024):   # two sizes: the per-call direction-table set-up cancels in the slope
        t = time.perf_counter()
        O.qmc_accumulate(m, a, 256, R, 42, 0, n, fdmc.NOISE_BB, bb=bb, ctrl_model=m, ctrl_payoff=g, E_ctrl=E,
                         beta=1.0)
        dts.append(time.perf_counter() - t)
    per_point = max(1e-9, (dts[1] - dts[0]) / 768)
    n2 = max(1024, int(target_s / per_point) // 64 * 64)
    t = time.perf_counter()
    O.qmc_accumulate(m, a, 256, R, 42, 0, n2, fdmc.NOISE_BB, bb=bb, ctrl_model=m, ctrl_payoff=g, E_ctrl=E, beta=1.0)
    dt = time.perf_counter() - t
    ps = R * n2 * 256
    return {"value": ps / dt, "unit": "path-steps/s", "cores": O.num_threads(), "kind": "port",
            "sample": "R=32 x %d points x 256 steps (BB + CV), %.1f s" % (n2, dt)}


def c3_gemm(eng, reps=5):
    """Config-3 PCA transform: 2^18 points x 16 replicas, dW = Z M^T with
    K = N = 1024 on MFMA (FP16 and BF16); HIP-event time of the GEMM alone."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import fdmc
    from c3_setup import c3_parameters
    import numpy as np

    sig, C = c3_parameters()
    # Kronecker PCA through the facade's closed-form time factor is not exposed to
    # Python; the product of the two factors is formed here in FP64.
    m = 64
    h = 1.0 / m
    k = np.arange(1, m + 1)
    lam_t = h / (4 * np.sin((2 * k - 1) * np.pi / (2 * (2 * m + 1))) ** 2)
    Vt = 2 / np.sqrt(2 * m + 1) * np.sin(np.outer(k, 2 * k - 1) * np.pi / (2 * m + 1))
    lc, Vc = np.linalg.eigh(C)
    lam = np.outer(lam_t, lc[::-1]).ravel()
    order = np.argsort(-lam, kind="stable")
    V = np.kron(Vt, Vc[:, ::-1])[:, order] * np.sqrt(np.maximum(lam[order], 0))
    M = V.copy()
    M[16:] -= V[:-16]
    pca = eng.pca_upload(1024, M.T.ravel())
    model = fdmc.MultiGBMParams(16, [100.0] * 16, list(sig), 0.05, 1.0)
    pay = fdmc.Basket(100.0, 0.05, 1.0, 16)
    out = {}
    for bf16 in (False, True):
        mode = fdmc.NOISE_PCA | (fdmc.PCA_BF16 if bf16 else 0)
        eng.qmc_accumulate(model, pay, 64, 16, 42, 0, 1 << 18, mode, dev_pca=pca)
        eng.timing_enable(True)
        eng.timing_read(reset=True)
        t0 = time.perf_counter()
        for _ in range(reps):
            sums = eng.qmc_accumulate(model, pay, 64, 16, 42, 0, 1 << 18, mode, dev_pca=pca)
        wall = (time.perf_counter() - t0) / reps
        n, ms = eng.timing_read(reset=True)
        eng.timing_enable(False)
        flops = 2.0 * 16 * (1 << 18) * 1024 * 1024
        tf = flops / (ms / n * 1e-3) / 1e12
        out["bf16" if bf16 else "fp16"] = {"gemm_ms": ms / n, "tflops": tf, "frac_mfma_peak": tf / MFMA_PEAK_TFLOPS,
                                           "pass_ms": 1e3 * wall,
                                           "price": sum(r[0] for r in sums) / (16 << 18)}
    pca.free()
    return out


def c4_heston_mlmc(eng):
    """Config 4: Heston MLMC, M = 4, L <= 8, eps = 5e-4, pilot 1e4, European call."""
    import fdmc

    hes = fdmc.HestonParams(100.0, 0.05, 2.0, 0.04, 0.5, -0.7, 0.04, 1.0).compute_cholesky()
    cfg = fdmc.MLMCConfig(M=4, max_L=8, pilot_n=10000, seed=42)
    eng.run_mlmc_cuda(hes, fdmc.European(100, 0.05, 1), 5e-4, cfg)
    ts = []
    for _ in range(3):
        r = eng.run_mlmc_cuda(hes, fdmc.European(100, 0.05, 1), 5e-4, cfg)
        ts.append(r.time_s)
    return {"eps": 5e-4, "time_to_eps_s": sorted(ts)[1], "price": r.price, "std_error": r.std_error,
            "levels": r.levels, "path_steps": r.path_steps, "path_steps_per_s": r.path_steps / sorted(ts)[1]}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--points", type=int, default=1 << 20)
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build runs on one GPU only; --gpus %d is not supported" % args.gpus, file=sys.stderr)
        sys.exit(2)

    import fdmc

    n_steps, R, N = 256, 32, args.points
    m = fdmc.GBMParams(100.0, 0.05, 0.3, 1.0)
    asian, geo = fdmc.Asian(100.0), fdmc.GeomAsian(100.0)
    E = fdmc.geom_asian_analytic(100, 100, 1, 0.05, 0.3, n_steps)
    eng = fdmc.Engine(0)
    bb = eng.bb_upload(fdmc.bb_precompute(n_steps, 1.0))

    def step():
        return eng.qmc_accumulate(m, asian, n_steps, R, 42, 0, N, fdmc.NOISE_BB, dev_bb=bb,
                                  ctrl_model=m, ctrl_payoff=geo, E_ctrl=E, beta=1.0)

    log("warmup", args.warmup)
    for _ in range(args.warmup):
        sums = step()
    eng.synchronize()
    eng.timing_enable(True)
    eng.timing_read(reset=True)
    log("timed steps", args.steps)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        sums = step()
    eng.synchronize()
    wall = time.perf_counter() - t0
    launches, kms = eng.timing_read(reset=True)
    eng.timing_enable(False)
    path_steps = R * N * n_steps
    ms_per_step = 1e3 * wall / args.steps
    value = path_steps * args.steps / wall
    k_avg_ms = kms / max(1, launches)
    achieved = path_steps * FLOPS_PER_PATH_STEP / (k_avg_ms * 1e-3) / 1e12
    price = sum(s[2] for s in sums) / (R * N)
    log("kernel avg %.3f ms, wall/step %.3f ms, price %.6f" % (k_avg_ms, ms_per_step, price))

    # time to eps = 1e-3 (price), full doubling loop from 2^14 points per replica
    cfg = fdmc.QMCConfig(R=32, N0=1 << 14, N_max=1 << 20)
    tte = []
    for _ in range(4):
        r = eng.run_qmc_cv_cuda(m, m, asian, geo, E, 1.0, 1e-3, n_steps, cfg, fdmc.NOISE_BB, dev_bb=bb)
        tte.append(r.time_s)
    tte = sorted(tte[1:])[1]
    # same, with price + adjoint Greeks (delta, vega, rho) and the CV on the Greeks
    dE = fdmc.geom_asian_greeks(100, 100, 1, 0.05, 0.3, n_steps)
    ttg = []
    for _ in range(4):
        rg, g = eng.run_qmc_greeks(m, asian, 1e-3, n_steps, cfg, fdmc.NOISE_BB, bb, ctrl_model=m, ctrl_payoff=geo,
                                   E_ctrl=E, dE_ctrl=dE, beta=1.0)
        ttg.append(rg.time_s)
    ttg = sorted(ttg[1:])[1]
    bb.free()
    log("c3 gemm")
    c3 = c3_gemm(eng)
    log("c4 heston mlmc")
    c4 = c4_heston_mlmc(eng)
    eng.close()

    out = {
        "metric": METRIC, "value": value, "unit": "path-steps/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": ms_per_step, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic (scrambled Sobol points, seeded salts)",
        "config": {"workload": "RQMC arithmetic-Asian call, GBM S0=K=100 r=0.05 sigma=0.3 T=1, D=256 Brownian "
                               "bridge, R=32 replicas x %d points, geometric-Asian control variate" % N,
                   "path_steps_per_step": path_steps, "replicas": R, "points_per_replica": N, "dates": n_steps},
        "roofline": {"bound": "valu", "achieved": achieved, "peak": VALU_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / VALU_PEAK_TFLOPS, "traffic": None,
                     "kernel": "k_path<GBM,Asian,CV_SHARED,GeomAsian,SOBOL_BB,LB=4>",
                     "kernel_avg_ms": k_avg_ms, "flops_per_path_step": FLOPS_PER_PATH_STEP},
        "time_to_eps": {"eps": 1e-3, "price_s": tte, "price": r.price, "std_error": r.std_error,
                        "n_samples": r.n_samples, "price_and_greeks_s": ttg,
                        "greeks": {"delta": g.delta, "vega": g.vega, "rho": g.rho, "delta_se": g.delta_se,
                                   "vega_se": g.vega_se, "rho_se": g.rho_se}},
        "c3_pca_gemm": c3,
        "c4_heston_mlmc": c4,
    }
    if not args.no_cpu:
        log("cpu baseline")
        out["cpu_baseline"] = cpu_baseline()
    print(json.dumps(out))


if __name__ == "__main__":
    main()
