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
    bb.free()
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
                        "n_samples": r.n_samples},
    }
    if not args.no_cpu:
        log("cpu baseline")
        out["cpu_baseline"] = cpu_baseline()
    print(json.dumps(out))


if __name__ == "__main__":
    main()
