"""Write the golden fixtures of tests/golden/.

sobol_scrambled.json  -- (salt, dim, index) -> scrambled uint32, from the
    oracle's row-by-row restatement of hh_scramble over the Gray-code Sobol
    point (methods_cuda.cu:781-825).  Regression vectors for the oracle and the
    GPU generator; tests/test_hh_seeding.py checks them against a numpy
    restatement of the seeding that shares no code with the oracle.
report_prices.json -- the reference prices printed in the reference's own
    reports (informes_finales/informe_completo.txt, lines cited per entry):
    fixed-seed 500k-path MC runs of the examples' parameters, statistical
    (+-0.01-0.02) pins of model and payoff semantics.

Usage: python tests/golden/make_golden.py  (needs oracle/_build/liboracle.so)
"""
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
import pyoracle as O  # noqa: E402

TRIPLES = [(42, 0, 0), (42, 0, 1), (42, 1, 2), (42, 255, 1023), (73, 17, 12345), (125, 63, 1 << 20),
           (0, 0, 0xFFFFFFFF), (0xFFFFFFFF, 19999, 7), (42 + 31, 128, 99999), (2026, 511, 65535),
           (42 + 5 * 16 + 3, 3, 777), (1, 2, 3), (123456789, 1000, 4000000000)]

REPORT = [
    {"example": "ejemplo01_european", "source": "informes_finales/informe_completo.txt:8", "price": 10.450584,
     "kind": "black_scholes", "model": {"kind": "gbm", "S0": 100, "mu": 0.05, "sigma": 0.2, "T": 1},
     "payoff": {"kind": "european", "K": 100, "r": 0.05, "T": 1}},
    {"example": "ejemplo02_asian", "source": "informes_finales/informe_completo.txt:95", "price": 6.128472,
     "kind": "mc500k", "n_steps": 64, "model": {"kind": "gbm", "S0": 100, "mu": 0.05, "sigma": 0.2, "T": 1},
     "payoff": {"kind": "asian", "K": 100}},
    {"example": "ejemplo03_lookback", "source": "informes_finales/informe_completo.txt:170", "price": 18.155249,
     "kind": "mc500k", "n_steps": 64, "model": {"kind": "gbm", "S0": 100, "mu": 0.05, "sigma": 0.2, "T": 1},
     "payoff": {"kind": "lookback", "sigma": 0.2}},
    {"example": "ejemplo04_barrier", "source": "informes_finales/informe_completo.txt:233", "price": 1.159241,
     "kind": "mc500k", "n_steps": 64, "model": {"kind": "gbm", "S0": 100, "mu": 0.05, "sigma": 0.2, "T": 1},
     "payoff": {"kind": "barrier", "K": 100, "B": 120, "sigma": 0.2, "r": 0.05, "T": 1}},
    {"example": "ejemplo05_heston", "source": "informes_finales/informe_completo.txt:296", "price": 10.163574,
     "kind": "mc500k", "n_steps": 256,
     "model": {"kind": "heston", "S0": 100, "mu": 0.05, "kappa": 2.0, "theta": 0.04, "xi": 0.5, "rho": -0.9,
               "v0": 0.04, "T": 1},
     "payoff": {"kind": "european", "K": 100, "r": 0.05, "T": 1}},
    {"example": "ejemplo06_dupire", "source": "informes_finales/informe_completo.txt:339", "price": 8.943956,
     "kind": "mc500k", "n_steps": 256,
     "model": {"kind": "dupire", "S0": 100, "mu": 0.05, "sigma0": 0.2, "alpha": 0.5, "beta_d": 0.7, "T": 1},
     "payoff": {"kind": "european", "K": 100, "r": 0.05, "T": 1}},
    {"example": "ejemplo07_basket_uncorr", "source": "informes_finales/informe_completo.txt:414", "price": 4.872409,
     "kind": "mc5k", "n_steps": 20,
     "model": {"kind": "multi_dupire", "n": 1000, "S0": 100, "mu": 0.05, "sigma0": 0.2, "alpha": 0.5, "beta_d": 0.7,
               "T": 1, "uncorrelated": True},
     "payoff": {"kind": "basket", "K": 100, "r": 0.05, "T": 1, "n": 1000}},
    {"example": "ejemplo11_otm_is", "source": "informes_finales/informe_completo.txt:605", "price": 0.028643,
     "kind": "black_scholes", "model": {"kind": "gbm", "S0": 100, "mu": 0.05, "sigma": 0.2, "T": 1},
     "payoff": {"kind": "european", "K": 180, "r": 0.05, "T": 1}},
    # run_is_cuda rows of ejemplo11 (z* = 2.78, default MCConfig): price, std error, N
    {"example": "ejemplo11_otm_is", "source": "informes_finales/informe_completo.txt:611", "kind": "is_run",
     "eps": 0.002, "price": 0.0286, "std_error": 0.0004, "n_samples": 10000,
     "model": {"kind": "gbm", "S0": 100, "mu": 0.05, "sigma": 0.2, "T": 1},
     "payoff": {"kind": "european", "K": 180, "r": 0.05, "T": 1}},
    {"example": "ejemplo11_otm_is", "source": "informes_finales/informe_completo.txt:629", "kind": "is_run",
     "eps": 0.0002, "price": 0.0288, "std_error": 0.0002, "n_samples": 32222,
     "model": {"kind": "gbm", "S0": 100, "mu": 0.05, "sigma": 0.2, "T": 1},
     "payoff": {"kind": "european", "K": 180, "r": 0.05, "T": 1}},
]


def main():
    sob = [{"salt": s, "dim": d, "index": i, "value": int(O.sobol_scrambled(s, d, i))} for s, d, i in TRIPLES]
    with open(os.path.join(HERE, "sobol_scrambled.json"), "w") as f:
        json.dump(sob, f, indent=1)
    with open(os.path.join(HERE, "report_prices.json"), "w") as f:
        json.dump(REPORT, f, indent=1)
    print("wrote", len(sob), "Sobol vectors and", len(REPORT), "report prices")


if __name__ == "__main__":
    main()
