"""Python mirror of the reference's host interface over the engine's C ABI.

The reference exposes C++ free functions (methods_cuda.cuh:54-134) over the
variant types of models.hpp / payoffs.hpp.  This module mirrors those names
(``GBMParams``, ``European``, ``run_qmc_cuda`` ...) on top of ``include/fdmc.h``
through ctypes so that tests and benchmarks read like the reference's own calls.
The compute always runs in ``lib/libfdmc.so`` (HIP, gfx950): if the library is
missing, loading fails loudly -- there is no CPU fallback.
"""
import ctypes as C
import math
import os
from collections import namedtuple

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libfdmc.so")

# ---------------------------------------------------------------- enums
MODEL_GBM, MODEL_HESTON, MODEL_DUPIRE, MODEL_MULTI_DUPIRE, MODEL_LOCALVOL_QUAD, MODEL_MULTI_GBM = range(6)
PAYOFF_EUROPEAN, PAYOFF_ASIAN, PAYOFF_GEOM_ASIAN, PAYOFF_LOOKBACK, PAYOFF_BARRIER, PAYOFF_BASKET = range(6)
NOISE_RAW, NOISE_BB, NOISE_PCA, NOISE_PRNG = 0, 1, 2, 3
FLAG_PRNG_FALLBACK, FLAG_NOT_CONVERGED = 1, 2
HESTON_EXP_EULER, HESTON_MILSTEIN = 0, 1  # Heston variance schemes
PCA_BF16 = 0x100  # OR into the mode of qmc_accumulate
BGK_BETA = 0.5826  # payoffs.hpp:7
D_MAX_SOBOL = 20000  # models.hpp:9


class NoiseMode:  # methods_cuda.cuh:35
    Raw, BrownianBridge, PCA = NOISE_RAW, NOISE_BB, NOISE_PCA


# ---------------------------------------------------------------- C structs (fdmc.h)
class CModel(C.Structure):
    _fields_ = [("kind", C.c_int32), ("n_assets", C.c_int32),
                ("S0", C.c_double), ("mu", C.c_double), ("T", C.c_double), ("sigma", C.c_double),
                ("kappa", C.c_double), ("theta", C.c_double), ("xi", C.c_double), ("rho", C.c_double),
                ("v0", C.c_double), ("L", C.c_double * 4),
                ("sigma0", C.c_double), ("alpha", C.c_double), ("beta_d", C.c_double),
                ("lv_a", C.c_double), ("lv_b", C.c_double), ("lv_c", C.c_double), ("lv_sref", C.c_double),
                ("uncorrelated", C.c_int32), ("heston_scheme", C.c_int32),
                ("S0_assets", C.POINTER(C.c_double)), ("sigma_assets", C.POINTER(C.c_double)),
                ("L_assets", C.POINTER(C.c_double))]


class CPayoff(C.Structure):
    _fields_ = [("kind", C.c_int32), ("n_assets", C.c_int32),
                ("K", C.c_double), ("r", C.c_double), ("T", C.c_double), ("B", C.c_double),
                ("sigma", C.c_double), ("weights", C.POINTER(C.c_double))]


class CMCConfig(C.Structure):
    _fields_ = [("N_batch", C.c_int64), ("pilot_n", C.c_int32), ("seed", C.c_uint32)]


class CQMCConfig(C.Structure):
    _fields_ = [("R", C.c_int32), ("max_doublings", C.c_int32), ("N0", C.c_int64),
                ("N_max", C.c_int64), ("salt_base", C.c_uint32), ("pca_precision", C.c_int32)]


class CMLMCConfig(C.Structure):
    _fields_ = [("M", C.c_int32), ("max_L", C.c_int32), ("pilot_n", C.c_int32), ("seed", C.c_uint32)]


class CResult(C.Structure):
    _fields_ = [("price", C.c_double), ("std_error", C.c_double), ("n_samples", C.c_int64),
                ("time_s", C.c_double), ("path_steps", C.c_int64), ("levels", C.c_int32),
                ("flags", C.c_uint32)]


class CGreeks(C.Structure):
    _fields_ = [("delta", C.c_double), ("vega", C.c_double), ("rho", C.c_double),
                ("delta_se", C.c_double), ("vega_se", C.c_double), ("rho_se", C.c_double)]


class CCVPilot(C.Structure):
    _fields_ = [("beta", C.c_double), ("var_plain", C.c_double), ("var_cv", C.c_double)]


class CBookContract(C.Structure):
    _fields_ = [("kind", C.c_int32), ("n_steps", C.c_int32), ("K", C.c_double), ("B", C.c_double),
                ("T", C.c_double), ("r", C.c_double), ("sigma_bgk", C.c_double)]


class CBookConfig(C.Structure):
    _fields_ = [("R", C.c_int32), ("pilot_n", C.c_int32), ("salt_base", C.c_uint32), ("greeks", C.c_int32),
                ("max_iter", C.c_int32), ("iters_per_graph", C.c_int32), ("variant", C.c_int32),
                ("wg_per_cu", C.c_int32)]


class CBookResult(C.Structure):
    _fields_ = [("price", C.c_double), ("std_error", C.c_double), ("delta", C.c_double), ("vega", C.c_double),
                ("rho", C.c_double), ("delta_se", C.c_double), ("vega_se", C.c_double), ("rho_se", C.c_double),
                ("n_samples", C.c_int64), ("path_steps", C.c_int64), ("levels", C.c_int32),
                ("iterations", C.c_int32), ("flags", C.c_uint32), ("pad", C.c_int32)]


class CBookStats(C.Structure):
    _fields_ = [("time_s", C.c_double), ("tune_s", C.c_double), ("path_steps", C.c_int64), ("passes", C.c_int32),
                ("graph_replays", C.c_int32), ("variant", C.c_int32), ("items_target", C.c_int32),
                ("wg_per_cu", C.c_int32), ("pad", C.c_int32)]


MCResult = namedtuple("MCResult", "price std_error n_samples time_s path_steps levels flags")
CVPilot = namedtuple("CVPilot", "beta var_plain var_cv")
Greeks = namedtuple("Greeks", "delta vega rho delta_se vega_se rho_se")
BookResult = namedtuple("BookResult", "price std_error delta vega rho delta_se vega_se rho_se n_samples "
                        "path_steps levels iterations flags")
BookStats = namedtuple("BookStats", "time_s tune_s path_steps passes graph_replays variant items_target wg_per_cu")


def _result(r):
    return MCResult(r.price, r.std_error, r.n_samples, r.time_s, r.path_steps, r.levels, r.flags)


def _darr(vals):
    if vals is None:
        return None
    a = (C.c_double * len(vals))(*[float(v) for v in vals])
    return a


# ---------------------------------------------------------------- models (models.hpp:15-68)
class _Model:
    kind = None

    def to_c(self):
        m = CModel()
        m.kind = self.kind
        m.S0, m.mu, m.T = getattr(self, "S0", 100.0), self.mu, self.T
        self._fill(m)
        return m

    def _fill(self, m):
        pass


class GBMParams(_Model):
    kind = MODEL_GBM

    def __init__(self, S0=100.0, mu=0.05, sigma=0.20, T=1.0):
        self.S0, self.mu, self.sigma, self.T = S0, mu, sigma, T

    def _fill(self, m):
        m.sigma = self.sigma


class HestonParams(_Model):
    kind = MODEL_HESTON

    def __init__(self, S0=100.0, mu=0.05, kappa=2.0, theta=0.04, xi=0.5, rho=-0.9, v0=0.04, T=1.0,
                 scheme=HESTON_EXP_EULER):
        # scheme: HESTON_EXP_EULER is the reference's full-truncation step
        # (methods_cuda.cu:276-284); HESTON_MILSTEIN adds xi^2/4 (dW2^2 - h)
        self.S0, self.mu, self.kappa, self.theta, self.xi = S0, mu, kappa, theta, xi
        self.rho, self.v0, self.T = rho, v0, T
        self.scheme = scheme
        self.L = [0.0, 0.0, 0.0, 0.0]

    def compute_cholesky(self):  # models.hpp:39-42
        self.L = [1.0, 0.0, self.rho, math.sqrt(1.0 - self.rho * self.rho)]
        return self

    def _fill(self, m):
        m.kappa, m.theta, m.xi, m.rho, m.v0 = self.kappa, self.theta, self.xi, self.rho, self.v0
        m.heston_scheme = self.scheme
        for i in range(4):
            m.L[i] = self.L[i]


class DupireLocalParams(_Model):
    kind = MODEL_DUPIRE

    def __init__(self, S0=100.0, mu=0.05, sigma0=0.20, alpha=0.5, beta_d=0.7, T=1.0):
        self.S0, self.mu, self.sigma0, self.alpha, self.beta_d, self.T = S0, mu, sigma0, alpha, beta_d, T

    def _fill(self, m):
        m.sigma0, m.alpha, m.beta_d = self.sigma0, self.alpha, self.beta_d


class LocalVolQuadParams(_Model):
    """sigma(S,t) = a + b exp(-c t) ln(S/s_ref)^2 (additive model kind)."""
    kind = MODEL_LOCALVOL_QUAD

    def __init__(self, S0=100.0, mu=0.05, a=0.2, b=0.1, c=1.0, s_ref=100.0, T=1.0):
        self.S0, self.mu, self.a, self.b, self.c, self.s_ref, self.T = S0, mu, a, b, c, s_ref, T

    def _fill(self, m):
        m.lv_a, m.lv_b, m.lv_c, m.lv_sref = self.a, self.b, self.c, self.s_ref


class MultiDupireParams(_Model):
    kind = MODEL_MULTI_DUPIRE

    def __init__(self, n=100, mu=0.05, sigma0=0.20, alpha=0.5, beta_d=0.7, T=1.0,
                 uncorrelated=False, S0=None, L=None):
        self.n, self.mu, self.sigma0, self.alpha, self.beta_d, self.T = n, mu, sigma0, alpha, beta_d, T
        self.uncorrelated = uncorrelated
        self.S0v = list(S0) if S0 is not None else [100.0] * n
        self.Lv = list(L) if L is not None else None

    def to_c(self):
        m = CModel()
        m.kind, m.n_assets, m.mu, m.T = self.kind, self.n, self.mu, self.T
        m.S0 = self.S0v[0] if self.S0v else 100.0
        m.sigma0, m.alpha, m.beta_d = self.sigma0, self.alpha, self.beta_d
        m.uncorrelated = 1 if self.uncorrelated else 0
        self._keep = (_darr(self.S0v), _darr(self.Lv))
        m.S0_assets = C.cast(self._keep[0], C.POINTER(C.c_double))
        if self._keep[1] is not None:
            m.L_assets = C.cast(self._keep[1], C.POINTER(C.c_double))
        return m


class MultiGBMParams(_Model):
    """n correlated GBM assets with per-asset sigma (additive model kind)."""
    kind = MODEL_MULTI_GBM

    def __init__(self, n, S0, sigmas, mu=0.05, T=1.0, L=None, uncorrelated=False):
        self.n, self.mu, self.T = n, mu, T
        self.S0v, self.sig, self.Lv = list(S0), list(sigmas), (list(L) if L is not None else None)
        self.uncorrelated = uncorrelated

    def to_c(self):
        m = CModel()
        m.kind, m.n_assets, m.mu, m.T = self.kind, self.n, self.mu, self.T
        m.S0 = self.S0v[0]
        m.uncorrelated = 1 if self.uncorrelated else 0
        self._keep = (_darr(self.S0v), _darr(self.sig), _darr(self.Lv))
        m.S0_assets = C.cast(self._keep[0], C.POINTER(C.c_double))
        m.sigma_assets = C.cast(self._keep[1], C.POINTER(C.c_double))
        if self._keep[2] is not None:
            m.L_assets = C.cast(self._keep[2], C.POINTER(C.c_double))
        return m


# ---------------------------------------------------------------- payoffs (payoffs.hpp:15-22)
class _Payoff:
    kind = None

    def to_c(self):
        p = CPayoff()
        p.kind = self.kind
        for f in ("K", "r", "T", "B", "sigma"):
            if hasattr(self, f):
                setattr(p, f, getattr(self, f))
        return p


class European(_Payoff):
    kind = PAYOFF_EUROPEAN

    def __init__(self, K, r, T):
        self.K, self.r, self.T = K, r, T


class Asian(_Payoff):
    kind = PAYOFF_ASIAN

    def __init__(self, K):
        self.K = K


class GeomAsian(_Payoff):
    kind = PAYOFF_GEOM_ASIAN

    def __init__(self, K):
        self.K = K


class Lookback(_Payoff):
    kind = PAYOFF_LOOKBACK

    def __init__(self, sigma):
        self.sigma = sigma


class Barrier(_Payoff):
    kind = PAYOFF_BARRIER

    def __init__(self, K, B, sigma, r, T):
        self.K, self.B, self.sigma, self.r, self.T = K, B, sigma, r, T


class Basket(_Payoff):
    kind = PAYOFF_BASKET

    def __init__(self, K, r, T, n_assets, weights=None):
        self.K, self.r, self.T, self.n_assets = K, r, T, n_assets
        self.weights = list(weights) if weights is not None else None

    def to_c(self):
        p = super().to_c()
        p.n_assets = self.n_assets
        self._keep = _darr(self.weights)
        if self._keep is not None:
            p.weights = C.cast(self._keep, C.POINTER(C.c_double))
        return p


# ---------------------------------------------------------------- configs (methods_cuda.cuh:10-32)
class MCConfig:
    def __init__(self, N_batch=1 << 16, pilot_n=10000, seed=123):
        self.N_batch, self.pilot_n, self.seed = N_batch, pilot_n, seed

    def to_c(self):
        return CMCConfig(self.N_batch, self.pilot_n, self.seed)


class QMCConfig:
    def __init__(self, R=32, max_doublings=20, N0=1 << 14, N_max=0, salt_base=42, pca_bf16=False):
        self.R, self.max_doublings, self.N0, self.N_max, self.salt_base = R, max_doublings, N0, N_max, salt_base
        self.pca_bf16 = pca_bf16

    def to_c(self):
        return CQMCConfig(self.R, self.max_doublings, self.N0, self.N_max, self.salt_base, 1 if self.pca_bf16 else 0)


class MLMCConfig:
    def __init__(self, M=2, max_L=10, pilot_n=400, seed=42):
        self.M, self.max_L, self.pilot_n, self.seed = M, max_L, pilot_n, seed

    def to_c(self):
        return CMLMCConfig(self.M, self.max_L, self.pilot_n, self.seed)


# ---------------------------------------------------------------- contract book
BOOK_LEVELS = 10          # levels 0..9: up to 512 dates
BOOK_BATCHED, BOOK_PLAIN = 0, 1


class BookContract:
    """One contract of a book (payoff kind, dates, strike, barrier, maturity)."""

    def __init__(self, kind, n_steps, K=100.0, B=0.0, T=1.0, r=0.05, sigma_bgk=0.2):
        self.kind, self.n_steps, self.K, self.B, self.T, self.r, self.sigma_bgk = kind, n_steps, K, B, T, r, sigma_bgk

    def to_c(self):
        return CBookContract(self.kind, self.n_steps, self.K, self.B, self.T, self.r, self.sigma_bgk)


class BookConfig:
    def __init__(self, R=16, pilot_n=400, salt_base=42, greeks=True, max_iter=50, iters_per_graph=4, variant=-1,
                 wg_per_cu=0):
        self.R, self.pilot_n, self.salt_base, self.greeks = R, pilot_n, salt_base, greeks
        self.max_iter, self.iters_per_graph = max_iter, iters_per_graph
        self.variant, self.wg_per_cu = variant, wg_per_cu

    def to_c(self):
        # greeks: False/0 none, True/1 adjoint, 2 forward tangents
        return CBookConfig(self.R, self.pilot_n, self.salt_base, int(self.greeks), self.max_iter,
                           self.iters_per_graph, self.variant, self.wg_per_cu)


def local_vol_book(n=1024, seed=2026, r=0.05):
    """Seeded book under the quadratic local vol sigma(S,t) = 0.2 + 0.1 e^{-t} ln(S/100)^2:
    payoff kind uniform over European / Asian / lookback / barrier, K ~ U[70, 130],
    T in {0.25, 0.5, 1, 2}, dates 2^k for k in 4..9 (16..512), barrier B = 1.3 K,
    BGK vol 0.2 (the surface's level).  Returns (model, contracts)."""
    rng = np.random.default_rng(seed)
    kinds = [PAYOFF_EUROPEAN, PAYOFF_ASIAN, PAYOFF_LOOKBACK, PAYOFF_BARRIER]
    kind = rng.integers(0, 4, n)
    K = rng.uniform(70.0, 130.0, n)
    T = rng.choice(np.array([0.25, 0.5, 1.0, 2.0]), n)
    steps = 2 ** rng.integers(4, 10, n)
    model = LocalVolQuadParams(S0=100.0, mu=r, a=0.2, b=0.1, c=1.0, s_ref=100.0, T=1.0)
    book = [BookContract(kinds[int(kind[i])], int(steps[i]), float(K[i]), 1.3 * float(K[i]), float(T[i]), r, 0.2)
            for i in range(n)]
    return model, book


# ---------------------------------------------------------------- library binding
_P = C.c_void_p
_SIGS = {
    "fdmc_abi_version": (C.c_int32, []),
    "fdmc_last_error": (C.c_char_p, []),
    "fdmc_context_create": (C.c_int, [C.c_int32, C.POINTER(_P)]),
    "fdmc_context_destroy": (C.c_int, [_P]),
    "fdmc_context_stream": (C.c_int, [_P, C.POINTER(_P)]),
    "fdmc_synchronize": (C.c_int, [_P]),
    "fdmc_timing_enable": (C.c_int, [_P, C.c_int32]),
    "fdmc_timing_read": (C.c_int, [_P, C.POINTER(C.c_int64), C.POINTER(C.c_double), C.c_int32]),
    "fdmc_bb_upload": (C.c_int, [_P, C.c_int32, C.c_double, C.POINTER(C.c_int32), C.POINTER(C.c_int32),
                                 C.POINTER(C.c_int32), C.POINTER(C.c_double), C.POINTER(C.c_double),
                                 C.POINTER(C.c_double), C.POINTER(_P)]),
    "fdmc_bb_free": (C.c_int, [_P]),
    "fdmc_pca_upload": (C.c_int, [_P, C.c_int32, C.POINTER(C.c_double), C.POINTER(_P)]),
    "fdmc_pca_free": (C.c_int, [_P]),
    "fdmc_run_mc": (C.c_int, [_P, C.POINTER(CModel), C.POINTER(CPayoff), C.c_double, C.c_int32,
                              C.POINTER(CMCConfig), C.POINTER(CResult)]),
    "fdmc_run_is": (C.c_int, [_P, C.POINTER(CModel), C.POINTER(CPayoff), C.c_double, C.c_double,
                              C.POINTER(CMCConfig), C.POINTER(CResult)]),
    "fdmc_run_mc_fixed": (C.c_int, [_P, C.POINTER(CModel), C.POINTER(CPayoff), C.c_int32, C.c_int64,
                                    C.c_uint32, C.POINTER(C.c_double), C.POINTER(C.c_double)]),
    "fdmc_run_mlmc": (C.c_int, [_P, C.POINTER(CModel), C.POINTER(CPayoff), C.c_double,
                                C.POINTER(CMLMCConfig), C.POINTER(CResult)]),
    "fdmc_run_qmc": (C.c_int, [_P, C.POINTER(CModel), C.POINTER(CPayoff), C.c_double, C.c_int32,
                               C.POINTER(CQMCConfig), C.c_int32, _P, _P, C.POINTER(CResult)]),
    "fdmc_run_mlqmc": (C.c_int, [_P, C.POINTER(CModel), C.POINTER(CPayoff), C.c_double,
                                 C.POINTER(CMLMCConfig), C.POINTER(CQMCConfig), C.c_int32,
                                 C.POINTER(_P), C.c_int32, C.POINTER(_P), C.c_int32, C.POINTER(CResult)]),
    "fdmc_cv_pilot": (C.c_int, [_P, C.POINTER(CModel), C.POINTER(CModel), C.POINTER(CPayoff),
                                C.POINTER(CPayoff), C.c_double, C.c_int32, C.c_int32, C.c_uint32,
                                C.POINTER(CCVPilot)]),
    "fdmc_run_mc_cv": (C.c_int, [_P, C.POINTER(CModel), C.POINTER(CModel), C.POINTER(CPayoff),
                                 C.POINTER(CPayoff), C.c_double, C.c_double, C.c_double, C.c_int32,
                                 C.POINTER(CMCConfig), C.POINTER(CResult)]),
    "fdmc_run_qmc_cv": (C.c_int, [_P, C.POINTER(CModel), C.POINTER(CModel), C.POINTER(CPayoff),
                                  C.POINTER(CPayoff), C.c_double, C.c_double, C.c_double, C.c_int32,
                                  C.POINTER(CQMCConfig), C.c_int32, _P, C.POINTER(CResult)]),
    "fdmc_qmc_accumulate": (C.c_int, [_P, C.POINTER(CModel), C.POINTER(CPayoff), C.POINTER(CModel),
                                      C.POINTER(CPayoff), C.c_double, C.c_double, C.c_int32, C.c_int32,
                                      C.c_uint32, C.c_int64, C.c_int64, C.c_int32, _P, _P,
                                      C.POINTER(C.c_double), C.POINTER(C.c_int32)]),
    "fdmc_qmc_accumulate_noalloc": (C.c_int, [_P, C.POINTER(CModel), C.POINTER(CPayoff), C.POINTER(CModel),
                                              C.POINTER(CPayoff), C.c_double, C.c_double, C.c_int32, C.c_int32,
                                              C.c_uint32, C.c_int64, C.c_int64, C.c_int32, _P, _P,
                                              C.POINTER(C.c_int32)]),
    "fdmc_levels_accumulate": (C.c_int, [_P, C.POINTER(CModel), C.POINTER(CPayoff), C.c_int32, C.c_int32,
                                         C.c_int32, C.c_int32, C.c_uint32, C.c_int32, C.POINTER(C.c_int32),
                                         C.POINTER(C.c_int32), C.POINTER(C.c_int64), C.POINTER(C.c_int64),
                                         C.POINTER(_P), C.c_int32, C.POINTER(_P), C.c_int32,
                                         C.POINTER(C.c_double)]),
    "fdmc_greeks_accumulate": (C.c_int, [_P, C.POINTER(CModel), C.POINTER(CPayoff), C.POINTER(CModel),
                                         C.POINTER(CPayoff), C.c_double, C.c_double, C.c_int32, C.c_int32,
                                         C.c_uint32, C.c_int64, C.c_int64, C.c_int32, _P,
                                         C.POINTER(C.c_double), C.POINTER(C.c_int32)]),
    "fdmc_greeks_accumulate_pca": (C.c_int, [_P, C.POINTER(CModel), C.POINTER(CPayoff), C.c_int32, C.c_int32,
                                             C.c_uint32, C.c_int64, C.c_int64, C.c_int32, _P,
                                             C.POINTER(C.c_double), C.POINTER(C.c_int32)]),
    "fdmc_run_qmc_greeks": (C.c_int, [_P, C.POINTER(CModel), C.POINTER(CPayoff), C.POINTER(CModel),
                                      C.POINTER(CPayoff), C.c_double, C.POINTER(C.c_double), C.c_double,
                                      C.c_double, C.c_int32, C.POINTER(CQMCConfig), C.c_int32, _P,
                                      C.POINTER(CResult), C.POINTER(CGreeks)]),
    "fdmc_pca_transform": (C.c_int, [_P, _P, C.c_uint32, C.c_int64, C.c_int64, C.c_int32, _P]),
    "fdmc_pca_products": (C.c_int, [_P, _P, C.c_uint32, C.c_int64, C.c_int64, C.c_int32, C.c_int32, C.c_float,
                                    _P, _P]),
    "fdmc_sobol_scrambled_u32": (C.c_int, [_P, C.c_uint32, C.c_int32, C.c_int64, C.c_int64, _P]),
    "fdmc_sobol_normal": (C.c_int, [_P, C.c_uint32, C.c_int32, C.c_int64, C.c_int64, _P]),
    "fdmc_icdf_u32": (C.c_int, [_P, _P, C.c_int64, _P]),
    "fdmc_book_create": (C.c_int, [_P, C.POINTER(CModel), C.POINTER(CBookContract), C.c_int32,
                                   C.POINTER(CBookConfig), C.POINTER(_P)]),
    "fdmc_book_run": (C.c_int, [_P, C.c_double, C.c_int32, C.POINTER(CBookResult), C.POINTER(CBookStats)]),
    "fdmc_book_level_sums": (C.c_int, [_P, C.c_int32, C.c_int32, C.c_int64, C.c_int64, C.POINTER(C.c_double)]),
    "fdmc_book_free": (C.c_int, [_P]),
}

_lib = None


def load_library(path=None):
    """Load libfdmc.so and declare every signature; raises if it is missing."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = path or LIB_PATH
    if not os.path.exists(p):
        raise RuntimeError("libfdmc.so not built at %s: run __graft_entry__.build()" % p)
    lib = C.CDLL(p)
    for name, (res, args) in _SIGS.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


class FdmcError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("fdmc status %d: %s" % (code, msg))
        self.code = code


def _check(lib, rc):
    if rc != 0:
        msg = lib.fdmc_last_error()
        raise FdmcError(rc, msg.decode() if msg else "")


class DeviceBB:
    def __init__(self, eng, handle, N):
        self.eng, self.handle, self.N = eng, handle, N

    def free(self):
        if self.handle:
            self.eng.lib.fdmc_bb_free(self.handle)
            self.handle = None


class DevicePCA:
    def __init__(self, eng, handle, m):
        self.eng, self.handle, self.m = eng, handle, m

    def free(self):
        if self.handle:
            self.eng.lib.fdmc_pca_free(self.handle)
            self.handle = None


class Engine:
    """One device context; methods mirror methods_cuda.cuh:56-134."""

    def __init__(self, device=0, lib_path=None):
        self.lib = load_library(lib_path)
        h = _P()
        _check(self.lib, self.lib.fdmc_context_create(device, C.byref(h)))
        self.ctx = h

    def close(self):
        if self.ctx:
            self.lib.fdmc_context_destroy(self.ctx)
            self.ctx = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def stream(self):
        s = _P()
        _check(self.lib, self.lib.fdmc_context_stream(self.ctx, C.byref(s)))
        return s.value

    def synchronize(self):
        _check(self.lib, self.lib.fdmc_synchronize(self.ctx))

    def timing_enable(self, on=True):
        _check(self.lib, self.lib.fdmc_timing_enable(self.ctx, 1 if on else 0))

    def timing_read(self, reset=True):
        n, ms = C.c_int64(), C.c_double()
        _check(self.lib, self.lib.fdmc_timing_read(self.ctx, C.byref(n), C.byref(ms), 1 if reset else 0))
        return n.value, ms.value

    # -- BB / PCA
    def bb_upload(self, bb):
        """bb: object with N, T, map_idx, left_idx, right_idx, weight_left, weight_right, std_dev."""
        N = bb.N
        ia = lambda v: (C.c_int32 * N)(*[int(x) for x in v])
        da = lambda v: (C.c_double * N)(*[float(x) for x in v])
        h = _P()
        _check(self.lib, self.lib.fdmc_bb_upload(self.ctx, N, bb.T, ia(bb.map_idx), ia(bb.left_idx),
                                                  ia(bb.right_idx), da(bb.weight_left), da(bb.weight_right),
                                                  da(bb.std_dev), C.byref(h)))
        return DeviceBB(self, h, N)

    def pca_upload(self, m, M_colmajor):
        arr = (C.c_double * (m * m))(*[float(x) for x in M_colmajor])
        h = _P()
        _check(self.lib, self.lib.fdmc_pca_upload(self.ctx, m, arr, C.byref(h)))
        return DevicePCA(self, h, m)

    # -- estimators
    def run_mc_cuda(self, model, payoff, eps, n_steps, cfg=None):
        cfg = cfg or MCConfig()
        r = CResult()
        m, p, c = model.to_c(), payoff.to_c(), cfg.to_c()
        _check(self.lib, self.lib.fdmc_run_mc(self.ctx, C.byref(m), C.byref(p), eps, n_steps, C.byref(c),
                                              C.byref(r)))
        return _result(r)

    def run_is_cuda(self, model, payoff, z_star, eps, cfg=None):
        """Importance-sampled GBM European call (run_is_cuda)."""
        cfg = cfg or MCConfig()
        r = CResult()
        m, p, c = model.to_c(), payoff.to_c(), cfg.to_c()
        _check(self.lib, self.lib.fdmc_run_is(self.ctx, C.byref(m), C.byref(p), z_star, eps, C.byref(c),
                                              C.byref(r)))
        return _result(r)

    def run_mc_fixed(self, model, payoff, n_steps, n_paths, seed):
        m, p = model.to_c(), payoff.to_c()
        mean, var = C.c_double(), C.c_double()
        _check(self.lib, self.lib.fdmc_run_mc_fixed(self.ctx, C.byref(m), C.byref(p), n_steps, n_paths, seed,
                                                    C.byref(mean), C.byref(var)))
        return mean.value, var.value

    def run_mlmc_cuda(self, model, payoff, eps, cfg=None):
        cfg = cfg or MLMCConfig()
        r = CResult()
        m, p, c = model.to_c(), payoff.to_c(), cfg.to_c()
        _check(self.lib, self.lib.fdmc_run_mlmc(self.ctx, C.byref(m), C.byref(p), eps, C.byref(c), C.byref(r)))
        return _result(r)

    def run_qmc_cuda(self, model, payoff, eps, n_steps, cfg=None, mode=NOISE_RAW, dev_bb=None, dev_pca=None):
        cfg = cfg or QMCConfig()
        r = CResult()
        m, p, c = model.to_c(), payoff.to_c(), cfg.to_c()
        _check(self.lib, self.lib.fdmc_run_qmc(self.ctx, C.byref(m), C.byref(p), eps, n_steps, C.byref(c), mode,
                                               dev_bb.handle if dev_bb else None,
                                               dev_pca.handle if dev_pca else None, C.byref(r)))
        return _result(r)

    def run_mlqmc_cuda(self, model, payoff, eps, ml_cfg=None, qmc_cfg=None, mode=NOISE_RAW, bb_list=(),
                       pca_list=()):
        ml_cfg, qmc_cfg = ml_cfg or MLMCConfig(), qmc_cfg or QMCConfig()
        r = CResult()
        m, p, a, b = model.to_c(), payoff.to_c(), ml_cfg.to_c(), qmc_cfg.to_c()
        bbs = (_P * max(1, len(bb_list)))(*[x.handle if x else None for x in bb_list])
        pcs = (_P * max(1, len(pca_list)))(*[x.handle if x else None for x in pca_list])
        _check(self.lib, self.lib.fdmc_run_mlqmc(self.ctx, C.byref(m), C.byref(p), eps, C.byref(a), C.byref(b),
                                                 mode, bbs, len(bb_list), pcs, len(pca_list), C.byref(r)))
        return _result(r)

    def cv_pilot(self, main_model, ctrl_model, main_payoff, ctrl_payoff, E_ctrl, n_steps, N_pilot, seed=0):
        out = CCVPilot()
        a, b, c, d = main_model.to_c(), ctrl_model.to_c(), main_payoff.to_c(), ctrl_payoff.to_c()
        _check(self.lib, self.lib.fdmc_cv_pilot(self.ctx, C.byref(a), C.byref(b), C.byref(c), C.byref(d), E_ctrl,
                                                n_steps, N_pilot, seed, C.byref(out)))
        return CVPilot(out.beta, out.var_plain, out.var_cv)

    def run_mc_cv_cuda(self, main_model, ctrl_model, main_payoff, ctrl_payoff, E_ctrl, beta, eps, n_steps,
                       cfg=None):
        cfg = cfg or MCConfig()
        r = CResult()
        a, b, c, d, k = main_model.to_c(), ctrl_model.to_c(), main_payoff.to_c(), ctrl_payoff.to_c(), cfg.to_c()
        _check(self.lib, self.lib.fdmc_run_mc_cv(self.ctx, C.byref(a), C.byref(b), C.byref(c), C.byref(d), E_ctrl,
                                                 beta, eps, n_steps, C.byref(k), C.byref(r)))
        return _result(r)

    def run_qmc_cv_cuda(self, main_model, ctrl_model, main_payoff, ctrl_payoff, E_ctrl, beta, eps, n_steps,
                        cfg=None, mode=NOISE_RAW, dev_bb=None):
        cfg = cfg or QMCConfig()
        r = CResult()
        a, b, c, d, k = main_model.to_c(), ctrl_model.to_c(), main_payoff.to_c(), ctrl_payoff.to_c(), cfg.to_c()
        _check(self.lib, self.lib.fdmc_run_qmc_cv(self.ctx, C.byref(a), C.byref(b), C.byref(c), C.byref(d),
                                                  E_ctrl, beta, eps, n_steps, C.byref(k), mode,
                                                  dev_bb.handle if dev_bb else None, C.byref(r)))
        return _result(r)

    def qmc_accumulate(self, model, payoff, n_steps, R, salt_base, offset, n_points, mode=NOISE_RAW,
                       dev_bb=None, dev_pca=None, ctrl_model=None, ctrl_payoff=None, E_ctrl=0.0, beta=0.0):
        m, p = model.to_c(), payoff.to_c()
        cm = C.byref(ctrl_model.to_c()) if ctrl_model is not None else None
        cp = C.byref(ctrl_payoff.to_c()) if ctrl_payoff is not None else None
        sums = (C.c_double * (R * 66))()   # baskets with deltas: 2 + 2 n
        nacc = C.c_int32()
        _check(self.lib, self.lib.fdmc_qmc_accumulate(self.ctx, C.byref(m), C.byref(p), cm, cp, E_ctrl, beta,
                                                      n_steps, R, salt_base, offset, n_points, mode,
                                                      dev_bb.handle if dev_bb else None,
                                                      dev_pca.handle if dev_pca else None, sums, C.byref(nacc)))
        na = nacc.value
        return [[sums[r * na + i] for i in range(na)] for r in range(R)]

    def levels_accumulate(self, model, payoff, M, reqs, qmc=False, mode=NOISE_RAW, R=1, salt_or_seed=42,
                          bb_list=(), pca_list=()):
        """reqs: list of (level, replica, first, count); returns n x 4 sums.
        mode may carry PCA_BF16 with NOISE_PCA."""
        n = len(reqs)
        lv = (C.c_int32 * n)(*[q[0] for q in reqs])
        rp = (C.c_int32 * n)(*[q[1] for q in reqs])
        fi = (C.c_int64 * n)(*[q[2] for q in reqs])
        co = (C.c_int64 * n)(*[q[3] for q in reqs])
        bbs = (_P * max(1, len(bb_list)))(*[x.handle if x else None for x in bb_list])
        pcs = (_P * max(1, len(pca_list)))(*[x.handle if x else None for x in pca_list])
        out = (C.c_double * (4 * n))()
        m, p = model.to_c(), payoff.to_c()
        _check(self.lib, self.lib.fdmc_levels_accumulate(self.ctx, C.byref(m), C.byref(p), M, 1 if qmc else 0, mode,
                                                         R, salt_or_seed, n, lv, rp, fi, co, bbs, len(bb_list),
                                                         pcs, len(pca_list), out))
        return [[out[4 * i + k] for k in range(4)] for i in range(n)]

    def greeks_accumulate(self, model, payoff, n_steps, R, salt_or_seed, offset, n_points, mode=NOISE_RAW,
                          dev_bb=None, ctrl_model=None, ctrl_payoff=None, E_ctrl=0.0, beta=0.0):
        """One fused price + adjoint-Greeks pass; R rows of [price sums..., sum/sum^2 of delta, vega, rho]."""
        m, p = model.to_c(), payoff.to_c()
        cm = C.byref(ctrl_model.to_c()) if ctrl_model is not None else None
        cp = C.byref(ctrl_payoff.to_c()) if ctrl_payoff is not None else None
        sums = (C.c_double * (R * 66))()   # baskets: 2 + 2 n (n <= 32)
        nacc = C.c_int32()
        _check(self.lib, self.lib.fdmc_greeks_accumulate(self.ctx, C.byref(m), C.byref(p), cm, cp, E_ctrl, beta,
                                                         n_steps, R, salt_or_seed, offset, n_points, mode,
                                                         dev_bb.handle if dev_bb else None, sums, C.byref(nacc)))
        na = nacc.value
        return [[sums[r * na + i] for i in range(na)] for r in range(R)]

    def greeks_accumulate_pca(self, model, payoff, n_steps, R, salt_base, offset, n_points, dev_pca, bf16=False):
        """One PCA pass of a GBM basket with per-asset pathwise deltas; R rows of
        [sum Y, sum Y^2, then {sum, sum^2} of each asset's delta]."""
        m, p = model.to_c(), payoff.to_c()
        sums = (C.c_double * (R * 66))()
        nacc = C.c_int32()
        mode = NOISE_PCA | (PCA_BF16 if bf16 else 0)
        _check(self.lib, self.lib.fdmc_greeks_accumulate_pca(self.ctx, C.byref(m), C.byref(p), n_steps, R,
                                                             salt_base, offset, n_points, mode, dev_pca.handle,
                                                             sums, C.byref(nacc)))
        na = nacc.value
        return [[sums[r * na + i] for i in range(na)] for r in range(R)]

    def run_qmc_greeks(self, model, payoff, eps, n_steps, cfg=None, mode=NOISE_RAW, dev_bb=None, ctrl_model=None,
                       ctrl_payoff=None, E_ctrl=0.0, dE_ctrl=None, beta=0.0):
        cfg = cfg or QMCConfig()
        r, g = CResult(), CGreeks()
        m, p, c = model.to_c(), payoff.to_c(), cfg.to_c()
        cm = C.byref(ctrl_model.to_c()) if ctrl_model is not None else None
        cp = C.byref(ctrl_payoff.to_c()) if ctrl_payoff is not None else None
        dE = (C.c_double * 3)(*(dE_ctrl or (0.0, 0.0, 0.0)))
        _check(self.lib, self.lib.fdmc_run_qmc_greeks(self.ctx, C.byref(m), C.byref(p), cm, cp, E_ctrl, dE, beta, eps,
                                                      n_steps, C.byref(c), mode, dev_bb.handle if dev_bb else None,
                                                      C.byref(r), C.byref(g)))
        return _result(r), Greeks(g.delta, g.vega, g.rho, g.delta_se, g.vega_se, g.rho_se)

    def pca_transform(self, dev_pca, salt, offset, n_points, bf16, d_out):
        _check(self.lib, self.lib.fdmc_pca_transform(self.ctx, dev_pca.handle, salt, offset, n_points,
                                                     1 if bf16 else 0, d_out))

    def pca_products(self, dev_pca, salt, offset, n_points, bf16, c0, sig, d_out):
        """GBM-basket slab products of the PCA path, [mp / 128][n_points][len(sig)] FP32 at d_out."""
        sv = (C.c_float * len(sig))(*sig)
        _check(self.lib, self.lib.fdmc_pca_products(self.ctx, dev_pca.handle, salt, offset, n_points,
                                                    1 if bf16 else 0, len(sig), c0, sv, d_out))

    # -- generator hooks (device pointers)
    def sobol_scrambled_u32(self, salt, D, n, offset, d_out):
        _check(self.lib, self.lib.fdmc_sobol_scrambled_u32(self.ctx, salt, D, n, offset, d_out))

    def sobol_normal(self, salt, D, n, offset, d_out):
        _check(self.lib, self.lib.fdmc_sobol_normal(self.ctx, salt, D, n, offset, d_out))

    def icdf_u32(self, d_in, n, d_out):
        _check(self.lib, self.lib.fdmc_icdf_u32(self.ctx, d_in, n, d_out))

    # -- contract book
    def book_create(self, model, contracts, cfg=None):
        return Book(self, model, contracts, cfg or BookConfig())


class Book:
    """A device-resident contract book (fdmc_book_*)."""

    def __init__(self, eng, model, contracts, cfg):
        self.eng, self.cfg, self.n = eng, cfg, len(contracts)
        arr = (CBookContract * self.n)(*[k.to_c() for k in contracts])
        m, c = model.to_c(), cfg.to_c()
        h = _P()
        _check(eng.lib, eng.lib.fdmc_book_create(eng.ctx, C.byref(m), arr, self.n, C.byref(c), C.byref(h)))
        self.handle = h

    def run(self, eps, mode=BOOK_BATCHED):
        res = (CBookResult * self.n)()
        st = CBookStats()
        _check(self.eng.lib, self.eng.lib.fdmc_book_run(self.handle, eps, mode, res, C.byref(st)))
        out = [BookResult(x.price, x.std_error, x.delta, x.vega, x.rho, x.delta_se, x.vega_se, x.rho_se,
                          x.n_samples, x.path_steps, x.levels, x.iterations, x.flags) for x in res]
        stats = BookStats(st.time_s, st.tune_s, st.path_steps, st.passes, st.graph_replays, st.variant,
                          st.items_target, st.wg_per_cu)
        return out, stats

    def level_sums(self, c, l, offset, n):
        s = (C.c_double * (4 * self.cfg.R))()
        _check(self.eng.lib, self.eng.lib.fdmc_book_level_sums(self.handle, c, l, offset, n, s))
        return [[s[r * 4 + k] for k in range(4)] for r in range(self.cfg.R)]

    def free(self):
        if self.handle:
            self.eng.lib.fdmc_book_free(self.handle)
            self.handle = None


# ---------------------------------------------------------------- host numerics (utils.cpp)
class BBData:
    """bb_precompute (utils.cpp:54-101), FP64."""

    def __init__(self, N, T):
        assert N > 0 and (N & (N - 1)) == 0
        self.N, self.T = N, T
        self.map_idx, self.left_idx, self.right_idx = [0] * N, [0] * N, [0] * N
        self.weight_left, self.weight_right, self.std_dev = [0.0] * N, [0.0] * N, [0.0] * N
        t = [i * T / N for i in range(N + 1)]
        self.map_idx[0], self.std_dev[0] = N, math.sqrt(T)
        step, levels = 1, N.bit_length() - 1
        for lev in range(levels):
            stride = N >> lev
            half = stride >> 1
            for j in range(1 << lev):
                L, R = j * stride, (j + 1) * stride
                M = L + half
                self.map_idx[step], self.left_idx[step], self.right_idx[step] = M, L, R
                self.weight_left[step] = (t[R] - t[M]) / (t[R] - t[L])
                self.weight_right[step] = (t[M] - t[L]) / (t[R] - t[L])
                self.std_dev[step] = math.sqrt((t[M] - t[L]) * (t[R] - t[M]) / (t[R] - t[L]))
                step += 1


def bb_precompute(N, T):
    return BBData(N, T)


def norm_cdf(x):
    return 0.5 * math.erfc(-x / math.sqrt(2.0))


def bs_call(S0, K, T, r, sigma):  # utils.cpp:21-27
    if T <= 0 or sigma <= 0:
        return max(S0 - K * math.exp(-r * T), 0.0)
    sT = math.sqrt(T)
    d1 = (math.log(S0 / K) + (r + 0.5 * sigma * sigma) * T) / (sigma * sT)
    return S0 * norm_cdf(d1) - K * math.exp(-r * T) * norm_cdf(d1 - sigma * sT)


def geom_asian_analytic(S0, K, T, mu, sigma, n):  # utils.cpp:34-47 (undiscounted)
    m_G = math.log(S0) + (mu - 0.5 * sigma * sigma) * T * (n + 1) / (2.0 * n)
    s2 = sigma * sigma * T * (n + 1) * (2 * n + 1) / (6.0 * n * n)
    s = math.sqrt(s2)
    if s < 1e-12:
        return max(math.exp(m_G) - K, 0.0)
    d1 = (m_G + s2 - math.log(K)) / s
    return math.exp(m_G + 0.5 * s2) * norm_cdf(d1) - K * norm_cdf(d1 - s)


def geom_asian_greeks(S0, K, T, mu, sigma, n):
    """(delta, vega, rho) of geom_asian_analytic by central differences (FP64)."""
    f = geom_asian_analytic
    hS, hs, hm = 1e-4 * S0, 1e-5, 1e-5
    return ((f(S0 + hS, K, T, mu, sigma, n) - f(S0 - hS, K, T, mu, sigma, n)) / (2 * hS),
            (f(S0, K, T, mu, sigma + hs, n) - f(S0, K, T, mu, sigma - hs, n)) / (2 * hs),
            (f(S0, K, T, mu + hm, sigma, n) - f(S0, K, T, mu - hm, sigma, n)) / (2 * hm))
