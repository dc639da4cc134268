/*
 * fdmc.h -- C ABI of the MI355X (gfx950) Monte Carlo path-simulation engine.
 *
 * Plain C types only (no HIP, torch or C++ types).  Every run_* entry point
 * replaces one function of the reference's methods_cuda.cuh; the citation is
 * given beside each declaration.  The C++ facade methods_hip.hpp re-exposes the
 * reference's own signatures (std::variant models/payoffs, MCResult, ...) on top
 * of this ABI.
 *
 * Conventions
 *  - Every function returns an fdmc_status.  On failure a message is available
 *    from fdmc_last_error() (thread-local, valid until the next call on the
 *    same thread).  No C++ exception crosses this boundary.
 *  - Inputs are validated before any device work is queued.
 *  - run_* calls are synchronous: they return after the device work finished.
 *  - A context owns one HIP stream (default priority), the device tables
 *    (Sobol direction numbers, scrambled direction-number caches) and grow-only
 *    scratch buffers.  Use one context per host thread.
 *  - Pointers named d_* are device pointers; all others are host pointers.
 */
#ifndef FDMC_H
#define FDMC_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define FDMC_ABI_VERSION 1

typedef enum {
    FDMC_OK = 0,
    FDMC_E_INVALID = 1,     /* bad argument (reference: std::runtime_error from validation) */
    FDMC_E_UNSUPPORTED = 2, /* model/payoff/mode combination the engine rejects */
    FDMC_E_RANGE = 3,       /* size or offset beyond a documented limit */
    FDMC_E_DEVICE = 4,      /* HIP runtime error */
    FDMC_E_ALLOC = 5,       /* device or host allocation failed */
    FDMC_E_INTERNAL = 6
} fdmc_status;

/* ---- models (reference models.hpp:15-71; kinds 4 and 5 are additive) ---- */
typedef enum {
    FDMC_MODEL_GBM = 0,          /* GBMParams          models.hpp:15-21 */
    FDMC_MODEL_HESTON = 1,       /* HestonParams       models.hpp:23-43 */
    FDMC_MODEL_DUPIRE = 2,       /* DupireLocalParams  models.hpp:45-54 */
    FDMC_MODEL_MULTI_DUPIRE = 3, /* MultiDupireParams  models.hpp:56-68 */
    FDMC_MODEL_LOCALVOL_QUAD = 4,/* sigma(S,t) = a + b exp(-c t) ln(S/s_ref)^2 */
    FDMC_MODEL_MULTI_GBM = 5     /* n correlated GBM assets with per-asset sigma */
} fdmc_model_kind;

typedef struct fdmc_model {
    int32_t kind;
    int32_t n_assets;            /* MULTI_*: number of assets */
    double S0, mu, T;
    double sigma;                /* GBM */
    double kappa, theta, xi, rho, v0; /* HESTON */
    double L[4];                 /* HESTON: 2x2 Cholesky, row-major (compute_cholesky) */
    double sigma0, alpha, beta_d;/* DUPIRE, MULTI_DUPIRE */
    double lv_a, lv_b, lv_c, lv_sref; /* LOCALVOL_QUAD */
    int32_t uncorrelated;        /* MULTI_*: nonzero -> identity correlation */
    int32_t heston_scheme;       /* HESTON: FDMC_HESTON_EXP_EULER (default) or FDMC_HESTON_MILSTEIN */
    const double* S0_assets;     /* MULTI_*: n_assets values */
    const double* sigma_assets;  /* MULTI_GBM: n_assets values */
    const double* L_assets;      /* MULTI_*: n x n lower Cholesky, row-major; NULL -> identity */
} fdmc_model;
/* Heston variance schemes.  EXP_EULER is the reference's full-truncation step
 * V' = theta + e^{-kappa h}(V - theta) + xi sqrt(V+) dW2 (methods_cuda.cu:276-284);
 * MILSTEIN adds the Milstein term xi^2/4 (dW2^2 - h) to it (no counterpart in
 * the reference, whose comment names the scheme Milstein but omits the term). */
#define FDMC_HESTON_EXP_EULER 0
#define FDMC_HESTON_MILSTEIN 1

/* ---- payoffs (reference payoffs.hpp:15-22) ---- */
typedef enum {
    FDMC_PAYOFF_EUROPEAN = 0,    /* max(S_T-K,0) e^{-rT} */
    FDMC_PAYOFF_ASIAN = 1,       /* max(mean S_k - K, 0), undiscounted */
    FDMC_PAYOFF_GEOM_ASIAN = 2,  /* max(exp(mean log S_k) - K, 0), undiscounted */
    FDMC_PAYOFF_LOOKBACK = 3,    /* S_T - min S_k (BGK-shifted), undiscounted */
    FDMC_PAYOFF_BARRIER = 4,     /* up-and-out call with BGK shift, e^{-rT} */
    FDMC_PAYOFF_BASKET = 5       /* max(sum w_a S_T^a - K, 0) e^{-rT} */
} fdmc_payoff_kind;

typedef struct fdmc_payoff {
    int32_t kind;
    int32_t n_assets;            /* BASKET */
    double K, r, T;              /* r and T only set the discount */
    double B;                    /* BARRIER level */
    double sigma;                /* LOOKBACK/BARRIER: BGK volatility */
    const double* weights;       /* BASKET: n_assets weights; NULL -> 1/n */
} fdmc_payoff;

/* ---- configurations (reference methods_cuda.cuh:10-32; trailing fields additive) ---- */
typedef struct fdmc_mc_config {  /* MCConfig */
    int64_t N_batch;             /* kept for API parity; paths per launch cap (0 = no cap) */
    int32_t pilot_n;             /* default 10000 */
    uint32_t seed;               /* default 123 */
} fdmc_mc_config;

typedef struct fdmc_qmc_config { /* QMCConfig */
    int32_t R;                   /* replicas, default 32 */
    int32_t max_doublings;       /* default 20 */
    int64_t N0;                  /* first points per replica (power of 2), default 1<<14 */
    int64_t N_max;               /* cap on points per replica (0 = 2^32) */
    uint32_t salt_base;          /* replica salt = salt_base + r, default 42 */
    int32_t pca_precision;       /* PCA GEMM operands: 0 = FP16 (default), 1 = BF16 */
} fdmc_qmc_config;

typedef struct fdmc_mlmc_config {/* MLMCConfig */
    int32_t M;                   /* refinement factor, default 2 */
    int32_t max_L;               /* default 10 */
    int32_t pilot_n;             /* default 400 */
    uint32_t seed;               /* PRNG key for run_mlmc, default 42 */
} fdmc_mlmc_config;

typedef enum { FDMC_NOISE_RAW = 0, FDMC_NOISE_BB = 1, FDMC_NOISE_PCA = 2,
               FDMC_NOISE_PRNG = 3 /* accumulate hooks only: Philox paths */ } fdmc_noise_mode;
/* qmc_accumulate: OR into mode to run the PCA GEMM on BF16 instead of FP16 operands. */
#define FDMC_PCA_BF16 0x100

#define FDMC_FLAG_PRNG_FALLBACK 1u   /* QMC dimension > 20000: pseudo-random points were used */
#define FDMC_FLAG_NOT_CONVERGED 2u   /* stopping rule not met within the configured limits (book: or an estimate that is not finite) */

typedef struct fdmc_result {     /* MCResult (methods_cuda.cuh:41-46) + additive fields */
    double price;
    double std_error;
    int64_t n_samples;
    double time_s;
    int64_t path_steps;          /* path-steps simulated (all levels, replicas) */
    int32_t levels;              /* MLMC/MLQMC: finest level L + 1; QMC: doublings done */
    uint32_t flags;
} fdmc_result;

/* Pathwise adjoint Greeks (no counterpart in the reference): delta = dV/dS0,
 * vega = dV/d(volatility parameter: sigma, Dupire sigma0, local-vol level a),
 * rho = dV/d(mu = r); each with its standard error across replicas. */
typedef struct fdmc_greeks {
    double delta, vega, rho;
    double delta_se, vega_se, rho_se;
} fdmc_greeks;

typedef struct fdmc_cv_pilot_result { double beta, var_plain, var_cv; } fdmc_cv_pilot_result; /* CVPilot */

/* ---- contract book (no counterpart in the reference, which prices one
 * contract per run_mlqmc_cuda call, methods_cuda.cuh:90-96) ----
 * Every contract of a book shares one one-factor model (GBM, Dupire or the
 * quadratic local vol) and is priced by multilevel QMC with refinement M = 2 and
 * a fixed finest level: level l simulates 2^l Euler steps over [0, T] (the
 * coarse path of level l > 0 takes 2^{l-1} steps on summed fine increments),
 * l = 0 .. log2(n_steps), driven by the Brownian bridge of Hong-Hickernell
 * scrambled Sobol points with salt
 *     salt_base + (c * FDMC_BOOK_LEVELS + l) * R + r
 * (contract c, replica r; contract 0 has run_mlqmc_cuda's 42 + l R + r,
 * methods_cuda.cu:1845).  Sample allocation, merges and the per-point variance
 * S2_l follow run_mlqmc_cuda (methods_cuda.cu:1680-2019) level by level (a
 * top-up adds at least a tenth of the level's points); a contract stops when
 * sum_l S2_l / (N_l R) <= eps^2 / 2 (its bias is that of its own date grid, so
 * there is no bias test). */
#define FDMC_BOOK_LEVELS 10           /* levels 0..9: n_steps up to 512 */

typedef struct fdmc_book_contract {
    int32_t kind;                /* FDMC_PAYOFF_EUROPEAN, _ASIAN, _LOOKBACK or _BARRIER */
    int32_t n_steps;             /* monitoring dates: a power of two, 1 .. 512 */
    double K;                    /* strike (unused by the lookback) */
    double B;                    /* up-and-out barrier (barrier only) */
    double T;                    /* maturity; the book holds at most 8 distinct values */
    double r;                    /* discount rate of European and barrier payoffs */
    double sigma_bgk;            /* BGK continuity correction vol (lookback, barrier) */
} fdmc_book_contract;

typedef struct fdmc_book_config {
    int32_t R;                   /* replicas, default 16 */
    int32_t pilot_n;             /* points per replica and level in the first pass, default 400 */
    uint32_t salt_base;          /* default 42 */
    int32_t greeks;              /* delta, vega, rho of Lipschitz payoffs: 0 none, 1 pathwise adjoint
                                    (default; reverse sweep recomputing states from the Sobol index),
                                    2 forward tangents (same derivative, one pass) */
    int32_t max_iter;            /* passes per contract (pilot included), default 50 */
    int32_t iters_per_graph;     /* batched mode: passes captured per hipGraph, default 4 */
    int32_t variant;             /* launch shape: -1 autotune on the first run (default), else the
                                    kernel variant of fdmc_book_stats.variant ... */
    int32_t wg_per_cu;           /* ... and persistent workgroups per CU (variant >= 0 only) */
} fdmc_book_config;

typedef enum { FDMC_BOOK_BATCHED = 0,  /* all contracts per pass, device-side planning, hipGraph */
               FDMC_BOOK_PLAIN = 1     /* one contract at a time, host-driven passes */
} fdmc_book_mode;

typedef struct fdmc_book_result {
    double price, std_error;
    double delta, vega, rho;     /* NaN for barrier contracts (not Lipschitz) or greeks = 0 */
    double delta_se, vega_se, rho_se;
    int64_t n_samples;           /* sum over levels of N_l R */
    int64_t path_steps;          /* sum over levels of N_l R 2^l */
    int32_t levels;              /* log2(n_steps) + 1 */
    int32_t iterations;          /* passes that added points */
    uint32_t flags;              /* FDMC_FLAG_NOT_CONVERGED */
    int32_t pad;
} fdmc_book_result;

typedef struct fdmc_book_stats {
    double time_s;               /* wall time of the run, device work included */
    double tune_s;               /* autotuning time (first run only; not in time_s) */
    int64_t path_steps;          /* all contracts */
    int32_t passes;              /* device passes (batched) or host iterations (plain) */
    int32_t graph_replays;       /* batched mode */
    int32_t variant;             /* autotuned kernel: the compiler's register allocation (0) or the
                                    three-waves-per-SIMD compilation (1) */
    int32_t items_target;        /* noise dimensions per lane and work item chosen by the autotuner */
    int32_t wg_per_cu;           /* persistent workgroups per CU chosen by the autotuner */
    int32_t pad;
} fdmc_book_stats;

typedef struct fdmc_context fdmc_context;
typedef struct fdmc_bb fdmc_bb;    /* DeviceBBData  (methods_cuda.cuh:52) */
typedef struct fdmc_pca fdmc_pca;  /* DevicePCAData (methods_cuda.cuh:53) */
typedef struct fdmc_book fdmc_book;

/* ---- library / context ---- */
int32_t fdmc_abi_version(void);
const char* fdmc_last_error(void);
fdmc_status fdmc_context_create(int32_t device, fdmc_context** out);
fdmc_status fdmc_context_destroy(fdmc_context* ctx);
/* The context's HIP stream as an opaque handle (hipStream_t), for event timing. */
fdmc_status fdmc_context_stream(fdmc_context* ctx, void** stream_out);
fdmc_status fdmc_synchronize(fdmc_context* ctx);
/* Kernel timing: when enabled, every fused path/level kernel launch is bracketed
 * by HIP events on the context stream; read the count and summed milliseconds
 * since the last reset (no counterpart in the reference; used by bench.py). */
fdmc_status fdmc_timing_enable(fdmc_context* ctx, int32_t enable);
fdmc_status fdmc_timing_read(fdmc_context* ctx, int64_t* launches, double* total_ms, int32_t reset);

/* ---- Brownian bridge / PCA data (methods_cuda.cuh:56-63) ---- */
/* bb_upload: tables as produced by bb_precompute (utils.cpp:54-101); N a power of 2
 * with the dyadic coarse-to-fine schedule (checked). */
fdmc_status fdmc_bb_upload(fdmc_context* ctx, int32_t N, double T,
                           const int32_t* map_idx, const int32_t* left_idx,
                           const int32_t* right_idx, const double* weight_left,
                           const double* weight_right, const double* std_dev,
                           fdmc_bb** out);
fdmc_status fdmc_bb_free(fdmc_bb* bb);
/* pca_upload: M (m x m, column-major, dW = M Z) as produced by pca_compute
 * (utils.cpp:131-167).  Stored on the device in FP16 and BF16. */
fdmc_status fdmc_pca_upload(fdmc_context* ctx, int32_t m, const double* M_colmajor,
                            fdmc_pca** out);
fdmc_status fdmc_pca_free(fdmc_pca* pca);

/* ---- estimators ---- */
/* run_mc_cuda (methods_cuda.cuh:71-72). */
fdmc_status fdmc_run_mc(fdmc_context* ctx, const fdmc_model* model, const fdmc_payoff* payoff,
                        double eps, int32_t n_steps, const fdmc_mc_config* cfg,
                        fdmc_result* out);
/* run_is_cuda (methods_cuda.cuh:127-128): importance-sampled GBM European call;
 * increments (z + z_star / sqrt(n)) sqrt(h) over n = max(4, ceil(T / eps)) steps,
 * payoffs weighted by the likelihood ratio.  Pilot of cfg->pilot_n paths
 * (Philox seed cfg->seed), then max(ceil(2 Var / eps^2), pilot_n) paths (seed + 1). */
fdmc_status fdmc_run_is(fdmc_context* ctx, const fdmc_model* model, const fdmc_payoff* payoff,
                        double z_star, double eps, const fdmc_mc_config* cfg, fdmc_result* out);
/* run_mc_fixed (methods_cuda.cuh:131-134): mean and variance of the mean. */
fdmc_status fdmc_run_mc_fixed(fdmc_context* ctx, const fdmc_model* model,
                              const fdmc_payoff* payoff, int32_t n_steps, int64_t n_paths,
                              uint32_t seed, double* mean_out, double* var_of_mean_out);
/* run_mlmc_cuda (methods_cuda.cuh:75-76). */
fdmc_status fdmc_run_mlmc(fdmc_context* ctx, const fdmc_model* model, const fdmc_payoff* payoff,
                          double eps, const fdmc_mlmc_config* cfg, fdmc_result* out);
/* run_qmc_cuda (methods_cuda.cuh:82-87). */
fdmc_status fdmc_run_qmc(fdmc_context* ctx, const fdmc_model* model, const fdmc_payoff* payoff,
                         double eps, int32_t n_steps, const fdmc_qmc_config* cfg,
                         int32_t mode, const fdmc_bb* bb, const fdmc_pca* pca,
                         fdmc_result* out);
/* run_mlqmc_cuda (methods_cuda.cuh:90-96). */
fdmc_status fdmc_run_mlqmc(fdmc_context* ctx, const fdmc_model* model, const fdmc_payoff* payoff,
                           double eps, const fdmc_mlmc_config* ml_cfg,
                           const fdmc_qmc_config* qmc_cfg, int32_t mode,
                           const fdmc_bb* const* bb_list, int32_t n_bb,
                           const fdmc_pca* const* pca_list, int32_t n_pca,
                           fdmc_result* out);
/* cv_pilot (methods_cuda.cuh:103-108). */
fdmc_status fdmc_cv_pilot(fdmc_context* ctx, const fdmc_model* main_model,
                          const fdmc_model* ctrl_model, const fdmc_payoff* main_payoff,
                          const fdmc_payoff* ctrl_payoff, double E_ctrl, int32_t n_steps,
                          int32_t N_pilot, uint32_t seed, fdmc_cv_pilot_result* out);
/* run_mc_cv_cuda (methods_cuda.cuh:110-116). */
fdmc_status fdmc_run_mc_cv(fdmc_context* ctx, const fdmc_model* main_model,
                           const fdmc_model* ctrl_model, const fdmc_payoff* main_payoff,
                           const fdmc_payoff* ctrl_payoff, double E_ctrl, double beta,
                           double eps, int32_t n_steps, const fdmc_mc_config* cfg,
                           fdmc_result* out);
/* run_qmc_cv_cuda (methods_cuda.cuh:118-124); mode/bb are additive (the reference
 * is Raw-only, methods_cuda.cu:2241-2242). */
fdmc_status fdmc_run_qmc_cv(fdmc_context* ctx, const fdmc_model* main_model,
                            const fdmc_model* ctrl_model, const fdmc_payoff* main_payoff,
                            const fdmc_payoff* ctrl_payoff, double E_ctrl, double beta,
                            double eps, int32_t n_steps, const fdmc_qmc_config* cfg,
                            int32_t mode, const fdmc_bb* bb, fdmc_result* out);

/* One fixed QMC pass (the body of accumulate_range, methods_cuda.cu:1549-1622) for
 * R replicas over points [offset, offset + n_points) of each replica's scrambled
 * sequence (salt = salt_base + r).  sums_out: R x n_acc doubles per replica:
 *   n_acc = 2 {sum Y, sum Y^2} without a control, or 7 with one
 *   {Y, Y^2, Ycv, Ycv^2, Yc, Yc^2, Y*Yc} (methods_cuda.cu:628-639); baskets of
 *   n <= 32 assets give 2 + 2 n (the delta slots are zero here), larger ones 2:
 *   a buffer of R x 66 doubles always suffices.
 * ctrl_model/ctrl_payoff may be NULL (no control variate). */
fdmc_status fdmc_qmc_accumulate(fdmc_context* ctx, const fdmc_model* model,
                                const fdmc_payoff* payoff, const fdmc_model* ctrl_model,
                                const fdmc_payoff* ctrl_payoff, double E_ctrl, double beta,
                                int32_t n_steps, int32_t R, uint32_t salt_base,
                                int64_t offset, int64_t n_points, int32_t mode,
                                const fdmc_bb* bb, const fdmc_pca* pca, double* sums_out,
                                int32_t* n_acc_out);

/* fdmc_qmc_accumulate that allocates nothing (no counterpart in the reference):
 * for callers that cannot allocate device or pinned memory -- code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc.  One-factor
 * models, Raw or BB noise on Sobol points (D = n_steps <= 20000), an optional
 * control leg; the same kernels and summation order as fdmc_qmc_accumulate, so
 * the sums are identical.  The scrambled direction records (R x D x 256 B) and the
 * per-workgroup FP64 partials live in a static device array of
 * FDMC_NOALLOC_SCRATCH_BYTES per device (held until the process ends); a pass
 * that needs more is rejected with FDMC_E_RANGE.  d_sums_out: device memory of
 * R x n_acc doubles (n_acc = 2, or 7 with a control), written on the context's
 * stream; the call returns after the device work is done.  Calls that overlap on
 * one device (from several threads or contexts) share the array and race. */
#define FDMC_NOALLOC_SCRATCH_BYTES (32u << 20)
fdmc_status fdmc_qmc_accumulate_noalloc(fdmc_context* ctx, const fdmc_model* model,
                                        const fdmc_payoff* payoff, const fdmc_model* ctrl_model,
                                        const fdmc_payoff* ctrl_payoff, double E_ctrl, double beta,
                                        int32_t n_steps, int32_t R, uint32_t salt_base,
                                        int64_t offset, int64_t n_points, int32_t mode,
                                        const fdmc_bb* bb, double* d_sums_out, int32_t* n_acc_out);

/* One coupled fine/coarse pass (the body of launch_level, methods_cuda.cu:1347-1395
 * and 1800-1924) for n_req (level, replica, sample range) requests, run as one
 * persistent launch.  qmc = 0: Philox paths (seed = salt_or_seed, stream
 * 0x100 + level); qmc = 1: Sobol points of salt salt_or_seed + level*R + replica
 * (mode RAW, BB with bb_list[level], or PCA with pca_list[level], m = M^level,
 * optionally | FDMC_PCA_BF16; PCA runs one GEMM and one launch per level and
 * group of consecutive replicas with equal ranges, at most 8 GiB of normals and
 * rows at a time: a larger request runs in chunks of its range).  sums_out: n_req x
 * {sum dY, sum dY^2, sum Yf, sum Yf^2} (methods_cuda.cu:514-517). */
fdmc_status fdmc_levels_accumulate(fdmc_context* ctx, const fdmc_model* model,
                                   const fdmc_payoff* payoff, int32_t M, int32_t qmc,
                                   int32_t mode, int32_t R, uint32_t salt_or_seed, int32_t n_req,
                                   const int32_t* levels, const int32_t* replicas,
                                   const int64_t* first, const int64_t* count,
                                   const fdmc_bb* const* bb_list, int32_t n_bb,
                                   const fdmc_pca* const* pca_list, int32_t n_pca,
                                   double* sums_out);

/* ---- adjoint Greeks (additive; the reference prices only) ----
 * Fused price + reverse-mode pathwise Greeks for one-factor GBM / Dupire /
 * local-vol models and Lipschitz payoffs (European, Asian, geometric Asian,
 * lookback), optionally with the geometric-Asian control on the same path.
 * Path states are recomputed in reverse from the Sobol index (or Philox
 * counter) with S checkpoints every 16 steps; no path is stored.
 * greeks_accumulate: one pass; sums_out R x n_acc with n_acc = 2 (+5 with a
 * control) price sums followed by {sum g, sum g^2} for delta, vega, rho;
 * for GBM baskets n_acc = 2 + 2 n with {sum, sum^2} of each asset's delta.
 * mode: RAW, BB or FDMC_NOISE_PRNG (Philox stream seed salt_or_seed + r). */
fdmc_status fdmc_greeks_accumulate(fdmc_context* ctx, const fdmc_model* model,
                                   const fdmc_payoff* payoff, const fdmc_model* ctrl_model,
                                   const fdmc_payoff* ctrl_payoff, double E_ctrl, double beta,
                                   int32_t n_steps, int32_t R, uint32_t salt_or_seed,
                                   int64_t offset, int64_t n_points, int32_t mode,
                                   const fdmc_bb* bb, double* sums_out, int32_t* n_acc_out);
/* greeks_accumulate for baskets of GBM assets (n <= 32) under PCA noise
 * (replaces the basket branch of run_qmc_cuda's PCA pass, methods_cuda.cu:1151-1166
 * and 1593, which prices only): per replica the price sums and {sum, sum^2} of each
 * asset's pathwise delta disc 1{B > K} w_a S_T^a / S0^a, from the same MFMA GEMM
 * (product epilogue when the asset count is a power of two and m a multiple of
 * 256, else dW rows) as fdmc_qmc_accumulate.  mode: FDMC_NOISE_PCA, optionally
 * | FDMC_PCA_BF16; pca: the uploaded m x m increment matrix (m = n_steps * n).
 * sums_out: R x (2 + 2 n) doubles. */
fdmc_status fdmc_greeks_accumulate_pca(fdmc_context* ctx, const fdmc_model* model,
                                       const fdmc_payoff* payoff, int32_t n_steps, int32_t R,
                                       uint32_t salt_base, int64_t offset, int64_t n_points,
                                       int32_t mode, const fdmc_pca* pca, double* sums_out,
                                       int32_t* n_acc_out);
/* run_qmc_cuda / run_qmc_cv_cuda with Greeks: same doubling loop and stopping
 * rule on the price; the control's expectation E_ctrl enters the Greeks through
 * dE_ctrl[3] = its delta, vega, rho (NULL: no control). */
fdmc_status fdmc_run_qmc_greeks(fdmc_context* ctx, const fdmc_model* model,
                                const fdmc_payoff* payoff, const fdmc_model* ctrl_model,
                                const fdmc_payoff* ctrl_payoff, double E_ctrl,
                                const double* dE_ctrl, double beta, double eps, int32_t n_steps,
                                const fdmc_qmc_config* cfg, int32_t mode, const fdmc_bb* bb,
                                fdmc_result* out, fdmc_greeks* greeks);

/* PCA noise of one replica as the path kernels see it: scrambled-Sobol normals
 * rounded to FP16 (bf16 = 0) or BF16 (bf16 = 1), times M^T on MFMA with FP32
 * accumulation (phase1_pca, methods_cuda.cu:1151-1166).  d_out: [n_points][m]
 * FP32 device rows; m must be a multiple of 128. */
fdmc_status fdmc_pca_transform(fdmc_context* ctx, const fdmc_pca* pca, uint32_t salt, int64_t offset,
                               int64_t n_points, int32_t bf16, float* d_out);

/* The same PCA noise pushed through the Euler factors of an n_assets GBM basket
 * the way the basket path forms them on matrix cores (phase1_pca,
 * methods_cuda.cu:1151-1166, then the GBM case of kernel_multi_dupire's steps,
 * methods_cuda.cu:526-551, fused into the GEMM's epilogue): column k of dW
 * belongs to asset k % n_assets, and per point, asset and 128-column slab the
 * output is the product of c0 + sig[asset] * dW over the slab's columns of that
 * asset below m (columns m .. of the padding count as 1).  d_out: [mp / 128]
 * [n_points][n_assets] FP32 device memory.  Needs mp (m rounded up to 128) a
 * multiple of 256 and n_assets a power of two in [2, 32].  A diagnostic of the
 * fused path (the basket kernels consume these products directly). */
fdmc_status fdmc_pca_products(fdmc_context* ctx, const fdmc_pca* pca, uint32_t salt, int64_t offset,
                              int64_t n_points, int32_t bf16, int32_t n_assets, float c0, const float* sig,
                              float* d_out);

/* ---- scrambled Sobol generator (methods_cuda.cu:753-844) ----
 * d_out is [D][n] (dimension-major, as kernel_gen_hh_scrambled_sobol_normal writes).
 * Point p of dimension d is the Hong-Hickernell-scrambled Gray-code Sobol point
 * with index offset + p, replica salt `salt`.  offset + n must fit in 32 bits. */
fdmc_status fdmc_sobol_scrambled_u32(fdmc_context* ctx, uint32_t salt, int32_t D,
                                     int64_t n, int64_t offset, uint32_t* d_out);
/* Same points mapped to N(0,1) (u = (s + 1/2) 2^-32, own FP32 inverse CDF). */
fdmc_status fdmc_sobol_normal(fdmc_context* ctx, uint32_t salt, int32_t D, int64_t n,
                              int64_t offset, float* d_out);
/* Map n uint32 values through the device inverse CDF (test hook for the ICDF). */
fdmc_status fdmc_icdf_u32(fdmc_context* ctx, const uint32_t* d_in, int64_t n, float* d_out);

/* ---- contract book (replaces one run_mlqmc_cuda call per contract,
 * methods_cuda.cuh:90-96, with one pass over every contract) ----
 * fdmc_book_create copies the contracts, builds the scrambled direction records
 * of every (contract, level, replica) and the bridge tables on the device.
 * model: GBM, DUPIRE or LOCALVOL_QUAD (its T is ignored: each contract has its own). */
fdmc_status fdmc_book_create(fdmc_context* ctx, const fdmc_model* model,
                             const fdmc_book_contract* contracts, int32_t n_contracts,
                             const fdmc_book_config* cfg, fdmc_book** out);
/* Price (and Greeks) of every contract to standard error eps / sqrt(2) of the
 * price.  mode: fdmc_book_mode; both modes run the same kernels and give
 * bit-identical results.  results: n_contracts entries; stats may be NULL.
 * The first call autotunes the kernel variant (reported in stats->tune_s)
 * unless the config fixes it.  With fdmc_timing_enable on, the batched mode
 * runs its passes without graph capture and brackets every pass kernel
 * (k_book) with HIP events, for fdmc_timing_read; results are unchanged. */
fdmc_status fdmc_book_run(fdmc_book* book, double eps, int32_t mode, fdmc_book_result* results,
                          fdmc_book_stats* stats);
/* Per-replica sums of one pass of contract c, level l over points
 * [offset, offset + n) of every replica: sums_out[r * 4 + k], k = price
 * difference P_f - P_c, then its delta, vega and rho (test hook). */
fdmc_status fdmc_book_level_sums(fdmc_book* book, int32_t c, int32_t l, int64_t offset, int64_t n,
                                 double* sums_out);
fdmc_status fdmc_book_free(fdmc_book* book);

#ifdef __cplusplus
}
#endif
#endif /* FDMC_H */
