/*
 * oracle.h -- CPU restatement of the reference's path-simulation algorithms.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load this library, and only as the checker
 * (or, for bench.py, as the timed CPU baseline).  The product library never
 * links or calls it.
 *
 * Plain C11 (+ optional OpenMP).  Path arithmetic is FP32 and sums are FP64, as
 * in the reference kernels (methods_cuda.cu:271-551); the inverse normal CDF is
 * FP64 (Halley-refined, accurate to ~1e-15) so that GPU normals can be checked
 * against it.  Each function cites the reference lines it restates.
 *
 * Pins (see tests/test_oracle_*.py): Sobol raw points vs scipy.stats.qmc.Sobol;
 * scrambling vs the F2-linearity identity and the lower-triangular bijection;
 * ndtri vs scipy.special.ndtri; BB covariance vs min(t_i,t_j); PCA vs the
 * closed-form eigenpairs and M M^T = h I; BS / geometric-Asian closed forms vs
 * scipy; Philox4x32-10 vs the Random123 known-answer vectors; path kernels vs
 * hand-computed 1-4 step paths and the special-case reductions (Dupire
 * alpha=0,beta=1 -> GBM, Heston xi=0 -> GBM, basket n=1 -> European).
 */
#ifndef FDMC_ORACLE_H
#define FDMC_ORACLE_H

#include <stdint.h>
#include "../include/fdmc.h"

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------- Sobol + Hong-Hickernell (methods_cuda.cu:753-825) ---------------- */
uint32_t orc_splitmix32(uint32_t x);
/* 32 direction words of Sobol dimension dim (< 20000), v[j] = m_{j+1} << (31-j). */
int orc_sobol_dirvectors(int dim, uint32_t v[32]);
/* Unscrambled point with Gray-code index `index` (what curand_init(dv, index)
 * followed by one curand() returns, methods_cuda.cu:815-817). */
uint32_t orc_sobol_raw(int dim, uint32_t index);
/* hh_scramble, row by row exactly as methods_cuda.cu:781-795. */
uint32_t orc_hh_scramble(uint32_t raw, uint32_t base);
/* Scrambled point for (salt, dim, index): base = splitmix32(salt) ^ dim (:819-820). */
uint32_t orc_sobol_scrambled(uint32_t salt, int dim, uint32_t index);
/* Scrambled direction words Lv[j] = L v_j and shift e, so that
 * scrambled(index) = e ^ XOR_{j: bit j of gray(index)} Lv[j]. */
void orc_scrambled_dirs(uint32_t salt, int dim, uint32_t Lv[32], uint32_t* shift);
/* Fill out[d*n + p] for p in [0,n), d in [0,D) (layout of :799-825). */
void orc_sobol_scrambled_block(uint32_t salt, int D, uint32_t offset, int64_t n, uint32_t* out);

/* Reference uniform map ((float)s + 1) * 2^-32 in FP32 (:823); hits 1.0f for
 * s >= 2^32-128 (defect documented, not used downstream). */
float orc_u_reference(uint32_t s);
/* The engine's map: u = (s + 1/2) 2^-32, exact in FP64, strictly inside (0,1). */
double orc_u_open(uint32_t s);
/* Inverse standard normal CDF in FP64. */
double orc_ndtri(double u);
/* Normal from a uint32 through orc_u_open (computed in the folded lower half). */
double orc_normal_u32(uint32_t s);
void orc_sobol_normal_block(uint32_t salt, int D, uint32_t offset, int64_t n, float* out);

/* ---------------- counter-based PRNG (Salmon et al., SC'11) ---------------- */
void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
/* Normal number j of stream (seed, tag, path): Box-Muller on Philox words. */
void orc_prng_normals(uint32_t seed, uint32_t tag, uint64_t path, int64_t count, float* out);

/* ---------------- Brownian bridge (utils.cpp:54-124, methods_cuda.cu:561-594) ---------------- */
typedef struct orc_bb {
    int N;
    double T;
    int32_t* map_idx;
    int32_t* left_idx;
    int32_t* right_idx;
    double* wl;
    double* wr;
    double* sd;
} orc_bb;
int orc_bb_precompute(int N, double T, orc_bb* out);
void orc_bb_release(orc_bb* bb);
/* FP64 bb_apply for one path: z[N] (BB order) -> dW[N] (time order). */
void orc_bb_apply_f64(const orc_bb* bb, const double* z, double* dW);
/* FP32 kernel_bb_transform for one path (weights cast to float as bb_upload does). */
void orc_bb_apply_f32(const orc_bb* bb, const float* z, float* dW);

/* ---------------- PCA (utils.cpp:131-167) ---------------- */
/* Symmetric eigendecomposition by cyclic Jacobi: A (n x n row-major, destroyed),
 * eigenvalues descending in w, eigenvectors as columns of V (row-major). */
int orc_eigh_jacobi(int n, double* A, double* w, double* V);
/* M (m x m, column-major): dW = M z, as pca_compute. */
int orc_pca_compute(int m, double T, double* M_colmajor);
/* Kronecker PCA of (C_time(n_steps,T) (x) Corr(n)) with dimension index k*n + a;
 * M is D x D column-major (D = n_steps * n) and dW_{k,a} = (M z)_{k*n+a}. */
int orc_pca_kron_compute(int n_steps, double T, int n, const double* corr, double* M_colmajor);

/* ---------------- closed forms (utils.cpp:14-47) ---------------- */
double orc_bs_call(double S0, double K, double T, double r, double sigma);
double orc_geom_asian_analytic(double S0, double K, double T, double mu, double sigma, int n);

/* ---------------- path kernels (methods_cuda.cu:205-551) ---------------- */
/* Flattened FP32 parameters (make_params, methods_cuda.cu:986-1051). */
typedef struct orc_params {
    int model_kind, payoff_kind, n_assets;
    float S0, mu, T, h, discount, sigma;
    float kappa, theta, xi, rho, v0, L10, L11;
    float mil_q;          /* Heston Milstein coefficient xi^2/4 (0: the reference's scheme) */
    float sigma0, alpha, beta_d;
    float lv_a, lv_b, lv_c, lv_sref;
    float K, B;
    int bgk_on;
    float sigma_bgk;
    const double* S0_assets;
    const double* sigma_assets;
    const double* weights;
} orc_params;
int orc_make_params(const fdmc_model* m, const fdmc_payoff* p, int n_steps, orc_params* out);
/* kernel_mc per path: dW holds n_steps*nf scaled increments, index k*nf + c. */
float orc_path_single(const orc_params* P, const float* dW, int n_steps);
/* Terminal S and V of Heston paths from explicit (dW1, dW2) pairs (test hook). */
void orc_heston_terminal(const orc_params* P, const float* dW, int n_steps, int n_paths,
                         float* S_T, float* V_T);
/* kernel_mlmc per path (with running updates for every model, fixing the Heston
 * omission at methods_cuda.cu:464-470).  in: scaled dW (1 factor) or unscaled Z
 * (Heston), n_fine*nf values. */
void orc_path_coupled(const orc_params* P, const float* in, int n_fine, int M,
                      float* pf, float* pc);
/* Basket of n GBM or Dupire assets with correlated increments:
 * dW_corr[k*n+a] already correlated and scaled. */
float orc_path_basket(const orc_params* P, const float* dW, int n_steps);
/* CV pair on one path (kernel_gbm_asian_cv / kernel_dupire_gbm_cv):
 * returns main and control payoffs. */
void orc_path_cv(const orc_params* Pm, const orc_params* Pc, const float* dW, int n_steps,
                 float* y_main, float* y_ctrl);

/* ---------------- estimators (methods_cuda.cu:1173-2278) ---------------- */
/* Per-replica sums of one QMC pass over [offset, offset+n) (accumulate_range). */
int orc_qmc_accumulate(const fdmc_model* model, const fdmc_payoff* payoff,
                       const fdmc_model* ctrl_model, const fdmc_payoff* ctrl_payoff,
                       double E_ctrl, double beta, int n_steps, int R, uint32_t salt_base,
                       int64_t offset, int64_t n, int mode, const orc_bb* bb,
                       const double* pca_M, double* sums, int* n_acc);
int orc_run_qmc(const fdmc_model* model, const fdmc_payoff* payoff, double eps, int n_steps,
                const fdmc_qmc_config* cfg, int mode, const orc_bb* bb, const double* pca_M,
                fdmc_result* out);
int orc_run_qmc_cv(const fdmc_model* main_model, const fdmc_model* ctrl_model,
                   const fdmc_payoff* main_payoff, const fdmc_payoff* ctrl_payoff,
                   double E_ctrl, double beta, double eps, int n_steps,
                   const fdmc_qmc_config* cfg, int mode, const orc_bb* bb, fdmc_result* out);
int orc_run_mc_fixed(const fdmc_model* model, const fdmc_payoff* payoff, int n_steps,
                     int64_t n_paths, uint32_t seed, double* mean, double* var_of_mean);
int orc_run_mc(const fdmc_model* model, const fdmc_payoff* payoff, double eps, int n_steps,
               const fdmc_mc_config* cfg, fdmc_result* out);
/* kernel_is_gbm sums {Y, Y^2} over paths [0, n) of Philox stream (seed, tag);
 * zs is the per-step shift. */
int orc_is_sums(const fdmc_model* model, const fdmc_payoff* payoff, int n_steps, float zs, uint32_t seed,
                uint32_t tag, int64_t n, double sums[2]);
int orc_run_is(const fdmc_model* model, const fdmc_payoff* payoff, double z_star, double eps,
               const fdmc_mc_config* cfg, fdmc_result* out);
/* Level sums {dY, dY^2, Yf, Yf^2} of level l over paths [first, first+n). */
int orc_mlmc_level_sums(const fdmc_model* model, const fdmc_payoff* payoff, int M, int l,
                        uint32_t seed, int64_t first, int64_t n, double sums[4]);
int orc_run_mlmc(const fdmc_model* model, const fdmc_payoff* payoff, double eps,
                 const fdmc_mlmc_config* cfg, fdmc_result* out);
int orc_mlqmc_level_sums(const fdmc_model* model, const fdmc_payoff* payoff, int M, int l,
                         int R, uint32_t salt_base, int64_t offset, int64_t n, int mode,
                         const orc_bb* bb, const double* pca_M /* column-major, PCA mode */,
                         double* sums /* R x 4 */);
int orc_run_mlqmc(const fdmc_model* model, const fdmc_payoff* payoff, double eps,
                  const fdmc_mlmc_config* ml, const fdmc_qmc_config* qc, int mode,
                  const orc_bb* bbs, int n_bb, const double* const* pcas, int n_pca,
                  fdmc_result* out);
int orc_cv_pilot(const fdmc_model* main_model, const fdmc_model* ctrl_model,
                 const fdmc_payoff* main_payoff, const fdmc_payoff* ctrl_payoff,
                 double E_ctrl, int n_steps, int N_pilot, uint32_t seed,
                 fdmc_cv_pilot_result* out);
int orc_run_mc_cv(const fdmc_model* main_model, const fdmc_model* ctrl_model,
                  const fdmc_payoff* main_payoff, const fdmc_payoff* ctrl_payoff,
                  double E_ctrl, double beta, double eps, int n_steps,
                  const fdmc_mc_config* cfg, fdmc_result* out);

/* ---------------- contract book (fixed-level MLQMC with pathwise Greeks) ----------------
 * out[0] = P_f - P_c of one coupled path of n_fine = 2^l scaled increments dW
 * (coarse: n_fine/2 steps on summed pairs; level 0: no coarse leg), out[1..3] its
 * derivatives wrt S0, the volatility parameter and mu = r (forward mode). */
void orc_book_path(const fdmc_model* m, const fdmc_book_contract* k, const float* dW, int n_fine,
                   float out[4]);
/* Per-replica sums (R x 4) of level l of contract c over points [offset, offset+n). */
int orc_book_level_sums(const fdmc_model* m, const fdmc_book_contract* k, int c, int l, int R,
                        uint32_t salt_base, int64_t offset, int64_t n, double* sums);
/* The whole allocation loop of one contract (fdmc_book_run's per-contract result). */
int orc_book_contract_run(const fdmc_model* m, const fdmc_book_contract* k, int c, double eps,
                          const fdmc_book_config* cfg, fdmc_book_result* out);

const char* orc_last_error(void);
int orc_num_threads(void);

#ifdef __cplusplus
}
#endif
#endif
