/*
 * oracle_drivers.c -- estimator loops of the reference (MC pilot + N rule, QMC
 * doubling, Giles MLMC and MLQMC allocation, control-variate drivers) driven by
 * the engine's noise sources.  TEST INFRASTRUCTURE (see oracle.h).
 *
 * Noise sources (identical definitions in the HIP library):
 *  - QMC: replica salt -> Hong-Hickernell scrambled Sobol -> orc_normal_u32;
 *    dimension D > 20000 falls back to Philox normals (seed = salt, tag 0x51,
 *    path = point index), as the reference falls back to XORWOW (:1521, :1767).
 *  - MC / MLMC / CV: Philox streams keyed by (seed, tag, path):
 *    run_mc pilot (seed, 0), main (seed + 1, 0); MLMC level l (seed, 0x100 + l);
 *    cv pilot (seed, 0x200); run_mc_cv main (seed + 100, 0x201);
 *    run_is pilot (seed, 0x300), main (seed + 1, 0x300).
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#ifdef _OPENMP
#include <omp.h>
#endif

void orc_set_error(const char* msg);

#define D_MAX_SOBOL 20000
#define TAG_QMC_FALLBACK 0x51u
#define TAG_MLMC 0x100u
#define TAG_CV_PILOT 0x200u
#define TAG_CV_MAIN 0x201u
#define TAG_IS 0x300u

int orc_num_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

static double now_s(void) {
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return ts.tv_sec + 1e-9 * ts.tv_nsec;
}

static int noise_dim(const fdmc_model* m) {
    switch (m->kind) {
    case FDMC_MODEL_HESTON: return 2;
    case FDMC_MODEL_MULTI_DUPIRE:
    case FDMC_MODEL_MULTI_GBM: return m->n_assets;
    default: return 1;
    }
}
static int is_basket(const fdmc_model* m) {
    return m->kind == FDMC_MODEL_MULTI_DUPIRE || m->kind == FDMC_MODEL_MULTI_GBM;
}

/* dW_k <- L dW_k for every step (correlate_multi_dupire, methods_cuda.cu:1133-1148). */
static void correlate(const fdmc_model* m, int n_steps, float* dW) {
    int n = m->n_assets;
    if (m->uncorrelated || !m->L_assets) return;
    float tmp[1024];
    for (int k = 0; k < n_steps; k++) {
        for (int a = 0; a < n; a++) {
            double s = 0.0;
            for (int b = 0; b <= a; b++) s += m->L_assets[a * n + b] * (double)dW[k * n + b];
            tmp[a] = (float)s;
        }
        memcpy(dW + (size_t)k * n, tmp, sizeof(float) * (size_t)n);
    }
}

/* Per-replica scrambled direction tables for D dimensions. */
typedef struct { uint32_t* Lv; uint32_t* e; int D; } dir_table;
static void dir_table_make(uint32_t salt, int D, dir_table* t) {
    t->D = D;
    t->Lv = (uint32_t*)malloc(sizeof(uint32_t) * 32 * (size_t)D);
    t->e = (uint32_t*)malloc(sizeof(uint32_t) * (size_t)D);
    for (int d = 0; d < D; d++) orc_scrambled_dirs(salt, d, t->Lv + 32 * (size_t)d, t->e + d);
}
static void dir_table_free(dir_table* t) { free(t->Lv); free(t->e); }
static void sobol_normals_point(const dir_table* t, uint32_t idx, float* z) {
    uint32_t g = idx ^ (idx >> 1);
    for (int d = 0; d < t->D; d++) {
        const uint32_t* L = t->Lv + 32 * (size_t)d;
        uint32_t x = t->e[d];
        for (uint32_t gg = g; gg; gg &= gg - 1) x ^= L[__builtin_ctz(gg)];
        z[d] = (float)orc_normal_u32(x);
    }
}

/* Evaluate one path from its normals z (length D = n_steps * nf); returns the
 * number of outputs written to y (1, or 2 with a control leg). */
typedef struct {
    const fdmc_model* model;
    orc_params P, Pc;
    int has_ctrl, nf, n_steps, mode, basket;
    float sqrt_h;
    const orc_bb* bb;
    const double* pca_M;
} path_ctx;

static int path_ctx_init(path_ctx* c, const fdmc_model* model, const fdmc_payoff* payoff,
                         const fdmc_model* cm, const fdmc_payoff* cp, int n_steps, int mode,
                         const orc_bb* bb, const double* pca_M) {
    memset(c, 0, sizeof(*c));
    c->model = model;
    if (orc_make_params(model, payoff, n_steps, &c->P)) return -1;
    c->has_ctrl = (cm && cp);
    if (c->has_ctrl && orc_make_params(cm, cp, n_steps, &c->Pc)) return -1;
    c->nf = noise_dim(model);
    c->n_steps = n_steps;
    c->mode = mode;
    c->basket = is_basket(model);
    c->sqrt_h = sqrtf(c->P.h);
    c->bb = bb;
    c->pca_M = pca_M;
    if (mode == FDMC_NOISE_BB && (c->nf != 1 || !bb || bb->N != n_steps)) {
        orc_set_error("BB mode needs a one-factor model and bb->N == n_steps");
        return -1;
    }
    if (mode == FDMC_NOISE_PCA && !pca_M) { orc_set_error("PCA mode needs M"); return -1; }
    return 0;
}

static void transform(const path_ctx* c, const float* z, float* dW) {
    int D = c->n_steps * c->nf;
    if (c->mode == FDMC_NOISE_BB) {
        orc_bb_apply_f32(c->bb, z, dW);
    } else if (c->mode == FDMC_NOISE_PCA) {
        for (int i = 0; i < D; i++) {
            double s = 0.0;
            for (int j = 0; j < D; j++) s += c->pca_M[(size_t)j * D + i] * (double)z[j];
            dW[i] = (float)s;
        }
    } else {
        for (int i = 0; i < D; i++) dW[i] = z[i] * c->sqrt_h;
        if (c->basket) correlate(c->model, c->n_steps, dW);
    }
}

static void eval_path(const path_ctx* c, const float* dW, float* ym, float* yc) {
    if (c->basket) *ym = orc_path_basket(&c->P, dW, c->n_steps);
    else *ym = orc_path_single(&c->P, dW, c->n_steps);
    if (c->has_ctrl) *yc = orc_path_single(&c->Pc, dW, c->n_steps);
}

static void add_sums(double* s, int has_ctrl, float ym, float yc, float beta, float E) {
    double a = ym;
    s[0] += a; s[1] += a * a;
    if (has_ctrl) {
        float ycv = ym - beta * (yc - E);
        double v = ycv, g = yc;
        s[2] += v; s[3] += v * v; s[4] += g; s[5] += g * g; s[6] += a * g;
    }
}

int orc_qmc_accumulate(const fdmc_model* model, const fdmc_payoff* payoff,
                       const fdmc_model* ctrl_model, const fdmc_payoff* ctrl_payoff,
                       double E_ctrl, double beta, int n_steps, int R, uint32_t salt_base,
                       int64_t offset, int64_t n, int mode, const orc_bb* bb,
                       const double* pca_M, double* sums, int* n_acc) {
    path_ctx c;
    if (path_ctx_init(&c, model, payoff, ctrl_model, ctrl_payoff, n_steps, mode, bb, pca_M))
        return -1;
    int D = n_steps * c.nf, na = c.has_ctrl ? 7 : 2;
    if (offset < 0 || offset + n > 4294967296LL) { orc_set_error("offset range"); return -1; }
    *n_acc = na;
    float beta_f = (float)beta, E_f = (float)E_ctrl;
    for (int r = 0; r < R; r++) {
        uint32_t salt = salt_base + (uint32_t)r;
        int use_sobol = D <= D_MAX_SOBOL;
        dir_table t = {0};
        if (use_sobol) dir_table_make(salt, D, &t);
        double tot[7] = {0};
#pragma omp parallel
        {
            double loc[7] = {0};
            float* z = (float*)malloc(sizeof(float) * (size_t)D);
            float* dW = (float*)malloc(sizeof(float) * (size_t)D);
#pragma omp for schedule(static)
            for (int64_t p = 0; p < n; p++) {
                uint32_t idx = (uint32_t)(offset + p);
                if (use_sobol) sobol_normals_point(&t, idx, z);
                else orc_prng_normals(salt, TAG_QMC_FALLBACK, idx, D, z);
                transform(&c, z, dW);
                float ym = 0, yc = 0;
                eval_path(&c, dW, &ym, &yc);
                add_sums(loc, c.has_ctrl, ym, yc, beta_f, E_f);
            }
#pragma omp critical
            for (int i = 0; i < 7; i++) tot[i] += loc[i];
            free(z); free(dW);
        }
        for (int i = 0; i < na; i++) sums[r * na + i] = tot[i];
        if (use_sobol) dir_table_free(&t);
    }
    return 0;
}

static void mean_var(const double* v, int R, double* m, double* var) {
    double s = 0.0;
    for (int r = 0; r < R; r++) s += v[r];
    s /= R;
    double q = 0.0;
    for (int r = 0; r < R; r++) q += (v[r] - s) * (v[r] - s);
    *m = s;
    *var = (R > 1) ? q / (R - 1) : 0.0;
}

/* run_qmc_cuda / run_qmc_cv_cuda doubling loop (methods_cuda.cu:1624-1665,
 * 2203-2273); the first size is cfg->N0 instead of a free-memory bound. */
static int qmc_loop(const fdmc_model* model, const fdmc_payoff* payoff, const fdmc_model* cm,
                    const fdmc_payoff* cp, double E, double beta, double eps, int n_steps,
                    const fdmc_qmc_config* cfg, int mode, const orc_bb* bb, const double* M,
                    fdmc_result* out) {
    double t0 = now_s();
    int R = cfg->R, nf = noise_dim(model);
    int64_t N = cfg->N0 > 0 ? cfg->N0 : (1LL << 14), done = 0;
    int64_t cap = cfg->N_max > 0 ? cfg->N_max : 4294967296LL;
    int na = (cm && cp) ? 7 : 2, col = (cm && cp) ? 2 : 0;
    double* tot = (double*)calloc((size_t)R * na, sizeof(double));
    double* part = (double*)calloc((size_t)R * na, sizeof(double));
    double* means = (double*)calloc((size_t)R, sizeof(double));
    double vom = 1e30, gm = 0.0;
    int64_t total = 0;
    int doublings = 0, conv = 0, rc = 0;
    memset(out, 0, sizeof(*out));
    for (int d = 0; d < cfg->max_doublings; d++) {
        if (N > 4294967296LL) { orc_set_error("Sobol offset beyond 32 bits"); rc = -1; break; }
        int got = 0;
        if (orc_qmc_accumulate(model, payoff, cm, cp, E, beta, n_steps, R, cfg->salt_base, done,
                               N - done, mode, bb, M, part, &got)) { rc = -1; break; }
        for (int i = 0; i < R * na; i++) tot[i] += part[i];
        out->path_steps += (int64_t)R * (N - done) * n_steps;
        done = N;
        doublings++;
        for (int r = 0; r < R; r++) means[r] = tot[r * na + col] / (double)N;
        double v;
        mean_var(means, R, &gm, &v);
        vom = v / R;
        total = (int64_t)R * N;
        if (vom < eps * eps / 2.0) { conv = 1; break; }
        if (2 * N > cap) break;
        N *= 2;
    }
    (void)nf;
    out->price = gm;
    out->std_error = sqrt(vom);
    out->n_samples = total;
    out->levels = doublings;
    out->flags = (conv ? 0u : FDMC_FLAG_NOT_CONVERGED) |
                 (n_steps * noise_dim(model) > D_MAX_SOBOL ? FDMC_FLAG_PRNG_FALLBACK : 0u);
    out->time_s = now_s() - t0;
    free(tot); free(part); free(means);
    return rc;
}

int orc_run_qmc(const fdmc_model* model, const fdmc_payoff* payoff, double eps, int n_steps,
                const fdmc_qmc_config* cfg, int mode, const orc_bb* bb, const double* pca_M,
                fdmc_result* out) {
    return qmc_loop(model, payoff, NULL, NULL, 0.0, 0.0, eps, n_steps, cfg, mode, bb, pca_M, out);
}

int orc_run_qmc_cv(const fdmc_model* main_model, const fdmc_model* ctrl_model,
                   const fdmc_payoff* main_payoff, const fdmc_payoff* ctrl_payoff,
                   double E_ctrl, double beta, double eps, int n_steps,
                   const fdmc_qmc_config* cfg, int mode, const orc_bb* bb, fdmc_result* out) {
    return qmc_loop(main_model, main_payoff, ctrl_model, ctrl_payoff, E_ctrl, beta, eps, n_steps,
                    cfg, mode, bb, NULL, out);
}

/* Sums over PRNG paths [first, first+n) of stream (seed, tag): MC and CV. */
static int prng_sums(const fdmc_model* model, const fdmc_payoff* payoff, const fdmc_model* cm,
                     const fdmc_payoff* cp, double E, double beta, int n_steps, uint32_t seed,
                     uint32_t tag, int64_t first, int64_t n, double* sums) {
    path_ctx c;
    if (path_ctx_init(&c, model, payoff, cm, cp, n_steps, FDMC_NOISE_RAW, NULL, NULL)) return -1;
    int D = n_steps * c.nf;
    float beta_f = (float)beta, E_f = (float)E;
    double tot[7] = {0};
#pragma omp parallel
    {
        double loc[7] = {0};
        float* z = (float*)malloc(sizeof(float) * (size_t)D);
        float* dW = (float*)malloc(sizeof(float) * (size_t)D);
#pragma omp for schedule(static)
        for (int64_t p = 0; p < n; p++) {
            orc_prng_normals(seed, tag, (uint64_t)(first + p), D, z);
            transform(&c, z, dW);
            float ym = 0, yc = 0;
            eval_path(&c, dW, &ym, &yc);
            add_sums(loc, c.has_ctrl, ym, yc, beta_f, E_f);
        }
#pragma omp critical
        for (int i = 0; i < 7; i++) tot[i] += loc[i];
        free(z); free(dW);
    }
    memcpy(sums, tot, sizeof(tot));
    return 0;
}

/* run_mc_fixed_impl (methods_cuda.cu:1173-1265): mean and variance of the mean. */
int orc_run_mc_fixed(const fdmc_model* model, const fdmc_payoff* payoff, int n_steps,
                     int64_t n_paths, uint32_t seed, double* mean, double* var) {
    double s[7];
    if (n_paths <= 0) { orc_set_error("n_paths must be positive"); return -1; }
    if (prng_sums(model, payoff, NULL, NULL, 0, 0, n_steps, seed, 0u, 0, n_paths, s)) return -1;
    double m = s[0] / n_paths;
    double v = s[1] / n_paths - m * m;
    *mean = m;
    *var = (v > 0 ? v : 0.0) / n_paths;
    return 0;
}

/* kernel_is_gbm (methods_cuda.cu:685-707) over paths [0, n) of the Philox
 * stream (seed, tag): z ~ N(0,1), dW = (z + zs) sqrt(h), GBM Euler steps,
 * payoff max(S_T - K, 0) e^{-rT} times exp(-zs sum z - n_steps zs^2 / 2). */
int orc_is_sums(const fdmc_model* model, const fdmc_payoff* payoff, int n_steps, float zs, uint32_t seed,
                uint32_t tag, int64_t n, double sums[2]) {
    orc_params P;
    if (model->kind != FDMC_MODEL_GBM || payoff->kind != FDMC_PAYOFF_EUROPEAN) {
        orc_set_error("IS: GBM European only");
        return -1;
    }
    if (orc_make_params(model, payoff, n_steps, &P)) return -1;
    float sq = sqrtf(P.h);
    double tot[2] = {0, 0};
#pragma omp parallel
    {
        double loc[2] = {0, 0};
        float* z = (float*)malloc(sizeof(float) * (size_t)n_steps);
        float* dW = (float*)malloc(sizeof(float) * (size_t)n_steps);
#pragma omp for schedule(static)
        for (int64_t p = 0; p < n; p++) {
            orc_prng_normals(seed, tag, (uint64_t)p, n_steps, z);
            float zsum = 0.0f;
            for (int k = 0; k < n_steps; k++) {
                zsum += z[k];
                dW[k] = (z[k] + zs) * sq;
            }
            float lr = expf(-zs * zsum - 0.5f * zs * zs * (float)n_steps);
            float y = orc_path_single(&P, dW, n_steps) * lr;
            loc[0] += y;
            loc[1] += (double)y * y;
        }
#pragma omp critical
        { tot[0] += loc[0]; tot[1] += loc[1]; }
        free(z); free(dW);
    }
    sums[0] = tot[0];
    sums[1] = tot[1];
    return 0;
}

/* run_is_cuda (methods_cuda.cu:2317-2365). */
int orc_run_is(const fdmc_model* model, const fdmc_payoff* payoff, double z_star, double eps,
               const fdmc_mc_config* cfg, fdmc_result* out) {
    double t0 = now_s(), s[2];
    memset(out, 0, sizeof(*out));
    double nd = ceil(model->T / eps);
    int n_steps = nd < 4.0 ? 4 : (int)nd;
    float zs = (float)(z_star / sqrt((double)n_steps));
    if (orc_is_sums(model, payoff, n_steps, zs, cfg->seed, TAG_IS, cfg->pilot_n, s)) return -1;
    double pm = s[0] / cfg->pilot_n;
    double pv = s[1] / cfg->pilot_n - pm * pm;
    int64_t N = (int64_t)ceil(2.0 * (pv > 0 ? pv : 0.0) / (eps * eps));
    if (N < cfg->pilot_n) N = cfg->pilot_n;
    N = (N + 1) & ~(int64_t)1;
    if (orc_is_sums(model, payoff, n_steps, zs, cfg->seed + 1u, TAG_IS, N, s)) return -1;
    double m = s[0] / N;
    double v = s[1] / N - m * m;
    out->price = m;
    out->std_error = sqrt((v > 0 ? v : 0.0) / N);
    out->n_samples = N;
    out->path_steps = (int64_t)(N + cfg->pilot_n) * n_steps;
    out->time_s = now_s() - t0;
    return 0;
}

/* run_mc_cuda (methods_cuda.cu:1272-1298). */
int orc_run_mc(const fdmc_model* model, const fdmc_payoff* payoff, double eps, int n_steps,
               const fdmc_mc_config* cfg, fdmc_result* out) {
    double t0 = now_s(), pm, pv;
    memset(out, 0, sizeof(*out));
    if (orc_run_mc_fixed(model, payoff, n_steps, cfg->pilot_n, cfg->seed, &pm, &pv)) return -1;
    double sample_var = pv * cfg->pilot_n;
    int64_t N = (int64_t)ceil(2.0 * sample_var / (eps * eps));
    if (N < cfg->pilot_n) N = cfg->pilot_n;
    double s[7];
    if (prng_sums(model, payoff, NULL, NULL, 0, 0, n_steps, cfg->seed + 1u, 0u, 0, N, s)) return -1;
    double m = s[0] / N;
    double v = s[1] / N - m * m;
    out->price = m;
    out->std_error = sqrt((v > 0 ? v : 0.0) / N);
    out->n_samples = N;
    out->path_steps = (int64_t)(N + cfg->pilot_n) * n_steps;
    out->time_s = now_s() - t0;
    return 0;
}

/* ---------------- multilevel ---------------- */
static int level_steps(int M, int l) {
    int n = 1;
    for (int i = 0; i < l; i++) n *= M;
    return n;
}

static void coupled_sums_from_noise(const orc_params* P, int nf, int nfine, int M, int mode,
                                    const orc_bb* bb, const double* pca_M, float* z, float* dW,
                                    double* s) {
    int D = nfine * nf;
    const float* in = z;
    if (P->model_kind != FDMC_MODEL_HESTON) {
        if (mode == FDMC_NOISE_BB) {
            orc_bb_apply_f32(bb, z, dW);
        } else if (mode == FDMC_NOISE_PCA) {
            /* phase1_pca's dW = M z in FP64 (methods_cuda.cu:1907-1913), rounded to FP32 */
            for (int i = 0; i < D; i++) {
                double a = 0.0;
                for (int j = 0; j < D; j++) a += pca_M[(size_t)j * D + i] * (double)z[j];
                dW[i] = (float)a;
            }
        } else {
            float sq = sqrtf((float)((double)P->T / nfine));
            for (int i = 0; i < D; i++) dW[i] = z[i] * sq;
        }
        in = dW;
    }
    float pf, pc;
    orc_path_coupled(P, in, nfine, M, &pf, &pc);
    double dy = (double)(pf - pc), yf = pf;
    s[0] += dy; s[1] += dy * dy; s[2] += yf; s[3] += yf * yf;
}

int orc_mlmc_level_sums(const fdmc_model* model, const fdmc_payoff* payoff, int M, int l,
                        uint32_t seed, int64_t first, int64_t n, double sums[4]) {
    orc_params P;
    int nfine = level_steps(M, l), nf = noise_dim(model);
    if (orc_make_params(model, payoff, 1, &P)) return -1;
    int D = nfine * nf;
    double tot[4] = {0};
#pragma omp parallel
    {
        double loc[4] = {0};
        float* z = (float*)malloc(sizeof(float) * (size_t)D);
        float* dW = (float*)malloc(sizeof(float) * (size_t)D);
#pragma omp for schedule(static)
        for (int64_t p = 0; p < n; p++) {
            orc_prng_normals(seed, TAG_MLMC + (uint32_t)l, (uint64_t)(first + p), D, z);
            coupled_sums_from_noise(&P, nf, nfine, M, FDMC_NOISE_RAW, NULL, NULL, z, dW, loc);
        }
#pragma omp critical
        for (int i = 0; i < 4; i++) tot[i] += loc[i];
        free(z); free(dW);
    }
    memcpy(sums, tot, sizeof(tot));
    return 0;
}

/* run_mlmc_cuda (methods_cuda.cu:1305-1487). */
int orc_run_mlmc(const fdmc_model* model, const fdmc_payoff* payoff, double eps,
                 const fdmc_mlmc_config* cfg, fdmc_result* out) {
    double t0 = now_s();
    int M = cfg->M, maxL = cfg->max_L, L = 2, nf = noise_dim(model);
    if (is_basket(model)) { orc_set_error("MLMC does not support baskets"); return -1; }
    double E[32] = {0}, V[32] = {0};
    int64_t N[32] = {0};
    double s[4];
    memset(out, 0, sizeof(*out));
    for (int l = 0; l <= L; l++) {
        if (orc_mlmc_level_sums(model, payoff, M, l, cfg->seed, 0, cfg->pilot_n, s)) return -1;
        double em = s[0] / cfg->pilot_n;
        double v = s[1] / cfg->pilot_n - em * em;
        E[l] = em; V[l] = v > 0 ? v : 0.0; N[l] = cfg->pilot_n;
        out->path_steps += (int64_t)cfg->pilot_n * level_steps(M, l);
    }
    int conv = 0, iter = 0;
    while (!conv && L <= maxL && iter++ < 50) {
        double st = 0.0;
        for (int l = 0; l <= L; l++) st += sqrt(V[l] * pow((double)M, l));
        for (int l = 0; l <= L; l++) {
            double Cl = pow((double)M, l);
            int64_t Nopt = (int64_t)ceil(2.0 / (eps * eps) * sqrt(V[l] / Cl) * st);
            if (Nopt < 100) Nopt = 100;
            if (Nopt > N[l]) {
                int64_t extra = Nopt - N[l];
                if (orc_mlmc_level_sums(model, payoff, M, l, cfg->seed, N[l], extra, s)) return -1;
                int64_t Nn = N[l] + extra;
                double ts = E[l] * N[l] + s[0];
                double t2 = (V[l] + E[l] * E[l]) * N[l] + s[1];
                E[l] = ts / Nn;
                double v = t2 / Nn - E[l] * E[l];
                V[l] = v > 0 ? v : 0.0;
                N[l] = Nn;
                out->path_steps += extra * level_steps(M, l);
            }
        }
        double bias = fabs(E[L]) / (M - 1);
        if (L >= 1) {
            double b2 = fabs(E[L - 1]) * M / (double)(M * M - M);
            if (b2 > bias) bias = b2;
        }
        conv = bias < eps / sqrt(2.0);
        if (!conv && L < maxL) {
            L++;
            if (orc_mlmc_level_sums(model, payoff, M, L, cfg->seed, 0, cfg->pilot_n, s)) return -1;
            double em = s[0] / cfg->pilot_n;
            double v = s[1] / cfg->pilot_n - em * em;
            E[L] = em; V[L] = v > 0 ? v : 0.0; N[L] = cfg->pilot_n;
            out->path_steps += (int64_t)cfg->pilot_n * level_steps(M, L);
        }
    }
    (void)nf;
    double price = 0.0, vs = 0.0;
    int64_t tot = 0;
    for (int l = 0; l <= L; l++) {
        price += E[l];
        vs += N[l] > 0 ? V[l] / N[l] : 0.0;
        tot += N[l];
    }
    out->price = price;
    out->std_error = sqrt(vs);
    out->n_samples = tot;
    out->levels = L + 1;
    out->flags = conv ? 0u : FDMC_FLAG_NOT_CONVERGED;
    out->time_s = now_s() - t0;
    return 0;
}

int orc_mlqmc_level_sums(const fdmc_model* model, const fdmc_payoff* payoff, int M, int l,
                         int R, uint32_t salt_base, int64_t offset, int64_t n, int mode,
                         const orc_bb* bb, const double* pca_M, double* sums) {
    orc_params P;
    int nfine = level_steps(M, l), nf = noise_dim(model);
    if (orc_make_params(model, payoff, 1, &P)) return -1;
    int D = nfine * nf;
    int use_sobol = D <= D_MAX_SOBOL;
    if (mode == FDMC_NOISE_BB && (nf != 1 || !bb || bb->N != nfine)) {
        orc_set_error("MLQMC BB needs a one-factor model and bb->N == M^l");
        return -1;
    }
    if (mode == FDMC_NOISE_PCA && (nf != 1 || !pca_M)) {
        orc_set_error("MLQMC PCA needs a one-factor model and an M^l x M^l matrix");
        return -1;
    }
    for (int r = 0; r < R; r++) {
        uint32_t salt = salt_base + (uint32_t)(l * R + r);
        dir_table t = {0};
        if (use_sobol) dir_table_make(salt, D, &t);
        double tot[4] = {0};
#pragma omp parallel
        {
            double loc[4] = {0};
            float* z = (float*)malloc(sizeof(float) * (size_t)D);
            float* dW = (float*)malloc(sizeof(float) * (size_t)D);
#pragma omp for schedule(static)
            for (int64_t p = 0; p < n; p++) {
                uint32_t idx = (uint32_t)(offset + p);
                if (use_sobol) sobol_normals_point(&t, idx, z);
                else orc_prng_normals(salt, TAG_QMC_FALLBACK, idx, D, z);
                coupled_sums_from_noise(&P, nf, nfine, M, mode, bb, pca_M, z, dW, loc);
            }
#pragma omp critical
            for (int i = 0; i < 4; i++) tot[i] += loc[i];
            free(z); free(dW);
        }
        memcpy(sums + 4 * r, tot, sizeof(tot));
        if (use_sobol) dir_table_free(&t);
    }
    return 0;
}

/* run_mlqmc_cuda (methods_cuda.cu:1680-2019). */
int orc_run_mlqmc(const fdmc_model* model, const fdmc_payoff* payoff, double eps,
                  const fdmc_mlmc_config* ml, const fdmc_qmc_config* qc, int mode,
                  const orc_bb* bbs, int n_bb, const double* const* pcas, int n_pca,
                  fdmc_result* out) {
    double t0 = now_s();
    int M = ml->M, maxL = ml->max_L, L = 2, R = qc->R;
    if (is_basket(model)) { orc_set_error("MLQMC does not support baskets"); return -1; }
    if (mode != FDMC_NOISE_RAW && noise_dim(model) != 1) {
        orc_set_error("MLQMC BB/PCA need a one-factor model"); return -1;
    }
    double E[32] = {0}, S2[32] = {0};
    int64_t N[32] = {0}, off[32] = {0};
    double* s = (double*)malloc(sizeof(double) * 4 * (size_t)R);
    double* rm = (double*)malloc(sizeof(double) * (size_t)R);
    memset(out, 0, sizeof(*out));
    int64_t Np = ml->pilot_n;
#define LEVEL_RUN(l, n, em, vv)                                                                  \
    do {                                                                                         \
        const orc_bb* bl = (mode == FDMC_NOISE_BB && (l) < n_bb) ? &bbs[l] : NULL;              \
        const double* pl = (mode == FDMC_NOISE_PCA && (l) < n_pca) ? pcas[l] : NULL;            \
        if (mode == FDMC_NOISE_BB && !bl) { orc_set_error("missing bb_list entry"); return -1; } \
        if (mode == FDMC_NOISE_PCA && !pl) { orc_set_error("missing pca_list entry"); return -1; } \
        if (orc_mlqmc_level_sums(model, payoff, M, (l), R, qc->salt_base, off[l], (n), mode, bl, \
                                 pl, s)) return -1;                                              \
        off[l] += (n);                                                                           \
        for (int r = 0; r < R; r++) rm[r] = s[4 * r] / (double)(n);                              \
        mean_var(rm, R, &(em), &(vv));                                                           \
        out->path_steps += (int64_t)R * (n) * level_steps(M, (l));                               \
    } while (0)
    for (int l = 0; l <= L; l++) {
        double em, vv;
        LEVEL_RUN(l, Np, em, vv);
        E[l] = em; S2[l] = vv * (double)Np; N[l] = Np;
    }
    int conv = 0, iter = 0;
    while (!conv && L <= maxL && iter++ < 50) {
        double st = 0.0;
        for (int l = 0; l <= L; l++) st += sqrt(S2[l] * pow((double)M, l));
        for (int l = 0; l <= L; l++) {
            double Cl = pow((double)M, l);
            int64_t Nopt = (int64_t)ceil(2.0 / ((double)R * eps * eps) * sqrt(S2[l] / Cl) * st);
            if (Nopt < 100) Nopt = 100;
            if (Nopt > N[l]) {
                int64_t extra = Nopt - N[l];
                double em, vv;
                LEVEL_RUN(l, extra, em, vv);
                double wo = (double)N[l], wn = (double)extra;
                E[l] = (E[l] * wo + em * wn) / (wo + wn);
                S2[l] = (S2[l] * wo + vv * wn * wn) / (wo + wn);
                N[l] += extra;
            }
        }
        double bias = fabs(E[L]) / (M - 1 > 1 ? M - 1 : 1);
        if (L >= 1) {
            double b2 = fabs(E[L - 1]) * M / (double)(M * (M - 1) > 1 ? M * (M - 1) : 1);
            if (b2 > bias) bias = b2;
        }
        conv = bias < eps / sqrt(2.0);
        if (!conv && L < maxL) {
            L++;
            double em, vv;
            LEVEL_RUN(L, Np, em, vv);
            E[L] = em; S2[L] = vv * (double)Np; N[L] = Np;
        }
    }
#undef LEVEL_RUN
    double price = 0.0, vs = 0.0;
    int64_t tot = 0;
    for (int l = 0; l <= L; l++) {
        price += E[l];
        vs += S2[l] / ((double)N[l] * (double)R);
        tot += N[l] * R;
    }
    out->price = price;
    out->std_error = sqrt(vs);
    out->n_samples = tot;
    out->levels = L + 1;
    out->flags = conv ? 0u : FDMC_FLAG_NOT_CONVERGED;
    out->time_s = now_s() - t0;
    free(s); free(rm);
    return 0;
}

/* cv_pilot (methods_cuda.cu:2026-2076). */
int orc_cv_pilot(const fdmc_model* main_model, const fdmc_model* ctrl_model,
                 const fdmc_payoff* main_payoff, const fdmc_payoff* ctrl_payoff,
                 double E_ctrl, int n_steps, int N_pilot, uint32_t seed,
                 fdmc_cv_pilot_result* out) {
    double hs[7];
    if (prng_sums(main_model, main_payoff, ctrl_model, ctrl_payoff, E_ctrl, 0.0, n_steps, seed,
                  TAG_CV_PILOT, 0, N_pilot, hs)) return -1;
    double N = (double)N_pilot;
    double Ea = hs[0] / N, Ea2 = hs[1] / N, Eg = hs[4] / N, Eg2 = hs[5] / N, Eag = hs[6] / N;
    double vm = Ea2 - Ea * Ea; if (vm < 0) vm = 0;
    double vc = Eg2 - Eg * Eg; if (vc < 1e-30) vc = 1e-30;
    double cov = Eag - Ea * Eg;
    double b = cov / vc;
    if (b < 0) b = 0;
    if (b > 5) b = 5;
    double vcv = vm - cov * cov / vc;
    out->beta = b;
    out->var_plain = vm;
    out->var_cv = vcv > 0 ? vcv : 0.0;
    return 0;
}

/* run_mc_cv_cuda (methods_cuda.cu:2078-2163). */
int orc_run_mc_cv(const fdmc_model* main_model, const fdmc_model* ctrl_model,
                  const fdmc_payoff* main_payoff, const fdmc_payoff* ctrl_payoff,
                  double E_ctrl, double beta, double eps, int n_steps,
                  const fdmc_mc_config* cfg, fdmc_result* out) {
    double t0 = now_s(), hs[7];
    memset(out, 0, sizeof(*out));
    if (prng_sums(main_model, main_payoff, ctrl_model, ctrl_payoff, E_ctrl, beta, n_steps,
                  cfg->seed, TAG_CV_PILOT, 0, cfg->pilot_n, hs)) return -1;
    double mu = hs[2] / cfg->pilot_n;
    double v = hs[3] / cfg->pilot_n - mu * mu;
    int64_t N = (int64_t)ceil(2.0 * (v > 0 ? v : 0.0) / (eps * eps));
    if (N < cfg->pilot_n) N = cfg->pilot_n;
    if (prng_sums(main_model, main_payoff, ctrl_model, ctrl_payoff, E_ctrl, beta, n_steps,
                  cfg->seed + 100u, TAG_CV_MAIN, 0, N, hs)) return -1;
    double m = hs[2] / N;
    double vf = hs[3] / N - m * m;
    out->price = m;
    out->std_error = sqrt((vf > 0 ? vf : 0.0) / N);
    out->n_samples = N;
    out->path_steps = (int64_t)(N + cfg->pilot_n) * n_steps;
    out->time_s = now_s() - t0;
    return 0;
}

/* ======================================================================
 * Contract book: every contract priced by multilevel QMC with M = 2 and its
 * finest level fixed at its own date grid (levels 0 .. log2(n_steps)), with
 * pathwise Greeks of the level differences.  Allocation and merges are
 * run_mlqmc_cuda's (methods_cuda.cu:1680-2019) applied level by level; the
 * salt of (contract c, level l, replica r) is salt_base + (c * 10 + l) R + r,
 * so contract 0 has the reference's 42 + l R + r (methods_cuda.cu:1845).
 *
 * The Greeks are computed here in FORWARD mode (tangents dS/dS0, dS/dvol,
 * dS/dmu carried along the path), an independent restatement of the engine's
 * reverse-mode adjoint: both are the pathwise derivative of the same FP32
 * Euler recursion, so they agree to rounding.
 * ====================================================================== */

/* One Euler step of a one-factor model and its partial derivatives
 * (methods_cuda.cu:271-291 for the step; the partials are its derivatives). */
static float book_step(const orc_params* P, float S, float dw, float t, float h, float* fS,
                       float* fvol, float* fmu, float* fS0) {
    *fmu = S * h;
    *fS0 = 0.0f;
    switch (P->model_kind) {
    case FDMC_MODEL_DUPIRE: {
        float e = expf(-P->alpha * t);
        float pw = powf(S / P->S0, P->beta_d - 1.0f);
        float sl = P->sigma0 * e * pw;
        *fS = 1.0f + P->mu * h + dw * sl * P->beta_d;
        *fvol = e * pw * S * dw;
        *fS0 = -sl * (P->beta_d - 1.0f) / P->S0 * S * dw;
        return S + P->mu * S * h + sl * S * dw;
    }
    case FDMC_MODEL_LOCALVOL_QUAD: {
        float l = logf(S / P->lv_sref);
        float be = P->lv_b * expf(-P->lv_c * t);
        float sl = P->lv_a + be * l * l;
        *fS = 1.0f + P->mu * h + dw * (sl + 2.0f * be * l);
        *fvol = S * dw;
        return S + P->mu * S * h + sl * S * dw;
    }
    default:   /* GBM */
        *fS = 1.0f + P->mu * h + P->sigma * dw;
        *fvol = S * dw;
        return S + P->mu * S * h + P->sigma * S * dw;
    }
}

/* One leg (n steps of size h; pair = 1 sums consecutive increments for the
 * coarse path) -> payoff y and its derivatives g[0..2] (S0, vol, mu = r). */
static void book_leg(const orc_params* P, int kind, const float* dW, int n, int pair, float h,
                     float bgk, float T, float* y, float g[3]) {
    float S = P->S0, d[3] = {1.0f, 0.0f, 0.0f};
    float sum = 0.0f, ds[3] = {0.0f, 0.0f, 0.0f};
    float mn = P->S0, dmn[3] = {1.0f, 0.0f, 0.0f}, mx = 0.0f;
    for (int j = 0; j < n; j++) {
        float dw = pair ? dW[2 * j] + dW[2 * j + 1] : dW[j];
        float fS, fv, fm, f0;
        float Sn = book_step(P, S, dw, (float)j * h, h, &fS, &fv, &fm, &f0);
        d[0] = fS * d[0] + f0;
        d[1] = fS * d[1] + fv;
        d[2] = fS * d[2] + fm;
        S = Sn;
        if (kind == FDMC_PAYOFF_ASIAN) {
            sum += S;
            for (int k = 0; k < 3; k++) ds[k] += d[k];
        } else if (kind == FDMC_PAYOFF_LOOKBACK) {
            if (S < mn) { mn = S; for (int k = 0; k < 3; k++) dmn[k] = d[k]; }
        } else if (kind == FDMC_PAYOFF_BARRIER) {
            mx = fmaxf(mx, S);
        }
    }
    g[0] = g[1] = g[2] = 0.0f;
    switch (kind) {
    case FDMC_PAYOFF_EUROPEAN:
        *y = fmaxf(S - P->K, 0.0f) * P->discount;
        if (S > P->K) for (int k = 0; k < 3; k++) g[k] = P->discount * d[k];
        g[2] -= T * *y;
        break;
    case FDMC_PAYOFF_ASIAN: {
        float A = sum / (float)n;
        *y = fmaxf(A - P->K, 0.0f);
        if (A > P->K) for (int k = 0; k < 3; k++) g[k] = ds[k] / (float)n;
        break;
    }
    case FDMC_PAYOFF_LOOKBACK:     /* 0 for a path that left the domain, as in payoff_eval */
        *y = (S > 0.0f) ? S - mn * bgk : 0.0f;
        if (S > 0.0f) for (int k = 0; k < 3; k++) g[k] = d[k] - bgk * dmn[k];
        break;
    default:  /* barrier: not Lipschitz, no pathwise Greeks */
        *y = (mx / bgk >= P->B) ? 0.0f : fmaxf(S - P->K, 0.0f) * P->discount;
        break;
    }
}

static float book_bgk(const fdmc_book_contract* k, float h) {
    if (k->kind != FDMC_PAYOFF_LOOKBACK && k->kind != FDMC_PAYOFF_BARRIER) return 1.0f;
    return expf(-0.5826f * (float)k->sigma_bgk * sqrtf(h));   /* payoffs.hpp:8 */
}

static int book_params(const fdmc_model* m, const fdmc_book_contract* k, orc_params* P) {
    fdmc_payoff p;
    memset(&p, 0, sizeof p);
    p.kind = k->kind == FDMC_PAYOFF_LOOKBACK ? FDMC_PAYOFF_LOOKBACK : FDMC_PAYOFF_EUROPEAN;
    p.K = k->K; p.r = k->r; p.T = k->T; p.sigma = k->sigma_bgk;
    if (orc_make_params(m, &p, 1, P)) return -1;
    P->K = (float)k->K;
    P->B = (float)k->B;
    P->discount = (k->kind == FDMC_PAYOFF_EUROPEAN || k->kind == FDMC_PAYOFF_BARRIER)
                      ? (float)exp(-k->r * k->T) : 1.0f;
    P->T = (float)k->T;
    return 0;
}

void orc_book_path(const fdmc_model* m, const fdmc_book_contract* k, const float* dW, int nf,
                   float out[4]) {
    orc_params P;
    book_params(m, k, &P);
    float hf = (float)(k->T / nf), yf, gf[3], yc = 0.0f, gc[3] = {0.0f, 0.0f, 0.0f};
    book_leg(&P, k->kind, dW, nf, 0, hf, book_bgk(k, hf), (float)k->T, &yf, gf);
    if (nf >= 2) {
        float hc = (float)(k->T / (nf / 2));
        book_leg(&P, k->kind, dW, nf / 2, 1, hc, book_bgk(k, hc), (float)k->T, &yc, gc);
    }
    out[0] = yf - yc;
    /* a leg that left the Euler scheme's domain (its states NaN after S <= 0)
       contributes 0 to the Greeks; the other leg counts as usual */
    float mf = 0.0f, mc = 0.0f;
    for (int i = 0; i < 3; i++) { mf += fabsf(gf[i]); mc += fabsf(gc[i]); }
    for (int i = 0; i < 3; i++) out[1 + i] = ((mf < INFINITY) ? gf[i] : 0.0f) - ((mc < INFINITY) ? gc[i] : 0.0f);
}

static int book_levels(const fdmc_book_contract* k) {
    int L = 0;
    while ((1 << L) < k->n_steps) L++;
    return L;
}

int orc_book_level_sums(const fdmc_model* m, const fdmc_book_contract* k, int c, int l, int R,
                        uint32_t salt_base, int64_t offset, int64_t n, double* sums) {
    int nf = 1 << l;
    if (l < 0 || l >= FDMC_BOOK_LEVELS || (1 << l) > k->n_steps) {
        orc_set_error("book level out of range"); return -1;
    }
    orc_bb bb;
    if (orc_bb_precompute(nf, k->T, &bb)) { orc_set_error("bb_precompute failed"); return -1; }
    for (int r = 0; r < R; r++) {
        uint32_t salt = salt_base + (uint32_t)((c * FDMC_BOOK_LEVELS + l) * R + r);
        dir_table t;
        dir_table_make(salt, nf, &t);
        double tot[4] = {0, 0, 0, 0};
#pragma omp parallel
        {
            double loc[4] = {0, 0, 0, 0};
            float* z = (float*)malloc(sizeof(float) * (size_t)nf);
            float* dW = (float*)malloc(sizeof(float) * (size_t)nf);
#pragma omp for schedule(static)
            for (int64_t p = 0; p < n; p++) {
                float o[4];
                sobol_normals_point(&t, (uint32_t)(offset + p), z);
                orc_bb_apply_f32(&bb, z, dW);
                orc_book_path(m, k, dW, nf, o);
                for (int i = 0; i < 4; i++) loc[i] += (double)o[i];
            }
#pragma omp critical
            for (int i = 0; i < 4; i++) tot[i] += loc[i];
            free(z); free(dW);
        }
        memcpy(sums + 4 * r, tot, sizeof tot);
        dir_table_free(&t);
    }
    orc_bb_release(&bb);
    return 0;
}

int orc_book_contract_run(const fdmc_model* m, const fdmc_book_contract* k, int c, double eps,
                          const fdmc_book_config* cfg, fdmc_book_result* out) {
    const int R = cfg->R, L = book_levels(k);
    double E[FDMC_BOOK_LEVELS][4], S2[FDMC_BOOK_LEVELS][4];
    int64_t N[FDMC_BOOK_LEVELS] = {0}, extra[FDMC_BOOK_LEVELS];
    double* s = (double*)malloc(sizeof(double) * 4 * (size_t)R);
    double* rm = (double*)malloc(sizeof(double) * (size_t)R);
    memset(out, 0, sizeof(*out));
    memset(E, 0, sizeof E);
    memset(S2, 0, sizeof S2);
    int conv = 0, iter = 0;
    while (iter < cfg->max_iter) {
        int any = 0;
        if (iter == 0) {
            for (int l = 0; l <= L; l++) extra[l] = cfg->pilot_n;
            any = 1;
        } else {
            /* stop once the price variance sum_l S2_l / (N_l R) is below eps^2/2
             * (with every N_l at its Giles target it is, so this never waits
             * longer than "no level needs more points") */
            double var = 0.0, st = 0.0;
            for (int l = 0; l <= L; l++) var += S2[l][0] / ((double)N[l] * R);
            if (var <= 0.5 * eps * eps) { conv = 1; break; }
            for (int l = 0; l <= L; l++) st += sqrt(S2[l][0] * (double)(1 << l));
            for (int l = 0; l <= L; l++) {
                int64_t Nopt = (int64_t)ceil(2.0 / ((double)R * eps * eps) *
                                             sqrt(S2[l][0] / (double)(1 << l)) * st);
                if (Nopt < 100) Nopt = 100;
                extra[l] = Nopt > N[l] ? Nopt - N[l] : 0;
                /* a top-up adds at least a tenth of the level's points (fdmc_book_run) */
                if (extra[l] > 0 && extra[l] < (N[l] + 9) / 10) extra[l] = (N[l] + 9) / 10;
                if (N[l] + extra[l] > 4294967296LL) extra[l] = 4294967296LL - N[l];
                if (extra[l] > 0) any = 1;
            }
        }
        if (!any) { conv = 1; break; }
        for (int l = 0; l <= L; l++) {
            if (extra[l] <= 0) continue;
            if (orc_book_level_sums(m, k, c, l, R, cfg->salt_base, N[l], extra[l], s)) {
                free(s); free(rm); return -1;
            }
            double wo = (double)N[l], wn = (double)extra[l];
            for (int q = 0; q < 4; q++) {
                double em, vv;
                for (int r = 0; r < R; r++) rm[r] = s[4 * r + q] / wn;
                mean_var(rm, R, &em, &vv);
                if (N[l] == 0) {
                    E[l][q] = em;
                    S2[l][q] = vv * wn;
                } else {
                    E[l][q] = (E[l][q] * wo + em * wn) / (wo + wn);
                    S2[l][q] = (S2[l][q] * wo + vv * wn * wn) / (wo + wn);
                }
            }
            N[l] += extra[l];
            out->path_steps += extra[l] * R * (int64_t)(1 << l);
        }
        iter++;
    }
    double v[4] = {0, 0, 0, 0}, val[4] = {0, 0, 0, 0};
    for (int l = 0; l <= L; l++) {
        for (int q = 0; q < 4; q++) {
            val[q] += E[l][q];
            v[q] += S2[l][q] / ((double)N[l] * R);
        }
        out->n_samples += N[l] * R;
    }
    out->price = val[0];
    out->std_error = sqrt(v[0]);
    int lip = k->kind != FDMC_PAYOFF_BARRIER && cfg->greeks;
    out->delta = lip ? val[1] : NAN;
    out->vega = lip ? val[2] : NAN;
    out->rho = lip ? val[3] : NAN;
    out->delta_se = lip ? sqrt(v[1]) : NAN;
    out->vega_se = lip ? sqrt(v[2]) : NAN;
    out->rho_se = lip ? sqrt(v[3]) : NAN;
    out->levels = L + 1;
    out->iterations = iter;
    /* an estimate that is not finite never counts as converged */
    for (int q = 0; q < 4; q++)
        if (!(fabs(val[q]) < INFINITY && v[q] < INFINITY)) conv = 0;
    out->flags = conv ? 0u : FDMC_FLAG_NOT_CONVERGED;
    free(s); free(rm);
    return 0;
}
