/*
 * oracle_paths.c -- FP32 path arithmetic of the reference kernels, one path at a
 * time.  TEST INFRASTRUCTURE (see oracle.h).
 */
#include "oracle.h"

#include <math.h>
#include <stdio.h>
#include <string.h>

#define BGK_BETA_F 0.5826f   /* payoffs.hpp:8 */

static __thread char g_err[512];
const char* orc_last_error(void) { return g_err; }
void orc_set_error(const char* msg) {
    snprintf(g_err, sizeof g_err, "%s", msg);
}

/* make_params (methods_cuda.cu:986-1051) plus the additive model kinds. */
int orc_make_params(const fdmc_model* m, const fdmc_payoff* p, int n_steps, orc_params* P) {
    memset(P, 0, sizeof(*P));
    P->model_kind = m->kind;
    P->payoff_kind = p->kind;
    switch (p->kind) {
    case FDMC_PAYOFF_EUROPEAN:
        P->K = (float)p->K; P->discount = (float)exp(-p->r * p->T); break;
    case FDMC_PAYOFF_ASIAN:
    case FDMC_PAYOFF_GEOM_ASIAN:
        P->K = (float)p->K; P->discount = 1.0f; break;
    case FDMC_PAYOFF_LOOKBACK:
        P->discount = 1.0f; P->bgk_on = 1; P->sigma_bgk = (float)p->sigma; break;
    case FDMC_PAYOFF_BARRIER:
        P->K = (float)p->K; P->B = (float)p->B; P->discount = (float)exp(-p->r * p->T);
        P->bgk_on = 1; P->sigma_bgk = (float)p->sigma; break;
    case FDMC_PAYOFF_BASKET:
        P->K = (float)p->K; P->discount = (float)exp(-p->r * p->T);
        P->n_assets = p->n_assets; P->weights = p->weights; break;
    default:
        orc_set_error("unknown payoff kind"); return -1;
    }
    P->T = (float)m->T;
    P->h = P->T / (float)n_steps;
    P->mu = (float)m->mu;
    switch (m->kind) {
    case FDMC_MODEL_GBM:
        P->S0 = (float)m->S0; P->sigma = (float)m->sigma; break;
    case FDMC_MODEL_HESTON:
        P->S0 = (float)m->S0; P->kappa = (float)m->kappa; P->theta = (float)m->theta;
        P->xi = (float)m->xi; P->rho = (float)m->rho; P->v0 = (float)m->v0;
        P->L10 = (float)m->L[2]; P->L11 = (float)m->L[3];
        P->mil_q = m->heston_scheme == FDMC_HESTON_MILSTEIN ? 0.25f * P->xi * P->xi : 0.0f; break;
    case FDMC_MODEL_DUPIRE:
        P->S0 = (float)m->S0; P->sigma0 = (float)m->sigma0; P->alpha = (float)m->alpha;
        P->beta_d = (float)m->beta_d; P->sigma = (float)m->sigma0; break;
    case FDMC_MODEL_LOCALVOL_QUAD:
        P->S0 = (float)m->S0; P->lv_a = (float)m->lv_a; P->lv_b = (float)m->lv_b;
        P->lv_c = (float)m->lv_c; P->lv_sref = (float)m->lv_sref; break;
    case FDMC_MODEL_MULTI_DUPIRE:
        P->S0 = (float)(m->S0_assets ? m->S0_assets[0] : 100.0);
        P->sigma0 = (float)m->sigma0; P->alpha = (float)m->alpha; P->beta_d = (float)m->beta_d;
        P->n_assets = m->n_assets; P->S0_assets = m->S0_assets; break;
    case FDMC_MODEL_MULTI_GBM:
        P->S0 = (float)(m->S0_assets ? m->S0_assets[0] : 100.0);
        P->n_assets = m->n_assets; P->S0_assets = m->S0_assets;
        P->sigma_assets = m->sigma_assets; break;
    default:
        orc_set_error("unknown model kind"); return -1;
    }
    return 0;
}

/* ---- steppers (methods_cuda.cu:271-291) ---- */
static inline float euler_gbm(const orc_params* P, float S, float dw, float h) {
    return S + P->mu * S * h + P->sigma * S * dw;
}
static inline float euler_dupire(const orc_params* P, float S, float dw, float t, float h) {
    float sigma_loc = P->sigma0 * expf(-P->alpha * t) * powf(S / P->S0, P->beta_d - 1.0f);
    return S + P->mu * S * h + sigma_loc * S * dw;
}
static inline float euler_lvq(const orc_params* P, float S, float dw, float t, float h) {
    float l = logf(S / P->lv_sref);
    float sigma_loc = P->lv_a + P->lv_b * expf(-P->lv_c * t) * l * l;
    return S + P->mu * S * h + sigma_loc * S * dw;
}
static inline void euler_heston(const orc_params* P, float* S, float* V, float dw1, float dw2,
                                float h) {
    float Vp = fmaxf(*V, 0.0f);
    float sq = sqrtf(Vp);
    *S = *S + P->mu * *S * h + sq * *S * dw1;
    float em = expf(-P->kappa * h);
    float v = P->theta + em * (*V - P->theta) + P->xi * sq * dw2;
    *V = fmaf(P->mil_q, dw2 * dw2 - h, v);  /* optional Milstein term; exact no-op at mil_q = 0 */
}

/* ---- payoff machinery (methods_cuda.cu:298-352) ---- */
static inline int path_dep(int pk) {
    return pk == FDMC_PAYOFF_ASIAN || pk == FDMC_PAYOFF_GEOM_ASIAN ||
           pk == FDMC_PAYOFF_LOOKBACK || pk == FDMC_PAYOFF_BARRIER;
}
static inline float running_init(const orc_params* P) {
    return P->payoff_kind == FDMC_PAYOFF_LOOKBACK ? P->S0 : 0.0f;
}
static inline void running_update(const orc_params* P, float* run, float S) {
    switch (P->payoff_kind) {
    case FDMC_PAYOFF_ASIAN: *run += S; break;
    case FDMC_PAYOFF_GEOM_ASIAN: *run += logf(S); break;
    case FDMC_PAYOFF_LOOKBACK: *run = fminf(*run, S); break;
    case FDMC_PAYOFF_BARRIER: *run = fmaxf(*run, S); break;
    default: break;
    }
}
static inline void apply_bgk(const orc_params* P, float* run, float h) {
    if (!P->bgk_on) return;
    float corr = expf(-BGK_BETA_F * P->sigma_bgk * sqrtf(h));
    if (P->payoff_kind == FDMC_PAYOFF_LOOKBACK) *run *= corr;
    else if (P->payoff_kind == FDMC_PAYOFF_BARRIER) *run /= corr;
}
static inline float payoff_eval(const orc_params* P, float S_T, float run, int n) {
    switch (P->payoff_kind) {
    case FDMC_PAYOFF_EUROPEAN: return fmaxf(S_T - P->K, 0.0f) * P->discount;
    case FDMC_PAYOFF_ASIAN: return fmaxf(run / n - P->K, 0.0f);
    case FDMC_PAYOFF_GEOM_ASIAN: return fmaxf(expf(run / n) - P->K, 0.0f);
    /* methods_cuda.cu:346-347 returns S_T - running, NaN for a path that left the
       Euler scheme's domain (S <= 0 under a local vol): such a path pays 0 here,
       as it does under the calls through fmaxf; for S_T > 0 nothing changes */
    case FDMC_PAYOFF_LOOKBACK: return (S_T > 0.0f) ? S_T - run : 0.0f;
    case FDMC_PAYOFF_BARRIER: return (run >= P->B) ? 0.0f : fmaxf(S_T - P->K, 0.0f) * P->discount;
    default: return fmaxf(S_T - P->K, 0.0f) * P->discount;
    }
}

static inline void one_factor_step(const orc_params* P, float* S, float dw, int k, float h) {
    switch (P->model_kind) {
    case FDMC_MODEL_GBM: *S = euler_gbm(P, *S, dw, h); break;
    case FDMC_MODEL_DUPIRE: *S = euler_dupire(P, *S, dw, (float)k * h, h); break;
    case FDMC_MODEL_LOCALVOL_QUAD: *S = euler_lvq(P, *S, dw, (float)k * h, h); break;
    default: break;
    }
}

/* kernel_mc + d_step (methods_cuda.cu:388-437). */
float orc_path_single(const orc_params* P, const float* dW, int n_steps) {
    float S = P->S0, V = P->v0, run = running_init(P);
    int pd = path_dep(P->payoff_kind);
    for (int k = 0; k < n_steps; k++) {
        if (P->model_kind == FDMC_MODEL_HESTON) {
            float z1 = dW[2 * k], z2 = dW[2 * k + 1];
            float dw2 = P->L10 * z1 + P->L11 * z2;
            euler_heston(P, &S, &V, z1, dw2, P->h);
        } else {
            one_factor_step(P, &S, dW[k], k, P->h);
        }
        if (pd) running_update(P, &run, S);
    }
    if (pd) apply_bgk(P, &run, P->h);
    return payoff_eval(P, S, run, n_steps);
}

/* Terminal (S, V) of n_paths Heston paths through the same step as
 * orc_path_single; dW holds each path's n_steps (dW1, dW2) pairs of scaled,
 * already correlated increments, path-major.  Test hook for the variance
 * scheme's strong order. */
void orc_heston_terminal(const orc_params* P, const float* dW, int n_steps, int n_paths,
                         float* S_T, float* V_T) {
    for (int p = 0; p < n_paths; p++) {
        const float* d = dW + (size_t)p * n_steps * 2;
        float S = P->S0, V = P->v0;
        for (int k = 0; k < n_steps; k++) euler_heston(P, &S, &V, d[2 * k], d[2 * k + 1], P->h);
        S_T[p] = S;
        V_T[p] = V;
    }
}

/* kernel_mlmc (methods_cuda.cu:445-518), with the running update applied for
 * Heston too (the reference skips it there, :464-470 and :484-490). */
void orc_path_coupled(const orc_params* P, const float* in, int n_fine, int M, float* pf,
                      float* pc) {
    int n_coarse = (n_fine >= M) ? n_fine / M : 0;
    float h_f = (float)((double)P->T / n_fine);
    float h_c = (float)((double)P->T / (n_coarse > 0 ? n_coarse : 1));
    float sq_hf = sqrtf(h_f);
    float Sf = P->S0, Vf = P->v0, Sc = Sf, Vc = Vf;
    float rf = running_init(P), rc = rf, acc1 = 0.0f, acc2 = 0.0f;
    int pd = path_dep(P->payoff_kind), ck = 0;
    for (int k = 0; k < n_fine; k++) {
        if (P->model_kind == FDMC_MODEL_HESTON) {
            float z1 = in[2 * k], z2 = in[2 * k + 1];
            float dw1 = z1 * sq_hf;
            float dw2 = (P->L10 * z1 + P->L11 * z2) * sq_hf;
            acc1 += dw1; acc2 += dw2;
            euler_heston(P, &Sf, &Vf, dw1, dw2, h_f);
        } else {
            float dw = in[k];
            acc1 += dw;
            one_factor_step(P, &Sf, dw, k, h_f);
        }
        if (pd) running_update(P, &rf, Sf);
        if ((k + 1) % M == 0) {
            if (P->model_kind == FDMC_MODEL_HESTON) {
                float Vcp = fmaxf(Vc, 0.0f), sq = sqrtf(Vcp);
                Sc = Sc + P->mu * Sc * h_c + sq * Sc * acc1;
                float emc = expf(-P->kappa * h_c);
                float vc = P->theta + emc * (Vc - P->theta) + P->xi * sq * acc2;
                Vc = fmaf(P->mil_q, acc2 * acc2 - h_c, vc);
                acc1 = acc2 = 0.0f;
            } else {
                one_factor_step(P, &Sc, acc1, ck, h_c);
                acc1 = 0.0f;
            }
            if (pd) running_update(P, &rc, Sc);
            ++ck;
        }
    }
    if (pd) { apply_bgk(P, &rf, h_f); apply_bgk(P, &rc, h_c); }
    *pf = payoff_eval(P, Sf, rf, n_fine);
    *pc = (n_coarse > 0) ? payoff_eval(P, Sc, rc, n_coarse) : 0.0f;
}

/* kernel_multi_dupire (methods_cuda.cu:526-551) generalised to weighted GBM baskets.
 * The Dupire local vol of every asset is taken relative to the first asset's S0,
 * as the reference does (c_p.S0 inside d_euler_dupire). */
float orc_path_basket(const orc_params* P, const float* dW, int n_steps) {
    int n = P->n_assets;
    float acc = 0.0f;
    for (int a = 0; a < n; a++) {
        float S = P->S0_assets ? (float)P->S0_assets[a] : P->S0;
        float sg = P->sigma_assets ? (float)P->sigma_assets[a] : 0.0f;
        for (int k = 0; k < n_steps; k++) {
            float dw = dW[k * n + a];
            if (P->model_kind == FDMC_MODEL_MULTI_DUPIRE)
                S = euler_dupire(P, S, dw, (float)k * P->h, P->h);
            else
                S = S + P->mu * S * P->h + sg * S * dw;
        }
        acc += P->weights ? (float)P->weights[a] * S : S;
    }
    float B = P->weights ? acc : acc / n;
    return fmaxf(B - P->K, 0.0f) * P->discount;
}

/* kernel_gbm_asian_cv / kernel_dupire_gbm_cv (methods_cuda.cu:603-677): both
 * legs driven by the same increments; each leg uses its own parameters. */
void orc_path_cv(const orc_params* Pm, const orc_params* Pc, const float* dW, int n_steps,
                 float* ym, float* yc) {
    *ym = orc_path_single(Pm, dW, n_steps);
    *yc = orc_path_single(Pc, dW, n_steps);
}
