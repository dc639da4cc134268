/*
 * oracle_numerics.c -- Brownian bridge, PCA (own Jacobi eigensolver) and closed
 * forms.  TEST INFRASTRUCTURE (see oracle.h).
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

/* ---------------- Brownian bridge: bb_precompute (utils.cpp:54-101) ---------------- */
int orc_bb_precompute(int N, double T, orc_bb* bb) {
    if (N <= 0 || (N & (N - 1)) != 0) return -1;
    memset(bb, 0, sizeof(*bb));
    bb->N = N;
    bb->T = T;
    bb->map_idx = (int32_t*)calloc((size_t)N, sizeof(int32_t));
    bb->left_idx = (int32_t*)calloc((size_t)N, sizeof(int32_t));
    bb->right_idx = (int32_t*)calloc((size_t)N, sizeof(int32_t));
    bb->wl = (double*)calloc((size_t)N, sizeof(double));
    bb->wr = (double*)calloc((size_t)N, sizeof(double));
    bb->sd = (double*)calloc((size_t)N, sizeof(double));
    double* t = (double*)malloc(sizeof(double) * (size_t)(N + 1));
    for (int i = 0; i <= N; i++) t[i] = i * T / N;
    bb->map_idx[0] = N;          /* first entry: terminal W_N with sd sqrt(T) */
    bb->sd[0] = sqrt(T);
    int step = 1, levels = 0;
    for (int tmp = N; tmp > 1; tmp >>= 1) levels++;
    for (int lev = 0; lev < levels; lev++) {
        int np = 1 << lev, stride = N >> lev, half = stride >> 1;
        for (int j = 0; j < np; j++) {
            int L = j * stride, R = (j + 1) * stride, M = L + half;
            bb->map_idx[step] = M;
            bb->left_idx[step] = L;
            bb->right_idx[step] = R;
            bb->wl[step] = (t[R] - t[M]) / (t[R] - t[L]);
            bb->wr[step] = (t[M] - t[L]) / (t[R] - t[L]);
            bb->sd[step] = sqrt((t[M] - t[L]) * (t[R] - t[M]) / (t[R] - t[L]));
            step++;
        }
    }
    free(t);
    return 0;
}

void orc_bb_release(orc_bb* bb) {
    free(bb->map_idx); free(bb->left_idx); free(bb->right_idx);
    free(bb->wl); free(bb->wr); free(bb->sd);
    memset(bb, 0, sizeof(*bb));
}

/* bb_apply (utils.cpp:104-124) for one path. */
void orc_bb_apply_f64(const orc_bb* bb, const double* z, double* dW) {
    int N = bb->N;
    double* W = (double*)calloc((size_t)(N + 1), sizeof(double));
    W[bb->map_idx[0]] = bb->sd[0] * z[0];
    for (int s = 1; s < N; s++)
        W[bb->map_idx[s]] = bb->wl[s] * W[bb->left_idx[s]] + bb->wr[s] * W[bb->right_idx[s]] +
                            bb->sd[s] * z[s];
    for (int k = 0; k < N; k++) dW[k] = W[k + 1] - W[k];
    free(W);
}

/* kernel_bb_transform (methods_cuda.cu:561-594) for one path, FP32 with the
 * weights rounded to float as bb_upload does (methods_cuda.cu:938-946). */
void orc_bb_apply_f32(const orc_bb* bb, const float* z, float* dW) {
    int N = bb->N;
    float Wst[1025];
    float* W = (N + 1 <= 1025) ? Wst : (float*)malloc(sizeof(float) * (size_t)(N + 1));
    W[0] = 0.0f;
    W[bb->map_idx[0]] = (float)bb->sd[0] * z[0];
    for (int s = 1; s < N; s++) {
        float wl = (float)bb->wl[s], wr = (float)bb->wr[s], sd = (float)bb->sd[s];
        W[bb->map_idx[s]] = wl * W[bb->left_idx[s]] + wr * W[bb->right_idx[s]] + sd * z[s];
    }
    for (int k = 0; k < N; k++) dW[k] = W[k + 1] - W[k];
    if (W != Wst) free(W);
}

/* ---------------- PCA ---------------- */
/* Cyclic Jacobi (Golub & Van Loan, Matrix Computations, 4th ed., Alg. 8.5.3)
 * with a threshold sweep; converges quadratically for symmetric A. */
int orc_eigh_jacobi(int n, double* A, double* w, double* V) {
    for (int i = 0; i < n; i++)
        for (int j = 0; j < n; j++) V[i * n + j] = (i == j) ? 1.0 : 0.0;
    for (int sweep = 0; sweep < 100; sweep++) {
        double off = 0.0, tot = 0.0;
        for (int i = 0; i < n; i++)
            for (int j = 0; j < n; j++) {
                double a2 = A[i * n + j] * A[i * n + j];
                tot += a2;
                if (i != j) off += a2;
            }
        if (off <= 1e-30 * tot) break;
        for (int p = 0; p < n - 1; p++)
            for (int q = p + 1; q < n; q++) {
                double apq = A[p * n + q];
                if (fabs(apq) < 1e-300) continue;
                double theta = (A[q * n + q] - A[p * n + p]) / (2.0 * apq);
                double t = (theta >= 0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta * theta + 1.0));
                double c = 1.0 / sqrt(t * t + 1.0), s = t * c;
                for (int k = 0; k < n; k++) {   /* A <- A J */
                    double akp = A[k * n + p], akq = A[k * n + q];
                    A[k * n + p] = c * akp - s * akq;
                    A[k * n + q] = s * akp + c * akq;
                }
                for (int k = 0; k < n; k++) {   /* A <- J^T A */
                    double apk = A[p * n + k], aqk = A[q * n + k];
                    A[p * n + k] = c * apk - s * aqk;
                    A[q * n + k] = s * apk + c * aqk;
                }
                for (int k = 0; k < n; k++) {   /* V <- V J */
                    double vkp = V[k * n + p], vkq = V[k * n + q];
                    V[k * n + p] = c * vkp - s * vkq;
                    V[k * n + q] = s * vkp + c * vkq;
                }
            }
    }
    /* sort descending (selection sort on columns) */
    for (int i = 0; i < n; i++) w[i] = A[i * n + i];
    for (int i = 0; i < n; i++) {
        int b = i;
        for (int j = i + 1; j < n; j++)
            if (w[j] > w[b]) b = j;
        if (b != i) {
            double tw = w[i]; w[i] = w[b]; w[b] = tw;
            for (int k = 0; k < n; k++) {
                double tv = V[k * n + i]; V[k * n + i] = V[k * n + b]; V[k * n + b] = tv;
            }
        }
    }
    return 0;
}

/* pca_compute (utils.cpp:131-167): C_ij = h min(i+1, j+1), eigenpairs sorted
 * descending, M_cum = V sqrt(Lambda), then rows differenced. */
int orc_pca_compute(int m, double T, double* M) {
    double h = T / m;
    double* A = (double*)malloc(sizeof(double) * (size_t)m * m);
    double* V = (double*)malloc(sizeof(double) * (size_t)m * m);
    double* w = (double*)malloc(sizeof(double) * (size_t)m);
    for (int i = 0; i < m; i++)
        for (int j = 0; j < m; j++) A[i * m + j] = h * (double)((i < j ? i : j) + 1);
    orc_eigh_jacobi(m, A, w, V);
    for (int k = 0; k < m; k++) {
        double sl = sqrt(w[k] > 0 ? w[k] : 0.0);
        for (int i = 0; i < m; i++) {
            double cur = V[i * m + k] * sl;
            double prev = (i > 0) ? V[(i - 1) * m + k] * sl : 0.0;
            M[(size_t)k * m + i] = cur - prev;   /* column-major */
        }
    }
    free(A); free(V); free(w);
    return 0;
}

typedef struct { double lam; int i, j; } kron_pair;
static int cmp_pair(const void* a, const void* b) {
    double x = ((const kron_pair*)a)->lam, y = ((const kron_pair*)b)->lam;
    return (x < y) - (x > y);
}

int orc_pca_kron_compute(int n_steps, double T, int n, const double* corr, double* M) {
    double h = T / n_steps;
    int D = n_steps * n;
    double* At = (double*)malloc(sizeof(double) * (size_t)n_steps * n_steps);
    double* Vt = (double*)malloc(sizeof(double) * (size_t)n_steps * n_steps);
    double* wt = (double*)malloc(sizeof(double) * (size_t)n_steps);
    double* Ac = (double*)malloc(sizeof(double) * (size_t)n * n);
    double* Vc = (double*)malloc(sizeof(double) * (size_t)n * n);
    double* wc = (double*)malloc(sizeof(double) * (size_t)n);
    for (int i = 0; i < n_steps; i++)
        for (int j = 0; j < n_steps; j++) At[i * n_steps + j] = h * (double)((i < j ? i : j) + 1);
    memcpy(Ac, corr, sizeof(double) * (size_t)n * n);
    orc_eigh_jacobi(n_steps, At, wt, Vt);
    orc_eigh_jacobi(n, Ac, wc, Vc);
    kron_pair* pr = (kron_pair*)malloc(sizeof(kron_pair) * (size_t)D);
    for (int i = 0; i < n_steps; i++)
        for (int j = 0; j < n; j++) {
            pr[i * n + j].lam = wt[i] * wc[j];
            pr[i * n + j].i = i;
            pr[i * n + j].j = j;
        }
    qsort(pr, (size_t)D, sizeof(kron_pair), cmp_pair);
    for (int col = 0; col < D; col++) {
        double sl = sqrt(pr[col].lam > 0 ? pr[col].lam : 0.0);
        int i = pr[col].i, j = pr[col].j;
        for (int k = 0; k < n_steps; k++)
            for (int a = 0; a < n; a++) {
                double cur = Vt[k * n_steps + i] * Vc[a * n + j] * sl;
                double prev = (k > 0) ? Vt[(k - 1) * n_steps + i] * Vc[a * n + j] * sl : 0.0;
                M[(size_t)col * D + (size_t)k * n + a] = cur - prev;
            }
    }
    free(At); free(Vt); free(wt); free(Ac); free(Vc); free(wc); free(pr);
    return 0;
}

/* ---------------- closed forms (utils.cpp:14-47) ---------------- */
static double norm_cdf(double x) { return 0.5 * erfc(-x / sqrt(2.0)); }

double orc_bs_call(double S0, double K, double T, double r, double sigma) {
    if (T <= 0.0 || sigma <= 0.0) {
        double v = S0 - K * exp(-r * T);
        return v > 0 ? v : 0.0;
    }
    double sT = sqrt(T);
    double d1 = (log(S0 / K) + (r + 0.5 * sigma * sigma) * T) / (sigma * sT);
    double d2 = d1 - sigma * sT;
    return S0 * norm_cdf(d1) - K * exp(-r * T) * norm_cdf(d2);
}

double orc_geom_asian_analytic(double S0, double K, double T, double mu, double sigma, int n) {
    double m_G = log(S0) + (mu - 0.5 * sigma * sigma) * T * (n + 1) / (2.0 * n);
    double s2 = sigma * sigma * T * (n + 1) * (2.0 * n + 1) / (6.0 * n * (double)n);
    double s = sqrt(s2);
    if (s < 1e-12) {
        double v = exp(m_G) - K;
        return v > 0 ? v : 0.0;
    }
    double d1 = (m_G + s2 - log(K)) / s;
    double d2 = d1 - s;
    return exp(m_G + 0.5 * s2) * norm_cdf(d1) - K * norm_cdf(d2);
}
