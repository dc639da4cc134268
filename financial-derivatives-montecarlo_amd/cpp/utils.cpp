// utils.cpp -- host numerics of the C++ facade (reference utils.cpp:14-210).
//
// PCA: the Brownian covariance C_ij = h min(i+1, j+1) has closed-form
// eigenpairs (lambda_k = h / (4 sin^2((2k-1) pi / (2(2m+1)))),
// v_k(i) = 2/sqrt(2m+1) sin((2k-1) i pi / (2m+1))), used directly in FP64; the
// correlation factor of the Kronecker PCA goes through eigh_symmetric (cyclic
// Jacobi with a threshold).  The reference calls Eigen's SelfAdjointEigenSolver
// (utils.cpp:145); eigenvector signs are implementation-defined in both.
#include "utils.hpp"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <numeric>
#include <stdexcept>

namespace {
double norm_cdf(double x) { return 0.5 * std::erfc(-x / std::sqrt(2.0)); }
constexpr double kPi = 3.14159265358979323846;
}  // namespace

double bs_call(double S0, double K, double T, double r, double sigma) {
    if (T <= 0.0 || sigma <= 0.0) return std::max(S0 - K * std::exp(-r * T), 0.0);
    const double sT = std::sqrt(T);
    const double d1 = (std::log(S0 / K) + (r + 0.5 * sigma * sigma) * T) / (sigma * sT);
    return S0 * norm_cdf(d1) - K * std::exp(-r * T) * norm_cdf(d1 - sigma * sT);
}

double geom_asian_analytic(double S0, double K, double T, double mu, double sigma, int n) {
    // ln G ~ N(m, s2): m = ln S0 + (mu - sigma^2/2) T (n+1)/(2n), s2 = sigma^2 T (n+1)(2n+1)/(6 n^2)
    const double m = std::log(S0) + (mu - 0.5 * sigma * sigma) * T * (n + 1) / (2.0 * n);
    const double s2 = sigma * sigma * T * (n + 1) * (2.0 * n + 1) / (6.0 * n * (double)n);
    const double s = std::sqrt(s2);
    if (s < 1e-12) return std::max(std::exp(m) - K, 0.0);
    const double d1 = (m + s2 - std::log(K)) / s;
    return std::exp(m + 0.5 * s2) * norm_cdf(d1) - K * norm_cdf(d1 - s);
}

BBData bb_precompute(int N, double T) {
    if (N <= 0 || (N & (N - 1))) throw std::runtime_error("bb_precompute: N must be a power of two");
    BBData bb{N, T, std::vector<int>(N, 0), std::vector<int>(N, 0), std::vector<int>(N, 0),
              std::vector<double>(N, 0.0), std::vector<double>(N, 0.0), std::vector<double>(N, 0.0)};
    auto t = [&](int i) { return i * T / N; };
    bb.map_idx[0] = N;               // the terminal value comes first
    bb.std_dev[0] = std::sqrt(T);
    int e = 1;
    for (int stride = N; stride > 1; stride >>= 1) {          // coarse to fine
        for (int L = 0; L < N; L += stride) {
            const int R = L + stride, M = L + stride / 2;
            bb.map_idx[e] = M;
            bb.left_idx[e] = L;
            bb.right_idx[e] = R;
            bb.weight_left[e] = (t(R) - t(M)) / (t(R) - t(L));
            bb.weight_right[e] = (t(M) - t(L)) / (t(R) - t(L));
            bb.std_dev[e] = std::sqrt((t(M) - t(L)) * (t(R) - t(M)) / (t(R) - t(L)));
            ++e;
        }
    }
    return bb;
}

void bb_apply(const BBData& bb, double* Z, double* dW, int n_sim) {
    const int N = bb.N;
    std::vector<double> W(N + 1);
    for (int s = 0; s < n_sim; s++) {
        const double* z = Z + (size_t)s * N;
        std::fill(W.begin(), W.end(), 0.0);
        W[bb.map_idx[0]] = bb.std_dev[0] * z[0];
        for (int j = 1; j < N; j++)
            W[bb.map_idx[j]] = bb.weight_left[j] * W[bb.left_idx[j]] + bb.weight_right[j] * W[bb.right_idx[j]] +
                               bb.std_dev[j] * z[j];
        for (int k = 0; k < N; k++) dW[(size_t)s * N + k] = W[k + 1] - W[k];
    }
}

void eigh_symmetric(int n, const std::vector<double>& A_in, std::vector<double>& w, std::vector<double>& V) {
    std::vector<double> A = A_in;
    V.assign((size_t)n * n, 0.0);
    for (int i = 0; i < n; i++) V[(size_t)i * n + i] = 1.0;
    auto a = [&](int i, int j) -> double& { return A[(size_t)i * n + j]; };
    for (int sweep = 0; sweep < 64; sweep++) {
        double off = 0.0, diag = 0.0;
        for (int i = 0; i < n; i++) {
            diag += a(i, i) * a(i, i);
            for (int j = i + 1; j < n; j++) off += 2.0 * a(i, j) * a(i, j);
        }
        if (off <= 1e-30 * (off + diag)) break;
        const double thresh = sweep < 3 ? 0.2 * std::sqrt(off) / (n * (double)n) : 0.0;
        for (int p = 0; p < n; p++)
            for (int q = p + 1; q < n; q++) {
                const double apq = a(p, q);
                if (std::fabs(apq) <= thresh || apq == 0.0) continue;
                // rotation that annihilates a(p, q)
                const double tau = (a(q, q) - a(p, p)) / (2.0 * apq);
                const double t = std::copysign(1.0, tau) / (std::fabs(tau) + std::hypot(1.0, tau));
                const double c = 1.0 / std::hypot(1.0, t), s = t * c;
                for (int k = 0; k < n; k++) {
                    const double kp = a(k, p), kq = a(k, q);
                    a(k, p) = c * kp - s * kq;
                    a(k, q) = s * kp + c * kq;
                }
                for (int k = 0; k < n; k++) {
                    const double pk = a(p, k), qk = a(q, k);
                    a(p, k) = c * pk - s * qk;
                    a(q, k) = s * pk + c * qk;
                }
                for (int k = 0; k < n; k++) {
                    const double kp = V[(size_t)k * n + p], kq = V[(size_t)k * n + q];
                    V[(size_t)k * n + p] = c * kp - s * kq;
                    V[(size_t)k * n + q] = s * kp + c * kq;
                }
            }
    }
    std::vector<int> ord(n);
    std::iota(ord.begin(), ord.end(), 0);
    std::stable_sort(ord.begin(), ord.end(), [&](int x, int y) { return a(x, x) > a(y, y); });
    std::vector<double> V2((size_t)n * n);
    w.resize(n);
    for (int c = 0; c < n; c++) {
        w[c] = a(ord[c], ord[c]);
        for (int k = 0; k < n; k++) V2[(size_t)k * n + c] = V[(size_t)k * n + ord[c]];
    }
    V.swap(V2);
}

namespace {
// Closed-form eigenpairs of h min(i+1, j+1) (descending): V[i*m + k].
void brownian_eig(int m, double h, std::vector<double>& lam, std::vector<double>& V) {
    lam.resize(m);
    V.resize((size_t)m * m);
    const double c = 2.0 / std::sqrt(2.0 * m + 1.0);
    for (int k = 1; k <= m; k++) {
        const double s = std::sin((2.0 * k - 1.0) * kPi / (2.0 * (2.0 * m + 1.0)));
        lam[k - 1] = h / (4.0 * s * s);
        for (int i = 1; i <= m; i++)
            V[(size_t)(i - 1) * m + (k - 1)] = c * std::sin((2.0 * k - 1.0) * i * kPi / (2.0 * m + 1.0));
    }
}
}  // namespace

PCAData pca_compute(int m, double T) {
    if (m <= 0) throw std::runtime_error("pca_compute: m must be positive");
    PCAData p{m, T, std::vector<double>((size_t)m * m), std::vector<float>((size_t)m * m)};
    std::vector<double> lam, V;
    brownian_eig(m, T / m, lam, V);
    // M_cum = V sqrt(Lambda); dW rows are differences of consecutive W rows
    for (int k = 0; k < m; k++) {
        const double sl = std::sqrt(std::max(lam[k], 0.0));
        for (int i = 0; i < m; i++) {
            const double cur = V[(size_t)i * m + k] * sl;
            const double prev = i > 0 ? V[(size_t)(i - 1) * m + k] * sl : 0.0;
            p.M_pca[(size_t)k * m + i] = cur - prev;
        }
    }
    for (size_t i = 0; i < p.M_pca.size(); i++) p.M_pca_f32[i] = (float)p.M_pca[i];
    return p;
}

PCAData pca_kron_compute(int n_steps, double T, const std::vector<double>& corr, int n) {
    if ((int)corr.size() != n * n) throw std::runtime_error("pca_kron_compute: corr must be n x n");
    const int D = n_steps * n;
    PCAData p{D, T, std::vector<double>((size_t)D * D), std::vector<float>((size_t)D * D)};
    std::vector<double> lt, Vt, lc, Vc;
    brownian_eig(n_steps, T / n_steps, lt, Vt);
    eigh_symmetric(n, corr, lc, Vc);
    struct Pair { double lam; int i, j; };
    std::vector<Pair> pr;
    for (int i = 0; i < n_steps; i++)
        for (int j = 0; j < n; j++) pr.push_back({lt[i] * lc[j], i, j});
    std::stable_sort(pr.begin(), pr.end(), [](const Pair& x, const Pair& y) { return x.lam > y.lam; });
    for (int col = 0; col < D; col++) {
        const double sl = std::sqrt(std::max(pr[col].lam, 0.0));
        const int i = pr[col].i, j = pr[col].j;
        for (int k = 0; k < n_steps; k++)
            for (int a = 0; a < n; a++) {
                const double cur = Vt[(size_t)k * n_steps + i] * Vc[(size_t)a * n + j] * sl;
                const double prev = k > 0 ? Vt[(size_t)(k - 1) * n_steps + i] * Vc[(size_t)a * n + j] * sl : 0.0;
                p.M_pca[(size_t)col * D + (size_t)k * n + a] = cur - prev;
            }
    }
    for (size_t i = 0; i < p.M_pca.size(); i++) p.M_pca_f32[i] = (float)p.M_pca[i];
    return p;
}

double estimar_c1_richardson(SimFn sim_fn, double T, int M_rich, int N_pilot, unsigned seed) {
    const int n_fine = std::max(4 * M_rich, 4), n_coarse = n_fine / M_rich;
    const double pf = sim_fn(n_fine, N_pilot, seed);
    const double pc = sim_fn(n_coarse, N_pilot, seed + 1);
    const double h = T / n_fine;
    const double c = std::max(std::fabs(pc - pf) / (h * (M_rich - 1)), 1e-12);
    return 1.0 / c;
}

double RunningStats::std_error() const { return n > 0 ? std::sqrt(variance() / n) : 0.0; }

void print_table(const std::vector<TableRow>& rows, double price_ref, double epsilon,
                 const std::string& example_name) {
    std::printf("\n[%s]   epsilon = %g\n", example_name.c_str(), epsilon);
    std::printf("  Referencia: %.6f\n\n", price_ref);
    std::printf("  %-22s%10s%10s%12s%9s%10s%6s\n", "Metodo", "Precio", "StdErr", "N", "T(s)", "|Error|", "OK?");
    std::printf("  %s\n", std::string(79, '-').c_str());
    for (const auto& r : rows) {
        const double err = std::fabs(r.price - price_ref);
        std::printf("  %-22s%10.4f%10.4f%12lld%9.3f%10.4f%6s\n", r.method.c_str(), r.price, r.std_error,
                    r.n_samples, r.time_s, err, err < 2.0 * epsilon ? "SI" : "NO");
    }
    std::printf("\n");
}
