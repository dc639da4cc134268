// utils.hpp -- host numerics of the engine's C++ facade (reference utils.hpp:14-102):
// closed forms, Brownian-bridge and PCA precomputation, Richardson step
// selection, running statistics and the results table.
#pragma once

#include <cstdint>
#include <functional>
#include <string>
#include <vector>

#include "models.hpp"
#include "payoffs.hpp"

double bs_call(double S0, double K, double T, double r, double sigma);
// Undiscounted discrete geometric-Asian call under GBM, n dates t_k = k T / n.
double geom_asian_analytic(double S0, double K, double T, double mu, double sigma, int n_steps);

struct BBData {
    int N;
    double T;
    std::vector<int> map_idx, left_idx, right_idx;
    std::vector<double> weight_left, weight_right, std_dev;
};
BBData bb_precompute(int N, double T);                       // N a power of two
void bb_apply(const BBData& bb, double* Z, double* dW, int n_sim);

struct PCAData {
    int m;
    double T;
    std::vector<double> M_pca;      // m x m column-major, dW = M z
    std::vector<float> M_pca_f32;
};
PCAData pca_compute(int m, double T);
// Additive: PCA of the correlated increments of n assets over n_steps dates,
// covariance C_time (x) Corr with dimension index k * n + a.
PCAData pca_kron_compute(int n_steps, double T, const std::vector<double>& corr, int n);

// Symmetric eigendecomposition (Householder tridiagonalisation + implicit QL):
// eigenvalues descending in w, eigenvectors in the columns of V (row-major n x n).
void eigh_symmetric(int n, const std::vector<double>& A, std::vector<double>& w, std::vector<double>& V);

using SimFn = std::function<double(int n_steps, long long n_paths, unsigned seed)>;
// Returns 1 / c where bias(h) ~ c h, from runs at n_fine = max(4 M, 4) and n_fine / M
// (the reference's naming, utils.cpp:174-185).
double estimar_c1_richardson(SimFn sim_fn, double T, int M_rich, int N_pilot, unsigned seed = 0);

struct RunningStats {
    double mean = 0.0, M2 = 0.0;
    long long n = 0;
    void update(double x) {
        ++n;
        const double d = x - mean;
        mean += d / n;
        M2 += d * (x - mean);
    }
    double variance() const { return n > 1 ? M2 / (n - 1) : 0.0; }
    double std_error() const;
};

struct TableRow {
    std::string method;
    double price, std_error;
    long long n_samples;
    double time_s;
};
// Same layout as the reference's table (utils.cpp:192-210): OK when |price - ref| < 2 eps.
void print_table(const std::vector<TableRow>& rows, double price_ref, double epsilon,
                 const std::string& example_name);
