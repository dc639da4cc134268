// models.hpp -- model parameter types of the engine's C++ facade.
//
// Same type names, fields and defaults as the reference's models.hpp
// (models.hpp:15-101) so that callers compile unchanged; two alternatives are
// appended to the variant (LocalVolQuadParams, MultiGBMParams), which keeps the
// index of every existing alternative.
#pragma once

#include <cmath>
#include <type_traits>
#include <variant>
#include <vector>

inline constexpr int CONST_DIM_MAX = 64;      // kept for source compatibility (unused)
inline constexpr int D_MAX_SOBOL = 20000;     // Sobol dimensions; above: pseudo-random points

struct GBMParams {          // dS = mu S dt + sigma S dW
    double S0 = 100.0, mu = 0.05, sigma = 0.20, T = 1.0;
};

struct HestonParams {       // dS = mu S dt + sqrt(v) S dW1; dv = kappa (theta - v) dt + xi sqrt(v) dW2
    double S0 = 100.0, mu = 0.05, kappa = 2.0, theta = 0.04, xi = 0.5, rho = -0.9, v0 = 0.04, T = 1.0;
    double L[2][2] = {};    // Cholesky factor of [[1, rho], [rho, 1]]
    // variance scheme: 0 = the reference's full-truncation exponential Euler,
    // 1 = with the Milstein term xi^2/4 (dW2^2 - h) (no counterpart in the reference)
    int scheme = 0;
    void compute_cholesky() {
        L[0][0] = 1.0;
        L[0][1] = 0.0;
        L[1][0] = rho;
        L[1][1] = std::sqrt(1.0 - rho * rho);
    }
};

struct DupireLocalParams {  // sigma_loc(S, t) = sigma0 exp(-alpha t) (S / S0)^(beta_d - 1)
    double S0 = 100.0, mu = 0.05, sigma0 = 0.20, alpha = 0.5, beta_d = 0.7, T = 1.0;
};

struct MultiDupireParams {  // n assets with the Dupire local vol, correlated through L
    int n = 100;
    double mu = 0.05, sigma0 = 0.20, alpha = 0.5, beta_d = 0.7, T = 1.0;
    bool uncorrelated = false;
    std::vector<double> S0;  // n values
    std::vector<double> L;   // n x n lower Cholesky, row-major (empty: identity)
};

// Additive: sigma(S, t) = a + b exp(-c t) ln(S / s_ref)^2.
struct LocalVolQuadParams {
    double S0 = 100.0, mu = 0.05, a = 0.2, b = 0.1, c = 1.0, s_ref = 100.0, T = 1.0;
};

// Additive: n correlated GBM assets with per-asset volatilities.
struct MultiGBMParams {
    int n = 1;
    double mu = 0.05, T = 1.0;
    bool uncorrelated = false;
    std::vector<double> S0, sigma, L;
};

using ModelVariant = std::variant<GBMParams, HestonParams, DupireLocalParams, MultiDupireParams,
                                  LocalVolQuadParams, MultiGBMParams>;

inline int model_noise_dim(const ModelVariant& mv) {
    return std::visit([](const auto& m) -> int {
        using T = std::decay_t<decltype(m)>;
        if constexpr (std::is_same_v<T, HestonParams>) return 2;
        else if constexpr (std::is_same_v<T, MultiDupireParams> || std::is_same_v<T, MultiGBMParams>) return m.n;
        else return 1;
    }, mv);
}

inline double model_T(const ModelVariant& mv) {
    return std::visit([](const auto& m) { return m.T; }, mv);
}

inline double model_S0(const ModelVariant& mv) {
    return std::visit([](const auto& m) -> double {
        using T = std::decay_t<decltype(m)>;
        if constexpr (std::is_same_v<T, MultiDupireParams> || std::is_same_v<T, MultiGBMParams>)
            return m.S0.empty() ? 100.0 : m.S0[0];
        else
            return m.S0;
    }, mv);
}
