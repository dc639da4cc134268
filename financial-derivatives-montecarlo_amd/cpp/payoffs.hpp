// payoffs.hpp -- payoff types of the engine's C++ facade.
//
// Same type names, fields and helpers as the reference's payoffs.hpp
// (payoffs.hpp:7-50).  Asian, GeomAsian and Lookback are undiscounted there and
// here.  Basket gains an optional weight vector (empty: equal weights 1/n).
#pragma once

#include <cmath>
#include <type_traits>
#include <variant>
#include <vector>

// Broadie-Glasserman-Kou continuity correction constant.
inline constexpr double BGK_BETA = 0.5826;
inline constexpr float BGK_BETA_F = 0.5826f;

struct European { double K, r, T; };          // max(S_T - K, 0) e^{-rT}
struct Asian { double K; };                   // max(mean_k S_k - K, 0)
struct GeomAsian { double K; };               // max(exp(mean_k ln S_k) - K, 0)
struct Lookback { double sigma; };            // S_T - min_k S_k (BGK-shifted)
struct Barrier { double K, B, sigma, r, T; }; // up-and-out call (BGK-shifted barrier)
struct Basket {                               // max(sum_a w_a S_T^a - K, 0) e^{-rT}
    double K, r, T;
    int n_assets;
    std::vector<double> weights{};
};

using PayoffVariant = std::variant<European, Asian, GeomAsian, Lookback, Barrier, Basket>;

inline bool payoff_needs_full_path(const PayoffVariant& pv) {
    return std::holds_alternative<Asian>(pv) || std::holds_alternative<GeomAsian>(pv) ||
           std::holds_alternative<Lookback>(pv) || std::holds_alternative<Barrier>(pv);
}

inline double payoff_discount(const PayoffVariant& pv) {
    return std::visit([](const auto& p) -> double {
        using T = std::decay_t<decltype(p)>;
        if constexpr (std::is_same_v<T, European> || std::is_same_v<T, Barrier> || std::is_same_v<T, Basket>)
            return std::exp(-p.r * p.T);
        else
            return 1.0;
    }, pv);
}
