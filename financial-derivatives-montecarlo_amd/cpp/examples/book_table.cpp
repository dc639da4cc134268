// book_table.cpp -- the contract book through the C++ facade: a seeded book of
// local-vol contracts (European / Asian / lookback / barrier, 16..512 dates)
// priced with adjoint Greeks in one batched run and one contract at a time,
// printed as a table.  Exit status 0 when both modes agree exactly and every
// price reached standard error eps / sqrt(2).
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <random>
#include <vector>

#include "../methods_hip.hpp"

int main(int argc, char** argv) {
    const double eps = argc > 1 ? std::atof(argv[1]) : 0.01;
    const int n = argc > 2 ? std::atoi(argv[2]) : 64;
    LocalVolQuadParams lv;                           // sigma = 0.2 + 0.1 e^-t ln(S/100)^2
    std::mt19937_64 rng(2026);
    std::uniform_real_distribution<double> strike(70.0, 130.0);
    const double mats[4] = {0.25, 0.5, 1.0, 2.0};
    std::vector<BookContract> book;
    const char* names[4] = {"European", "Asian", "Lookback", "Barrier"};
    std::vector<int> kinds;
    for (int i = 0; i < n; ++i) {
        const int kind = (int)(rng() % 4);
        const double K = strike(rng), T = mats[rng() % 4];
        const int steps = 16 << (rng() % 6);
        PayoffVariant p = European{K, lv.mu, T};
        if (kind == 1) p = Asian{K};
        if (kind == 2) p = Lookback{0.2};
        if (kind == 3) p = Barrier{K, 1.3 * K, 0.2, lv.mu, T};
        book.push_back({p, steps, T});
        kinds.push_back(kind);
    }
    auto t0 = std::chrono::steady_clock::now();
    const std::vector<BookResult> bat = run_book_cuda(lv, book, eps, 16, BookMode::Batched);
    auto t1 = std::chrono::steady_clock::now();
    const std::vector<BookResult> pla = run_book_cuda(lv, book, eps, 16, BookMode::Plain);
    auto t2 = std::chrono::steady_clock::now();

    std::printf("%4s %-9s %5s %8s %5s %12s %10s %9s %9s %9s %5s\n", "c", "payoff", "dates", "K", "T", "price", "se",
                "delta", "vega", "rho", "pass");
    bool ok = true;
    for (int i = 0; i < n; ++i) {
        const BookResult& r = bat[i];
        const BookResult& q = pla[i];
        const bool same = r.price.price == q.price.price && r.price.std_error == q.price.std_error &&
                          r.price.n_samples == q.price.n_samples &&
                          (r.greeks.delta == q.greeks.delta || (std::isnan(r.greeks.delta) && std::isnan(q.greeks.delta)));
        ok = ok && same && r.converged && r.price.std_error <= eps / std::sqrt(2.0) * (1 + 1e-12);
        double K = 0.0;
        if (auto* e = std::get_if<European>(&book[i].payoff)) K = e->K;
        if (auto* a = std::get_if<Asian>(&book[i].payoff)) K = a->K;
        if (auto* b = std::get_if<Barrier>(&book[i].payoff)) K = b->K;
        std::printf("%4d %-9s %5d %8.3f %5.2f %12.6f %10.2e %9.5f %9.4f %9.4f %5d%s\n", i, names[kinds[i]],
                    book[i].n_steps, K, book[i].T, r.price.price, r.price.std_error, r.greeks.delta, r.greeks.vega,
                    r.greeks.rho, r.iterations,
                    !same ? "  (plain differs)" : !r.converged ? "  (not converged)" : "");
    }
    const double tb = std::chrono::duration<double>(t1 - t0).count();
    const double tp = std::chrono::duration<double>(t2 - t1).count();
    std::printf("batched %.3f s, plain %.3f s (book set-up included), batched == plain: %s\n", tb, tp,
                ok ? "yes" : "NO");
    return ok ? 0 : 1;
}
