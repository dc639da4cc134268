// heston_table.cpp -- drop-in check of the C++ facade on the reference's Heston
// example (examples/ejemplo05_heston.cpp): fixed-step reference price, the
// Richardson estimate of the Euler bias constant, the step count it implies
// (capped at 2^11 dates), then MC, MLMC, QMC Raw and MLQMC Raw through the
// reference's function names, plus MLMC with the optional Milstein variance
// term.  A second table prices the GBM Asian of examples/ejemplo02_asian.cpp on
// 2048 dates (that example's cap) with bridge and PCA noise.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "../methods_hip.hpp"

namespace {

int pow2_at_least(int n) {
    int p = 1;
    while (p < n) p <<= 1;
    return p;
}

// Single-level rows estimate the price at their own step count: within twice
// eps (the bias budget) plus three combined standard errors of the reference.
int outside(const std::vector<TableRow>& rows, double ref, double ref_se, double eps) {
    int bad = 0;
    for (const auto& r : rows) {
        const double tol = 2.0 * eps + 3.0 * std::sqrt(r.std_error * r.std_error + ref_se * ref_se);
        if (std::abs(r.price - ref) >= tol) {
            std::printf("  outside: %s %.6f vs %.6f (tol %.4f)\n", r.method.c_str(), r.price, ref, tol);
            ++bad;
        }
    }
    return bad;
}

}  // namespace

int main(int argc, char** argv) {
    const double eps = argc > 1 ? std::atof(argv[1]) : 0.05;
    HestonParams hes;                               // S0=100, mu=0.05, kappa=2, theta=0.04, xi=0.5, rho=-0.9
    hes.compute_cholesky();
    const double K = 100.0;
    ModelVariant mv = hes;
    PayoffVariant pv = European{K, hes.mu, hes.T};

    const long long n_ref = 500000;
    const auto ref = run_mc_fixed(mv, pv, 256, n_ref, 99u);
    const double ref_se = std::sqrt(ref.second / (double)n_ref);
    auto sim = [&](int ns, long long np, unsigned s) { return run_mc_fixed(mv, pv, ns, np, s).first; };
    const double c1 = estimar_c1_richardson(sim, hes.T, 4, 20000);
    int n_steps = pow2_at_least(std::max(1, (int)std::ceil(std::sqrt(2.0) * hes.T / (eps * std::abs(c1)))));
    n_steps = std::min(n_steps, 1 << 11);
    std::printf("c1 = %.4f, n_steps = %d\n", c1, n_steps);

    MCConfig mc_cfg;
    MLMCConfig ml_cfg;
    ml_cfg.M = 2;
    ml_cfg.max_L = 10;
    QMCConfig qmc_cfg;
    std::vector<TableRow> rows;
    auto add = [&](const char* name, MCResult r) { rows.push_back({name, r.price, r.std_error, r.n_samples, r.time_s}); };
    add("MC", run_mc_cuda(mv, pv, eps, n_steps, mc_cfg));
    add("MLMC", run_mlmc_cuda(mv, pv, eps, ml_cfg));
    add("QMC Raw", run_qmc_cuda(mv, pv, eps, n_steps, qmc_cfg, NoiseMode::Raw));
    add("MLQMC Raw", run_mlqmc_cuda(mv, pv, eps, ml_cfg, qmc_cfg, NoiseMode::Raw));
    HestonParams mil = hes;
    mil.scheme = 1;
    add("MLMC Milstein", run_mlmc_cuda(ModelVariant{mil}, pv, eps, ml_cfg));
    print_table(rows, ref.first, eps, "Heston European (facade)");
    int bad = outside(rows, ref.first, ref_se, eps);

    // GBM Asian on 2048 dates: bridge and PCA noise at the example's step cap
    GBMParams gbm;
    const int n_long = 1 << 11;
    const double eps_a = 0.01;
    ModelVariant ma = gbm;
    PayoffVariant pa = Asian{K};
    const auto ref_a = run_mc_fixed(ma, pa, n_long, n_ref, 7u);
    const double ref_a_se = std::sqrt(ref_a.second / (double)n_ref);
    DeviceBBData* bb = bb_upload(bb_precompute(n_long, gbm.T));
    DevicePCAData* pca = pca_upload(pca_compute(n_long, gbm.T));
    std::vector<TableRow> rows_a;
    auto add_a = [&](const char* name, MCResult r) {
        rows_a.push_back({name, r.price, r.std_error, r.n_samples, r.time_s});
    };
    add_a("QMC BB 2048", run_qmc_cuda(ma, pa, eps_a, n_long, qmc_cfg, NoiseMode::BrownianBridge, bb));
    add_a("QMC PCA 2048", run_qmc_cuda(ma, pa, eps_a, n_long, qmc_cfg, NoiseMode::PCA, nullptr, pca));
    bb_free(bb);
    pca_free(pca);
    print_table(rows_a, ref_a.first, eps_a, "Asian GBM, 2048 dates (facade)");
    bad += outside(rows_a, ref_a.first, ref_a_se, eps_a);
    std::printf("facade rows outside tolerance: %d\n", bad);
    return bad ? 1 : 0;
}
