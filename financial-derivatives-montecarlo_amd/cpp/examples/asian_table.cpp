// asian_table.cpp -- drop-in check of the C++ facade: the call sequence of the
// reference's Asian example (examples/ejemplo02_asian.cpp) -- fixed-step
// reference price, MC, MLMC, QMC Raw/BB/PCA and MLQMC Raw/BB/PCA through the
// reference's function names -- with the include switched to methods_hip.hpp,
// then the importance-sampled deep-OTM call of examples/ejemplo11_otm_is.cpp.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <stdexcept>
#include <vector>

#include "../methods_hip.hpp"

int main(int argc, char** argv) {
    GBMParams gbm;                                  // S0=100, mu=0.05, sigma=0.2, T=1
    const double K = 100.0;
    const double eps = argc > 1 ? std::atof(argv[1]) : 0.01;
    const int n_steps = 64, L_MAX = 10;
    ModelVariant mv = gbm;
    PayoffVariant pv = Asian{K};
    const double price_ref = run_mc_fixed(mv, pv, 64, 500000, 99u).first;

    DeviceBBData* dev_bb = bb_upload(bb_precompute(n_steps, gbm.T));
    DevicePCAData* dev_pca = pca_upload(pca_compute(n_steps, gbm.T));
    std::vector<DeviceBBData*> bb_list(L_MAX + 1);
    std::vector<DevicePCAData*> pca_list(L_MAX + 1);
    for (int l = 0; l <= L_MAX; ++l) {
        bb_list[l] = bb_upload(bb_precompute(1 << l, gbm.T));
        pca_list[l] = pca_upload(pca_compute(1 << l, gbm.T));
    }

    MCConfig mc_cfg;
    MLMCConfig ml_cfg;
    ml_cfg.M = 2;
    ml_cfg.max_L = L_MAX;
    QMCConfig qmc_cfg;
    qmc_cfg.N0 = 1 << 12;
    std::vector<TableRow> rows;
    auto add = [&](const char* name, MCResult r) { rows.push_back({name, r.price, r.std_error, r.n_samples, r.time_s}); };
    add("MC", run_mc_cuda(mv, pv, eps, n_steps, mc_cfg));
    add("MLMC", run_mlmc_cuda(mv, pv, eps, ml_cfg));
    add("QMC Raw", run_qmc_cuda(mv, pv, eps, n_steps, qmc_cfg, NoiseMode::Raw));
    add("QMC BB", run_qmc_cuda(mv, pv, eps, n_steps, qmc_cfg, NoiseMode::BrownianBridge, dev_bb));
    add("QMC PCA", run_qmc_cuda(mv, pv, eps, n_steps, qmc_cfg, NoiseMode::PCA, nullptr, dev_pca));
    add("MLQMC Raw", run_mlqmc_cuda(mv, pv, eps, ml_cfg, qmc_cfg, NoiseMode::Raw));
    add("MLQMC BB", run_mlqmc_cuda(mv, pv, eps, ml_cfg, qmc_cfg, NoiseMode::BrownianBridge, bb_list));
    add("MLQMC PCA", run_mlqmc_cuda(mv, pv, eps, ml_cfg, qmc_cfg, NoiseMode::PCA, {}, pca_list));
    print_table(rows, price_ref, eps, "Asian GBM (facade)");

    // Single-level rows estimate the 64-date price of the reference run (whose own
    // standard error is ~0.01); multilevel rows converge to the continuous-average
    // limit, ~0.06 lower, as in the reference's reports.
    int bad = 0;
    for (const auto& r : rows) {
        const bool multilevel = r.method.rfind("ML", 0) == 0;
        const double tol = multilevel ? 0.1 : 2.0 * eps + 3.0 * std::sqrt(r.std_error * r.std_error + 1e-4);
        bad += std::abs(r.price - price_ref) >= tol;
    }
    bb_free(dev_bb);
    pca_free(dev_pca);
    for (auto* b : bb_list) bb_free(b);
    for (auto* p : pca_list) pca_free(p);

    // deep out-of-the-money call by importance sampling (ejemplo11_otm_is.cpp:14-53)
    const double K_otm = 180.0, eps_is = 0.002;
    const double z_star = (std::log(K_otm / gbm.S0) - (gbm.mu - 0.5 * gbm.sigma * gbm.sigma) * gbm.T) /
                          (gbm.sigma * std::sqrt(gbm.T));
    const MCResult is = run_is_cuda(gbm, European{K_otm, gbm.mu, gbm.T}, z_star, eps_is);
    const double bs_otm = bs_call(gbm.S0, K_otm, gbm.T, gbm.mu, gbm.sigma);
    std::printf("IS (z*=%.2f): %.6f +- %.6f, N = %lld, Black-Scholes %.6f\n", z_star, is.price, is.std_error,
                is.n_samples, bs_otm);
    bad += std::abs(is.price - bs_otm) >= 3.0 * is.std_error + 1e-4;
    std::printf("facade rows outside tolerance: %d\n", bad);
    return bad ? 1 : 0;
}
