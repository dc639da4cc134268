// utils_check.cpp -- prints facade host-numerics results as JSON for
// tests/test_facade.py (CPU only).
#include <cstdio>
#include <vector>

#include "utils.hpp"

int main() {
    std::printf("{\"bs\": %.15g, \"geo\": %.15g", bs_call(100, 100, 1, 0.05, 0.2),
                geom_asian_analytic(100, 95, 1, 0.05, 0.3, 64));
    PCAData p = pca_compute(8, 1.0);
    std::printf(", \"pca8\": [");
    for (size_t i = 0; i < p.M_pca.size(); i++) std::printf("%s%.17g", i ? "," : "", p.M_pca[i]);
    std::vector<double> corr = {1.0, 0.3, 0.2, 0.3, 1.0, 0.4, 0.2, 0.4, 1.0};
    PCAData k = pca_kron_compute(4, 1.0, corr, 3);
    std::printf("], \"kron\": [");
    for (size_t i = 0; i < k.M_pca.size(); i++) std::printf("%s%.17g", i ? "," : "", k.M_pca[i]);
    BBData bb = bb_precompute(16, 2.0);
    std::printf("], \"bb_map\": [");
    for (int i = 0; i < 16; i++) std::printf("%s%d", i ? "," : "", bb.map_idx[i]);
    std::printf("], \"bb_sd\": [");
    for (int i = 0; i < 16; i++) std::printf("%s%.17g", i ? "," : "", bb.std_dev[i]);
    std::vector<double> A = {4, 1, 2, 1, 3, 0, 2, 0, 5}, w, V;
    eigh_symmetric(3, A, w, V);
    std::printf("], \"eig\": [%.17g,%.17g,%.17g]}\n", w[0], w[1], w[2]);
    return 0;
}
