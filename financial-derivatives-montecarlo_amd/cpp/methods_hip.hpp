// methods_hip.hpp -- C++ facade of the MI355X engine with the reference's API
// (methods_cuda.cuh:10-134): same configuration structs, result struct, BB/PCA
// handles and run_* entry points, so that the reference's callers compile by
// including this header instead.  Everything runs through the C ABI
// (include/fdmc.h) on a per-thread engine context (device 0 unless
// set_device() is called first).  Failures throw std::runtime_error, as the
// reference does.  Trailing fields of the configs are additive.
#pragma once

#include <utility>
#include <vector>

#include "models.hpp"
#include "payoffs.hpp"
#include "utils.hpp"

struct MCConfig {
    long long N_batch = 1LL << 16;   // kept for API parity: one launch covers every path
    int pilot_n = 10000;
    unsigned seed = 123u;
};

struct QMCConfig {
    int R = 32;                      // replicas, each with its own Hong-Hickernell scramble
    int max_doublings = 20;
    long long N0 = 1LL << 14;        // additive: first points per replica (power of two)
    long long N_max = 0;             // additive: cap on points per replica (0 = 2^32)
    unsigned salt_base = 42u;        // additive: replica salt = salt_base + r (+ l R in MLQMC)
    int pca_precision = 0;           // additive: PCA GEMM operands 0 = FP16, 1 = BF16
};

struct MLMCConfig {
    int M = 2;
    int max_L = 10;
    int pilot_n = 400;
    unsigned seed = 42u;             // additive: Philox key of the MLMC levels
};

enum class NoiseMode { Raw, BrownianBridge, PCA };

struct MCResult {
    double price = 0.0;
    double std_error = 0.0;
    long long n_samples = 0;
    double time_s = 0.0;
};

// Additive: pathwise-adjoint Greeks with standard errors across replicas.
struct GreeksResult {
    double delta = 0.0, vega = 0.0, rho = 0.0;
    double delta_se = 0.0, vega_se = 0.0, rho_se = 0.0;
};

struct DeviceBBData;
struct DevicePCAData;

DeviceBBData* bb_upload(const BBData& bb);
void bb_free(DeviceBBData*);
DevicePCAData* pca_upload(const PCAData& pca);
void pca_free(DevicePCAData*);

MCResult run_mc_cuda(const ModelVariant& model, const PayoffVariant& payoff, double eps, int n_steps,
                     const MCConfig& cfg = {});
MCResult run_mlmc_cuda(const ModelVariant& model, const PayoffVariant& payoff, double eps,
                       const MLMCConfig& cfg = {});
MCResult run_qmc_cuda(const ModelVariant& model, const PayoffVariant& payoff, double eps, int n_steps,
                      const QMCConfig& cfg = {}, NoiseMode mode = NoiseMode::Raw, DeviceBBData* dev_bb = nullptr,
                      DevicePCAData* dev_pca = nullptr);
MCResult run_mlqmc_cuda(const ModelVariant& model, const PayoffVariant& payoff, double eps,
                        const MLMCConfig& ml_cfg = {}, const QMCConfig& qmc_cfg = {},
                        NoiseMode mode = NoiseMode::Raw, std::vector<DeviceBBData*> bb_list = {},
                        std::vector<DevicePCAData*> pca_list = {});

struct CVPilot { double beta, var_plain, var_cv; };
CVPilot cv_pilot(const ModelVariant& main_model, const ModelVariant& ctrl_model, const PayoffVariant& main_payoff,
                 const PayoffVariant& ctrl_payoff, double E_ctrl, int n_steps, int N_pilot, unsigned seed = 0);
MCResult run_mc_cv_cuda(const ModelVariant& main_model, const ModelVariant& ctrl_model,
                        const PayoffVariant& main_payoff, const PayoffVariant& ctrl_payoff, double E_ctrl,
                        double beta, double eps, int n_steps, const MCConfig& cfg = {});
MCResult run_qmc_cv_cuda(const ModelVariant& main_model, const ModelVariant& ctrl_model,
                         const PayoffVariant& main_payoff, const PayoffVariant& ctrl_payoff, double E_ctrl,
                         double beta, double eps, int n_steps, const QMCConfig& cfg = {},
                         NoiseMode mode = NoiseMode::Raw, DeviceBBData* dev_bb = nullptr);

// Importance-sampled GBM European call (drift shift z_star over the horizon).
MCResult run_is_cuda(const GBMParams& model, const European& payoff, double z_star, double eps,
                     const MCConfig& cfg = {});

std::pair<double, double> run_mc_fixed(const ModelVariant& model, const PayoffVariant& payoff, int n_steps,
                                       long long n_paths, unsigned seed);

// Additive: price + adjoint Greeks under the QMC doubling loop (optionally with a
// control leg whose expectation E_ctrl and its sensitivities dE_ctrl are known).
MCResult run_qmc_greeks_cuda(const ModelVariant& model, const PayoffVariant& payoff, double eps, int n_steps,
                             GreeksResult& greeks, const QMCConfig& cfg = {}, NoiseMode mode = NoiseMode::Raw,
                             DeviceBBData* dev_bb = nullptr);

// Additive: a book of contracts under one one-factor model (GBM, Dupire or the
// quadratic local vol), each priced by multilevel QMC with M = 2 and levels
// 0..log2(n_steps) on the Brownian bridge (run_mlqmc_cuda's estimator per
// contract, salt 42 + (10 c + l) R + r), with pathwise Greeks; every contract
// runs in the same device passes.  No counterpart in the reference.
struct BookContract {
    PayoffVariant payoff;   // European, Asian, Lookback or Barrier
    int n_steps;            // dates: a power of two, 1..512
    double T;               // maturity (<= 0: European/Barrier use their own T)
};
struct BookResult {
    MCResult price;
    GreeksResult greeks;    // NaN for barriers and with greeks = 0
    int iterations = 0;     // passes that added points
    bool converged = false;
};
enum class BookMode { Batched, Plain };
// greeks: 0 none, 1 pathwise adjoint (reverse sweep), 2 forward tangents.
std::vector<BookResult> run_book_cuda(const ModelVariant& model, const std::vector<BookContract>& contracts,
                                      double eps, int R = 16, BookMode mode = BookMode::Batched, int greeks = 1);

// Select the HIP device used by this thread's engine context (before the first call).
void set_device(int device);
