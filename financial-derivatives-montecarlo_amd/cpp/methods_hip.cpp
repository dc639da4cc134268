// methods_hip.cpp -- the reference's C++ API (methods_cuda.cuh) over the C ABI.
#include "methods_hip.hpp"

#include <memory>
#include <stdexcept>
#include <string>
#include <type_traits>

#include "../../include/fdmc.h"

struct DeviceBBData { fdmc_bb* h; int N; };
struct DevicePCAData { fdmc_pca* h; int m; };

namespace {

// Per-thread context; destroyed at thread exit.
struct Ctx {
    fdmc_context* c = nullptr;
    int device = 0;
    ~Ctx() {
        if (c) fdmc_context_destroy(c);
    }
};
thread_local Ctx g_ctx;

fdmc_context* ctx() {
    if (!g_ctx.c) {
        if (fdmc_context_create(g_ctx.device, &g_ctx.c) != FDMC_OK)
            throw std::runtime_error(std::string("fdmc: ") + fdmc_last_error());
    }
    return g_ctx.c;
}

void check(fdmc_status s) {
    if (s != FDMC_OK) throw std::runtime_error(std::string("fdmc: ") + fdmc_last_error());
}

// Model descriptor with storage for the per-asset arrays.
struct CModel {
    fdmc_model m{};
    std::vector<double> s0, sig, L;
};

CModel to_c(const ModelVariant& mv) {
    CModel r;
    std::visit([&](const auto& m) {
        using T = std::decay_t<decltype(m)>;
        r.m.mu = m.mu;
        r.m.T = m.T;
        if constexpr (std::is_same_v<T, GBMParams>) {
            r.m.kind = FDMC_MODEL_GBM; r.m.S0 = m.S0; r.m.sigma = m.sigma;
        } else if constexpr (std::is_same_v<T, HestonParams>) {
            r.m.kind = FDMC_MODEL_HESTON; r.m.S0 = m.S0; r.m.kappa = m.kappa; r.m.theta = m.theta;
            r.m.xi = m.xi; r.m.rho = m.rho; r.m.v0 = m.v0; r.m.heston_scheme = m.scheme;
            r.m.L[0] = m.L[0][0]; r.m.L[1] = m.L[0][1]; r.m.L[2] = m.L[1][0]; r.m.L[3] = m.L[1][1];
        } else if constexpr (std::is_same_v<T, DupireLocalParams>) {
            r.m.kind = FDMC_MODEL_DUPIRE; r.m.S0 = m.S0; r.m.sigma0 = m.sigma0; r.m.alpha = m.alpha;
            r.m.beta_d = m.beta_d;
        } else if constexpr (std::is_same_v<T, LocalVolQuadParams>) {
            r.m.kind = FDMC_MODEL_LOCALVOL_QUAD; r.m.S0 = m.S0; r.m.lv_a = m.a; r.m.lv_b = m.b;
            r.m.lv_c = m.c; r.m.lv_sref = m.s_ref;
        } else if constexpr (std::is_same_v<T, MultiDupireParams>) {
            r.m.kind = FDMC_MODEL_MULTI_DUPIRE; r.m.n_assets = m.n; r.m.sigma0 = m.sigma0; r.m.alpha = m.alpha;
            r.m.beta_d = m.beta_d; r.m.uncorrelated = m.uncorrelated ? 1 : 0;
            r.s0 = m.S0; r.L = m.L;
            r.m.S0 = r.s0.empty() ? 100.0 : r.s0[0];
        } else {
            r.m.kind = FDMC_MODEL_MULTI_GBM; r.m.n_assets = m.n; r.m.uncorrelated = m.uncorrelated ? 1 : 0;
            r.s0 = m.S0; r.sig = m.sigma; r.L = m.L;
            r.m.S0 = r.s0.empty() ? 100.0 : r.s0[0];
        }
    }, mv);
    r.m.S0_assets = r.s0.empty() ? nullptr : r.s0.data();
    r.m.sigma_assets = r.sig.empty() ? nullptr : r.sig.data();
    r.m.L_assets = r.L.empty() ? nullptr : r.L.data();
    return r;
}

struct CPayoff {
    fdmc_payoff p{};
    std::vector<double> w;
};

CPayoff to_c(const PayoffVariant& pv) {
    CPayoff r;
    std::visit([&](const auto& p) {
        using T = std::decay_t<decltype(p)>;
        if constexpr (std::is_same_v<T, European>) {
            r.p.kind = FDMC_PAYOFF_EUROPEAN; r.p.K = p.K; r.p.r = p.r; r.p.T = p.T;
        } else if constexpr (std::is_same_v<T, Asian>) {
            r.p.kind = FDMC_PAYOFF_ASIAN; r.p.K = p.K;
        } else if constexpr (std::is_same_v<T, GeomAsian>) {
            r.p.kind = FDMC_PAYOFF_GEOM_ASIAN; r.p.K = p.K;
        } else if constexpr (std::is_same_v<T, Lookback>) {
            r.p.kind = FDMC_PAYOFF_LOOKBACK; r.p.sigma = p.sigma;
        } else if constexpr (std::is_same_v<T, Barrier>) {
            r.p.kind = FDMC_PAYOFF_BARRIER; r.p.K = p.K; r.p.B = p.B; r.p.sigma = p.sigma; r.p.r = p.r;
            r.p.T = p.T;
        } else {
            r.p.kind = FDMC_PAYOFF_BASKET; r.p.K = p.K; r.p.r = p.r; r.p.T = p.T; r.p.n_assets = p.n_assets;
            r.w = p.weights;
        }
    }, pv);
    r.p.weights = r.w.empty() ? nullptr : r.w.data();
    return r;
}

fdmc_qmc_config to_c(const QMCConfig& q) {
    fdmc_qmc_config c{};
    c.R = q.R; c.max_doublings = q.max_doublings; c.N0 = q.N0; c.N_max = q.N_max; c.salt_base = q.salt_base;
    c.pca_precision = q.pca_precision;
    return c;
}
fdmc_mc_config to_c(const MCConfig& m) { return fdmc_mc_config{m.N_batch, m.pilot_n, m.seed}; }
fdmc_mlmc_config to_c(const MLMCConfig& m) { return fdmc_mlmc_config{m.M, m.max_L, m.pilot_n, m.seed}; }
int mode_c(NoiseMode m) {
    return m == NoiseMode::Raw ? FDMC_NOISE_RAW : m == NoiseMode::BrownianBridge ? FDMC_NOISE_BB : FDMC_NOISE_PCA;
}
MCResult res(const fdmc_result& r) { return MCResult{r.price, r.std_error, (long long)r.n_samples, r.time_s}; }

}  // namespace

void set_device(int device) {
    if (g_ctx.c) throw std::runtime_error("set_device: the engine context of this thread already exists");
    g_ctx.device = device;
}

DeviceBBData* bb_upload(const BBData& bb) {
    std::vector<int32_t> m(bb.map_idx.begin(), bb.map_idx.end()), l(bb.left_idx.begin(), bb.left_idx.end()),
        r(bb.right_idx.begin(), bb.right_idx.end());
    fdmc_bb* h = nullptr;
    check(fdmc_bb_upload(ctx(), bb.N, bb.T, m.data(), l.data(), r.data(), bb.weight_left.data(),
                         bb.weight_right.data(), bb.std_dev.data(), &h));
    return new DeviceBBData{h, bb.N};
}

void bb_free(DeviceBBData* d) {
    if (!d) return;
    fdmc_bb_free(d->h);
    delete d;
}

DevicePCAData* pca_upload(const PCAData& pca) {
    fdmc_pca* h = nullptr;
    check(fdmc_pca_upload(ctx(), pca.m, pca.M_pca.data(), &h));
    return new DevicePCAData{h, pca.m};
}

void pca_free(DevicePCAData* d) {
    if (!d) return;
    fdmc_pca_free(d->h);
    delete d;
}

MCResult run_mc_cuda(const ModelVariant& model, const PayoffVariant& payoff, double eps, int n_steps,
                     const MCConfig& cfg) {
    CModel m = to_c(model);
    CPayoff p = to_c(payoff);
    fdmc_mc_config c = to_c(cfg);
    fdmc_result r{};
    check(fdmc_run_mc(ctx(), &m.m, &p.p, eps, n_steps, &c, &r));
    return res(r);
}

MCResult run_mlmc_cuda(const ModelVariant& model, const PayoffVariant& payoff, double eps, const MLMCConfig& cfg) {
    CModel m = to_c(model);
    CPayoff p = to_c(payoff);
    fdmc_mlmc_config c = to_c(cfg);
    fdmc_result r{};
    check(fdmc_run_mlmc(ctx(), &m.m, &p.p, eps, &c, &r));
    return res(r);
}

MCResult run_qmc_cuda(const ModelVariant& model, const PayoffVariant& payoff, double eps, int n_steps,
                      const QMCConfig& cfg, NoiseMode mode, DeviceBBData* dev_bb, DevicePCAData* dev_pca) {
    CModel m = to_c(model);
    CPayoff p = to_c(payoff);
    fdmc_qmc_config c = to_c(cfg);
    fdmc_result r{};
    check(fdmc_run_qmc(ctx(), &m.m, &p.p, eps, n_steps, &c, mode_c(mode), dev_bb ? dev_bb->h : nullptr,
                       dev_pca ? dev_pca->h : nullptr, &r));
    return res(r);
}

MCResult run_mlqmc_cuda(const ModelVariant& model, const PayoffVariant& payoff, double eps,
                        const MLMCConfig& ml_cfg, const QMCConfig& qmc_cfg, NoiseMode mode,
                        std::vector<DeviceBBData*> bb_list, std::vector<DevicePCAData*> pca_list) {
    CModel m = to_c(model);
    CPayoff p = to_c(payoff);
    fdmc_mlmc_config a = to_c(ml_cfg);
    fdmc_qmc_config b = to_c(qmc_cfg);
    std::vector<const fdmc_bb*> bbs;
    for (auto* x : bb_list) bbs.push_back(x ? x->h : nullptr);
    std::vector<const fdmc_pca*> pcs;
    for (auto* x : pca_list) pcs.push_back(x ? x->h : nullptr);
    fdmc_result r{};
    check(fdmc_run_mlqmc(ctx(), &m.m, &p.p, eps, &a, &b, mode_c(mode), bbs.data(), (int32_t)bbs.size(),
                         pcs.data(), (int32_t)pcs.size(), &r));
    return res(r);
}

CVPilot cv_pilot(const ModelVariant& main_model, const ModelVariant& ctrl_model, const PayoffVariant& main_payoff,
                 const PayoffVariant& ctrl_payoff, double E_ctrl, int n_steps, int N_pilot, unsigned seed) {
    CModel a = to_c(main_model), b = to_c(ctrl_model);
    CPayoff c = to_c(main_payoff), d = to_c(ctrl_payoff);
    fdmc_cv_pilot_result o{};
    check(fdmc_cv_pilot(ctx(), &a.m, &b.m, &c.p, &d.p, E_ctrl, n_steps, N_pilot, seed, &o));
    return CVPilot{o.beta, o.var_plain, o.var_cv};
}

MCResult run_mc_cv_cuda(const ModelVariant& main_model, const ModelVariant& ctrl_model,
                        const PayoffVariant& main_payoff, const PayoffVariant& ctrl_payoff, double E_ctrl,
                        double beta, double eps, int n_steps, const MCConfig& cfg) {
    CModel a = to_c(main_model), b = to_c(ctrl_model);
    CPayoff c = to_c(main_payoff), d = to_c(ctrl_payoff);
    fdmc_mc_config k = to_c(cfg);
    fdmc_result r{};
    check(fdmc_run_mc_cv(ctx(), &a.m, &b.m, &c.p, &d.p, E_ctrl, beta, eps, n_steps, &k, &r));
    return res(r);
}

MCResult run_qmc_cv_cuda(const ModelVariant& main_model, const ModelVariant& ctrl_model,
                         const PayoffVariant& main_payoff, const PayoffVariant& ctrl_payoff, double E_ctrl,
                         double beta, double eps, int n_steps, const QMCConfig& cfg, NoiseMode mode,
                         DeviceBBData* dev_bb) {
    CModel a = to_c(main_model), b = to_c(ctrl_model);
    CPayoff c = to_c(main_payoff), d = to_c(ctrl_payoff);
    fdmc_qmc_config k = to_c(cfg);
    fdmc_result r{};
    check(fdmc_run_qmc_cv(ctx(), &a.m, &b.m, &c.p, &d.p, E_ctrl, beta, eps, n_steps, &k, mode_c(mode),
                          dev_bb ? dev_bb->h : nullptr, &r));
    return res(r);
}

MCResult run_is_cuda(const GBMParams& model, const European& payoff, double z_star, double eps,
                     const MCConfig& cfg) {
    CModel m = to_c(ModelVariant{model});
    CPayoff p = to_c(PayoffVariant{payoff});
    fdmc_mc_config c = to_c(cfg);
    fdmc_result r{};
    check(fdmc_run_is(ctx(), &m.m, &p.p, z_star, eps, &c, &r));
    return res(r);
}

std::pair<double, double> run_mc_fixed(const ModelVariant& model, const PayoffVariant& payoff, int n_steps,
                                       long long n_paths, unsigned seed) {
    CModel m = to_c(model);
    CPayoff p = to_c(payoff);
    double mean = 0.0, var = 0.0;
    check(fdmc_run_mc_fixed(ctx(), &m.m, &p.p, n_steps, n_paths, seed, &mean, &var));
    return {mean, var};
}

MCResult run_qmc_greeks_cuda(const ModelVariant& model, const PayoffVariant& payoff, double eps, int n_steps,
                             GreeksResult& greeks, const QMCConfig& cfg, NoiseMode mode, DeviceBBData* dev_bb) {
    CModel m = to_c(model);
    CPayoff p = to_c(payoff);
    fdmc_qmc_config c = to_c(cfg);
    fdmc_result r{};
    fdmc_greeks g{};
    check(fdmc_run_qmc_greeks(ctx(), &m.m, &p.p, nullptr, nullptr, 0.0, nullptr, 0.0, eps, n_steps, &c,
                              mode_c(mode), dev_bb ? dev_bb->h : nullptr, &r, &g));
    greeks = GreeksResult{g.delta, g.vega, g.rho, g.delta_se, g.vega_se, g.rho_se};
    return res(r);
}

std::vector<BookResult> run_book_cuda(const ModelVariant& model, const std::vector<BookContract>& contracts,
                                      double eps, int R, BookMode mode, int greeks) {
    CModel m = to_c(model);
    std::vector<fdmc_book_contract> k;
    k.reserve(contracts.size());
    for (const BookContract& c : contracts) {
        const CPayoff p = to_c(c.payoff);
        fdmc_book_contract x{};
        x.kind = p.p.kind;
        x.n_steps = c.n_steps;
        x.K = p.p.K;
        x.B = p.p.B;
        x.T = c.T > 0.0 ? c.T : p.p.T;
        x.r = p.p.r;
        x.sigma_bgk = p.p.sigma;
        k.push_back(x);
    }
    fdmc_book_config cfg{R, 400, 42u, greeks, 50, 4, -1, 0};
    fdmc_book* b = nullptr;
    check(fdmc_book_create(ctx(), &m.m, k.data(), (int32_t)k.size(), &cfg, &b));
    std::vector<fdmc_book_result> r(k.size());
    fdmc_book_stats st{};
    const fdmc_status s = fdmc_book_run(b, eps, mode == BookMode::Plain ? FDMC_BOOK_PLAIN : FDMC_BOOK_BATCHED,
                                        r.data(), &st);
    const std::string err = s == FDMC_OK ? std::string() : std::string(fdmc_last_error());
    fdmc_book_free(b);
    if (s != FDMC_OK) throw std::runtime_error(err);
    std::vector<BookResult> out(k.size());
    for (size_t i = 0; i < k.size(); i++) {
        out[i].price = MCResult{r[i].price, r[i].std_error, (long long)r[i].n_samples, st.time_s};
        out[i].greeks = GreeksResult{r[i].delta, r[i].vega, r[i].rho, r[i].delta_se, r[i].vega_se, r[i].rho_se};
        out[i].iterations = r[i].iterations;
        out[i].converged = (r[i].flags & FDMC_FLAG_NOT_CONVERGED) == 0;
    }
    return out;
}
