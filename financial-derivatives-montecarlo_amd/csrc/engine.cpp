// engine.cpp -- host side of the engine: context, validation, estimator loops
// and the C ABI of include/fdmc.h.
//
// The estimator loops keep the reference's control logic (MC pilot + N rule,
// QMC doubling, Giles MLMC/MLQMC allocation, control variates;
// methods_cuda.cu:1173-2278) but drive fused gfx950 kernels: one launch per
// pass covers every replica, noise is generated in registers, sums are reduced
// deterministically on the device, and each pass costs one small D2H copy.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <map>
#include <stdexcept>
#include <string>
#include <tuple>
#include <vector>

#include "engine_host.hpp"

#include "joe_kuo6_table.inc"

using namespace fdmc;

using namespace fdmc_host;

thread_local std::string fdmc_host::g_last_error;

namespace {
constexpr int D_MAX_SOBOL = 20000;  // models.hpp:9
constexpr uint32_t TAG_QMC_FALLBACK = 0x51u, TAG_MLMC = 0x100u, TAG_CV_PILOT = 0x200u,
                   TAG_CV_MAIN = 0x201u, TAG_IS = 0x300u;
}  // namespace

// Device memory of fdmc_qmc_accumulate_noalloc: FDMC_NOALLOC_SCRATCH_BYTES per
// device, held from the load of the code object until the process ends.  Host
// code never uses the array's name as an address: fdmc_context_create looks up
// this device's address with hipGetSymbolAddress and keeps it in the context.
static __device__ __align__(256) unsigned char scratch[FDMC_NOALLOC_SCRATCH_BYTES];

struct fdmc_bb {
    fdmc_context* ctx;
    int N;
    double T;
    float* d_w;   // [N][4] {wl, wr, sd, 0}
    // uniform date grid: every FP32 weight is 1/2 and the sd depends only on
    // the depth (sd_depth[0] terminal, [k] midpoints of schedule level k - 1)
    int uniform;
    float sd_depth[BB_SD_DEPTHS];
};

struct fdmc_pca {
    fdmc_context* ctx;
    int m;
    int mp;                 // m rounded up to a multiple of 128 (zero padded on the device)
    void* d_f16;
    void* d_bf16;
    std::vector<double> M;  // column-major FP64 copy
};

// ---------------------------------------------------------------- direction numbers
static void expand_dirvectors(std::vector<uint32_t>& v) {
    v.assign((size_t)FDMC_JK_NDIM * 32, 0u);
    for (int d = 0; d < FDMC_JK_NDIM; d++) {
        uint32_t m[32];
        if (d == 0) {
            for (int j = 0; j < 32; j++) m[j] = 1u;
        } else {
            const uint32_t poly = fdmc_jk_poly[d];
            const int s = 31 - __builtin_clz(poly);
            const uint32_t a = (poly >> 1) & ((1u << (s - 1)) - 1u);
            const uint32_t* mi = fdmc_jk_m + fdmc_jk_off[d];
            for (int j = 0; j < s && j < 32; j++) m[j] = mi[j];
            for (int j = s; j < 32; j++) {
                uint32_t x = m[j - s] ^ (m[j - s] << s);
                for (int k = 1; k < s; k++)
                    if ((a >> (s - 1 - k)) & 1u) x ^= m[j - k] << k;
                m[j] = x;
            }
        }
        for (int j = 0; j < 32; j++) v[(size_t)d * 32 + j] = m[j] << (31 - j);
    }
}

// ---------------------------------------------------------------- parameters
namespace {

int noise_dim(const fdmc_model& m) {
    switch (m.kind) {
    case FDMC_MODEL_HESTON: return 2;
    case FDMC_MODEL_MULTI_DUPIRE:
    case FDMC_MODEL_MULTI_GBM: return m.n_assets;
    default: return 1;
    }
}
bool is_basket(const fdmc_model& m) {
    return m.kind == FDMC_MODEL_MULTI_DUPIRE || m.kind == FDMC_MODEL_MULTI_GBM;
}

void validate_model(const fdmc_model* m) {
    if (!m) fail(FDMC_E_INVALID, "model is null");
    if (m->kind < 0 || m->kind > FDMC_MODEL_MULTI_GBM) fail(FDMC_E_INVALID, "unknown model kind");
    if (!(m->T > 0.0)) fail(FDMC_E_INVALID, "model T must be positive");
    if (!is_basket(*m) && !(m->S0 > 0.0)) fail(FDMC_E_INVALID, "model S0 must be positive");
    if (m->kind == FDMC_MODEL_HESTON && m->heston_scheme != FDMC_HESTON_EXP_EULER &&
        m->heston_scheme != FDMC_HESTON_MILSTEIN)
        fail(FDMC_E_INVALID, "unknown Heston variance scheme");
    if (is_basket(*m)) {
        if (m->n_assets <= 0) fail(FDMC_E_INVALID, "basket n_assets must be positive");
        if (!m->S0_assets) fail(FDMC_E_INVALID, "basket S0_assets is null");
        if (m->kind == FDMC_MODEL_MULTI_GBM && !m->sigma_assets)
            fail(FDMC_E_INVALID, "MULTI_GBM sigma_assets is null");
    }
}
void validate_payoff(const fdmc_payoff* p) {
    if (!p) fail(FDMC_E_INVALID, "payoff is null");
    if (p->kind < 0 || p->kind > FDMC_PAYOFF_BASKET) fail(FDMC_E_INVALID, "unknown payoff kind");
}

}  // namespace

// make_params (methods_cuda.cu:986-1051) for one leg.
LegParams make_leg(const fdmc_model& m, const fdmc_payoff& p) {
    LegParams L{};
    L.mu = (float)m.mu;
    L.S0 = (float)m.S0;
    switch (p.kind) {
    case FDMC_PAYOFF_EUROPEAN:
    case FDMC_PAYOFF_BASKET:
        L.K = (float)p.K; L.discount = (float)std::exp(-p.r * p.T); break;
    case FDMC_PAYOFF_BARRIER:
        L.K = (float)p.K; L.B = (float)p.B; L.discount = (float)std::exp(-p.r * p.T); break;
    case FDMC_PAYOFF_ASIAN:
    case FDMC_PAYOFF_GEOM_ASIAN:
        L.K = (float)p.K; L.discount = 1.0f; break;
    default:
        L.discount = 1.0f; break;
    }
    switch (m.kind) {
    case FDMC_MODEL_GBM: L.sigma = (float)m.sigma; break;
    case FDMC_MODEL_HESTON:
        L.kappa = (float)m.kappa; L.theta = (float)m.theta; L.xi = (float)m.xi; L.v0 = (float)m.v0;
        L.L10 = (float)m.L[2]; L.L11 = (float)m.L[3];
        L.mil_q = m.heston_scheme == FDMC_HESTON_MILSTEIN ? 0.25f * L.xi * L.xi : 0.0f; break;
    case FDMC_MODEL_DUPIRE:
        L.sigma0 = (float)m.sigma0; L.alpha = (float)m.alpha;
        L.beta_m1 = (float)m.beta_d - 1.0f; L.inv_S0 = 1.0f / (float)m.S0;
        L.sigma = (float)m.sigma0; break;
    case FDMC_MODEL_LOCALVOL_QUAD:
        L.lv_a = (float)m.lv_a; L.lv_b = (float)m.lv_b; L.lv_c = (float)m.lv_c;
        L.inv_sref = 1.0f / (float)m.lv_sref; break;
    default: break;
    }
    return L;
}

float bgk_factor(const fdmc_payoff& p, float h) {
    if (p.kind != FDMC_PAYOFF_LOOKBACK && p.kind != FDMC_PAYOFF_BARRIER) return 1.0f;
    return std::exp(-0.5826f * (float)p.sigma * std::sqrt(h));  // payoffs.hpp:8, :332-337
}

int model_kind_dev(const fdmc_model& m) {
    switch (m.kind) {
    case FDMC_MODEL_GBM: return MK_GBM;
    case FDMC_MODEL_HESTON: return m.heston_scheme == FDMC_HESTON_MILSTEIN ? MK_HESTON_MIL : MK_HESTON;
    case FDMC_MODEL_DUPIRE: return MK_DUPIRE;
    case FDMC_MODEL_LOCALVOL_QUAD: return MK_LVQ;
    case FDMC_MODEL_MULTI_DUPIRE: return MK_MULTI_DUPIRE;
    default: return MK_MULTI_GBM;
    }
}

namespace {
// Single-asset models price a Basket payoff as a European on S_T, as the
// reference's MC_TABLE does (methods_cuda.cu:345-347).
int payoff_kind_dev(const fdmc_payoff& p) {
    return p.kind == FDMC_PAYOFF_BASKET ? PK_EUROPEAN : p.kind;
}

bool same_model(const fdmc_model& a, const fdmc_model& b) {
    return a.kind == b.kind && a.S0 == b.S0 && a.mu == b.mu && a.T == b.T && a.sigma == b.sigma &&
           a.sigma0 == b.sigma0 && a.alpha == b.alpha && a.beta_d == b.beta_d && a.lv_a == b.lv_a &&
           a.lv_b == b.lv_b && a.lv_c == b.lv_c && a.lv_sref == b.lv_sref && a.kappa == b.kappa &&
           a.theta == b.theta && a.xi == b.xi && a.v0 == b.v0 &&
           (a.kind != FDMC_MODEL_HESTON || a.heston_scheme == b.heston_scheme);
}

// ---------------------------------------------------------------- device helpers
void ensure_host_sums(fdmc_context* c, size_t n) {
    if (n <= c->h_sums_cap) return;
    if (c->h_sums) (void)hipHostFree(c->h_sums);
    c->h_sums = nullptr;
    c->h_sums_cap = 0;
    HIPC(hipHostMalloc((void**)&c->h_sums, n * sizeof(double), hipHostMallocDefault));
    c->h_sums_cap = n;
}

// Scrambled direction records for the given salts and D (cached per context).
const uint32_t* get_sdir(fdmc_context* c, const std::vector<uint32_t>& salts, int D) {
    if (D == c->sdir_D && salts == c->sdir_key) return (const uint32_t*)c->sdir.p;
    const size_t n = salts.size() * (size_t)D * SDIR_WORDS;
    c->sdir.ensure(n * sizeof(uint32_t));
    c->salts.ensure(salts.size() * sizeof(uint32_t));
    HIPC(hipMemcpyAsync(c->salts.p, salts.data(), salts.size() * sizeof(uint32_t), hipMemcpyHostToDevice,
                        c->stream));
    HIPC(launch_build_sdir(c->d_dirv, (const uint32_t*)c->salts.p, (int)salts.size(), D,
                           (uint32_t*)c->sdir.p, c->stream));
    c->sdir_key = salts;
    c->sdir_D = D;
    return (const uint32_t*)c->sdir.p;
}

// ---------------------------------------------------------------- one fused pass
// Outputs of a pass that may allocate nothing (fdmc_qmc_accumulate_noalloc):
// direction records and partials in the static scratch array, sums reduced
// into the caller's device buffer.
struct NoAlloc {
    unsigned char* scratch = nullptr;   // device address of the scratch array
    size_t cap = 0;
    double* d_sums = nullptr;           // caller's device buffer, R x n_acc doubles
};

struct PassSpec {
    const fdmc_model* model = nullptr;
    const fdmc_payoff* payoff = nullptr;
    const fdmc_model* cmodel = nullptr;     // control leg or null
    const fdmc_payoff* cpayoff = nullptr;
    double E_ctrl = 0.0, beta = 0.0;
    int n_steps = 0;
    int R = 1;
    uint32_t salt_base = 0;       // QMC: replica salt = salt_base + r; PRNG: seed = salt_base + r
    int64_t first = 0, count = 0;
    int mode = FDMC_NOISE_RAW;
    bool prng = false;            // pseudo-random noise (MC, CV, QMC fallback)
    uint32_t tag = 0;
    const fdmc_bb* bb = nullptr;
    const fdmc_pca* pca = nullptr;
    int pca_bf16 = 0;             // PCA GEMM operands: 0 FP16 (as the reference), 1 BF16
    bool greeks = false;          // price + adjoint Greeks kernel (k_greeks)
};

int run_basket_pass(fdmc_context* c, const PassSpec& s, std::vector<double>& out);
int pca_rows(fdmc_context* c, const fdmc_pca* pca, const uint32_t* sdir, int D, int R, int64_t first,
             int64_t count, bool bf16);

// Runs one pass; writes R x nacc sums into out (into na->d_sums when na is
// given, out left untouched), returns nacc.
int run_pass(fdmc_context* c, const PassSpec& s, std::vector<double>& out, const NoAlloc* na = nullptr) {
    const fdmc_model& m = *s.model;
    const fdmc_payoff& p = *s.payoff;
    if (s.n_steps <= 0) fail(FDMC_E_INVALID, "n_steps must be positive");
    if (s.R <= 0 || s.R > 65535) fail(FDMC_E_INVALID, "R out of range");
    if (s.count <= 0) fail(FDMC_E_INVALID, "number of points must be positive");
    if (na && (is_basket(m) || s.prng || s.greeks || (s.mode != FDMC_NOISE_RAW && s.mode != FDMC_NOISE_BB)))
        fail(FDMC_E_UNSUPPORTED, "the no-allocation pass takes one-factor models with Raw or BB Sobol noise");
    if (is_basket(m)) return run_basket_pass(c, s, out);
    const int nf = noise_dim(m);
    const int64_t D64 = (int64_t)s.n_steps * nf;
    const bool use_prng = s.prng || D64 > D_MAX_SOBOL;
    if (!use_prng && s.first + s.count > (int64_t)4294967296LL)
        fail(FDMC_E_RANGE, "Sobol point index beyond 32 bits (offset + n > 2^32)");
    const int D = (int)D64;

    PathLaunch a{};
    const float T = (float)m.T;
    a.h = T / (float)s.n_steps;
    a.sqrt_h = std::sqrt(a.h);
    a.em = std::exp(-(float)m.kappa * a.h);
    a.leg[0] = make_leg(m, p);
    a.bgk[0] = bgk_factor(p, a.h);
    a.bgk[1] = 1.0f;
    int cv = CV_NONE, cpk = 0;
    if (s.cmodel && s.cpayoff) {
        a.leg[1] = make_leg(*s.cmodel, *s.cpayoff);
        a.bgk[1] = bgk_factor(*s.cpayoff, a.h);
        cpk = payoff_kind_dev(*s.cpayoff);
        if (same_model(m, *s.cmodel)) cv = CV_SHARED;
        else if (s.cmodel->kind == FDMC_MODEL_GBM) cv = CV_GBM;
        else fail(FDMC_E_UNSUPPORTED, "control leg must use the main model or a GBM model");
        a.beta_cv = (float)s.beta;
        a.E_ctrl = (float)s.E_ctrl;
    }
    a.n_steps = s.n_steps;
    a.nf = nf;
    a.D = D;
    a.T = T;
    const int mk = model_kind_dev(m), pk = payoff_kind_dev(p);
    if (s.greeks) {
        if (mk != MK_GBM && mk != MK_DUPIRE && mk != MK_LVQ)
            fail(FDMC_E_UNSUPPORTED, "adjoint Greeks need a one-factor GBM, Dupire or local-vol model");
        if (pk == PK_BARRIER)
            fail(FDMC_E_UNSUPPORTED, "adjoint Greeks need a Lipschitz payoff (the barrier indicator is not)");
    }

    int nz, lb = 0;
    const void* fn = nullptr;
    if (use_prng) {
        if (na) fail(FDMC_E_UNSUPPORTED, "the no-allocation pass needs Sobol points (dimension <= 20000)");
        if (s.mode != FDMC_NOISE_RAW && !s.prng)
            fail(FDMC_E_UNSUPPORTED, "BB/PCA need Sobol points (dimension <= 20000)");
        nz = NZ_PRNG;
        a.seed = s.salt_base;
        a.tag = s.prng ? s.tag : TAG_QMC_FALLBACK;
        fn = s.greeks ? greeks_kernel_prng(mk, pk, cv, cpk, 0) : path_kernel_prng(mk, pk, cv, cpk, 0);
    } else if (s.mode == FDMC_NOISE_BB) {
        if (nf != 1) fail(FDMC_E_UNSUPPORTED, "Brownian bridge needs a one-factor model");
        if (!s.bb) fail(FDMC_E_INVALID, "BB mode needs a DeviceBBData");
        if (s.bb->N != s.n_steps) fail(FDMC_E_INVALID, "BB data has N != n_steps");
        if (s.greeks && s.n_steps > SOBOL_LDS_MAX_D)
            fail(FDMC_E_UNSUPPORTED, "adjoint Greeks with a Brownian bridge support at most 1024 steps");
        nz = s.n_steps > SOBOL_LDS_MAX_D ? NZ_SOBOL_BB_GLOBAL : NZ_SOBOL_BB;
        lb = s.n_steps >= 16 ? 4 : 0;
        a.G = s.n_steps >> lb;
        a.bbw = s.bb->d_w;
        a.bb_uniform = s.bb->uniform;
        for (int k = 0; k < BB_SD_DEPTHS; k++) a.bb_sd[k] = s.bb->sd_depth[k];
        fn = s.greeks ? greeks_kernel_bb(mk, pk, cv, cpk, lb)
                      : path_kernel_bb(mk, pk, cv, cpk, lb, nz == NZ_SOBOL_BB_GLOBAL);
    } else if (s.mode == FDMC_NOISE_PCA) {
        if (nf != 1) fail(FDMC_E_UNSUPPORTED, "PCA noise needs a one-factor model (methods_cuda.cu:1702-1704)");
        if (s.greeks) fail(FDMC_E_UNSUPPORTED, "adjoint Greeks are available with Raw and BB noise");
        if (!s.pca || s.pca->m != D) fail(FDMC_E_INVALID, "PCA data must have m == n_steps * noise dimension");
        nz = NZ_ROWS;
        fn = path_kernel_rows(mk, pk, cv, cpk, 0);
    } else {
        nz = NZ_SOBOL_RAW;
        fn = s.greeks ? greeks_kernel_raw(mk, pk, cv, cpk, 0) : path_kernel_raw(mk, pk, cv, cpk, 0);
    }
    if (!fn) fail(FDMC_E_UNSUPPORTED, "model/payoff/control combination not supported");

    // scratch layout of a no-allocation pass: direction records, then partials
    const size_t na_sdir_bytes = ((size_t)s.R * D * SDIR_WORDS * sizeof(uint32_t) + 255) / 256 * 256;
    if (na) {
        if (na_sdir_bytes > na->cap)
            fail(FDMC_E_RANGE, "R x D too large for the no-allocation scratch (R * D * 256 B must fit in " +
                                   std::to_string(na->cap >> 20) + " MiB)");
        uint32_t* rec = (uint32_t*)na->scratch;
        HIPC(launch_build_sdir(c->d_dirv, nullptr, s.R, D, rec, c->stream, s.salt_base));
        a.sdir = rec;
    } else if (nz != NZ_PRNG) {
        std::vector<uint32_t> salts((size_t)s.R);
        for (int r = 0; r < s.R; r++) salts[r] = s.salt_base + (uint32_t)r;
        a.sdir = get_sdir(c, salts, D);
    }
    if (nz == NZ_ROWS) {
        const int64_t bytes = (int64_t)s.R * s.count * ((D + 127) / 128 * 128) * 6;
        if (bytes > (int64_t)16 << 30) fail(FDMC_E_RANGE, "PCA pass above 16 GiB of buffers: lower N_max or R");
        a.ld = pca_rows(c, s.pca, a.sdir, D, s.R, s.first, s.count, s.pca_bf16 != 0);
        a.rows = (const float*)c->pca_dw.p;
    }

    // Tiles of 256 points in absolute index space; a few tiles per workgroup.
    const int64_t t_first = s.first >> 8, t_last = (s.first + s.count - 1) >> 8;
    const int64_t ntiles = t_last - t_first + 1;
    const int64_t target_wgs = (int64_t)c->n_cu * 32;
    int64_t per_rep = std::max<int64_t>(1, target_wgs / s.R);
    int64_t tpw = std::max<int64_t>(1, (ntiles + per_rep - 1) / per_rep);
    if (tpw > (1 << 20)) tpw = 1 << 20;
    const int64_t nwg = (ntiles + tpw - 1) / tpw;
    if (nwg > 0x7FFFFFFF) fail(FDMC_E_RANGE, "too many workgroups");
    a.first = s.first;
    a.count = s.count;
    a.tiles_per_wg = (int)tpw;
    a.nwg_x = (int)nwg;
    const int nacc = ((cv == CV_NONE) ? 2 : 7) + (s.greeks ? 6 : 0);
    double* d_sums;
    if (na) {
        const size_t pbytes = (size_t)s.R * nwg * nacc * sizeof(double);
        if (na_sdir_bytes + pbytes > na->cap)
            fail(FDMC_E_RANGE, "pass too large for the no-allocation scratch (" + std::to_string(na->cap >> 20) +
                                   " MiB)");
        a.partials = (double*)(na->scratch + na_sdir_bytes);
        d_sums = na->d_sums;
    } else {
        a.partials = (double*)c->partials.ensure((size_t)s.R * nwg * nacc * sizeof(double));
        d_sums = (double*)c->sums.ensure((size_t)s.R * nacc * sizeof(double));
    }

    const size_t lds = s.greeks ? greeks_lds_bytes(nz, D, a.G, s.n_steps) : path_lds_bytes(nz, D, a.G);
    if (lds > LDS_DYNAMIC_MAX) fail(FDMC_E_UNSUPPORTED, "LDS footprint of this pass exceeds 160 KiB");
    if (lds > 65536) HIPC(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    void* args[] = {(void*)&a};
    const bool time_path = c->timing && nz != NZ_ROWS;   // PCA passes time their GEMM
    if (time_path) HIPC(hipEventRecord(c->ev0, c->stream));
    HIPC(hipLaunchKernel(fn, dim3((unsigned)nwg, (unsigned)s.R), dim3(256), args, lds, c->stream));
    if (time_path) HIPC(hipEventRecord(c->ev1, c->stream));
    HIPC(launch_reduce_partials(a.partials, s.R, (int)nwg, nacc, d_sums, c->stream));
    if (!na) {
        ensure_host_sums(c, (size_t)s.R * nacc);
        HIPC(hipMemcpyAsync(c->h_sums, d_sums, (size_t)s.R * nacc * sizeof(double), hipMemcpyDeviceToHost,
                            c->stream));
    }
    HIPC(hipStreamSynchronize(c->stream));
    if (c->timing) {
        float ms = 0.0f;
        HIPC(hipEventElapsedTime(&ms, c->ev0, c->ev1));
        c->timed_ms += ms;
        c->timed_launches++;
    }
    if (!na) out.assign(c->h_sums, c->h_sums + (size_t)s.R * nacc);
    return nacc;
}

// PCA noise for R replicas, points [first, first + count): Z (FP16/BF16,
// point-major) then dW = Z M^T on MFMA into c->pca_dw rows r * count + p.
// Returns the row stride (D rounded up to 128).
int pca_rows(fdmc_context* c, const fdmc_pca* pca, const uint32_t* sdir, int D, int R, int64_t first,
             int64_t count, bool bf16) {
    if (!pca || pca->m != D) fail(FDMC_E_INVALID, "PCA data must have m == n_steps * noise dimension");
    const int ld = pca->mp;
    const int64_t rows = (int64_t)R * count, mrows = (rows + 127) / 128 * 128;
    if (mrows > 0x7FFFFFFF) fail(FDMC_E_RANGE, "too many PCA rows in one pass");
    void* z = c->pca_z.ensure((size_t)mrows * ld * 2);
    float* dw = (float*)c->pca_dw.ensure((size_t)mrows * ld * sizeof(float));
    HIPC(launch_gen_z(sdir, D, ld, R, first, count, bf16, z, c->stream));
    if (c->timing) HIPC(hipEventRecord(c->ev0, c->stream));
    HIPC(launch_gemm_nt(z, bf16 ? pca->d_bf16 : pca->d_f16, dw, (int)mrows, ld, ld, bf16, c->stream));
    if (c->timing) HIPC(hipEventRecord(c->ev1, c->stream));
    return ld;
}

// PCA noise of a GBM basket for R replicas: Z as in pca_rows, then the GEMM
// with the product epilogue (ProdEpi) into c->pca_dw as [ld / 128][mrows][n]
// Euler-factor products.  Returns mrows, the rows per 128-column slab.
int64_t pca_prod(fdmc_context* c, const fdmc_pca* pca, const uint32_t* sdir, int D, int R, int64_t first,
                 int64_t count, bool bf16, const ProdEpi& epi) {
    if (!pca || pca->m != D) fail(FDMC_E_INVALID, "PCA data must have m == n_steps * noise dimension");
    const int ld = pca->mp;
    const int64_t rows = (int64_t)R * count, mrows = (rows + 255) / 256 * 256;
    if (mrows > 0x7FFFFFFF) fail(FDMC_E_RANGE, "too many PCA rows in one pass");
    void* z = c->pca_z.ensure((size_t)mrows * ld * 2);
    ProdEpi e = epi;
    e.P = (float*)c->pca_dw.ensure((size_t)(ld / 128) * mrows * e.n * sizeof(float));
    HIPC(launch_gen_z(sdir, D, ld, R, first, count, bf16, z, c->stream));
    if (c->timing) HIPC(hipEventRecord(c->ev0, c->stream));
    HIPC(launch_gemm_nt(z, bf16 ? pca->d_bf16 : pca->d_f16, nullptr, (int)mrows, ld, ld, bf16, c->stream, &e));
    if (c->timing) HIPC(hipEventRecord(c->ev1, c->stream));
    return mrows;
}

// Baskets of more than BASKET_MAX assets (k_basket_wide): raw Sobol or Philox
// normals, uncorrelated or Cholesky-correlated (n <= WIDE_CORR_MAX); prices only,
// so 2 accumulators per replica {sum Y, sum Y^2}.
int run_basket_wide(fdmc_context* c, const PassSpec& s, std::vector<double>& out) {
    const fdmc_model& m = *s.model;
    const fdmc_payoff& p = *s.payoff;
    const int n = m.n_assets;
    if (s.greeks) fail(FDMC_E_UNSUPPORTED, "basket deltas are available up to 32 assets");
    const int64_t D64 = (int64_t)s.n_steps * n;
    if (D64 > 0x7FFFFFFF) fail(FDMC_E_RANGE, "n_steps * n_assets too large");
    const bool prng = s.prng || D64 > D_MAX_SOBOL;
    if (!prng && s.mode != FDMC_NOISE_RAW)
        fail(FDMC_E_UNSUPPORTED, "baskets of more than 32 assets take Raw noise (as the reference's examples 07/08)");
    if (!prng && s.first + s.count > (int64_t)4294967296LL)
        fail(FDMC_E_RANGE, "Sobol point index beyond 32 bits (offset + n > 2^32)");
    const bool corr = !m.uncorrelated && m.L_assets;
    if (corr && n > WIDE_CORR_MAX) fail(FDMC_E_UNSUPPORTED, "correlated baskets: at most 256 assets");
    BasketWide b{};
    b.n = n;
    b.n_steps = s.n_steps;
    b.D = (int)D64;
    const float T = (float)m.T;
    b.h = T / (float)s.n_steps;
    b.sqrt_h = std::sqrt(b.h);
    b.mu = (float)m.mu;
    b.K = (float)p.K;
    b.discount = (float)std::exp(-p.r * p.T);
    b.sigma0 = (float)m.sigma0;
    b.alpha = (float)m.alpha;
    b.beta_m1 = (float)m.beta_d - 1.0f;
    b.inv_S0 = 1.0f / (float)m.S0_assets[0];
    // per-asset parameters and L in one device buffer: [S0 n][sig n][w n][L n x n]
    std::vector<float> par((size_t)3 * n + (corr ? (size_t)n * n : 0));
    for (int a = 0; a < n; a++) {
        par[a] = (float)m.S0_assets[a];
        par[n + a] = m.sigma_assets ? (float)m.sigma_assets[a] : 0.0f;
        par[2 * n + a] = p.weights ? (float)p.weights[a] : 1.0f / (float)n;
    }
    if (corr)
        for (size_t i = 0; i < (size_t)n * n; i++) par[3 * (size_t)n + i] = (float)m.L_assets[i];
    float* dpar = (float*)c->basket_L.ensure(par.size() * sizeof(float));
    HIPC(hipMemcpyAsync(dpar, par.data(), par.size() * sizeof(float), hipMemcpyHostToDevice, c->stream));
    b.S0 = dpar;
    b.sig = dpar + n;
    b.w = dpar + 2 * n;
    b.L = corr ? dpar + 3 * n : nullptr;
    const int mk = m.kind == FDMC_MODEL_MULTI_GBM ? MK_MULTI_GBM : MK_MULTI_DUPIRE;
    const int src = prng ? SRC_PRNG : SRC_SOBOL;
    if (prng) {
        b.seed = s.salt_base;
        b.tag = s.prng ? s.tag : TAG_QMC_FALLBACK;
    } else {
        std::vector<uint32_t> salts((size_t)s.R);
        for (int r = 0; r < s.R; r++) salts[r] = s.salt_base + (uint32_t)r;
        b.sdir = get_sdir(c, salts, b.D);
    }
    const void* fn = basket_wide_kernel(mk, src, corr, n);
    if (!fn) fail(FDMC_E_UNSUPPORTED, "wide basket kernel not available");
    const int wg = basket_wide_block(corr), tile = basket_wide_tile(corr);
    const int64_t t_first = s.first / tile, t_last = (s.first + s.count - 1) / tile;
    const int64_t ntiles = t_last - t_first + 1;
    // as many workgroups as are resident at once: residency depends on the
    // kernel's registers and LDS (the correlated kernel's on its block count:
    // three per CU for 100 assets), and all workgroups carry equal work, so a
    // second, partly filled round would leave most of the chip idle
    const size_t lds = basket_wide_lds_bytes(n, corr);
    if (lds > LDS_DYNAMIC_MAX) fail(FDMC_E_UNSUPPORTED, "wide basket LDS footprint too large");
    if (lds > 65536) HIPC(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    int resident = corr ? 4 : 16;
    {
        int o = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&o, fn, wg, lds) == hipSuccess && o > 0) resident = o;
    }
    const int64_t per_rep = std::max<int64_t>(1, (int64_t)c->n_cu * resident / s.R);
    const int64_t tpw = std::max<int64_t>(1, (ntiles + per_rep - 1) / per_rep);
    const int64_t nwg = (ntiles + tpw - 1) / tpw;
    if (nwg > 0x7FFFFFFF || tpw > 0x7FFFFFFF) fail(FDMC_E_RANGE, "too many workgroups");
    b.first = s.first;
    b.count = s.count;
    b.tiles_per_wg = (int)tpw;
    b.nwg_x = (int)nwg;
    b.partials = (double*)c->partials.ensure((size_t)s.R * nwg * 2 * sizeof(double));
    double* d_sums = (double*)c->sums.ensure((size_t)s.R * 2 * sizeof(double));
    void* args[] = {(void*)&b};
    HIPC(hipLaunchKernel(fn, dim3((unsigned)nwg, (unsigned)s.R), dim3(wg), args, lds, c->stream));
    HIPC(launch_reduce_partials(b.partials, s.R, (int)nwg, 2, d_sums, c->stream));
    ensure_host_sums(c, (size_t)s.R * 2);
    HIPC(hipMemcpyAsync(c->h_sums, d_sums, (size_t)s.R * 2 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPC(hipStreamSynchronize(c->stream));
    out.assign(c->h_sums, c->h_sums + (size_t)s.R * 2);
    return 2;
}

int run_basket_pass(fdmc_context* c, const PassSpec& s, std::vector<double>& out) {
    const fdmc_model& m = *s.model;
    const fdmc_payoff& p = *s.payoff;
    const int n = m.n_assets;
    if (p.kind != FDMC_PAYOFF_BASKET && p.kind != FDMC_PAYOFF_EUROPEAN)
        fail(FDMC_E_UNSUPPORTED, "basket models take a Basket (or European) payoff");
    if (s.cmodel) fail(FDMC_E_UNSUPPORTED, "control variates are not available for baskets");
    if (n > BASKET_MAX) return run_basket_wide(c, s, out);
    if (s.greeks && m.kind != FDMC_MODEL_MULTI_GBM)
        fail(FDMC_E_UNSUPPORTED, "basket deltas need GBM assets");
    int na = 4;
    while (na < n) na <<= 1;
    const int64_t D64 = (int64_t)s.n_steps * n;
    const bool prng = s.prng || D64 > D_MAX_SOBOL;
    if (!prng && s.first + s.count > (int64_t)4294967296LL)
        fail(FDMC_E_RANGE, "Sobol point index beyond 32 bits (offset + n > 2^32)");
    const int D = (int)D64;
    BasketLaunch b{};
    b.n = n;
    b.n_steps = s.n_steps;
    b.D = D;
    b.greeks = s.greeks ? 1 : 0;
    const float T = (float)m.T;
    b.h = T / (float)s.n_steps;
    b.sqrt_h = std::sqrt(b.h);
    b.mu = (float)m.mu;
    b.K = (float)p.K;
    b.discount = (float)std::exp(-p.r * p.T);
    b.sigma0 = (float)m.sigma0;
    b.alpha = (float)m.alpha;
    b.beta_m1 = (float)m.beta_d - 1.0f;
    b.inv_S0 = 1.0f / (float)m.S0_assets[0];
    for (int a = 0; a < n; a++) {
        b.S0[a] = (float)m.S0_assets[a];
        b.sig[a] = m.sigma_assets ? (float)m.sigma_assets[a] : 0.0f;
        b.w[a] = p.weights ? (float)p.weights[a] : 1.0f / (float)n;
    }
    const int mk = m.kind == FDMC_MODEL_MULTI_GBM ? MK_MULTI_GBM : MK_MULTI_DUPIRE;
    int src;
    if (prng) {
        if (n % 4) fail(FDMC_E_UNSUPPORTED, "pseudo-random baskets need a multiple of 4 assets");
        src = SRC_PRNG;
        b.seed = s.salt_base;
        b.tag = s.prng ? s.tag : TAG_QMC_FALLBACK;
    } else if (s.mode == FDMC_NOISE_PCA) {
        if (!s.pca || s.pca->m != D) fail(FDMC_E_INVALID, "PCA data must have m == n_steps * noise dimension");
        // GBM assets, a power-of-two count and the GEMM's 256-column tiling: the
        // Euler factors multiply inside the GEMM's epilogue (no dW in memory)
        const bool pow2 = n >= 2 && (n & (n - 1)) == 0;
        const bool tiled = s.pca && s.pca->mp % 256 == 0;
        if (m.kind == FDMC_MODEL_MULTI_GBM && pow2 && tiled) {
            src = SRC_PROD;
            na = n;
        } else {
            src = SRC_ROWS;
        }
    } else if (s.mode == FDMC_NOISE_RAW) {
        src = SRC_SOBOL;
    } else {
        fail(FDMC_E_UNSUPPORTED, "baskets support Raw and PCA noise (the bridge is one-factor)");
    }
    if (src != SRC_ROWS && !m.uncorrelated && m.L_assets) {
        std::vector<float> L((size_t)n * n);
        for (int i = 0; i < n * n; i++) L[i] = (float)m.L_assets[i];
        c->basket_L.ensure(L.size() * sizeof(float));
        HIPC(hipMemcpyAsync(c->basket_L.p, L.data(), L.size() * sizeof(float), hipMemcpyHostToDevice, c->stream));
        HIPC(hipStreamSynchronize(c->stream));
        b.L = (const float*)c->basket_L.p;
    }
    const void* fn = basket_kernel(mk, src, na);
    if (!fn) fail(FDMC_E_UNSUPPORTED, "basket kernel not available");
    const uint32_t* sdir = nullptr;
    if (src != SRC_PRNG) {
        std::vector<uint32_t> salts((size_t)s.R);
        for (int r = 0; r < s.R; r++) salts[r] = s.salt_base + (uint32_t)r;
        sdir = get_sdir(c, salts, D);
    }
    const int64_t t_first = s.first >> 8, t_last = (s.first + s.count - 1) >> 8;
    const int64_t ntiles = t_last - t_first + 1;
    const int64_t per_rep = std::max<int64_t>(1, (int64_t)c->n_cu * 32 / s.R);
    const int64_t tpw = std::max<int64_t>(1, (ntiles + per_rep - 1) / per_rep);
    const int64_t nwg = (ntiles + tpw - 1) / tpw;
    const int nacc = 2 + 2 * na;
    b.first = s.first;
    b.count = s.count;
    b.tiles_per_wg = (int)tpw;
    b.nwg_x = (int)nwg;
    double* part = (double*)c->partials.ensure((size_t)s.R * nwg * nacc * sizeof(double));
    double* d_sums = (double*)c->sums.ensure((size_t)s.R * nacc * sizeof(double));
    const size_t lds = basket_lds_bytes(src, D, na);
    if (lds > LDS_DYNAMIC_MAX) fail(FDMC_E_UNSUPPORTED, "basket LDS footprint too large");
    if (lds > 65536) HIPC(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    // replicas in groups so that the PCA buffers stay below ~8 GiB (dW rows)
    // or ~16 GiB (normals and slab products only)
    int group = s.R;
    if (src == SRC_ROWS) {
        const double per = (double)s.count * (D + 128) * 6.0;
        group = (int)std::max(1.0, std::min((double)s.R, 8.0 * (1 << 30) / per));
    } else if (src == SRC_PROD) {
        const double per = (double)s.count * s.pca->mp * (2.0 + 4.0 * n / 128.0);
        group = (int)std::max(1.0, std::min((double)s.R, 16.0 * (1 << 30) / per));
    }
    for (int r0 = 0; r0 < s.R; r0 += group) {
        const int g = std::min(group, s.R - r0);
        BasketLaunch bg = b;
        bg.partials = part + (size_t)r0 * nwg * nacc;
        if (src == SRC_ROWS) {
            bg.ld = pca_rows(c, s.pca, sdir + (size_t)r0 * D * SDIR_WORDS, D, g, s.first, s.count, s.pca_bf16 != 0);
            bg.dW = (const float*)c->pca_dw.p;
        } else if (src == SRC_PROD) {
            ProdEpi e{};
            e.c0 = 1.0f + b.mu * b.h;
            for (int a = 0; a < n; a++) e.sig[a] = b.sig[a];
            e.n = n;
            e.D = D;
            bg.prow = pca_prod(c, s.pca, sdir + (size_t)r0 * D * SDIR_WORDS, D, g, s.first, s.count,
                               s.pca_bf16 != 0, e);
            bg.npc = s.pca->mp / 128;
            bg.ld = s.pca->mp;
            bg.dW = (const float*)c->pca_dw.p;
        } else {
            bg.sdir = sdir ? sdir + (size_t)r0 * D * SDIR_WORDS : nullptr;
            bg.seed = b.seed + (uint32_t)r0;
        }
        void* args[] = {(void*)&bg};
        HIPC(hipLaunchKernel(fn, dim3((unsigned)nwg, (unsigned)g), dim3(256), args, lds, c->stream));
        if ((src == SRC_ROWS || src == SRC_PROD) && c->timing) {
            HIPC(hipStreamSynchronize(c->stream));
            float ms = 0.0f;
            HIPC(hipEventElapsedTime(&ms, c->ev0, c->ev1));
            c->timed_ms += ms;
            c->timed_launches++;
        }
    }
    HIPC(launch_reduce_partials(part, s.R, (int)nwg, nacc, d_sums, c->stream));
    ensure_host_sums(c, (size_t)s.R * nacc);
    HIPC(hipMemcpyAsync(c->h_sums, d_sums, (size_t)s.R * nacc * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPC(hipStreamSynchronize(c->stream));
    // compact to 2 + 2 n accumulators per replica
    const int nout = 2 + 2 * n;
    out.assign((size_t)s.R * nout, 0.0);
    for (int r = 0; r < s.R; r++)
        for (int i = 0; i < nout; i++) out[(size_t)r * nout + i] = c->h_sums[(size_t)r * nacc + i];
    return nout;
}

void mean_var(const std::vector<double>& v, double& m, double& var) {
    const int R = (int)v.size();
    double s = 0.0;
    for (double x : v) s += x;
    s /= R;
    double q = 0.0;
    for (double x : v) q += (x - s) * (x - s);
    m = s;
    var = (R > 1) ? q / (R - 1) : 0.0;
}

// Doubling loop of run_qmc_cuda / run_qmc_cv_cuda (methods_cuda.cu:1624-1665,
// 2203-2273): accumulate only the new points of each replica, stop when
// var_of_means < eps^2/2.
void qmc_loop(fdmc_context* c, PassSpec s, double eps, const fdmc_qmc_config& cfg, fdmc_result* out,
              fdmc_greeks* greeks = nullptr, const double* dE = nullptr) {
    const double t0 = now_s();
    if (cfg.R <= 0) fail(FDMC_E_INVALID, "QMCConfig.R must be positive");
    if (cfg.max_doublings <= 0) fail(FDMC_E_INVALID, "QMCConfig.max_doublings must be positive");
    if (!(eps > 0.0)) fail(FDMC_E_INVALID, "eps must be positive");
    int64_t N = cfg.N0 > 0 ? cfg.N0 : (1LL << 14);
    if (N & (N - 1)) fail(FDMC_E_INVALID, "QMCConfig.N0 must be a power of two");
    const int64_t cap = cfg.N_max > 0 ? cfg.N_max : (int64_t)4294967296LL;
    const bool ctrl = s.cmodel && s.cpayoff;
    const int col = ctrl ? 2 : 0;
    s.R = cfg.R;
    s.salt_base = cfg.salt_base;
    std::vector<double> tot, part, means((size_t)cfg.R);
    int64_t done = 0, total = 0;
    double vom = 1e30, gm = 0.0;
    int doublings = 0;
    bool conv = false;
    *out = fdmc_result{};
    const int D = s.n_steps * noise_dim(*s.model);
    // A PCA pass holds its normals and GEMM rows (6 bytes per coordinate and
    // replica) in HBM: the new points of a doubling go through in chunks of at
    // most 8 GiB of those buffers, whatever N grows to.
    int64_t chunk = INT64_MAX;
    if (s.mode == FDMC_NOISE_PCA) {
        const int64_t per_point = (int64_t)cfg.R * ((D + 127) / 128 * 128) * 6;
        chunk = std::max<int64_t>(256, (((int64_t)8 << 30) / per_point) & ~(int64_t)255);
    }
    for (int d = 0; d < cfg.max_doublings; d++) {
        if (N > (int64_t)4294967296LL && D <= D_MAX_SOBOL)
            fail(FDMC_E_RANGE, "N_per_replica exceeds the 32-bit Sobol offset");
        int nacc = 0;
        for (int64_t f = done; f < N; f += s.count) {
            s.first = f;
            s.count = std::min(chunk, N - f);
            nacc = run_pass(c, s, part);
            if (tot.empty()) tot.assign(part.size(), 0.0);
            for (size_t i = 0; i < part.size(); i++) tot[i] += part[i];
        }
        out->path_steps += (int64_t)cfg.R * (N - done) * s.n_steps;
        done = N;
        doublings++;
        for (int r = 0; r < cfg.R; r++) means[r] = tot[(size_t)r * nacc + col] / (double)N;
        double v;
        mean_var(means, gm, v);
        vom = v / cfg.R;
        total = (int64_t)cfg.R * N;
        if (vom < eps * eps / 2.0) { conv = true; break; }
        if (2 * N > cap) break;
        N *= 2;
    }
    if (greeks && !tot.empty()) {
        // per replica Greek means, then mean and standard error across replicas
        const int nacc = (int)(tot.size() / cfg.R), np = nacc - 6;
        double g[3], se[3];
        for (int k = 0; k < 3; k++) {
            for (int r = 0; r < cfg.R; r++) means[r] = tot[(size_t)r * nacc + np + 2 * k] / (double)done;
            double v;
            mean_var(means, g[k], v);
            se[k] = std::sqrt(v / cfg.R);
        }
        const double beta = s.beta;
        greeks->delta = g[0] + (dE ? beta * dE[0] : 0.0);
        greeks->vega = g[1] + (dE ? beta * dE[1] : 0.0);
        greeks->rho = g[2] + (dE ? beta * dE[2] : 0.0);
        greeks->delta_se = se[0];
        greeks->vega_se = se[1];
        greeks->rho_se = se[2];
    }
    out->price = gm;
    out->std_error = std::sqrt(vom);
    out->n_samples = total;
    out->levels = doublings;
    out->flags = (conv ? 0u : FDMC_FLAG_NOT_CONVERGED) | (D > D_MAX_SOBOL ? FDMC_FLAG_PRNG_FALLBACK : 0u);
    out->time_s = now_s() - t0;
}

// ---------------------------------------------------------------- multilevel
int level_steps(int M, int l) {
    int64_t n = 1;
    for (int i = 0; i < l; i++) {
        n *= M;
        if (n > (1LL << 30)) fail(FDMC_E_RANGE, "M^l exceeds 2^30 fine steps");
    }
    return (int)n;
}

// One (level, replica) request of a multilevel iteration.
struct LevelReq {
    int level = 0, replica = 0;
    int64_t first = 0, count = 0;
};

struct MultiSpec {
    const fdmc_model* model = nullptr;
    const fdmc_payoff* payoff = nullptr;
    int M = 2;
    bool qmc = false;                 // Sobol (MLQMC) or Philox (MLMC)
    int mode = FDMC_NOISE_RAW;        // MLQMC: RAW, BB or PCA
    int R = 1;                        // replicas per level (MLQMC)
    uint32_t salt_base = 42;          // MLQMC salt = salt_base + l R + r; MLMC: PRNG seed
    const fdmc_bb* const* bb_list = nullptr;
    int n_bb = 0;
    const fdmc_pca* const* pca_list = nullptr;
    int n_pca = 0;
    bool pca_bf16 = false;            // PCA GEMM operands: FP16 (as the reference) or BF16
};

// Scrambled records of all R replicas of level l (cached per context for the
// lifetime of one multilevel call through `cache`).
const uint32_t* level_sdir(fdmc_context* c, std::map<int, DevBuf>& cache, const MultiSpec& ms, int l, int D) {
    auto it = cache.find(l);
    if (it != cache.end()) return (const uint32_t*)it->second.p;
    DevBuf& b = cache[l];
    b.ensure((size_t)ms.R * D * SDIR_WORDS * sizeof(uint32_t));
    std::vector<uint32_t> salts((size_t)ms.R);
    for (int r = 0; r < ms.R; r++) salts[r] = ms.salt_base + (uint32_t)(l * ms.R + r);
    c->salts.ensure(salts.size() * sizeof(uint32_t));
    HIPC(hipMemcpyAsync(c->salts.p, salts.data(), salts.size() * sizeof(uint32_t), hipMemcpyHostToDevice,
                        c->stream));
    HIPC(launch_build_sdir(c->d_dirv, (const uint32_t*)c->salts.p, ms.R, D, (uint32_t*)b.p, c->stream));
    HIPC(hipStreamSynchronize(c->stream));  // the salts buffer is reused by the next level
    c->sdir_D = -1;
    return (const uint32_t*)b.p;
}

// One persistent multilevel launch over `ent` (any order; sorted here longest
// first); entry i's {sum dY, sum dY^2, sum Yf, sum Yf^2} go to out[which[i]].
void launch_levels(fdmc_context* c, const MultiSpec& ms, int nz, const std::vector<LevelEntry>& ent,
                   const std::vector<int>& which, std::vector<std::array<double, 4>>& out) {
    const fdmc_model& m = *ms.model;
    const fdmc_payoff& p = *ms.payoff;
    const int nf = noise_dim(m);
    const int mk = model_kind_dev(m), pk = payoff_kind_dev(p);
    std::vector<int> order(ent.size());
    for (size_t i = 0; i < order.size(); i++) order[i] = (int)i;
    std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return ent[x].D > ent[y].D; });
    std::vector<LevelEntry> sorted;
    std::vector<int> sw;
    int total = 0, max_D = 0, skel_n = 0;
    for (int o : order) {
        LevelEntry e = ent[o];
        const int64_t t0 = e.first >> 8, t1 = (e.first + e.count - 1) >> 8;
        const int64_t nt = t1 - t0 + 1;
        if (nt > (1LL << 30)) fail(FDMC_E_RANGE, "too many tiles in one multilevel launch");
        // about 512 noise dimensions of work per lane and item, and at least
        // 64 items per entry where it has that many tiles
        e.tpi = (int)std::max<int64_t>(1, std::min<int64_t>(512 / std::max(e.D, 1), nt / 64));
        const int64_t items = (nt + e.tpi - 1) / e.tpi;
        if (total + items > 0x7FFFFFFF) fail(FDMC_E_RANGE, "too many tiles in one multilevel launch");
        e.tile_begin = total;
        e.ntiles = (int)items;
        total += e.ntiles;
        max_D = std::max(max_D, e.D);
        if (nz == NZ_SOBOL_BB)   // skeleton (under 16 dates) or walk-stack entries per lane
            skel_n = std::max(skel_n, e.n_fine < 16 ? e.G + 1 : bb_walk_levels(e.G));
        sorted.push_back(e);
        sw.push_back(which[o]);
    }
    if ((nz == NZ_SOBOL_BB || nz == NZ_ROWS) && nf != 1) fail(FDMC_E_UNSUPPORTED, "BB/PCA need a one-factor model");
    const void* fn = nz == NZ_PRNG ? levels_kernel_prng(mk, pk)
                     : nz == NZ_SOBOL_BB ? levels_kernel_bb(mk, pk)
                     : nz == NZ_ROWS ? levels_kernel_rows(mk, pk) : levels_kernel_raw(mk, pk);
    if (!fn) fail(FDMC_E_UNSUPPORTED, "model/payoff combination not supported by the multilevel kernel");

    const size_t ent_bytes = sorted.size() * sizeof(LevelEntry);
    const size_t part_bytes = (size_t)total * 4 * sizeof(double);
    const size_t sums_bytes = sorted.size() * 4 * sizeof(double);
    char* pbuf = (char*)c->partials.ensure(part_bytes + ent_bytes + sums_bytes + 256);
    double* d_part = (double*)pbuf;
    LevelEntry* d_ent = (LevelEntry*)(pbuf + ((part_bytes + 63) & ~(size_t)63));
    double* d_out = (double*)((char*)d_ent + ((ent_bytes + 63) & ~(size_t)63));
    unsigned int* d_queue = (unsigned int*)c->sums.ensure(64);
    HIPC(hipMemcpyAsync(d_ent, sorted.data(), ent_bytes, hipMemcpyHostToDevice, c->stream));
    HIPC(launch_zero_u32(d_queue, 1, c->stream));

    LevelsLaunch a{};
    a.leg = make_leg(m, p);
    a.M = ms.M;
    a.nf = nf;
    a.n_entries = (int)sorted.size();
    a.total_items = total;
    a.entries = d_ent;
    a.queue = d_queue;
    a.partials = d_part;
    const size_t lds = levels_lds_bytes(nz, max_D, skel_n);
    if (lds > LDS_DYNAMIC_MAX) fail(FDMC_E_UNSUPPORTED, "multilevel LDS footprint too large");
    if (lds > 65536) HIPC(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    const int grid = (int)std::min<int64_t>(total, (int64_t)c->n_cu * 8);
    void* args[] = {(void*)&a, (void*)&max_D, (void*)&skel_n};
    if (c->timing) HIPC(hipEventRecord(c->ev0, c->stream));
    HIPC(hipLaunchKernel(fn, dim3(grid), dim3(256), args, lds, c->stream));
    if (c->timing) HIPC(hipEventRecord(c->ev1, c->stream));
    HIPC(launch_reduce_entries(d_part, d_ent, a.n_entries, 4, d_out, c->stream));
    ensure_host_sums(c, sorted.size() * 4);
    HIPC(hipMemcpyAsync(c->h_sums, d_out, sums_bytes, hipMemcpyDeviceToHost, c->stream));
    HIPC(hipStreamSynchronize(c->stream));
    if (c->timing) {
        float ms_ = 0.0f;
        HIPC(hipEventElapsedTime(&ms_, c->ev0, c->ev1));
        c->timed_ms += ms_;
        c->timed_launches++;
    }
    for (size_t i = 0; i < sorted.size(); i++)
        for (int k = 0; k < 4; k++) out[sw[i]][k] = c->h_sums[i * 4 + k];
}

// PCA requests (methods_cuda.cu:1907-1913): the normals of consecutive
// replicas of one level over the same point range go through one GEMM (FP16 or
// BF16 operands, FP32 accumulation) into the context's row buffer, then one
// multilevel launch reads the fine increments from those rows.
void run_levels_pca(fdmc_context* c, const MultiSpec& ms, const std::vector<LevelReq>& reqs,
                    const std::vector<LevelEntry>& ent, const std::vector<int>& which,
                    std::vector<std::array<double, 4>>& out) {
    std::vector<int> idx(ent.size());
    for (size_t i = 0; i < idx.size(); i++) idx[i] = (int)i;
    auto key = [&](int i) {
        const LevelReq& q = reqs[which[i]];
        return std::make_tuple(q.level, q.first, q.count, q.replica);
    };
    std::stable_sort(idx.begin(), idx.end(), [&](int x, int y) { return key(x) < key(y); });
    size_t i = 0;
    while (i < idx.size()) {
        const LevelEntry& e0 = ent[idx[i]];
        const LevelReq& q0 = reqs[which[idx[i]]];
        const fdmc_pca* pca = ms.pca_list[q0.level];
        // Z (2 B) and dW (4 B) per padded element: groups of at most 8 GiB
        const double per = (double)q0.count * pca->mp * 6.0;
        if (per > 8.0 * (1 << 30)) {
            // one request above 8 GiB (the reference batches such a level by free
            // memory, methods_cuda.cu:1791-1803): its point range runs in chunks of
            // a multiple of 256 points with at most 8 GiB of buffers each, whose
            // sums add up to the request's
            const int64_t chunk = std::max<int64_t>(
                256, (int64_t)(8.0 * (1 << 30) / ((double)pca->mp * 6.0)) / 256 * 256);
            const int w = which[idx[i]];
            std::array<double, 4> acc{0.0, 0.0, 0.0, 0.0};
            for (int64_t f = q0.first; f < q0.first + q0.count; f += chunk) {
                LevelEntry e = e0;
                e.first = f;
                e.count = std::min(chunk, q0.first + q0.count - f);
                e.ld = pca_rows(c, pca, e0.sdir, e0.D, 1, e.first, e.count, ms.pca_bf16);
                e.rows = (const float*)c->pca_dw.p;
                std::vector<LevelEntry> ge{e};
                std::vector<int> gw{w};
                launch_levels(c, ms, NZ_ROWS, ge, gw, out);
                for (int k = 0; k < 4; k++) acc[k] += out[w][k];
            }
            out[w] = acc;
            i++;
            continue;
        }
        const size_t gmax = (size_t)std::max(1.0, 8.0 * (1 << 30) / per);
        size_t j = i + 1;
        while (j < idx.size() && j - i < gmax) {
            const LevelReq& q = reqs[which[idx[j]]];
            if (q.level != q0.level || q.first != q0.first || q.count != q0.count ||
                q.replica != q0.replica + (int)(j - i))
                break;
            j++;
        }
        const int g = (int)(j - i);
        const int ld = pca_rows(c, pca, e0.sdir, e0.D, g, q0.first, q0.count, ms.pca_bf16);
        std::vector<LevelEntry> ge;
        std::vector<int> gw;
        for (size_t k = i; k < j; k++) {
            LevelEntry e = ent[idx[k]];
            e.rows = (const float*)c->pca_dw.p + (size_t)(k - i) * (size_t)q0.count * (size_t)ld;
            e.ld = ld;
            ge.push_back(e);
            gw.push_back(which[idx[k]]);
        }
        launch_levels(c, ms, NZ_ROWS, ge, gw, out);
        i = j;
    }
}

// Runs every request, one persistent launch per noise kind (PCA: per level
// and replica group); out[i] receives {sum dY, sum dY^2, sum Yf, sum Yf^2} of
// request i.
void run_levels(fdmc_context* c, const MultiSpec& ms, std::map<int, DevBuf>& cache,
                const std::vector<LevelReq>& reqs, std::vector<std::array<double, 4>>& out) {
    const fdmc_model& m = *ms.model;
    const fdmc_payoff& p = *ms.payoff;
    const int nf = noise_dim(m);
    out.assign(reqs.size(), {0.0, 0.0, 0.0, 0.0});
    // split by noise kind: Sobol (D <= 20000) and pseudo-random (MLMC or QMC fallback)
    for (int pass = 0; pass < 2; pass++) {
        std::vector<LevelEntry> ent;
        std::vector<int> which;
        for (size_t i = 0; i < reqs.size(); i++) {
            const LevelReq& q = reqs[i];
            const int nfine = level_steps(ms.M, q.level);
            const int64_t D64 = (int64_t)nfine * nf;
            const bool prng = !ms.qmc || D64 > D_MAX_SOBOL;
            if ((pass == 0) == prng) continue;
            if (q.count <= 0) continue;
            LevelEntry e{};
            e.level = q.level;
            e.n_fine = nfine;
            e.n_coarse = q.level == 0 ? 0 : nfine / ms.M;
            e.D = (int)D64;
            e.first = q.first;
            e.count = q.count;
            const double T = m.T;
            e.h_f = (float)(T / nfine);
            e.h_c = (float)(T / std::max(e.n_coarse, 1));
            e.sqrt_hf = std::sqrt(e.h_f);
            e.em_f = std::exp(-(float)m.kappa * e.h_f);
            e.em_c = std::exp(-(float)m.kappa * e.h_c);
            e.bgk_f = bgk_factor(p, e.h_f);
            e.bgk_c = bgk_factor(p, e.h_c);
            if (prng) {
                e.seed = ms.qmc ? ms.salt_base + (uint32_t)(q.level * ms.R + q.replica) : ms.salt_base;
                e.tag = ms.qmc ? TAG_QMC_FALLBACK : TAG_MLMC + (uint32_t)q.level;
            } else {
                if (q.first + q.count > (int64_t)4294967296LL)
                    fail(FDMC_E_RANGE, "MLQMC Sobol offset beyond 32 bits (methods_cuda.cu:1847-1851)");
                if (ms.mode == FDMC_NOISE_BB) {
                    if (q.level >= ms.n_bb || !ms.bb_list[q.level])
                        fail(FDMC_E_INVALID, "missing bb_list[" + std::to_string(q.level) + "]");
                    const fdmc_bb* b = ms.bb_list[q.level];
                    if (b->N != nfine)
                        fail(FDMC_E_INVALID, "bb_list[" + std::to_string(q.level) + "] has N=" +
                                                 std::to_string(b->N) + ", expected " + std::to_string(nfine));
                    e.bbw = b->d_w;
                    e.G = nfine >= 16 ? nfine / 16 : nfine;
                } else if (ms.mode == FDMC_NOISE_PCA) {  // methods_cuda.cu:1781-1789
                    if (q.level >= ms.n_pca || !ms.pca_list[q.level])
                        fail(FDMC_E_INVALID, "missing pca_list[" + std::to_string(q.level) + "]");
                    const fdmc_pca* pc = ms.pca_list[q.level];
                    if (pc->m != nfine)
                        fail(FDMC_E_INVALID, "pca_list[" + std::to_string(q.level) + "] has m=" +
                                                 std::to_string(pc->m) + ", expected " + std::to_string(nfine));
                }
            }
            ent.push_back(e);
            which.push_back((int)i);
        }
        // Sobol pass: direction records only once every request has been validated
        if (pass == 0)
            for (size_t k = 0; k < ent.size(); k++) {
                const LevelReq& q = reqs[which[k]];
                ent[k].sdir = level_sdir(c, cache, ms, q.level, ent[k].D) + (size_t)q.replica * ent[k].D * SDIR_WORDS;
            }
        if (ent.empty()) continue;
        if (pass == 0 && ms.mode == FDMC_NOISE_PCA) {
            run_levels_pca(c, ms, reqs, ent, which, out);
            continue;
        }
        const int nz = pass == 1 ? NZ_PRNG : (ms.mode == FDMC_NOISE_BB ? NZ_SOBOL_BB : NZ_SOBOL_RAW);
        if (pass == 1) {
            launch_levels(c, ms, nz, ent, which, out);
            continue;
        }
        // levels whose direction records fit in LDS run in one launch, longer
        // ones (records read from global memory) in another, so a fine level
        // does not move the short levels' records out of LDS
        std::vector<LevelEntry> part_ent[2];
        std::vector<int> part_which[2];
        for (size_t k = 0; k < ent.size(); k++) {
            const int g = ent[k].D > SOBOL_LDS_MAX_D ? 1 : 0;
            part_ent[g].push_back(ent[k]);
            part_which[g].push_back(which[k]);
        }
        for (int g = 0; g < 2; g++)
            if (!part_ent[g].empty()) launch_levels(c, ms, nz, part_ent[g], part_which[g], out);
    }
}

// run_mlmc_cuda (methods_cuda.cu:1305-1487): pilot on levels 0..2, Giles
// allocation N_l = ceil(2/eps^2 sqrt(V_l/C_l) sum sqrt(V_l C_l)), bias test on
// the two finest levels, L++ with a pilot.  Each iteration's extra samples of
// all levels run in one launch; level l's samples are Philox paths
// (seed, 0x100 + l, index) continuing from N_l.
void mlmc_loop(fdmc_context* c, const MultiSpec& ms, double eps, int max_L, int pilot_n, fdmc_result* out) {
    const double t0 = now_s();
    const int M = ms.M;
    if (M < 2) fail(FDMC_E_INVALID, "MLMCConfig.M must be >= 2");
    if (max_L < 2 || max_L > 30) fail(FDMC_E_INVALID, "MLMCConfig.max_L must be in [2, 30]");
    if (pilot_n <= 1) fail(FDMC_E_INVALID, "pilot_n must be > 1");
    if (!(eps > 0.0)) fail(FDMC_E_INVALID, "eps must be positive");
    std::map<int, DevBuf> cache;
    std::vector<double> E(max_L + 1, 0.0), V(max_L + 1, 0.0);
    std::vector<int64_t> N(max_L + 1, 0);
    *out = fdmc_result{};
    std::vector<std::array<double, 4>> s;
    auto steps = [&](int l, int64_t n) { out->path_steps += n * (int64_t)level_steps(M, l); };
    int L = 2;
    {
        std::vector<LevelReq> rq;
        for (int l = 0; l <= L; l++) rq.push_back({l, 0, 0, pilot_n});
        run_levels(c, ms, cache, rq, s);
        for (int l = 0; l <= L; l++) {
            const double em = s[l][0] / pilot_n;
            E[l] = em;
            V[l] = std::max(0.0, s[l][1] / pilot_n - em * em);
            N[l] = pilot_n;
            steps(l, pilot_n);
        }
    }
    bool conv = false;
    int iter = 0;
    while (!conv && L <= max_L && iter++ < 50) {
        double st = 0.0;
        for (int l = 0; l <= L; l++) st += std::sqrt(V[l] * std::pow((double)M, l));
        std::vector<LevelReq> rq;
        for (int l = 0; l <= L; l++) {
            const double Cl = std::pow((double)M, l);
            int64_t Nopt = (int64_t)std::ceil(2.0 / (eps * eps) * std::sqrt(V[l] / Cl) * st);
            Nopt = std::max<int64_t>(Nopt, 100);
            if (Nopt > N[l]) rq.push_back({l, 0, N[l], Nopt - N[l]});
        }
        if (!rq.empty()) {
            run_levels(c, ms, cache, rq, s);
            for (size_t i = 0; i < rq.size(); i++) {
                const int l = rq[i].level;
                const int64_t ex = rq[i].count, Nn = N[l] + ex;
                const double ts = E[l] * N[l] + s[i][0];
                const double t2 = (V[l] + E[l] * E[l]) * N[l] + s[i][1];
                E[l] = ts / Nn;
                V[l] = std::max(0.0, t2 / Nn - E[l] * E[l]);
                N[l] = Nn;
                steps(l, ex);
            }
        }
        double bias = std::fabs(E[L]) / (M - 1);
        if (L >= 1) bias = std::max(bias, std::fabs(E[L - 1]) * M / (double)(M * M - M));
        conv = bias < eps / std::sqrt(2.0);
        if (!conv && L < max_L) {
            L++;
            std::vector<LevelReq> one{{L, 0, 0, pilot_n}};
            run_levels(c, ms, cache, one, s);
            const double em = s[0][0] / pilot_n;
            E[L] = em;
            V[L] = std::max(0.0, s[0][1] / pilot_n - em * em);
            N[L] = pilot_n;
            steps(L, pilot_n);
        }
    }
    double price = 0.0, vs = 0.0;
    int64_t tot = 0;
    for (int l = 0; l <= L; l++) {
        price += E[l];
        vs += N[l] > 0 ? V[l] / N[l] : 0.0;
        tot += N[l];
    }
    for (auto& kv : cache) kv.second.release();
    out->price = price;
    out->std_error = std::sqrt(vs);
    out->n_samples = tot;
    out->levels = L + 1;
    out->flags = conv ? 0u : FDMC_FLAG_NOT_CONVERGED;
    out->time_s = now_s() - t0;
}

// run_mlqmc_cuda (methods_cuda.cu:1680-2019): per level the estimator is the
// mean of R replica means (salt = salt_base + l R + r, Sobol offsets continuing
// per (l, r)), sig2_l = S^2(replica means) N_l, N_opt carries 1/R.
void mlqmc_loop(fdmc_context* c, const MultiSpec& ms, double eps, int max_L, int pilot_n, fdmc_result* out) {
    const double t0 = now_s();
    const int M = ms.M, R = ms.R;
    if (M < 2) fail(FDMC_E_INVALID, "MLMCConfig.M must be >= 2");
    if (max_L < 2 || max_L > 30) fail(FDMC_E_INVALID, "MLMCConfig.max_L must be in [2, 30]");
    if (pilot_n <= 1) fail(FDMC_E_INVALID, "pilot_n must be > 1");
    if (R <= 0) fail(FDMC_E_INVALID, "QMCConfig.R must be positive");
    if (!(eps > 0.0)) fail(FDMC_E_INVALID, "eps must be positive");
    std::map<int, DevBuf> cache;
    std::vector<double> E(max_L + 1, 0.0), S2(max_L + 1, 0.0);
    std::vector<int64_t> N(max_L + 1, 0), off(max_L + 1, 0);
    *out = fdmc_result{};
    std::vector<std::array<double, 4>> s;
    // launch R requests for each (level, n) and return (mean, sample var) of replica means per level
    auto run = [&](const std::vector<std::pair<int, int64_t>>& lv, std::vector<std::pair<double, double>>& mv) {
        std::vector<LevelReq> rq;
        for (auto& x : lv)
            for (int r = 0; r < R; r++) rq.push_back({x.first, r, off[x.first], x.second});
        run_levels(c, ms, cache, rq, s);
        mv.clear();
        for (size_t k = 0; k < lv.size(); k++) {
            const int l = lv[k].first;
            const int64_t n = lv[k].second;
            std::vector<double> rm((size_t)R);
            for (int r = 0; r < R; r++) rm[r] = s[k * R + r][0] / (double)n;
            double m_, v_;
            mean_var(rm, m_, v_);
            mv.push_back({m_, v_});
            off[l] += n;
            out->path_steps += (int64_t)R * n * level_steps(M, l);
        }
    };
    int L = 2;
    std::vector<std::pair<double, double>> mv;
    {
        std::vector<std::pair<int, int64_t>> lv;
        for (int l = 0; l <= L; l++) lv.push_back({l, pilot_n});
        run(lv, mv);
        for (int l = 0; l <= L; l++) {
            E[l] = mv[l].first;
            S2[l] = mv[l].second * (double)pilot_n;
            N[l] = pilot_n;
        }
    }
    bool conv = false;
    int iter = 0;
    while (!conv && L <= max_L && iter++ < 50) {
        double st = 0.0;
        for (int l = 0; l <= L; l++) st += std::sqrt(S2[l] * std::pow((double)M, l));
        std::vector<std::pair<int, int64_t>> lv;
        for (int l = 0; l <= L; l++) {
            const double Cl = std::pow((double)M, l);
            int64_t Nopt = (int64_t)std::ceil(2.0 / ((double)R * eps * eps) * std::sqrt(S2[l] / Cl) * st);
            Nopt = std::max<int64_t>(Nopt, 100);
            if (Nopt > N[l]) lv.push_back({l, Nopt - N[l]});
        }
        if (!lv.empty()) {
            run(lv, mv);
            for (size_t k = 0; k < lv.size(); k++) {
                const int l = lv[k].first;
                const double wo = (double)N[l], wn = (double)lv[k].second;
                E[l] = (E[l] * wo + mv[k].first * wn) / (wo + wn);
                S2[l] = (S2[l] * wo + mv[k].second * wn * wn) / (wo + wn);
                N[l] += lv[k].second;
            }
        }
        double bias = std::fabs(E[L]) / std::max(M - 1, 1);
        if (L >= 1) bias = std::max(bias, std::fabs(E[L - 1]) * M / (double)std::max(M * (M - 1), 1));
        conv = bias < eps / std::sqrt(2.0);
        if (!conv && L < max_L) {
            L++;
            run({{L, pilot_n}}, mv);
            E[L] = mv[0].first;
            S2[L] = mv[0].second * (double)pilot_n;
            N[L] = pilot_n;
        }
    }
    double price = 0.0, vs = 0.0;
    int64_t tot = 0;
    for (int l = 0; l <= L; l++) {
        price += E[l];
        vs += S2[l] / ((double)N[l] * (double)R);
        tot += N[l] * R;
    }
    for (auto& kv : cache) kv.second.release();
    out->price = price;
    out->std_error = std::sqrt(vs);
    out->n_samples = tot;
    out->levels = L + 1;
    out->flags = conv ? 0u : FDMC_FLAG_NOT_CONVERGED;
    out->time_s = now_s() - t0;
}

}  // namespace

// ================================================================ C ABI
extern "C" {

int32_t fdmc_abi_version(void) { return FDMC_ABI_VERSION; }

const char* fdmc_last_error(void) { return g_last_error.c_str(); }

fdmc_status fdmc_context_create(int32_t device, fdmc_context** out) {
    return guarded([&] {
        if (!out) fail(FDMC_E_INVALID, "out is null");
        *out = nullptr;
        int n = 0;
        if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) fail(FDMC_E_DEVICE, "no HIP device available");
        if (device < 0 || device >= n) fail(FDMC_E_INVALID, "device index out of range");
        HIPC(hipSetDevice(device));
        auto* c = new fdmc_context;
        c->device = device;
        hipDeviceProp_t prop;
        if (hipGetDeviceProperties(&prop, device) == hipSuccess && prop.multiProcessorCount > 0)
            c->n_cu = prop.multiProcessorCount;
        try {
            HIPC(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
            // the no-allocation entry point's scratch array on this (current) device
            HIPC(hipGetSymbolAddress((void**)&c->noalloc_scratch, HIP_SYMBOL(scratch)));
            std::vector<uint32_t> dv;
            expand_dirvectors(dv);
            HIPC(hipMalloc((void**)&c->d_dirv, dv.size() * sizeof(uint32_t)));
            HIPC(hipMemcpy(c->d_dirv, dv.data(), dv.size() * sizeof(uint32_t), hipMemcpyHostToDevice));
        } catch (...) {
            fdmc_context_destroy(c);
            throw;
        }
        *out = c;
    });
}

fdmc_status fdmc_context_destroy(fdmc_context* c) {
    if (!c) return FDMC_OK;
    (void)hipSetDevice(c->device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    c->partials.release();
    c->sums.release();
    c->salts.release();
    c->sdir.release();
    c->gen_sdir.release();
    c->pca_z.release();
    c->pca_dw.release();
    c->basket_L.release();
    if (c->d_dirv) (void)hipFree(c->d_dirv);
    if (c->h_sums) (void)hipHostFree(c->h_sums);
    if (c->ev0) (void)hipEventDestroy(c->ev0);
    if (c->ev1) (void)hipEventDestroy(c->ev1);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
    return FDMC_OK;
}

fdmc_status fdmc_context_stream(fdmc_context* c, void** s) {
    return guarded([&] {
        if (!c || !s) fail(FDMC_E_INVALID, "null argument");
        *s = (void*)c->stream;
    });
}

fdmc_status fdmc_timing_enable(fdmc_context* c, int32_t enable) {
    return guarded([&] {
        if (!c) fail(FDMC_E_INVALID, "null context");
        if (enable && !c->ev0) {
            HIPC(hipEventCreate(&c->ev0));
            HIPC(hipEventCreate(&c->ev1));
        }
        c->timing = enable != 0;
    });
}

fdmc_status fdmc_timing_read(fdmc_context* c, int64_t* launches, double* total_ms, int32_t reset) {
    return guarded([&] {
        if (!c) fail(FDMC_E_INVALID, "null context");
        if (launches) *launches = c->timed_launches;
        if (total_ms) *total_ms = c->timed_ms;
        if (reset) {
            c->timed_launches = 0;
            c->timed_ms = 0.0;
        }
    });
}

fdmc_status fdmc_synchronize(fdmc_context* c) {
    return guarded([&] {
        if (!c) fail(FDMC_E_INVALID, "null context");
        HIPC(hipStreamSynchronize(c->stream));
    });
}

// ---- BB / PCA ------------------------------------------------------------------
fdmc_status fdmc_bb_upload(fdmc_context* c, int32_t N, double T, const int32_t* map_idx,
                           const int32_t* left_idx, const int32_t* right_idx, const double* wl,
                           const double* wr, const double* sd, fdmc_bb** out) {
    return guarded([&] {
        if (!c || !out || !map_idx || !left_idx || !right_idx || !wl || !wr || !sd)
            fail(FDMC_E_INVALID, "null argument");
        if (N <= 0 || (N & (N - 1))) fail(FDMC_E_INVALID, "BB N must be a power of two");
        // The fused kernels assume the dyadic coarse-to-fine schedule of
        // bb_precompute (utils.cpp:54-101); check it entry by entry.
        if (map_idx[0] != N) fail(FDMC_E_INVALID, "BB entry 0 must map to W_N");
        for (int j = 1; j < N; j++) {
            const int lev = 31 - __builtin_clz((unsigned)j);
            const int q = j - (1 << lev), st = N >> lev;
            if (left_idx[j] != q * st || right_idx[j] != q * st + st || map_idx[j] != q * st + st / 2)
                fail(FDMC_E_INVALID, "BB tables are not the dyadic schedule of bb_precompute");
        }
        std::vector<float> w((size_t)N * 4);
        for (int j = 0; j < N; j++) {
            w[4 * j + 0] = (float)wl[j];
            w[4 * j + 1] = (float)wr[j];
            w[4 * j + 2] = (float)sd[j];
            w[4 * j + 3] = 0.0f;
        }
        auto* b = new fdmc_bb{c, N, T, nullptr, 1, {}};
        b->sd_depth[0] = w[2];
        for (int j = 1; j < N && b->uniform; j++) {
            const int depth = 32 - __builtin_clz((unsigned)j);     // schedule level + 1
            if (depth >= BB_SD_DEPTHS || w[4 * j] != 0.5f || w[4 * j + 1] != 0.5f) b->uniform = 0;
            else if (j == (1 << (depth - 1))) b->sd_depth[depth] = w[4 * j + 2];
            else if (w[4 * j + 2] != b->sd_depth[depth]) b->uniform = 0;
        }
        if (hipMalloc((void**)&b->d_w, w.size() * sizeof(float)) != hipSuccess) {
            delete b;
            fail(FDMC_E_ALLOC, "hipMalloc failed");
        }
        HIPC(hipMemcpy(b->d_w, w.data(), w.size() * sizeof(float), hipMemcpyHostToDevice));
        *out = b;
    });
}

fdmc_status fdmc_bb_free(fdmc_bb* b) {
    if (!b) return FDMC_OK;
    if (b->d_w) (void)hipFree(b->d_w);
    delete b;
    return FDMC_OK;
}

fdmc_status fdmc_pca_upload(fdmc_context* c, int32_t m, const double* M, fdmc_pca** out) {
    return guarded([&] {
        if (!c || !M || !out) fail(FDMC_E_INVALID, "null argument");
        if (m <= 0) fail(FDMC_E_INVALID, "PCA m must be positive");
        const int mp = (m + 127) / 128 * 128;
        const size_t n = (size_t)mp * mp;
        std::vector<uint16_t> f16(n, 0), bf16(n, 0);
        // device copies are row-major (M[i][j] at i*mp + j), zero padded to mp:
        // the GEMM's Bt operand
        for (int i = 0; i < m; i++)
            for (int j = 0; j < m; j++) {
                const double v = M[(size_t)j * m + i];
                const size_t o = (size_t)i * mp + j;
                const _Float16 h = (_Float16)v;
                std::memcpy(&f16[o], &h, 2);
                const float f = (float)v;
                uint32_t u;
                std::memcpy(&u, &f, 4);
                u += 0x7FFFu + ((u >> 16) & 1u);  // round to nearest even
                bf16[o] = (uint16_t)(u >> 16);
            }
        auto* p = new fdmc_pca{c, m, mp, nullptr, nullptr, std::vector<double>(M, M + (size_t)m * m)};
        if (hipMalloc(&p->d_f16, n * 2) != hipSuccess || hipMalloc(&p->d_bf16, n * 2) != hipSuccess) {
            if (p->d_f16) (void)hipFree(p->d_f16);
            delete p;
            fail(FDMC_E_ALLOC, "hipMalloc failed");
        }
        HIPC(hipMemcpy(p->d_f16, f16.data(), n * 2, hipMemcpyHostToDevice));
        HIPC(hipMemcpy(p->d_bf16, bf16.data(), n * 2, hipMemcpyHostToDevice));
        *out = p;
    });
}

fdmc_status fdmc_pca_free(fdmc_pca* p) {
    if (!p) return FDMC_OK;
    if (p->d_f16) (void)hipFree(p->d_f16);
    if (p->d_bf16) (void)hipFree(p->d_bf16);
    delete p;
    return FDMC_OK;
}

// ---- estimators ------------------------------------------------------------------
fdmc_status fdmc_qmc_accumulate(fdmc_context* c, const fdmc_model* model, const fdmc_payoff* payoff,
                                const fdmc_model* cmodel, const fdmc_payoff* cpayoff, double E_ctrl,
                                double beta, int32_t n_steps, int32_t R, uint32_t salt_base,
                                int64_t offset, int64_t n_points, int32_t mode, const fdmc_bb* bb,
                                const fdmc_pca* pca, double* sums_out, int32_t* n_acc_out) {
    const bool pca_precision_bf16 = (mode & FDMC_PCA_BF16) != 0;
    mode &= ~FDMC_PCA_BF16;
    return guarded([&] {
        if (!c || !sums_out) fail(FDMC_E_INVALID, "null argument");
        validate_model(model);
        validate_payoff(payoff);
        if ((cmodel == nullptr) != (cpayoff == nullptr))
            fail(FDMC_E_INVALID, "control model and payoff must both be given");
        if (cmodel) { validate_model(cmodel); validate_payoff(cpayoff); }
        if (offset < 0) fail(FDMC_E_INVALID, "offset must be non-negative");
        PassSpec s;
        s.model = model; s.payoff = payoff; s.cmodel = cmodel; s.cpayoff = cpayoff;
        s.E_ctrl = E_ctrl; s.beta = beta; s.n_steps = n_steps; s.R = R; s.salt_base = salt_base;
        s.first = offset; s.count = n_points; s.mode = mode; s.bb = bb; s.pca = pca;
        s.pca_bf16 = (mode == FDMC_NOISE_PCA && pca_precision_bf16) ? 1 : 0;
        std::vector<double> v;
        const int nacc = run_pass(c, s, v);
        std::memcpy(sums_out, v.data(), v.size() * sizeof(double));
        if (n_acc_out) *n_acc_out = nacc;
    });
}

fdmc_status fdmc_qmc_accumulate_noalloc(fdmc_context* c, const fdmc_model* model, const fdmc_payoff* payoff,
                                        const fdmc_model* cmodel, const fdmc_payoff* cpayoff, double E_ctrl,
                                        double beta, int32_t n_steps, int32_t R, uint32_t salt_base,
                                        int64_t offset, int64_t n_points, int32_t mode, const fdmc_bb* bb,
                                        double* d_sums_out, int32_t* n_acc_out) {
    return guarded([&] {
        if (!c || !d_sums_out) fail(FDMC_E_INVALID, "null argument");
        if (!c->noalloc_scratch) fail(FDMC_E_DEVICE, "no-allocation scratch not available on this device");
        validate_model(model);
        validate_payoff(payoff);
        if ((cmodel == nullptr) != (cpayoff == nullptr))
            fail(FDMC_E_INVALID, "control model and payoff must both be given");
        if (cmodel) { validate_model(cmodel); validate_payoff(cpayoff); }
        if (offset < 0) fail(FDMC_E_INVALID, "offset must be non-negative");
        PassSpec s;
        s.model = model; s.payoff = payoff; s.cmodel = cmodel; s.cpayoff = cpayoff;
        s.E_ctrl = E_ctrl; s.beta = beta; s.n_steps = n_steps; s.R = R; s.salt_base = salt_base;
        s.first = offset; s.count = n_points; s.mode = mode; s.bb = bb;
        NoAlloc na;
        na.scratch = c->noalloc_scratch;
        na.cap = FDMC_NOALLOC_SCRATCH_BYTES;
        na.d_sums = d_sums_out;
        std::vector<double> unused;   // stays empty: sums go to d_sums_out
        const int nacc = run_pass(c, s, unused, &na);
        if (n_acc_out) *n_acc_out = nacc;
    });
}

fdmc_status fdmc_greeks_accumulate(fdmc_context* c, const fdmc_model* model, const fdmc_payoff* payoff,
                                   const fdmc_model* cmodel, const fdmc_payoff* cpayoff, double E_ctrl,
                                   double beta, int32_t n_steps, int32_t R, uint32_t salt_or_seed,
                                   int64_t offset, int64_t n_points, int32_t mode, const fdmc_bb* bb,
                                   double* sums_out, int32_t* n_acc_out) {
    return guarded([&] {
        if (!c || !sums_out) fail(FDMC_E_INVALID, "null argument");
        validate_model(model);
        validate_payoff(payoff);
        if ((cmodel == nullptr) != (cpayoff == nullptr))
            fail(FDMC_E_INVALID, "control model and payoff must both be given");
        if (cmodel) { validate_model(cmodel); validate_payoff(cpayoff); }
        if (offset < 0) fail(FDMC_E_INVALID, "offset must be non-negative");
        PassSpec s;
        s.model = model; s.payoff = payoff; s.cmodel = cmodel; s.cpayoff = cpayoff;
        s.E_ctrl = E_ctrl; s.beta = beta; s.n_steps = n_steps; s.R = R; s.salt_base = salt_or_seed;
        s.first = offset; s.count = n_points; s.bb = bb; s.greeks = true;
        if (mode == FDMC_NOISE_PRNG) { s.prng = true; s.tag = 0; s.mode = FDMC_NOISE_RAW; }
        else s.mode = mode;
        std::vector<double> v;
        const int nacc = run_pass(c, s, v);
        std::memcpy(sums_out, v.data(), v.size() * sizeof(double));
        if (n_acc_out) *n_acc_out = nacc;
    });
}

fdmc_status fdmc_greeks_accumulate_pca(fdmc_context* c, const fdmc_model* model, const fdmc_payoff* payoff,
                                       int32_t n_steps, int32_t R, uint32_t salt_base, int64_t offset,
                                       int64_t n_points, int32_t mode, const fdmc_pca* pca, double* sums_out,
                                       int32_t* n_acc_out) {
    const bool bf16 = (mode & FDMC_PCA_BF16) != 0;
    mode &= ~FDMC_PCA_BF16;
    return guarded([&] {
        if (!c || !sums_out || !pca) fail(FDMC_E_INVALID, "null argument");
        validate_model(model);
        validate_payoff(payoff);
        if (mode != FDMC_NOISE_PCA) fail(FDMC_E_INVALID, "mode must be FDMC_NOISE_PCA (optionally | FDMC_PCA_BF16)");
        if (model->kind != FDMC_MODEL_MULTI_GBM || model->n_assets > BASKET_MAX)
            fail(FDMC_E_UNSUPPORTED, "PCA Greeks are available for baskets of at most 32 GBM assets");
        if (offset < 0) fail(FDMC_E_INVALID, "offset must be non-negative");
        PassSpec s;
        s.model = model; s.payoff = payoff; s.n_steps = n_steps; s.R = R; s.salt_base = salt_base;
        s.first = offset; s.count = n_points; s.mode = FDMC_NOISE_PCA; s.pca = pca; s.pca_bf16 = bf16 ? 1 : 0;
        s.greeks = true;
        std::vector<double> v;
        const int nacc = run_pass(c, s, v);
        std::memcpy(sums_out, v.data(), v.size() * sizeof(double));
        if (n_acc_out) *n_acc_out = nacc;
    });
}

fdmc_status fdmc_run_qmc_greeks(fdmc_context* c, const fdmc_model* model, const fdmc_payoff* payoff,
                                const fdmc_model* cmodel, const fdmc_payoff* cpayoff, double E_ctrl,
                                const double* dE_ctrl, double beta, double eps, int32_t n_steps,
                                const fdmc_qmc_config* cfg, int32_t mode, const fdmc_bb* bb,
                                fdmc_result* out, fdmc_greeks* greeks) {
    return guarded([&] {
        if (!c || !cfg || !out || !greeks) fail(FDMC_E_INVALID, "null argument");
        validate_model(model);
        validate_payoff(payoff);
        if ((cmodel == nullptr) != (cpayoff == nullptr))
            fail(FDMC_E_INVALID, "control model and payoff must both be given");
        if (cmodel) { validate_model(cmodel); validate_payoff(cpayoff); }
        PassSpec s;
        s.model = model; s.payoff = payoff; s.cmodel = cmodel; s.cpayoff = cpayoff;
        s.E_ctrl = E_ctrl; s.beta = beta; s.n_steps = n_steps; s.mode = mode; s.bb = bb; s.greeks = true;
        qmc_loop(c, s, eps, *cfg, out, greeks, cmodel ? dE_ctrl : nullptr);
    });
}

fdmc_status fdmc_pca_transform(fdmc_context* c, const fdmc_pca* pca, uint32_t salt, int64_t offset,
                               int64_t n_points, int32_t bf16, float* d_out) {
    return guarded([&] {
        if (!c || !pca || !d_out) fail(FDMC_E_INVALID, "null argument");
        if (n_points <= 0 || offset < 0 || offset + n_points > (int64_t)4294967296LL)
            fail(FDMC_E_RANGE, "point range outside the 32-bit Sobol index");
        const int D = pca->m;
        if (D > D_MAX_SOBOL) fail(FDMC_E_RANGE, "PCA dimension above 20000");
        std::vector<uint32_t> salts{salt};
        const uint32_t* sdir = get_sdir(c, salts, D);
        const int ld = pca_rows(c, pca, sdir, D, 1, offset, n_points, bf16 != 0);
        HIPC(hipMemcpy2DAsync(d_out, (size_t)D * sizeof(float), c->pca_dw.p, (size_t)ld * sizeof(float),
                              (size_t)D * sizeof(float), (size_t)n_points, hipMemcpyDeviceToDevice, c->stream));
        HIPC(hipStreamSynchronize(c->stream));
        if (c->timing) {
            float ms = 0.0f;
            HIPC(hipEventElapsedTime(&ms, c->ev0, c->ev1));
            c->timed_ms += ms;
            c->timed_launches++;
        }
    });
}

fdmc_status fdmc_pca_products(fdmc_context* c, const fdmc_pca* pca, uint32_t salt, int64_t offset,
                              int64_t n_points, int32_t bf16, int32_t n_assets, float c0, const float* sig,
                              float* d_out) {
    return guarded([&] {
        if (!c || !pca || !sig || !d_out) fail(FDMC_E_INVALID, "null argument");
        if (n_points <= 0 || offset < 0 || offset + n_points > (int64_t)4294967296LL)
            fail(FDMC_E_RANGE, "point range outside the 32-bit Sobol index");
        if (n_assets < 2 || n_assets > BASKET_MAX || (n_assets & (n_assets - 1)))
            fail(FDMC_E_INVALID, "n_assets must be a power of two in [2, 32]");
        const int D = pca->m, ld = pca->mp;
        if (D > D_MAX_SOBOL) fail(FDMC_E_RANGE, "PCA dimension above 20000");
        if (ld % 256) fail(FDMC_E_INVALID, "the product epilogue needs m rounded up to 128 to be a multiple of 256");
        ProdEpi e{};
        e.c0 = c0;
        for (int a = 0; a < n_assets; a++) e.sig[a] = sig[a];
        e.n = n_assets;
        e.D = D;
        std::vector<uint32_t> salts{salt};
        const uint32_t* sdir = get_sdir(c, salts, D);
        const int64_t mrows = pca_prod(c, pca, sdir, D, 1, offset, n_points, bf16 != 0, e);
        const size_t row = (size_t)n_assets * sizeof(float);
        HIPC(hipMemcpy2DAsync(d_out, (size_t)n_points * row, c->pca_dw.p, (size_t)mrows * row, (size_t)n_points * row,
                              (size_t)(ld / 128), hipMemcpyDeviceToDevice, c->stream));
        HIPC(hipStreamSynchronize(c->stream));
        if (c->timing) {
            float ms = 0.0f;
            HIPC(hipEventElapsedTime(&ms, c->ev0, c->ev1));
            c->timed_ms += ms;
            c->timed_launches++;
        }
    });
}

fdmc_status fdmc_run_qmc(fdmc_context* c, const fdmc_model* model, const fdmc_payoff* payoff, double eps,
                         int32_t n_steps, const fdmc_qmc_config* cfg, int32_t mode, const fdmc_bb* bb,
                         const fdmc_pca* pca, fdmc_result* out) {
    return guarded([&] {
        if (!c || !cfg || !out) fail(FDMC_E_INVALID, "null argument");
        validate_model(model);
        validate_payoff(payoff);
        if (mode == FDMC_NOISE_BB && noise_dim(*model) != 1)
            fail(FDMC_E_UNSUPPORTED, "Brownian bridge needs a one-factor model");
        PassSpec s;
        s.model = model; s.payoff = payoff; s.n_steps = n_steps; s.mode = mode; s.bb = bb; s.pca = pca;
        s.pca_bf16 = cfg->pca_precision;
        qmc_loop(c, s, eps, *cfg, out);
    });
}

fdmc_status fdmc_run_qmc_cv(fdmc_context* c, const fdmc_model* mm, const fdmc_model* cm, const fdmc_payoff* mp,
                            const fdmc_payoff* cp, double E_ctrl, double beta, double eps, int32_t n_steps,
                            const fdmc_qmc_config* cfg, int32_t mode, const fdmc_bb* bb, fdmc_result* out) {
    return guarded([&] {
        if (!c || !cfg || !out) fail(FDMC_E_INVALID, "null argument");
        validate_model(mm); validate_model(cm); validate_payoff(mp); validate_payoff(cp);
        PassSpec s;
        s.model = mm; s.payoff = mp; s.cmodel = cm; s.cpayoff = cp; s.E_ctrl = E_ctrl; s.beta = beta;
        s.n_steps = n_steps; s.mode = mode; s.bb = bb;
        qmc_loop(c, s, eps, *cfg, out);
    });
}

fdmc_status fdmc_run_mc_fixed(fdmc_context* c, const fdmc_model* model, const fdmc_payoff* payoff,
                              int32_t n_steps, int64_t n_paths, uint32_t seed, double* mean_out,
                              double* var_out) {
    return guarded([&] {
        if (!c || !mean_out || !var_out) fail(FDMC_E_INVALID, "null argument");
        validate_model(model);
        validate_payoff(payoff);
        if (n_paths <= 0) fail(FDMC_E_INVALID, "n_paths must be positive");
        PassSpec s;
        s.model = model; s.payoff = payoff; s.n_steps = n_steps; s.prng = true; s.tag = 0;
        s.salt_base = seed; s.first = 0; s.count = n_paths;
        std::vector<double> v;
        run_pass(c, s, v);
        const double m = v[0] / (double)n_paths;
        const double var = v[1] / (double)n_paths - m * m;
        *mean_out = m;
        *var_out = std::max(0.0, var) / (double)n_paths;
    });
}

fdmc_status fdmc_run_mc(fdmc_context* c, const fdmc_model* model, const fdmc_payoff* payoff, double eps,
                        int32_t n_steps, const fdmc_mc_config* cfg, fdmc_result* out) {
    return guarded([&] {
        if (!c || !cfg || !out) fail(FDMC_E_INVALID, "null argument");
        validate_model(model);
        validate_payoff(payoff);
        if (!(eps > 0.0)) fail(FDMC_E_INVALID, "eps must be positive");
        if (cfg->pilot_n <= 1) fail(FDMC_E_INVALID, "pilot_n must be > 1");
        const double t0 = now_s();
        *out = fdmc_result{};
        PassSpec s;
        s.model = model; s.payoff = payoff; s.n_steps = n_steps; s.prng = true; s.tag = 0;
        s.salt_base = cfg->seed; s.first = 0; s.count = cfg->pilot_n;
        std::vector<double> v;
        run_pass(c, s, v);  // pilot (run_mc_cuda, methods_cuda.cu:1276-1280)
        const double pm = v[0] / cfg->pilot_n;
        const double pvar = std::max(0.0, v[1] / cfg->pilot_n - pm * pm);
        int64_t N = (int64_t)std::ceil(2.0 * pvar / (eps * eps));
        N = std::max<int64_t>(N, cfg->pilot_n);
        s.salt_base = cfg->seed + 1u;
        s.count = N;
        run_pass(c, s, v);
        const double m = v[0] / (double)N;
        const double var = std::max(0.0, v[1] / (double)N - m * m) / (double)N;
        out->price = m;
        out->std_error = std::sqrt(var);
        out->n_samples = N;
        out->path_steps = (N + cfg->pilot_n) * (int64_t)n_steps;
        out->time_s = now_s() - t0;
    });
}

// Sums {Y, Y^2} of n importance-sampled paths of the Philox stream (seed, TAG_IS).
static void is_sums(fdmc_context* c, ISLaunch a, uint32_t seed, int64_t n, double out[2]) {
    a.seed = seed;
    a.tag = TAG_IS;
    a.count = n;
    const int64_t ntiles = (n + 255) >> 8;
    const int64_t target = (int64_t)c->n_cu * 32;
    int64_t tpw = std::max<int64_t>(1, (ntiles + target - 1) / target);
    if (tpw > (1 << 20)) tpw = 1 << 20;
    const int64_t nwg = (ntiles + tpw - 1) / tpw;
    if (nwg > 0x7FFFFFFF) fail(FDMC_E_RANGE, "too many workgroups");
    a.tiles_per_wg = (int)tpw;
    a.partials = (double*)c->partials.ensure((size_t)nwg * 2 * sizeof(double));
    double* d_sums = (double*)c->sums.ensure(2 * sizeof(double));
    HIPC(launch_is_gbm(a, (int)nwg, c->stream));
    HIPC(launch_reduce_partials(a.partials, 1, (int)nwg, 2, d_sums, c->stream));
    ensure_host_sums(c, 2);
    HIPC(hipMemcpyAsync(c->h_sums, d_sums, 2 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPC(hipStreamSynchronize(c->stream));
    out[0] = c->h_sums[0];
    out[1] = c->h_sums[1];
}

// run_is_cuda (methods_cuda.cu:2317-2365): n = max(4, ceil(T / eps)) steps,
// per-step shift z* / sqrt(n); pilot of pilot_n paths (seed), then
// N = max(ceil(2 Var / eps^2), pilot_n) rounded up to even paths (seed + 1).
fdmc_status fdmc_run_is(fdmc_context* c, const fdmc_model* model, const fdmc_payoff* payoff, double z_star,
                        double eps, const fdmc_mc_config* cfg, fdmc_result* out) {
    return guarded([&] {
        if (!c || !model || !payoff || !cfg || !out) fail(FDMC_E_INVALID, "null argument");
        if (model->kind != FDMC_MODEL_GBM) fail(FDMC_E_UNSUPPORTED, "run_is: GBM models only (methods_cuda.cuh:127)");
        if (payoff->kind != FDMC_PAYOFF_EUROPEAN) fail(FDMC_E_UNSUPPORTED, "run_is: European payoffs only");
        validate_model(model);
        validate_payoff(payoff);
        if (!(eps > 0.0)) fail(FDMC_E_INVALID, "eps must be positive");
        if (cfg->pilot_n <= 1) fail(FDMC_E_INVALID, "pilot_n must be > 1");
        if (!std::isfinite(z_star)) fail(FDMC_E_INVALID, "z_star must be finite");
        const double t0 = now_s();
        *out = fdmc_result{};
        const double nd = std::max(4.0, std::ceil(model->T / eps));
        if (nd > (double)(1 << 30)) fail(FDMC_E_RANGE, "T / eps gives too many steps");
        const int n_steps = (int)nd;
        ISLaunch a{};
        a.S0 = (float)model->S0;
        a.mu = (float)model->mu;
        a.sigma = (float)model->sigma;
        a.h = (float)model->T / (float)n_steps;
        a.sqrt_h = std::sqrt(a.h);
        a.K = (float)payoff->K;
        a.discount = (float)std::exp(-payoff->r * payoff->T);
        a.zs = (float)(z_star / std::sqrt((double)n_steps));
        a.n_steps = n_steps;
        double s[2];
        is_sums(c, a, cfg->seed, cfg->pilot_n, s);
        const double pm = s[0] / cfg->pilot_n;
        const double pvar = std::max(0.0, s[1] / cfg->pilot_n - pm * pm);
        int64_t N = (int64_t)std::ceil(2.0 * pvar / (eps * eps));
        N = std::max<int64_t>(N, cfg->pilot_n);
        N = (N + 1) & ~(int64_t)1;
        is_sums(c, a, cfg->seed + 1u, N, s);
        const double m = s[0] / (double)N;
        const double var = std::max(0.0, s[1] / (double)N - m * m) / (double)N;
        out->price = m;
        out->std_error = std::sqrt(var);
        out->n_samples = N;
        out->path_steps = (N + cfg->pilot_n) * (int64_t)n_steps;
        out->time_s = now_s() - t0;
    });
}

fdmc_status fdmc_cv_pilot(fdmc_context* c, const fdmc_model* mm, const fdmc_model* cm, const fdmc_payoff* mp,
                          const fdmc_payoff* cp, double E_ctrl, int32_t n_steps, int32_t N_pilot, uint32_t seed,
                          fdmc_cv_pilot_result* out) {
    return guarded([&] {
        if (!c || !out) fail(FDMC_E_INVALID, "null argument");
        validate_model(mm); validate_model(cm); validate_payoff(mp); validate_payoff(cp);
        if (N_pilot <= 1) fail(FDMC_E_INVALID, "N_pilot must be > 1");
        PassSpec s;
        s.model = mm; s.payoff = mp; s.cmodel = cm; s.cpayoff = cp; s.E_ctrl = E_ctrl; s.beta = 0.0;
        s.n_steps = n_steps; s.prng = true; s.tag = TAG_CV_PILOT; s.salt_base = seed;
        s.first = 0; s.count = N_pilot;
        std::vector<double> hs;
        run_pass(c, s, hs);
        // methods_cuda.cu:2061-2075
        const double N = (double)N_pilot;
        const double Ea = hs[0] / N, Ea2 = hs[1] / N, Eg = hs[4] / N, Eg2 = hs[5] / N, Eag = hs[6] / N;
        const double vm = std::max(0.0, Ea2 - Ea * Ea);
        const double vc = std::max(1e-30, Eg2 - Eg * Eg);
        const double cov = Eag - Ea * Eg;
        out->beta = std::clamp(cov / vc, 0.0, 5.0);
        out->var_plain = vm;
        out->var_cv = std::max(0.0, vm - cov * cov / vc);
    });
}

fdmc_status fdmc_run_mc_cv(fdmc_context* c, const fdmc_model* mm, const fdmc_model* cm, const fdmc_payoff* mp,
                           const fdmc_payoff* cp, double E_ctrl, double beta, double eps, int32_t n_steps,
                           const fdmc_mc_config* cfg, fdmc_result* out) {
    return guarded([&] {
        if (!c || !cfg || !out) fail(FDMC_E_INVALID, "null argument");
        validate_model(mm); validate_model(cm); validate_payoff(mp); validate_payoff(cp);
        if (!(eps > 0.0)) fail(FDMC_E_INVALID, "eps must be positive");
        const double t0 = now_s();
        *out = fdmc_result{};
        PassSpec s;
        s.model = mm; s.payoff = mp; s.cmodel = cm; s.cpayoff = cp; s.E_ctrl = E_ctrl; s.beta = beta;
        s.n_steps = n_steps; s.prng = true; s.tag = TAG_CV_PILOT; s.salt_base = cfg->seed;
        s.first = 0; s.count = cfg->pilot_n;
        std::vector<double> hs;
        run_pass(c, s, hs);  // methods_cuda.cu:2116-2120
        const double mu = hs[2] / cfg->pilot_n;
        const double v = std::max(0.0, hs[3] / cfg->pilot_n - mu * mu);
        int64_t N = (int64_t)std::ceil(2.0 * v / (eps * eps));
        N = std::max<int64_t>(N, cfg->pilot_n);
        s.salt_base = cfg->seed + 100u;
        s.tag = TAG_CV_MAIN;
        s.count = N;
        run_pass(c, s, hs);
        const double m = hs[2] / (double)N;
        const double vf = std::max(0.0, hs[3] / (double)N - m * m) / (double)N;
        out->price = m;
        out->std_error = std::sqrt(vf);
        out->n_samples = N;
        out->path_steps = (N + cfg->pilot_n) * (int64_t)n_steps;
        out->time_s = now_s() - t0;
    });
}

fdmc_status fdmc_run_mlmc(fdmc_context* c, const fdmc_model* model, const fdmc_payoff* payoff, double eps,
                          const fdmc_mlmc_config* cfg, fdmc_result* out) {
    return guarded([&] {
        if (!c || !cfg || !out) fail(FDMC_E_INVALID, "null argument");
        validate_model(model);
        validate_payoff(payoff);
        if (is_basket(*model))  // methods_cuda.cu:1318-1320
            fail(FDMC_E_UNSUPPORTED, "run_mlmc: basket models are not supported (no multi-asset MLMC kernel)");
        MultiSpec ms;
        ms.model = model; ms.payoff = payoff; ms.M = cfg->M; ms.qmc = false; ms.salt_base = cfg->seed;
        mlmc_loop(c, ms, eps, cfg->max_L, cfg->pilot_n, out);
    });
}

// One coupled-level pass (test and bench hook): request i covers samples
// [first_i, first_i + count_i) of level level_i (replica replica_i for Sobol).
fdmc_status fdmc_levels_accumulate(fdmc_context* c, const fdmc_model* model, const fdmc_payoff* payoff,
                                   int32_t M, int32_t qmc, int32_t mode, int32_t R, uint32_t salt_or_seed,
                                   int32_t n_req, const int32_t* levels, const int32_t* replicas,
                                   const int64_t* first, const int64_t* count,
                                   const fdmc_bb* const* bb_list, int32_t n_bb,
                                   const fdmc_pca* const* pca_list, int32_t n_pca, double* sums_out) {
    return guarded([&] {
        if (!c || !levels || !first || !count || !sums_out || n_req <= 0) fail(FDMC_E_INVALID, "null argument");
        validate_model(model);
        validate_payoff(payoff);
        if (is_basket(*model)) fail(FDMC_E_UNSUPPORTED, "basket models are not supported by the multilevel kernel");
        const bool bf16 = (mode & FDMC_PCA_BF16) != 0;
        mode &= ~FDMC_PCA_BF16;
        if (mode != FDMC_NOISE_RAW && mode != FDMC_NOISE_BB && mode != FDMC_NOISE_PCA)
            fail(FDMC_E_INVALID, "unknown noise mode");
        MultiSpec ms;
        ms.model = model; ms.payoff = payoff; ms.M = M; ms.qmc = qmc != 0; ms.mode = mode; ms.R = std::max(R, 1);
        ms.salt_base = salt_or_seed; ms.bb_list = bb_list; ms.n_bb = n_bb;
        ms.pca_list = pca_list; ms.n_pca = n_pca; ms.pca_bf16 = bf16;
        if (M < 2) fail(FDMC_E_INVALID, "M must be >= 2");
        std::vector<LevelReq> rq;
        for (int i = 0; i < n_req; i++) {
            if (levels[i] < 0 || count[i] <= 0 || first[i] < 0) fail(FDMC_E_INVALID, "bad level request");
            const int r = replicas ? replicas[i] : 0;
            if (r < 0 || r >= ms.R) fail(FDMC_E_INVALID, "replica out of range");
            rq.push_back({levels[i], r, first[i], count[i]});
        }
        std::map<int, DevBuf> cache;
        std::vector<std::array<double, 4>> s;
        run_levels(c, ms, cache, rq, s);
        for (auto& kv : cache) kv.second.release();
        for (int i = 0; i < n_req; i++)
            for (int k = 0; k < 4; k++) sums_out[i * 4 + k] = s[i][k];
    });
}

fdmc_status fdmc_run_mlqmc(fdmc_context* c, const fdmc_model* model, const fdmc_payoff* payoff, double eps,
                           const fdmc_mlmc_config* ml, const fdmc_qmc_config* qc, int32_t mode,
                           const fdmc_bb* const* bbl, int32_t n_bb, const fdmc_pca* const* pcal, int32_t n_pca,
                           fdmc_result* out) {
    return guarded([&] {
        if (!c || !ml || !qc || !out) fail(FDMC_E_INVALID, "null argument");
        validate_model(model);
        validate_payoff(payoff);
        if (is_basket(*model))  // methods_cuda.cu:1695-1697
            fail(FDMC_E_UNSUPPORTED, "run_mlqmc: basket models are not supported");
        if (mode != FDMC_NOISE_RAW && noise_dim(*model) != 1)  // methods_cuda.cu:1702-1704
            fail(FDMC_E_UNSUPPORTED, "run_mlqmc: Brownian bridge/PCA need a one-factor model");
        MultiSpec ms;
        ms.model = model; ms.payoff = payoff; ms.M = ml->M; ms.qmc = true; ms.mode = mode; ms.R = qc->R;
        ms.salt_base = qc->salt_base; ms.bb_list = bbl; ms.n_bb = n_bb;
        ms.pca_list = pcal; ms.n_pca = n_pca; ms.pca_bf16 = qc->pca_precision != 0;
        mlqmc_loop(c, ms, eps, ml->max_L, ml->pilot_n, out);
    });
}

// ---- generator hooks ----------------------------------------------------------------
static fdmc_status sobol_out(fdmc_context* c, uint32_t salt, int32_t D, int64_t n, int64_t offset, void* d_out,
                             bool normal) {
    return guarded([&] {
        if (!c || !d_out) fail(FDMC_E_INVALID, "null argument");
        if (D <= 0 || D > D_MAX_SOBOL) fail(FDMC_E_RANGE, "D must be in [1, 20000]");
        if (n <= 0 || offset < 0) fail(FDMC_E_INVALID, "n must be positive, offset non-negative");
        if (offset + n > (int64_t)4294967296LL)
            fail(FDMC_E_RANGE, "offset + n exceeds the 32-bit Sobol index (methods_cuda.cu:834-837)");
        std::vector<uint32_t> salts{salt};
        const size_t recs = (size_t)D * SDIR_WORDS;
        c->gen_sdir.ensure(recs * sizeof(uint32_t));
        c->salts.ensure(sizeof(uint32_t));
        HIPC(hipMemcpyAsync(c->salts.p, salts.data(), sizeof(uint32_t), hipMemcpyHostToDevice, c->stream));
        HIPC(launch_build_sdir(c->d_dirv, (const uint32_t*)c->salts.p, 1, D, (uint32_t*)c->gen_sdir.p,
                               c->stream));
        c->sdir_D = -1;  // salts buffer reused: invalidate the pass cache key
        HIPC(launch_sobol_gen((const uint32_t*)c->gen_sdir.p, D, n, (uint32_t)offset, normal, d_out, c->stream));
        HIPC(hipStreamSynchronize(c->stream));
    });
}

fdmc_status fdmc_sobol_scrambled_u32(fdmc_context* c, uint32_t salt, int32_t D, int64_t n, int64_t offset,
                                     uint32_t* d_out) {
    return sobol_out(c, salt, D, n, offset, d_out, false);
}

fdmc_status fdmc_sobol_normal(fdmc_context* c, uint32_t salt, int32_t D, int64_t n, int64_t offset,
                              float* d_out) {
    return sobol_out(c, salt, D, n, offset, d_out, true);
}

fdmc_status fdmc_icdf_u32(fdmc_context* c, const uint32_t* d_in, int64_t n, float* d_out) {
    return guarded([&] {
        if (!c || !d_in || !d_out) fail(FDMC_E_INVALID, "null argument");
        if (n <= 0) fail(FDMC_E_INVALID, "n must be positive");
        HIPC(launch_icdf(d_in, n, d_out, c->stream));
        HIPC(hipStreamSynchronize(c->stream));
    });
}

}  // extern "C"
