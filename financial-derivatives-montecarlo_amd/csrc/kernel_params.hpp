// kernel_params.hpp -- POD launch parameters shared by host code and kernels.
//
// The reference flattens a model and a payoff into one global __constant__
// KernelParams (methods_cuda.cu:218-246), rewritten before every launch.  Here
// each launch carries its own parameters by value (kernel arguments), so
// launches on one stream never race on shared constant memory.
#pragma once

#include "icdf_coeffs.h"

#include <stddef.h>
#include <stdint.h>

namespace fdmc {

// MK_HESTON_MIL is Heston with the Milstein variance term, a kernel variant of its own
// so that the reference's scheme (MK_HESTON) carries no extra instructions.
enum ModelKind : int { MK_GBM = 0, MK_HESTON = 1, MK_DUPIRE = 2, MK_MULTI_DUPIRE = 3, MK_LVQ = 4,
                       MK_MULTI_GBM = 5, MK_HESTON_MIL = 6 };
constexpr bool is_heston(int mk) { return mk == MK_HESTON || mk == MK_HESTON_MIL; }
enum PayoffKind : int { PK_EUROPEAN = 0, PK_ASIAN = 1, PK_GEOM_ASIAN = 2, PK_LOOKBACK = 3,
                        PK_BARRIER = 4, PK_BASKET = 5 };
// Noise sources of the fused path kernels.
// NZ_SOBOL_BB_GLOBAL: a bridge of more than SOBOL_LDS_MAX_D dates, whose
// direction records are read from global memory (its own kernel variant, so
// the LDS-record bridge keeps its register allocation).
enum NoiseKind : int { NZ_SOBOL_RAW = 0, NZ_SOBOL_BB = 1, NZ_PRNG = 2, NZ_ROWS = 3, NZ_SOBOL_BB_GLOBAL = 4 };
// Control-variate leg: none, a second payoff on the same path, or a GBM leg.
enum CtrlKind : int { CV_NONE = 0, CV_SHARED = 1, CV_GBM = 2 };

// One (model, payoff) leg in FP32, as make_params builds it (methods_cuda.cu:986-1051).
struct LegParams {
    float S0, mu, sigma, discount, K, B;
    float theta, xi, L10, L11, v0, kappa;
    float sigma0, alpha, beta_m1, inv_S0;
    float lv_a, lv_b, lv_c, inv_sref;
    float mil_q;   // Heston Milstein coefficient xi^2/4, 0 for the reference's scheme
};

// Scrambled Sobol direction record of one (replica, dimension), 64 words:
//  [0..7]   TA[k] = shift ^ XOR_{j<3, bit j of k} Lv[j]   (lane Gray bits 0..2)
//  [8..15]  TB[k] = XOR_{j<3, bit j of k} Lv[j+3]         (lane Gray bits 3..5)
//  [16..19] WV[w] = XOR_{j in gray(64 w)} Lv[j]           (wave index within 256)
//  [32..63] Lv[0..31]
// so that for point index i = base + 64 w + lane (base a multiple of 256):
//  scrambled(i) = TA[g & 7] ^ TB[g >> 3] ^ WV[w] ^ XOR_{j in gray(base)} Lv[j],
//  g = gray(lane).
constexpr int SDIR_WORDS = 64;

constexpr int BB_SD_DEPTHS = 16;   // a uniform bridge of at most 2^15 steps has at most 16 depths
// Entries of the skeleton walk's per-lane stack over G = 2^m blocks: W at the
// right ends of aligned 2^0 .. 2^m-block intervals.
inline int bb_walk_levels(int G) {
    int m = 0;
    while ((1 << m) < G) m++;
    return m + 1;
}

struct PathLaunch {
    LegParams leg[2];       // main, control
    float h, sqrt_h, em;    // step, sqrt(step), e^{-kappa h}
    float bgk[2];           // exp(-0.5826 sigma_bgk sqrt(h)) per leg (1 when unused)
    float beta_cv, E_ctrl;
    float T;                // maturity (rho of the discount)
    int n_steps, nf, D;
    int G;                  // BB: coarse blocks (N / 16, or N when N < 16)
    uint32_t seed, tag;     // PRNG stream key
    const uint32_t* sdir;   // [R][D][SDIR_WORDS]
    const float* bbw;       // BB weights [N][4] = {wl, wr, sd, 0} in bb_precompute order
    int bb_uniform;         // BB on a uniform date grid: every weight 1/2 in FP32, one sd per depth
    float bb_sd[BB_SD_DEPTHS];  // then: [0] terminal sd, [k] sd of the midpoints of schedule level k - 1
    const float* L;         // basket: n x n lower Cholesky (row-major) or null
    const float* rows;      // NZ_ROWS: precomputed increments [R * count][ld] (PCA)
    int ld;
    int64_t first;          // first point index (inclusive)
    int64_t count;          // points per replica
    int tiles_per_wg;       // 256-point tiles per workgroup
    int nwg_x;              // workgroups per replica
    double* partials;       // [R][nwg_x][NACC]
};

// Coupled multilevel launch (MLMC / MLQMC): one entry per (level, replica)
// request; entries are cut into work items of tpi 256-sample tiles that a
// persistent grid drains through an atomic counter.  Item t of entry e writes
// its sums to partials[tile_begin + t] so the result does not depend on
// scheduling.
struct LevelEntry {
    int level, n_fine, n_coarse, D, G;
    int tile_begin, ntiles, tpi;   // first work item, work items, 256-sample tiles per item
    int64_t first, count;   // sample (PRNG path id / Sobol point index) range
    float h_f, h_c, sqrt_hf, em_f, em_c, bgk_f, bgk_c;
    int ld;                 // row stride of `rows`
    uint32_t seed, tag;     // PRNG stream (MLMC)
    const uint32_t* sdir;   // Sobol records of this (level, replica), or null
    const float* bbw;       // BB weights for N = n_fine, or null
    const float* rows;      // NZ_ROWS: fine increments of sample first + i at rows[i * ld], or null
};

struct LevelsLaunch {
    LegParams leg;
    int M, nf, n_entries, total_items;
    const LevelEntry* entries;  // device array, sorted by decreasing cost
    unsigned int* queue;        // work counter (zeroed before the launch)
    double* partials;           // [total_items][4], one row per work item
};

// Multi-asset basket launch (k_basket.hip).
enum BasketSrc : int { SRC_ROWS = 0, SRC_SOBOL = 1, SRC_PRNG = 2, SRC_PROD = 3 };
constexpr int BASKET_MAX = 32;

// Product epilogue of the PCA GEMM for baskets of GBM assets (k_gemm_nt256):
// the Euler factor of step k and asset a is c0 + sig[a] dW[row][k n + a]
// (c0 = 1 + mu h), and S_T^a = S0^a times the product of all of them, in any
// order; each 128-column slab (128 / n dates) leaves its product per row and
// asset in P[slab][row][a].  n is a power of two in [2, 32]; columns >= D
// (padding) contribute 1.
struct ProdEpi {
    float c0;
    float sig[BASKET_MAX];
    int n, D;
    float* P;                 // [ld / 128][Mrows][n]
};

struct BasketLaunch {
    int n, n_steps, D, greeks;
    float h, sqrt_h, mu, K, discount;
    float sigma0, alpha, beta_m1, inv_S0;     // MULTI_DUPIRE (vol normalised by the first S0)
    float S0[BASKET_MAX], sig[BASKET_MAX], w[BASKET_MAX];
    const float* L;           // n x n lower Cholesky (row-major) or null
    const float* dW;          // SRC_ROWS: [R * count][ld]; SRC_PROD: the GEMM's products [npc][prow][n]
    int ld;                   // row stride of dW (D rounded up to 128)
    int npc;                  // SRC_PROD: 128-column slabs (ld / 128)
    int64_t prow;             // SRC_PROD: rows per slab
    const uint32_t* sdir;     // SRC_SOBOL: [R][D][SDIR_WORDS]
    uint32_t seed, tag;       // SRC_PRNG
    int64_t first, count;
    int tiles_per_wg, nwg_x;
    double* partials;         // [R][nwg_x][2 + 2 n]
};

// Baskets of more than BASKET_MAX assets (k_basket.hip, k_basket_wide): per-asset
// parameters in device arrays.
constexpr int WIDE_WG = 128;        // lanes (paths) per workgroup (uncorrelated; correlated: 64)
constexpr int WIDE_CORR_MAX = 256;  // correlated: 16 column blocks of 16 assets in k_basket_corr
struct BasketWide {
    int n, n_steps, D;
    float h, sqrt_h, mu, K, discount;
    float sigma0, alpha, beta_m1, inv_S0;     // MULTI_DUPIRE
    const float* S0;          // [n]
    const float* sig;         // [n] (MULTI_GBM)
    const float* w;           // [n] weights
    const float* L;           // n x n lower Cholesky (row-major) or null
    const uint32_t* sdir;     // SRC_SOBOL: [R][D][SDIR_WORDS]
    uint32_t seed, tag;       // SRC_PRNG
    int64_t first, count;
    int tiles_per_wg, nwg_x;  // tiles of one workgroup's paths
    double* partials;         // [R][nwg_x][2]
};

// Importance-sampled GBM European call (k_is.hip).
struct ISLaunch {
    float S0, mu, sigma, h, sqrt_h, K, discount, zs;   // zs: per-step drift shift z* / sqrt(n)
    int n_steps, tiles_per_wg;
    uint32_t seed, tag;       // Philox stream
    int64_t count;            // paths 0 .. count - 1
    double* partials;         // [nwg][2]
};

// Static LDS of the inverse-CDF table (icdf_u32_tab) in every kernel that
// draws normals from Sobol points, and what that leaves of a CU's 160 KiB for
// a launch's dynamic LDS.
constexpr size_t ICDF_TAB_LDS_BYTES = (size_t)FDMC_ICDF_TAB_ROWS * 16;
constexpr size_t LDS_DYNAMIC_MAX = 160 * 1024 - ICDF_TAB_LDS_BYTES;

// Sobol direction records are staged in LDS up to this dimension; above it the
// path kernels read them from global memory (L2-resident).
constexpr int SOBOL_LDS_MAX_D = 1024;

// Dynamic LDS bytes of one k_path launch (layout documented in k_path.inl).
inline size_t path_lds_bytes(int nz, int D, int G) {
    size_t b = 0;
    if (nz != NZ_PRNG && nz != NZ_ROWS && nz != NZ_SOBOL_BB_GLOBAL && D <= SOBOL_LDS_MAX_D) b += (size_t)D * 20 * 4;
    if (nz == NZ_SOBOL_BB || nz == NZ_SOBOL_BB_GLOBAL)   // skeleton or walk stack
        b += (size_t)(D < 16 ? G + 1 : bb_walk_levels(G)) * 256 * 4;
    b = (b + 15) & ~(size_t)15;
    return b + 4 * 7 * sizeof(double);
}

// Dynamic LDS bytes of one k_greeks launch: k_path's layout plus one S
// checkpoint per lane and 16-step block.
inline size_t greeks_lds_bytes(int nz, int D, int G, int n_steps) {
    const int Gk = (nz == NZ_SOBOL_BB) ? G : (n_steps + 15) / 16;
    size_t b = 0;
    if (nz != NZ_PRNG && D <= SOBOL_LDS_MAX_D) b += (size_t)D * 20 * 4;
    if (nz == NZ_SOBOL_BB) b += (size_t)(G + 1) * 256 * 4;
    b += (size_t)Gk * 256 * 4;
    b = (b + 15) & ~(size_t)15;
    return b + 4 * 13 * sizeof(double);
}

// Dynamic LDS bytes of one k_levels launch (max over its entries).
inline size_t levels_lds_bytes(int nz, int max_D, int skel_n) {
    size_t b = 0;
    if (nz != NZ_PRNG && nz != NZ_ROWS && max_D <= SOBOL_LDS_MAX_D) b += (size_t)max_D * 20 * 4;
    if (nz == NZ_SOBOL_BB) b += (size_t)skel_n * 256 * 4;
    b = (b + 15) & ~(size_t)15;
    return b + 4 * 4 * sizeof(double) + 16;
}

// ---- contract book (k_book.hip) ---------------------------------------------
// Level l of a contract simulates 2^l fine steps; at most BOOK_NL levels.
constexpr int BOOK_NL = 10;
constexpr int BOOK_DMAX = 1 << (BOOK_NL - 1);     // 512 dates
constexpr int BOOK_GMAX = BOOK_DMAX / 16;         // 16-step blocks of the finest level
constexpr int BOOK_NACC = 4;                      // P_f - P_c, delta, vega, rho
constexpr int BOOK_MAX_ITEMS = 128;               // work items per entry and pass
constexpr int BOOK_MAX_T = 8;                     // distinct maturities (bridge tables)
constexpr int BOOK_VT_STRIDE = 2 * BOOK_DMAX - 1; // time factors of levels 0..9 per maturity
constexpr int BOOK_VT_PLANE = BOOK_MAX_T * BOOK_VT_STRIDE;   // planes: factors, x 1/16, x 1/8

struct BookContract {
    int kind;                 // PK_EUROPEAN / PK_ASIAN / PK_LOOKBACK / PK_BARRIER
    int L;                    // finest level (n_steps = 2^L)
    int t_idx;                // maturity slot of the bridge table
    int entry0;               // entry id of (l, r) is entry0 + l R + r
    float K, B, discount, Tf;
    double T;
    float bgk[BOOK_NL];       // BGK factor exp(-0.5826 sigma sqrt(T / 2^l)), 1 when unused
    int slot[BOOK_NL];        // index among the contracts with L >= l (Sobol records of level l)
};

struct BookEntry {            // one (contract, level, replica) of a pass
    int c, l, r, tpi;         // tpi: 256-point tiles per work item
    int tile_begin, ntiles;   // first work item, work items
    int64_t first, count;     // Sobol point range of this pass
};

struct BookState {            // per (contract, level), merged as run_mlqmc_cuda
    double E[BOOK_NACC], S2[BOOK_NACC];
    int64_t N, extra;
};

struct BookCtl {              // per contract
    int iter, done, conv, pad;
};

struct BookOut {              // per contract, read back by the host
    double v[8];              // price, se, delta, vega, rho, their se
    int64_t n_samples, path_steps;
    int iterations, flags;
};

struct BookLaunch {
    LegParams leg;            // model (K, B, discount overwritten per contract)
    int R, c_begin, n_contracts, n_entries;
    int pilot_n, max_iter, greeks, items_target;
    double eps;
    const BookContract* contracts;
    BookEntry* entries;       // this launch's entry array (batched: level-major; plain: one contract)
    const int* entry_pos;     // entry id -> position in entries
    BookState* state;         // [contracts][BOOK_NL]
    BookCtl* ctl;             // [contracts]
    const uint32_t* sdir[BOOK_NL];  // level l: [slots][R][2^l][SDIR_WORDS]
    const float* bbtab;       // bridge sd per depth: [maturities][BOOK_NL levels][BOOK_NL depths]
    const float* voltab;      // local-vol time factor per fine step: [3 planes][maturities][BOOK_VT_STRIDE]
    unsigned int* queue;
    int* total_items;
    int* active;
    double* partials;         // [items][BOOK_NACC]
    double* entry_sums;       // [n_entries][BOOK_NACC]
    float* scratch;           // per workgroup: checkpoints [2][BOOK_GMAX][256]
    BookOut* out;
};

// LDS of one book workgroup: direction records [BOOK_DMAX][20] words, the
// inverse-CDF table, the block reduction and the work-item slot.
constexpr size_t book_lds_static() {
    return (size_t)BOOK_DMAX * 20 * 4 + (size_t)FDMC_ICDF_TAB_ROWS * 16 + 4 * BOOK_NACC * 8 + 16;
}
static_assert(book_lds_static() <= 160 * 1024, "book LDS footprint exceeds a CU's LDS");

// Global scratch of one book workgroup: the adjoint's checkpoints of S (fine
// and coarse leg) at the block starts, [2][BOOK_GMAX][256] floats.
constexpr size_t book_scratch_floats() { return (size_t)2 * BOOK_GMAX * 256; }

}  // namespace fdmc
