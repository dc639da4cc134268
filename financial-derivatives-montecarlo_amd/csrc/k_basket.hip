// k_basket.hip -- multi-asset basket paths (reference kernel_multi_dupire,
// methods_cuda.cu:526-551, generalised to per-asset GBM and weights).
//
// One lane = one path with all NA asset prices in registers, stepped in time
// order (the reference loops assets outside, steps inside, re-reading dW with a
// stride of N paths).  Increments come from
//   SRC_ROWS : dW rows [row][k * n + a] produced by the PCA GEMM (FP32), or
//   SRC_SOBOL/SRC_PRNG : normals generated in the lane, scaled by sqrt(h) and
//              correlated with the lower Cholesky factor L (dW_k = L z_k sqrt h,
//              the batched GEMM of correlate_multi_dupire, :1133-1148), or
//   SRC_PROD : GBM assets through the PCA GEMM's product epilogue: the Euler
//              step S <- S (1 + mu h + sig dW) multiplies, so S_T^a is S0^a
//              times the products of factors the GEMM left per 128-column slab
//              (dW itself never reaches memory).
// Payoff max(sum_a w_a S_T^a - K, 0) e^{-rT}.  For GBM assets the pathwise
// adjoint of the Euler scheme gives dS_T^a / dS0^a = S_T^a / S0^a, hence the
// per-asset deltas disc 1{B > K} w_a S_T^a / S0^a accumulated beside the price.
#include "k_path.inl"

namespace fdmc {

template <int NA, int MK, int SRC>
__global__ __launch_bounds__(256) void k_basket(BasketLaunch b) {
    constexpr int NACC = 2 + 2 * NA;
    extern __shared__ __align__(16) unsigned char smem[];
    const int tid = threadIdx.x, lane = tid & 63;
    const int r = blockIdx.y;
    const int n = b.n, D = b.D;
    // inverse-CDF table (icdf_u32_tab), filled before the first barrier-separated use
    constexpr bool ITAB = SRC == SRC_SOBOL;
    __shared__ __align__(16) float4 itab[ITAB ? FDMC_ICDF_TAB_ROWS : 1];
    if constexpr (ITAB) {
        icdf_tab_load(itab, threadIdx.x, 256);
        __syncthreads();
    }
    const bool lds_mode = (SRC == SRC_SOBOL) && D <= SOBOL_LDS_MAX_D;
    uint32_t* rec = (uint32_t*)smem;
    double* red = (double*)(smem + ((size_t)(lds_mode ? D * 20 : 0) * 4 + 15) / 16 * 16);
    const uint32_t* grec = (SRC == SRC_SOBOL) ? b.sdir + (int64_t)r * D * SDIR_WORDS : nullptr;
    SobolLane sl;
    {
        const uint32_t g = (uint32_t)(lane ^ (lane >> 1));
        sl.rec = rec; sl.grec = grec; sl.ga = g & 7u; sl.gb = g >> 3; sl.w = (uint32_t)(tid >> 6);
        sl.hi_gray = 0; sl.lds_mode = lds_mode;
    }
    if (SRC == SRC_SOBOL && lds_mode) {
        for (int i = tid; i < D * 16; i += 256) {
            const int d = i >> 4, k = i & 15;
            rec[d * 20 + k] = grec[(int64_t)d * SDIR_WORDS + k];
        }
    }
    double acc[NACC];
#pragma unroll
    for (int i = 0; i < NACC; i++) acc[i] = 0.0;
    const int64_t last = b.first + b.count;
    const int64_t tile0 = (b.first >> 8) + (int64_t)blockIdx.x * b.tiles_per_wg;
    for (int it = 0; it < b.tiles_per_wg; it++) {
        const int64_t base = (tile0 + it) << 8;
        if (base >= last) break;
        if constexpr (SRC == SRC_SOBOL) {
            const uint32_t b32 = (uint32_t)base;
            if (lds_mode) {
                __syncthreads();
                sobol_tile_words(rec, grec, D, (uint32_t)(tile0 + it), it == 0, tid);
                __syncthreads();
            } else {
                const uint32_t hi = b32 + (sl.w << 6);
                sl.hi_gray = hi ^ (hi >> 1);
            }
        }
        const int64_t idx = base + tid;
        const bool valid = idx >= b.first && idx < last;
        const int64_t row = (int64_t)r * b.count + (valid ? idx - b.first : 0);
        float S[NA];
#pragma unroll
        for (int a = 0; a < NA; a++) S[a] = b.S0[a];
        if constexpr (SRC == SRC_PROD) {
            for (int pc = 0; pc < b.npc; pc++) {
                const float* src = b.dW + ((int64_t)pc * b.prow + row) * NA;   // n == NA here
#pragma unroll
                for (int a = 0; a < NA; a++) S[a] *= src[a];
            }
        } else
        for (int k = 0; k < b.n_steps; k++) {
            float dw[NA];
            if constexpr (SRC == SRC_ROWS) {
                const float* src = b.dW + row * b.ld + (int64_t)k * n;
#pragma unroll
                for (int a = 0; a < NA; a++) dw[a] = (a < n) ? src[a] : 0.0f;
            } else {
                float z[NA];
                if constexpr (SRC == SRC_SOBOL) {
#pragma unroll
                    for (int a = 0; a < NA; a++) z[a] = (a < n) ? icdf_u32_tab<true>(sobol_coord(sl, k * n + a), itab) * b.sqrt_h : 0.0f;
                } else {
#pragma unroll
                    for (int a0 = 0; a0 < NA; a0 += 4) {
                        float q[4];
                        const int j = k * n + a0;   // normal index; blocks of 4 (n % 4 == 0 checked)
                        prng_normals4(b.seed + (uint32_t)r, b.tag, (uint64_t)idx, (uint32_t)(j >> 2), q);
#pragma unroll
                        for (int c = 0; c < 4; c++) z[a0 + c] = (a0 + c < n) ? q[c] * b.sqrt_h : 0.0f;
                    }
                }
                if (b.L) {
#pragma unroll
                    for (int a = 0; a < NA; a++) {
                        float s = 0.0f;
#pragma unroll
                        for (int c = 0; c <= a; c++) s += b.L[a * n + c] * z[c];
                        dw[a] = (a < n) ? s : 0.0f;
                    }
                } else {
#pragma unroll
                    for (int a = 0; a < NA; a++) dw[a] = z[a];
                }
            }
            const float t = (float)k * b.h;
#pragma unroll
            for (int a = 0; a < NA; a++) {
                if constexpr (MK == MK_MULTI_GBM) {
                    S[a] = S[a] + b.mu * S[a] * b.h + b.sig[a] * S[a] * dw[a];
                } else {
                    const float sl_ = b.sigma0 * fast_exp(-b.alpha * t) * fast_exp2(b.beta_m1 * fast_log2(S[a] * b.inv_S0));
                    S[a] = S[a] + b.mu * S[a] * b.h + sl_ * S[a] * dw[a];
                }
            }
        }
        float B = 0.0f;
#pragma unroll
        for (int a = 0; a < NA; a++) B += (a < n) ? b.w[a] * S[a] : 0.0f;
        const float y = fmaxf(B - b.K, 0.0f) * b.discount;
        if (valid) {
            acc[0] += (double)y;
            acc[1] += (double)y * (double)y;
            if (b.greeks && MK == MK_MULTI_GBM) {
                const float ind = (B > b.K) ? b.discount : 0.0f;
#pragma unroll
                for (int a = 0; a < NA; a++) {
                    const double g = (double)(ind * b.w[a] * S[a] / b.S0[a]);
                    acc[2 + 2 * a] += g;
                    acc[3 + 2 * a] += g * g;
                }
            }
        }
    }
    __syncthreads();
    block_sum<NACC, 4>(acc, red);
    if (tid == 0) {
        double* out = b.partials + ((int64_t)r * b.nwg_x + blockIdx.x) * NACC;
#pragma unroll
        for (int i = 0; i < NACC; i++) out[i] = acc[i];
    }
}

// ---- baskets of more than BASKET_MAX assets ----------------------------------------
// Reference examples 07 (1000 uncorrelated local-vol assets, D = 20000) and 08
// (100 correlated assets).  A lane is one path; normal j = k n + a of a path is
// Sobol coordinate j of its point (direction records read from global memory,
// L2/MALL resident) or Philox normal j of the path.
//  - uncorrelated (128-lane workgroups): assets one after the other -- an asset's
//    path needs only its own increments, so a lane carries only the basket sum;
//  - correlated (n <= WIDE_CORR_MAX): k_basket_corr below, dates one after the
//    other with dW = Z L^T on FP32 MFMA.
template <int MK>
__device__ __forceinline__ float wide_step(const BasketWide& b, float S, float sg, float dw, int k) {
    if constexpr (MK == MK_MULTI_GBM) {
        return S + b.mu * S * b.h + sg * S * dw;
    } else {
        const float t = (float)k * b.h;
        const float sl = b.sigma0 * fast_exp(-b.alpha * t) * fast_exp2(b.beta_m1 * fast_log2(S * b.inv_S0));
        return S + b.mu * S * b.h + sl * S * dw;
    }
}

template <int MK, int SRC>
__global__ __launch_bounds__(WIDE_WG) void k_basket_wide(BasketWide b) {
    constexpr int WG = WIDE_WG;
    const int tid = threadIdx.x, lane = tid & 63;
    const int r = blockIdx.y;
    const int n = b.n, D = b.D;
    constexpr bool ITAB = SRC == SRC_SOBOL;
    __shared__ __align__(16) float4 itab[ITAB ? FDMC_ICDF_TAB_ROWS : 1];
    if constexpr (ITAB) icdf_tab_load(itab, tid, WG);
    __shared__ double red[(WG / 64) * 2];
    __syncthreads();
    SobolLane sl;
    {
        const uint32_t g = (uint32_t)(lane ^ (lane >> 1));
        sl.rec = nullptr;
        sl.grec = (SRC == SRC_SOBOL) ? b.sdir + (int64_t)r * D * SDIR_WORDS : nullptr;
        sl.ga = g & 7u; sl.gb = g >> 3; sl.w = (uint32_t)(tid >> 6);
        sl.hi_gray = 0; sl.lds_mode = false;
    }
    double acc[2] = {0.0, 0.0};
    const int64_t last = b.first + b.count;
    const int64_t tile0 = (b.first / WG) + (int64_t)blockIdx.x * b.tiles_per_wg;
    for (int it = 0; it < b.tiles_per_wg; it++) {
        const int64_t base = (tile0 + it) * WG;
        if (base >= last) break;   // the same for every lane
        const int64_t idx = base + tid;
        const bool valid = idx >= b.first && idx < last;
        if constexpr (SRC == SRC_SOBOL) {
            const uint32_t hi = (uint32_t)base + (sl.w << 6);
            sl.hi_gray = hi ^ (hi >> 1);
        }
        float B = 0.0f;
        if constexpr (SRC == SRC_SOBOL) {
            for (int a = 0; a < n; a++) {
                float S = b.S0[a];
                const float sg = (MK == MK_MULTI_GBM) ? b.sig[a] : 0.0f;
                for (int k = 0; k < b.n_steps; k++) {
                    const float z = icdf_u32_tab<true>(sobol_coord(sl, k * n + a), itab);
                    S = wide_step<MK>(b, S, sg, z * b.sqrt_h, k);
                }
                B += b.w[a] * S;
            }
        } else {
            // four assets at a time: normals k n + a0 .. k n + a0 + 3 of a date are
            // consecutive, so one Philox block (two when k n + a0 is not a multiple
            // of 4; the offset is the same for every lane) serves all four
            for (int a0 = 0; a0 < n; a0 += 4) {
                float S[4], sg[4];
#pragma unroll
                for (int t = 0; t < 4; t++) {
                    const int a = min(a0 + t, n - 1);
                    S[t] = b.S0[a];
                    sg[t] = (MK == MK_MULTI_GBM) ? b.sig[a] : 0.0f;
                }
                for (int k = 0; k < b.n_steps; k++) {
                    const int j0 = k * n + a0, off = j0 & 3;
                    float g0[4];
                    prng_normals4(b.seed + (uint32_t)r, b.tag, (uint64_t)idx, (uint32_t)(j0 >> 2), g0);
                    if (off == 0) {          // aligned (always so when n is a multiple of 4): no selection
#pragma unroll
                        for (int t = 0; t < 4; t++) S[t] = wide_step<MK>(b, S[t], sg[t], g0[t] * b.sqrt_h, k);
                    } else {
                        float g1[4];
                        prng_normals4(b.seed + (uint32_t)r, b.tag, (uint64_t)idx, (uint32_t)(j0 >> 2) + 1u, g1);
#pragma unroll
                        for (int t = 0; t < 4; t++) {
                            const int u = off + t;
                            float z = g0[0];
                            z = (u == 1) ? g0[1] : z;
                            z = (u == 2) ? g0[2] : z;
                            z = (u == 3) ? g0[3] : z;
                            z = (u == 4) ? g1[0] : z;
                            z = (u == 5) ? g1[1] : z;
                            z = (u == 6) ? g1[2] : z;
                            S[t] = wide_step<MK>(b, S[t], sg[t], z * b.sqrt_h, k);
                        }
                    }
                }
#pragma unroll
                for (int t = 0; t < 4; t++)
                    if (a0 + t < n) B += b.w[a0 + t] * S[t];
            }
        }
        const float y = fmaxf(B - b.K, 0.0f) * b.discount;
        if (valid) {
            acc[0] += (double)y;
            acc[1] += (double)y * (double)y;
        }
    }
    block_sum<2, WG / 64>(acc, red);
    if (tid == 0) {
        double* out = b.partials + ((int64_t)r * b.nwg_x + blockIdx.x) * 2;
        out[0] = acc[0];
        out[1] = acc[1];
    }
}

// ---- correlated baskets of more than BASKET_MAX assets on FP32 MFMA ---------
// The reference correlates a date's normals with a batched GEMM dW = Z L^T
// (correlate_multi_dupire, methods_cuda.cu:1133-1148).  Here each wave takes 16
// paths and, per date, forms their increments with v_mfma_f32_16x16x4_f32 (FP32
// operands, FP32 accumulation): A = the paths' scaled normals, B = L^T, in
// asset blocks of 16.  The paths' asset prices stay in the accumulator layout
// (lane l: asset 16 j + l % 16 of paths 4 (l / 16) + i, i = 0..3), so the Euler
// step is elementwise on the MFMA output and nothing per path touches memory.
// Only the block pairs (q, j) with q <= j carry work (L is lower triangular);
// the B fragments of all pairs sit in LDS as one float4 per lane (the four k
// steps of one 16-asset block).
//
// The A operand of one MFMA holds, in lane l, row (path) l % 16 and k slot
// l / 16.  The k order inside an asset block q is permuted so that a lane's
// four k steps t = 0..3 are the consecutive assets 16 q + 4 (l / 16) + t: one
// Philox call (four normals of consecutive index) or four consecutive Sobol
// dimensions per lane and block.  B follows the same order.
constexpr int CORR_TILE = 64;      // paths per tile: 4 waves x 16

typedef float f32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ int corr_pair(int q, int j, int NB) { return q * NB - ((q * (q - 1)) >> 1) + (j - q); }

template <int MK, int SRC, int NB>
__global__ __launch_bounds__(256) void k_basket_corr(BasketWide b) {
    extern __shared__ __align__(16) unsigned char smem[];
    float4* lt = (float4*)smem;                      // [pair][64 lanes]: L^T B fragments for t = 0..3
    constexpr bool ITAB = SRC == SRC_SOBOL;
    __shared__ __align__(16) float4 itab[ITAB ? FDMC_ICDF_TAB_ROWS : 1];
    if constexpr (ITAB) icdf_tab_load(itab, threadIdx.x, 256);
    __shared__ double red[4 * 2];
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int col = lane & 15, ks = lane >> 4;
    const int r = blockIdx.y;
    // NB = ceil(n / 16) exactly (the host picks the instantiation): the block
    // loops below have compile-time bounds, with no branch between the MFMAs
    // (uniform early exits there made the compiler shuttle the accumulators
    // between AGPRs and VGPRs at every join, about ten moves per MFMA)
    const int n = b.n;
    constexpr int nb = NB;
    // B fragments: lane (ks, col) of pair (q, j) holds L[16 j + col][16 q + 4 ks + t]
    // for t = 0..3, zero above the diagonal and outside n
    constexpr int NPAIRS = NB * (NB + 1) / 2;
    for (int e = tid; e < NPAIRS * 64; e += 256) {
        int p = e >> 6, q = 0;
        while (p >= NB - q) { p -= NB - q; q++; }
        const int j = q + p, ln = e & 63;
        const int arow = 16 * j + (ln & 15), c0 = 16 * q + 4 * (ln >> 4);
        float v[4];
#pragma unroll
        for (int t = 0; t < 4; t++) {
            const int c = c0 + t;
            v[t] = (arow < n && c <= arow) ? b.L[(int64_t)arow * n + c] : 0.0f;
        }
        lt[e] = make_float4(v[0], v[1], v[2], v[3]);
    }
    __syncthreads();
    const uint32_t* grec = (SRC == SRC_SOBOL) ? b.sdir + (int64_t)r * b.D * SDIR_WORDS : nullptr;
    double acc[2] = {0.0, 0.0};
    const int64_t last = b.first + b.count;
    const int64_t tile0 = (b.first / CORR_TILE) + (int64_t)blockIdx.x * b.tiles_per_wg;
    for (int it = 0; it < b.tiles_per_wg; it++) {
        const int64_t base = (tile0 + it) * CORR_TILE;
        if (base >= last) break;                     // the same for every lane
        const int64_t wbase = base + 16 * wv;
        // the A-layout path of this lane: row col
        const uint64_t pa = (uint64_t)(wbase + col);
        SobolLane sl;
        {
            const uint32_t x = (uint32_t)pa, g = x ^ (x >> 1);
            sl.rec = nullptr; sl.grec = grec; sl.ga = g & 7u; sl.gb = (g >> 3) & 7u; sl.w = 0;
            sl.hi_gray = g & ~63u; sl.lds_mode = false;
        }
        f32x4 S[NB];
#pragma unroll
        for (int j = 0; j < NB; j++) {
            const int a = 16 * j + col;
            const float s0 = a < n ? b.S0[a] : 1.0f;      // outside n: a finite dummy with weight 0
            S[j] = f32x4{s0, s0, s0, s0};
        }
        for (int k = 0; k < b.n_steps; k++) {
            f32x4 dw[NB];
#pragma unroll
            for (int j = 0; j < NB; j++) dw[j] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
            const int jk = k * n;
#pragma unroll
            for (int q = 0; q < NB; q++) {
                if (q >= nb) break;                  // uniform
                const int c0 = 16 * q + 4 * ks;      // this lane's four assets c0 .. c0 + 3
                float za[4] = {0.0f, 0.0f, 0.0f, 0.0f};
                if (c0 < n) {
                    if constexpr (SRC == SRC_SOBOL) {
#pragma unroll
                        for (int t = 0; t < 4; t++)
                            if (c0 + t < n) za[t] = icdf_u32_tab<true>(sobol_coord_glb(sl, jk + c0 + t), itab) * b.sqrt_h;
                    } else {
                        // normals jk + c0 + t of the path, in groups of four from
                        // index 4 g; the offset jk & 3 is the same for every lane
                        // (c0 is a multiple of 4), so it is computed from uniform
                        // values only and its test is a scalar branch: aligned
                        // (always so when n is a multiple of 4) one call and no
                        // selection
                        const int j0 = jk + c0, off = jk & 3;
                        float g0[4];
                        prng_normals4(b.seed + (uint32_t)r, b.tag, pa, (uint32_t)(j0 >> 2), g0);
                        if (off == 0) {
#pragma unroll
                            for (int t = 0; t < 4; t++)
                                if (c0 + t < n) za[t] = g0[t] * b.sqrt_h;
                        } else {
                            float g1[4];
                            prng_normals4(b.seed + (uint32_t)r, b.tag, pa, (uint32_t)(j0 >> 2) + 1u, g1);
#pragma unroll
                            for (int t = 0; t < 4; t++) {
                                const int u = off + t;
                                float z = g0[0];
                                z = (u == 1) ? g0[1] : z;
                                z = (u == 2) ? g0[2] : z;
                                z = (u == 3) ? g0[3] : z;
                                z = (u == 4) ? g1[0] : z;
                                z = (u == 5) ? g1[1] : z;
                                z = (u == 6) ? g1[2] : z;
                                if (c0 + t < n) za[t] = z * b.sqrt_h;
                            }
                        }
                    }
                }
                float4 bf[NB];
#pragma unroll
                for (int j = q; j < NB; j++)
                    if (j < nb) bf[j] = lt[corr_pair(q, j, NB) * 64 + lane];
                // k step outside, blocks inside: consecutive MFMAs are independent
#pragma unroll
                for (int t = 0; t < 4; t++)
#pragma unroll
                    for (int j = q; j < NB; j++) {
                        if (j >= nb) break;          // uniform
                        const float bv = t == 0 ? bf[j].x : t == 1 ? bf[j].y : t == 2 ? bf[j].z : bf[j].w;
                        dw[j] = __builtin_amdgcn_mfma_f32_16x16x4f32(za[t], bv, dw[j], 0, 0, 0);
                    }
                // block q's normals are drawn next to its MFMAs: hoisting every
                // block's Philox draws ahead of the products (the scheduler's
                // choice) doubles the registers and halves the waves per SIMD
                __builtin_amdgcn_sched_barrier(0);
            }
            // Euler step of every (path, asset) the lane holds (kernel_multi_dupire's
            // S + mu S h + sigma S dW, methods_cuda.cu:526-551)
            float eterm = 0.0f;
            if constexpr (MK != MK_MULTI_GBM) eterm = b.sigma0 * fast_exp(-b.alpha * ((float)k * b.h));
#pragma unroll
            for (int j = 0; j < NB; j++) {
                if (j >= nb) break;
                const int a = 16 * j + col;
                const float sg = (MK == MK_MULTI_GBM && a < n) ? b.sig[a] : 0.0f;
#pragma unroll
                for (int i = 0; i < 4; i++) {
                    const float Sv = S[j][i], d = dw[j][i];
                    if constexpr (MK == MK_MULTI_GBM) {
                        S[j][i] = Sv + b.mu * Sv * b.h + sg * Sv * d;
                    } else {
                        const float sl_ = eterm * fast_exp2(b.beta_m1 * fast_log2(Sv * b.inv_S0));
                        S[j][i] = Sv + b.mu * Sv * b.h + sl_ * Sv * d;
                    }
                }
            }
        }
        // basket value per path: the lane's column, then the 16 lanes of its row group
#pragma unroll
        for (int i = 0; i < 4; i++) {
            float B = 0.0f;
#pragma unroll
            for (int j = 0; j < NB; j++) {
                const int a = 16 * j + col;
                if (j < nb) B += (a < n ? b.w[a] : 0.0f) * S[j][i];
            }
#pragma unroll
            for (int o = 1; o < 16; o <<= 1) B += __shfl_xor(B, o, 64);
            const int64_t idx = wbase + 4 * ks + i;
            if (col == 0 && idx >= b.first && idx < last) {
                const float y = fmaxf(B - b.K, 0.0f) * b.discount;
                acc[0] += (double)y;
                acc[1] += (double)y * (double)y;
            }
        }
    }
    block_sum<2, 4>(acc, red);
    if (tid == 0) {
        double* out = b.partials + ((int64_t)r * b.nwg_x + blockIdx.x) * 2;
        out[0] = acc[0];
        out[1] = acc[1];
    }
}

// Column blocks of 16 assets the correlated kernel is compiled for: exactly
// ceil(n / 16), 3..16 for the 33..256 assets of this path.
static int corr_nb(int n) { return (n + 15) >> 4; }

template <int MK, int SRC>
static const void* pick_corr(int n) {
    switch (corr_nb(n)) {
    case 3: return (const void*)&k_basket_corr<MK, SRC, 3>;
    case 4: return (const void*)&k_basket_corr<MK, SRC, 4>;
    case 5: return (const void*)&k_basket_corr<MK, SRC, 5>;
    case 6: return (const void*)&k_basket_corr<MK, SRC, 6>;
    case 7: return (const void*)&k_basket_corr<MK, SRC, 7>;
    case 8: return (const void*)&k_basket_corr<MK, SRC, 8>;
    case 9: return (const void*)&k_basket_corr<MK, SRC, 9>;
    case 10: return (const void*)&k_basket_corr<MK, SRC, 10>;
    case 11: return (const void*)&k_basket_corr<MK, SRC, 11>;
    case 12: return (const void*)&k_basket_corr<MK, SRC, 12>;
    case 13: return (const void*)&k_basket_corr<MK, SRC, 13>;
    case 14: return (const void*)&k_basket_corr<MK, SRC, 14>;
    case 15: return (const void*)&k_basket_corr<MK, SRC, 15>;
    case 16: return (const void*)&k_basket_corr<MK, SRC, 16>;
    default: return nullptr;
    }
}

const void* basket_wide_kernel(int mk, int src, bool corr, int n) {
#define W(MKv, SRCv)                                                                              \
    if (mk == MKv && src == SRCv)                                                                 \
        return corr ? pick_corr<MKv, SRCv>(n) : (const void*)&k_basket_wide<MKv, SRCv>;
    W(MK_MULTI_GBM, SRC_SOBOL) W(MK_MULTI_GBM, SRC_PRNG) W(MK_MULTI_DUPIRE, SRC_SOBOL) W(MK_MULTI_DUPIRE, SRC_PRNG)
#undef W
    return nullptr;
}

int basket_wide_block(bool corr) { return corr ? 256 : WIDE_WG; }
int basket_wide_tile(bool corr) { return corr ? CORR_TILE : WIDE_WG; }
size_t basket_wide_lds_bytes(int n, bool corr) {
    if (!corr) return 0;
    const int nb = corr_nb(n);
    return (size_t)nb * (nb + 1) / 2 * 64 * sizeof(float4);
}

template <int MK, int SRC>
static const void* pick_na(int na) {
    if constexpr (SRC == SRC_PROD) {
        if (na == 2) return (const void*)&k_basket<2, MK, SRC>;
    }
    switch (na) {
    case 4: return (const void*)&k_basket<4, MK, SRC>;
    case 8: return (const void*)&k_basket<8, MK, SRC>;
    case 16: return (const void*)&k_basket<16, MK, SRC>;
    case 32: return (const void*)&k_basket<32, MK, SRC>;
    default: return nullptr;
    }
}

const void* basket_kernel(int mk, int src, int na) {
    if (mk == MK_MULTI_GBM && src == SRC_PROD) return pick_na<MK_MULTI_GBM, SRC_PROD>(na);
    if (src == SRC_PROD) return nullptr;
    if (mk == MK_MULTI_GBM) {
        if (src == SRC_ROWS) return pick_na<MK_MULTI_GBM, SRC_ROWS>(na);
        if (src == SRC_SOBOL) return pick_na<MK_MULTI_GBM, SRC_SOBOL>(na);
        return pick_na<MK_MULTI_GBM, SRC_PRNG>(na);
    }
    if (src == SRC_ROWS) return pick_na<MK_MULTI_DUPIRE, SRC_ROWS>(na);
    if (src == SRC_SOBOL) return pick_na<MK_MULTI_DUPIRE, SRC_SOBOL>(na);
    return pick_na<MK_MULTI_DUPIRE, SRC_PRNG>(na);
}

size_t basket_lds_bytes(int src, int D, int na) {
    return ((size_t)((src == SRC_SOBOL && D <= SOBOL_LDS_MAX_D) ? D * 20 : 0) * 4 + 15) / 16 * 16 +
           4 * (2 + 2 * na) * sizeof(double);
}

size_t basket_launch_size() { return sizeof(BasketLaunch); }

}  // namespace fdmc
