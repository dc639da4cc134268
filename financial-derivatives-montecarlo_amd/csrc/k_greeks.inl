// k_greeks.inl -- price + pathwise-adjoint Greeks in one fused kernel.
//
// The reference computes prices only.  For Lipschitz payoffs (European, Asian,
// geometric Asian, lookback) the pathwise derivatives of the discounted payoff
// with respect to S0 (delta), the volatility parameter (vega: sigma for GBM,
// sigma0 for Dupire, the level a of the quadratic local vol) and the drift /
// rate (rho: mu and r moved together) are computed in reverse mode:
//
//   forward : noise -> Euler steps; S is checkpointed in LDS at the start of
//             every 16-step block (G floats per lane), running payoff state kept;
//   payoff  : direct sensitivities dY/dS_k (1{A>K}/n for Asian,
//             1{e^G>K} e^G/(n S_k) for the geometric mean, disc 1{S_T>K} at
//             maturity, -bgk at the running minimum for the lookback);
//   reverse : blocks i = G-1 .. 0: regenerate the block's increments from the
//             Sobol index (or Philox counter) and the bridge skeleton, and run
//             the block forward from its checkpoint carrying J = dS_k/dS_start
//             and the tangents T_k = dS_k/dtheta of this block's steps; then
//             Sbar_start = sum_k direct_k J_k + Sbar_end J_B and
//             theta_bar += sum_k direct_k T_k + Sbar_end T_B (the chain rule
//             of Sbar_k = direct_k + Sbar_{k+1} dS_{k+1}/dS_k summed in the
//             other order: no per-step state is stored).
//
// No path is stored: memory per lane is G + (bridge skeleton) floats.
#pragma once

#include "k_path.inl"

namespace fdmc {

// Derivatives of one Euler step S' = f(S, dw, t) (step_1f).
template <int MK>
__device__ __forceinline__ void step_partials(const LegParams& P, float S, float dw, float t, float h,
                                              float& fS, float& fvol, float& fmu, float& fS0) {
    fmu = S * h;
    fS0 = 0.0f;
    if constexpr (MK == MK_GBM) {
        fS = 1.0f + P.mu * h + P.sigma * dw;
        fvol = S * dw;
    } else if constexpr (MK == MK_DUPIRE) {
        const float e = fast_exp(-P.alpha * t);
        const float pw = fast_exp2(P.beta_m1 * fast_log2(S * P.inv_S0));
        const float sl = P.sigma0 * e * pw;
        fS = 1.0f + P.mu * h + dw * sl * (P.beta_m1 + 1.0f);
        fvol = e * pw * S * dw;                        // d/d sigma0
        fS0 = -sl * P.beta_m1 * P.inv_S0 * S * dw;     // local-vol normalisation by S0
    } else {  // MK_LVQ
        const float l = fast_ln(S * P.inv_sref);
        const float be = P.lv_b * fast_exp(-P.lv_c * t);
        const float sl = P.lv_a + be * l * l;
        fS = 1.0f + P.mu * h + dw * (sl + 2.0f * be * l);
        fvol = S * dw;                                 // d/d a
    }
}

template <int MK, int PK, int CV, int CPK, int NZ, int LB>
// The compiler's own allocation (about 210 VGPRs, two waves per SIMD): the
// three-waves-per-SIMD build spills 34-46 VGPRs and ran the C2 adjoint pass
// 6 % slower (0.280 against 0.263 ms), measured on MI355X.
__global__ __launch_bounds__(256) void k_greeks(PathLaunch a) {
    constexpr int NP = (CV == CV_NONE) ? 2 : 7;
    constexpr int NACC = NP + 6;
    constexpr int B = (NZ == NZ_SOBOL_BB && LB == 0) ? 1 : 16;
    extern __shared__ __align__(16) unsigned char smem[];
    const int tid = threadIdx.x, lane = tid & 63;
    const int r = blockIdx.y;
    const int D = a.D, n = a.n_steps;
    // inverse-CDF table (icdf_u32_tab), filled before the first barrier-separated use
    constexpr bool ITAB = NZ != NZ_PRNG;
    __shared__ __align__(16) float4 itab[ITAB ? FDMC_ICDF_TAB_ROWS : 1];
    if constexpr (ITAB) {
        icdf_tab_load(itab, threadIdx.x, 256);
        __syncthreads();
    }
    const int G = (NZ == NZ_SOBOL_BB) ? a.G : (n + B - 1) / B;
    const int lgG = 31 - __builtin_clz(G > 0 ? G : 1);
    const bool lds_mode = (NZ != NZ_PRNG) && D <= SOBOL_LDS_MAX_D;
    uint32_t* rec = (uint32_t*)smem;
    const int rec_words = (NZ != NZ_PRNG && lds_mode) ? D * 20 : 0;
    float* skel = (float*)(smem + (size_t)rec_words * 4);
    const int skel_words = (NZ == NZ_SOBOL_BB) ? (a.G + 1) * 256 : 0;
    float* ckp = skel + skel_words;                  // [G][256] checkpoints
    double* red = (double*)(smem + (((size_t)(rec_words + skel_words + G * 256) * 4 + 15) & ~(size_t)15));

    const uint32_t* grec = (NZ != NZ_PRNG) ? a.sdir + (int64_t)r * D * SDIR_WORDS : nullptr;
    SobolLane sl;
    {
        const uint32_t g = (uint32_t)(lane ^ (lane >> 1));
        sl.rec = rec;
        sl.grec = grec;
        sl.ga = g & 7u;
        sl.gb = g >> 3;
        sl.w = (uint32_t)(tid >> 6);
        sl.hi_gray = 0;
        sl.lds_mode = lds_mode;
    }
    // a uniform bridge of 16-date blocks reads its records block-contiguous
    const int gslot = (NZ == NZ_SOBOL_BB && B == 16 && a.bb_uniform) ? G : 0;
    if (NZ != NZ_PRNG && lds_mode) rec_fill(rec, grec, D, gslot, tid);
    double acc[NACC];
#pragma unroll
    for (int i = 0; i < NACC; i++) acc[i] = 0.0;
    const LegParams& P = a.leg[0];
    // the leg for the 16-date bridge blocks' scaled increments (bb_block_uniform);
    // of the step's partial derivatives only d/d(vol) = S dw carries the scale,
    // so the vega of such paths is multiplied by 1/16 at the end
    const LegParams P16 = vol_scaled(P, 0.0625f);
    const float h = a.h;

    const int64_t last = a.first + a.count;
    const int64_t tile0 = (a.first >> 8) + (int64_t)blockIdx.x * a.tiles_per_wg;
    for (int it = 0; it < a.tiles_per_wg; it++) {
        const int64_t base = (tile0 + it) << 8;
        if (base >= last) break;
        if constexpr (NZ != NZ_PRNG) {
            const uint32_t b32 = (uint32_t)base;
            if (lds_mode) {
                __syncthreads();
                sobol_tile_words(rec, grec, D, (uint32_t)(tile0 + it), it == 0, tid, gslot);
                __syncthreads();
            } else {
                const uint32_t hi = b32 + (sl.w << 6);
                sl.hi_gray = hi ^ (hi >> 1);
            }
        }
        const int64_t idx = base + tid;
        const bool valid = idx >= a.first && idx < last;
        float* sk = skel + tid;
        float* ck = ckp + tid;

        // the rest of the tile, for a uniform-grid bridge (UNI) or bb_precompute's weights
        auto tile = [&](auto uniform_tag) {
            constexpr bool UNI = decltype(uniform_tag)::value;
            constexpr bool SC = NZ == NZ_SOBOL_BB && UNI && B == 16;
            const LegParams& PS = SC ? P16 : P;
            if constexpr (NZ == NZ_SOBOL_BB) {
                sk[0] = 0.0f;
                sk[G * 256] = bb_terminal<UNI, true>(a, sl, itab);
                for (int j = 1; j < G; j++) {
                    const int lev = 31 - __builtin_clz(j);
                    const int q = j - (1 << lev);
                    const int st = G >> lev;
                    const int L = q * st, R = L + st, M = L + (st >> 1);
                    sk[M * 256] = bb_mid<UNI, true>(a, sl, itab, sk[L * 256], sk[R * 256], j, lev + 1);
                }
            }
            // increments of block i (B steps; steps beyond n are zero)
            auto gen_block = [&](int i, float (&dw)[B]) {
                if constexpr (NZ == NZ_SOBOL_BB) {
                    if constexpr (B == 1) {
                        dw[0] = sk[(i + 1) * 256] - sk[i * 256];
                    } else if constexpr (UNI) {
                        bb_block_uniform<true, true>(a, sl, itab, G, lgG, i, sk[i * 256], sk[(i + 1) * 256], dw);
                    } else {
                        float W[B + 1];
                        W[0] = sk[i * 256];
                        W[B] = sk[(i + 1) * 256];
    #pragma unroll
                        for (int l = 0; l < 4; l++) {
                            const int j0 = (G << l) + (i << l);
    #pragma unroll
                            for (int q = 0; q < (1 << l); q++) {
                                const int s = B >> l;
                                const int Lp = q * s, Rp = Lp + s, Mp = Lp + (s >> 1);
                                W[Mp] = bb_mid<UNI, true>(a, sl, itab, W[Lp], W[Rp], j0 + q, lgG + l + 1);
                            }
                        }
    #pragma unroll
                        for (int k = 0; k < B; k++) dw[k] = W[k + 1] - W[k];
                    }
                } else if constexpr (NZ == NZ_SOBOL_RAW) {
    #pragma unroll
                    for (int k = 0; k < B; k++) {
                        const int g = i * B + k;
                        dw[k] = (g < n) ? icdf_u32_tab<true>(sobol_coord(sl, g), itab) * a.sqrt_h : 0.0f;
                    }
                } else {
    #pragma unroll
                    for (int b = 0; b < B / 4; b++) {
                        float z[4];
                        prng_normals4(a.seed + (uint32_t)r, a.tag, (uint64_t)idx, (uint32_t)(i * (B / 4) + b), z);
    #pragma unroll
                        for (int c = 0; c < 4; c++) dw[4 * b + c] = (i * B + 4 * b + c < n) ? z[c] * a.sqrt_h : 0.0f;
                    }
                }
            };

            // ---- forward
            float S = P.S0, run = run_init<PK>(P), crun = run_init<CPK>(a.leg[1]);
            int kmin = 0;
            // a block that ends by date n takes a copy of the step loop without
            // the per-step test (as selects it cost three per step); the last,
            // partial block of an n that is no multiple of B keeps it
            auto fwd_block = [&](int i, const float (&dw)[B], auto full) {
    #pragma unroll
                for (int k = 0; k < B; k++) {
                    const int g = i * B + k;
                    if (decltype(full)::value || g < n) {
                        S = step_1f<MK>(PS, S, dw[k], (float)g * h, h);
                        if constexpr (PK == PK_LOOKBACK) {
                            if (S < run) { run = S; kmin = g + 1; }
                        } else if constexpr (is_path_dep<PK>()) {
                            run_update<PK>(run, S);
                        }
                        if constexpr (CV == CV_SHARED) run_update<CPK>(crun, S);
                    }
                }
            };
            for (int i = 0; i < G; i++) {
                float dw[B];
                gen_block(i, dw);
                ck[i * 256] = S;
                if ((i + 1) * B <= n) fwd_block(i, dw, std::true_type{});
                else fwd_block(i, dw, std::false_type{});
            }
            const float ym = payoff<PK>(P, S, run, n, a.bgk[0]);
            float yc = 0.0f;
            if constexpr (CV == CV_SHARED) yc = payoff<CPK>(a.leg[1], S, crun, n, a.bgk[1]);

            // ---- direct sensitivities dY/dS_k (k >= 1)
            const float inv_n = 1.0f / (float)n;
            float c_each = 0.0f;        // added to every S_k
            float c_log = 0.0f;         // times 1/S_k for every S_k
            float c_last = 0.0f;        // added at S_n
            float c_min = 0.0f;         // added at S_kmin
            if constexpr (PK == PK_ASIAN) c_each = (run * inv_n > P.K) ? inv_n : 0.0f;
            if constexpr (PK == PK_GEOM_ASIAN) {
                const float e = fast_exp(run * inv_n);
                c_log = (e > P.K) ? e * inv_n : 0.0f;
            }
            if constexpr (PK == PK_EUROPEAN) c_last = (S > P.K) ? P.discount : 0.0f;
            if constexpr (PK == PK_LOOKBACK) {    // the payoff is 0 unless S_T > 0
                c_last = (S > 0.0f) ? 1.0f : 0.0f;
                c_min = (S > 0.0f) ? -a.bgk[0] : 0.0f;
            }
            if constexpr (CV == CV_SHARED && CPK == PK_GEOM_ASIAN) {
                const float e = fast_exp(crun * inv_n);
                c_log -= a.beta_cv * ((e > a.leg[1].K) ? e * inv_n : 0.0f);
            }

            // ---- reverse sweep, block by block.  The state is one-dimensional,
            // so a block's vector-Jacobian product needs no per-step storage: one
            // forward sweep from the block's checkpoint carries J = dS_k/dS_start
            // and the tangents T of S_k in the parameters through this block's
            // steps, and sums the direct terms d_k against them:
            //   adj(S_start)  = sum_k d_k J_k + adj(S_end) J_B,
            //   theta_bar    += sum_k d_k T_k + adj(S_end) T_B.
            float adj = 0.0f, vb = 0.0f, mb = 0.0f, s0b = 0.0f;   // adj: d/dS at the block's end
            for (int i = G - 1; i >= 0; i--) {
                float dw[B];
                gen_block(i, dw);
                float S = ck[i * 256];
                float J = 1.0f, Tv = 0.0f, Tm = 0.0f, T0 = 0.0f;
                float A = 0.0f, Av = 0.0f, Am = 0.0f, A0 = 0.0f;
    #pragma unroll
                for (int k = 0; k < B; k++) {
                    const int g = i * B + k;
                    if (g < n) {
                        float fS, fv, fm, f0;
                        step_partials<MK>(PS, S, dw[k], (float)g * h, h, fS, fv, fm, f0);
                        S = step_1f<MK>(PS, S, dw[k], (float)g * h, h);
                        J *= fS;
                        Tv = fmaf(fS, Tv, fv);
                        Tm = fmaf(fS, Tm, fm);
                        T0 = fmaf(fS, T0, f0);
                        float d = c_each + c_log / S;             // direct term of S_{g+1}
                        if (g + 1 == n) d += c_last;
                        if (g + 1 == kmin) d += c_min;
                        A = fmaf(d, J, A);
                        Av = fmaf(d, Tv, Av);
                        Am = fmaf(d, Tm, Am);
                        A0 = fmaf(d, T0, A0);
                    }
                }
                vb += fmaf(adj, Tv, Av);
                mb += fmaf(adj, Tm, Am);
                s0b += fmaf(adj, T0, A0);
                adj = fmaf(adj, J, A);
            }
            if (kmin == 0) adj += c_min;     // running minimum at S0
            if constexpr (SC) vb *= 0.0625f;
            float delta = adj + s0b;
            float rho = mb;
            if constexpr (PK == PK_EUROPEAN) rho -= a.T * ym;   // d/dr of the discount e^{-rT}
            // a path whose state left the Euler scheme's domain (S <= 0 under a
            // local vol: its log is NaN) pays 0 through fmaxf in the price and
            // contributes 0 to the Greeks as well
            if (!(fabsf(delta) + fabsf(vb) + fabsf(rho) < INFINITY)) delta = vb = rho = 0.0f;

            if (valid) {
                const double dm = (double)ym;
                acc[0] += dm;
                acc[1] += dm * dm;
                if constexpr (CV != CV_NONE) {
                    const float ycv = ym - a.beta_cv * (yc - a.E_ctrl);
                    const double dv = (double)ycv, dc = (double)yc;
                    acc[2] += dv;
                    acc[3] += dv * dv;
                    acc[4] += dc;
                    acc[5] += dc * dc;
                    acc[6] += dm * dc;
                }
                const double gd = delta, gv = vb, gr = rho;
                acc[NP + 0] += gd;
                acc[NP + 1] += gd * gd;
                acc[NP + 2] += gv;
                acc[NP + 3] += gv * gv;
                acc[NP + 4] += gr;
                acc[NP + 5] += gr * gr;
            }
        };
        if constexpr (NZ == NZ_SOBOL_BB) {
            if (a.bb_uniform) tile(std::true_type{});
            else tile(std::false_type{});
        } else {
            tile(std::false_type{});
        }
    }
    __syncthreads();
    block_sum<NACC, 4>(acc, red);
    if (tid == 0) {
        double* out = a.partials + ((int64_t)r * a.nwg_x + blockIdx.x) * NACC;
#pragma unroll
        for (int i = 0; i < NACC; i++) out[i] = acc[i];
    }
}

}  // namespace fdmc
