// k_levels.inl -- coupled fine/coarse multilevel kernel, one persistent launch
// for every level (and replica) of an MLMC / MLQMC iteration.
//
// The reference launches one kernel per level on its own stream, each with its
// own generator and buffers (methods_cuda.cu:1347-1395, 1762-1924).  Here a
// grid sized to the device drains a queue of work items of all requested
// (level, replica) entries, longest levels first; a workgroup takes the next
// item with one atomic.  An item is tpi consecutive 256-sample tiles of one
// entry, tpi chosen on the host so that items of short levels cost about as
// much as one tile of a long level (a 1-step level would otherwise be millions
// of 256-path items).  The per-path arithmetic is kernel_mlmc
// (methods_cuda.cu:445-518) with the fine path's running payoff updated for
// every model (the reference skips it for Heston, :464-470).
#pragma once

#include "k_path.inl"

namespace fdmc {

template <int MK, int PK>
struct Coupled {
    float Sf, Vf, Sc, Vc, rf, rc, a1, a2;
    int cnt, ck;
    __device__ __forceinline__ void init(const LegParams& P) {
        Sf = Sc = P.S0;
        Vf = Vc = P.v0;
        rf = rc = run_init<PK>(P);
        a1 = a2 = 0.0f;
        cnt = ck = 0;
    }
    // one-factor fine step with a scaled increment dw; coarse step every M.
    __device__ __forceinline__ void fine1(const LegParams& P, const LevelEntry& E, float dw, int k, int M) {
        a1 += dw;
        Sf = step_1f<MK>(P, Sf, dw, (float)k * E.h_f, E.h_f);
        if constexpr (is_path_dep<PK>()) run_update<PK>(rf, Sf);
        if (++cnt == M) {
            Sc = step_1f<MK>(P, Sc, a1, (float)ck * E.h_c, E.h_c);
            if constexpr (is_path_dep<PK>()) run_update<PK>(rc, Sc);
            a1 = 0.0f;
            cnt = 0;
            ++ck;
        }
    }
    // Heston fine step from unscaled normals z1, z2.
    __device__ __forceinline__ void fineH(const LegParams& P, const LevelEntry& E, float z1, float z2, int M) {
        const float dw1 = z1 * E.sqrt_hf;
        const float dw2 = (P.L10 * z1 + P.L11 * z2) * E.sqrt_hf;
        a1 += dw1;
        a2 += dw2;
        step_heston<MK == MK_HESTON_MIL>(P, Sf, Vf, dw1, dw2, E.h_f, E.em_f);
        if constexpr (is_path_dep<PK>()) run_update<PK, true>(rf, Sf);
        if (++cnt == M) {
            step_heston<MK == MK_HESTON_MIL>(P, Sc, Vc, a1, a2, E.h_c, E.em_c);
            if constexpr (is_path_dep<PK>()) run_update<PK, true>(rc, Sc);
            a1 = a2 = 0.0f;
            cnt = 0;
        }
    }
    __device__ __forceinline__ void finish(const LegParams& P, const LevelEntry& E, float& pf, float& pc) const {
        pf = payoff<PK>(P, Sf, rf, E.n_fine, E.bgk_f);
        pc = (E.n_coarse > 0) ? payoff<PK>(P, Sc, rc, E.n_coarse, E.bgk_c) : 0.0f;
    }
};

template <int MK, int PK, int NZ, int LB>
// Compiled for three waves per SIMD (168 VGPRs with the bridge, some spills):
// 15 % faster than two waves at 202 VGPRs on MI355X; the PRNG variants use
// far fewer registers and are unaffected.
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3))) void k_levels(LevelsLaunch a, int max_D, int skel_n) {
    extern __shared__ __align__(16) unsigned char smem[];
    const int tid = threadIdx.x, lane = tid & 63;
    constexpr bool SOBOL = NZ == NZ_SOBOL_RAW || NZ == NZ_SOBOL_BB;
    const bool lds_ok = SOBOL && max_D <= SOBOL_LDS_MAX_D;
    // inverse-CDF table (icdf_u32_tab), filled before the first barrier-separated use
    constexpr bool ITAB = SOBOL;
    __shared__ __align__(16) float4 itab[ITAB ? FDMC_ICDF_TAB_ROWS : 1];
    if constexpr (ITAB) {
        icdf_tab_load(itab, threadIdx.x, 256);
        __syncthreads();
    }
    uint32_t* rec = (uint32_t*)smem;
    const int rec_words = lds_ok ? max_D * 20 : 0;
    float* skel = (float*)(smem + (size_t)rec_words * 4);
    // bridge: the skeleton of levels under 16 dates or the walk stack, per lane
    const int skel_words = (NZ == NZ_SOBOL_BB) ? skel_n * 256 : 0;
    double* red = (double*)(smem + (((size_t)(rec_words + skel_words) * 4 + 15) & ~(size_t)15));
    int* s_item = (int*)(red + 4 * 4);

    SobolLane sl;
    {
        const uint32_t g = (uint32_t)(lane ^ (lane >> 1));
        sl.rec = rec;
        sl.grec = nullptr;
        sl.ga = g & 7u;
        sl.gb = g >> 3;
        sl.w = (uint32_t)(tid >> 6);
        sl.hi_gray = 0;
        sl.lds_mode = lds_ok;
    }
    int last_e = -1;
    const LegParams& P = a.leg;

    for (;;) {
        __syncthreads();
        if (tid == 0) *s_item = (int)atomicAdd(a.queue, 1u);
        __syncthreads();
        const int item = *s_item;
        if (item >= a.total_items) break;   // every workgroup leaves once the queue is empty
        int lo = 0, hi = a.n_entries - 1;
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (a.entries[mid].tile_begin <= item) lo = mid;
            else hi = mid - 1;
        }
        const LevelEntry E = a.entries[lo];
        // this item covers tiles [t0, t1) of the entry
        const int64_t tfirst = E.first >> 8;
        const int nt_all = (int)(((E.first + E.count - 1) >> 8) - tfirst + 1);
        const int t0 = (item - E.tile_begin) * E.tpi;
        const int t1 = min(t0 + E.tpi, nt_all);
        const int D = E.D;
        double acc[4] = {0.0, 0.0, 0.0, 0.0};
        for (int t = t0; t < t1; t++) {
        const int64_t base = (tfirst + t) << 8;
        const int64_t idx = base + tid;
        const bool valid = idx >= E.first && idx < E.first + E.count;

        if constexpr (SOBOL) {
            sl.grec = E.sdir;
            const uint32_t b32 = (uint32_t)base;
            if (lds_ok) {
                if (t > t0) __syncthreads();   // every lane is done with the previous tile's table
                if (lo != last_e) {
                    for (int i = tid; i < D * 16; i += 256) {
                        const int d = i >> 4, k = i & 15;
                        rec[d * 20 + k] = E.sdir[(int64_t)d * SDIR_WORDS + k];
                    }
                }
                sobol_tile_words(rec, E.sdir, D, (uint32_t)(base >> 8), t == t0, tid);
                __syncthreads();
            } else {
                const uint32_t hi2 = b32 + (sl.w << 6);
                sl.hi_gray = hi2 ^ (hi2 >> 1);
            }
        }
        last_e = lo;

        Coupled<MK, PK> cp;
        cp.init(P);
        const int M = a.M, nfine = E.n_fine;
        if constexpr (NZ == NZ_ROWS) {
            // fine increments precomputed by the PCA GEMM (methods_cuda.cu:1907-1913)
            const float* src = E.rows + (int64_t)(valid ? idx - E.first : 0) * E.ld;
            for (int k = 0; k < nfine; k++) cp.fine1(P, E, src[k], k, M);
        } else if constexpr (NZ == NZ_PRNG) {
            float z[4];
            if constexpr (is_heston(MK)) {
                for (int b = 0; 2 * b < nfine; b++) {
                    prng_normals4(E.seed, E.tag, (uint64_t)idx, (uint32_t)b, z);
                    cp.fineH(P, E, z[0], z[1], M);
                    if (2 * b + 1 < nfine) cp.fineH(P, E, z[2], z[3], M);
                }
            } else {
                for (int b = 0; 4 * b < nfine; b++) {
                    prng_normals4(E.seed, E.tag, (uint64_t)idx, (uint32_t)b, z);
#pragma unroll
                    for (int c = 0; c < 4; c++)
                        if (4 * b + c < nfine) cp.fine1(P, E, z[c] * E.sqrt_hf, 4 * b + c, M);
                }
            }
        } else if constexpr (NZ == NZ_SOBOL_RAW) {
            if constexpr (is_heston(MK)) {
                for (int k = 0; k < nfine; k++)
                    cp.fineH(P, E, icdf_u32_tab<true>(sobol_coord(sl, 2 * k), itab), icdf_u32_tab<true>(sobol_coord(sl, 2 * k + 1), itab), M);
            } else {
                for (int k = 0; k < nfine; k++) cp.fine1(P, E, icdf_u32_tab<true>(sobol_coord(sl, k), itab) * E.sqrt_hf, k, M);
            }
        } else {  // Brownian bridge over n_fine steps (one-factor models)
            const float4* bw = (const float4*)E.bbw;
            const int G = E.G;
            // W_M of dyadic node j from the ends of its interval, bb_precompute's weights
            auto mid = [&](float WL, float WR, int j) {
                const float4 c = bw[j];
                return c.x * WL + c.y * WR + c.z * icdf_u32_tab<true>(sobol_coord(sl, j), itab);
            };
            if (LB == 0 || nfine < 16) {
                // fewer than 16 dates: the whole skeleton in LDS, one step per entry
                float* sk = skel + tid;
                sk[0] = 0.0f;
                sk[G * 256] = bw[0].z * icdf_u32_tab<true>(sobol_coord(sl, 0), itab);
                for (int j = 1; j < G; j++) {
                    const int lev = 31 - __builtin_clz(j);
                    const int q = j - (1 << lev);
                    const int st = G >> lev;
                    const int L = q * st, R = L + st, Mi = L + (st >> 1);
                    sk[Mi * 256] = mid(sk[L * 256], sk[R * 256], j);
                }
                float left = 0.0f;
                for (int i = 0; i < G; i++) {
                    const float right = sk[(i + 1) * 256];
                    cp.fine1(P, E, right - left, i, M);
                    left = right;
                }
            } else {
                // G = 2^m blocks of B dates; the skeleton at the block boundaries is
                // walked left to right on a per-lane LDS stack rt[0..m] (rt[k]: W at
                // the right end of the aligned 2^k-block interval holding the current
                // boundary), each node generated from the same operands as the
                // top-down skeleton, so the LDS does not grow with the date count.
                constexpr int B = 1 << (LB > 0 ? LB : 1);
                const int m = 31 - __builtin_clz(G);
                float* rt = skel + tid;
                auto node = [&](int k, int L, float WL, float WR) {
                    return mid(WL, WR, (1 << (m - 1 - k)) + (L >> (k + 1)));
                };
                float Wp = 0.0f;
                rt[m * 256] = bw[0].z * icdf_u32_tab<true>(sobol_coord(sl, 0), itab);
                for (int k = m - 1; k >= 0; k--) rt[k * 256] = node(k, 0, Wp, rt[(k + 1) * 256]);
                for (int i = 0; i < G; i++) {
                    if (i > 0) {
                        Wp = rt[0];
                        const int c = __builtin_ctz(i);
                        rt[c * 256] = rt[(c + 1) * 256];
                        for (int k = c - 1; k >= 0; k--) rt[k * 256] = node(k, i, Wp, rt[(k + 1) * 256]);
                    }
                    float W[B + 1];
                    W[0] = Wp;
                    W[B] = rt[0];
#pragma unroll
                    for (int l = 0; l < LB; l++) {
                        const int j0 = (G << l) + (i << l);
#pragma unroll
                        for (int q = 0; q < (1 << l); q++) {
                            const int s = B >> l;
                            const int Lp = q * s, Rp = Lp + s, Mp = Lp + (s >> 1);
                            W[Mp] = mid(W[Lp], W[Rp], j0 + q);
                        }
                    }
#pragma unroll
                    for (int k = 0; k < B; k++) cp.fine1(P, E, W[k + 1] - W[k], i * B + k, M);
                }
            }
        }

        float pf, pc;
        cp.finish(P, E, pf, pc);
        if (valid) {
            const double dy = (double)(pf - pc), yf = (double)pf;
            acc[0] += dy;
            acc[1] += dy * dy;
            acc[2] += yf;
            acc[3] += yf * yf;
        }
        }  // tiles of the item
        block_sum<4, 4>(acc, red);
        if (tid == 0) {
            double* out = a.partials + (int64_t)item * 4;
#pragma unroll
            for (int i = 0; i < 4; i++) out[i] = acc[i];
        }
    }
}

}  // namespace fdmc
