// k_path.inl -- fused single-level path kernel (included by k_path_*.hip).
//
// One lane = one path; a 256-lane workgroup walks `tiles_per_wg` consecutive
// 256-point tiles of one replica (blockIdx.y).  Per path it fuses what the
// reference runs as 4-6 kernels with HBM round trips between them
// (methods_cuda.cu:1551-1621): scrambled Sobol coordinate -> inverse CDF ->
// Brownian bridge (or raw scaling) -> Euler steps -> running payoff (+ control
// leg) -> FP64 partial sums.  Nothing per path touches HBM.
//
// LDS (dynamic):
//   rec  [D][20] u32  TA[8] | TB[8] | C[4]   (Sobol, D <= SOBOL_LDS_MAX_D)
//   skel [G+1][256] f32 bridge skeleton, one column per lane (BB)
//   red  [4][7] f64   workgroup reduction
// Brownian bridge: the dyadic schedule of bb_precompute (utils.cpp:54-101) is
// split at block size B = 16: entries j < G (coarse levels) build W at the block
// boundaries in LDS; the 15 interior entries of block i are
// j = (G << l) + (i << l) + q (l = 0..3, q < 2^l) and are built in registers,
// immediately followed by the 16 Euler steps of the block.  Each W value is the
// same formula wl W_left + wr W_right + sd z as kernel_bb_transform (:561-594).
#pragma once

#include <type_traits>

#include "device_math.hpp"
#include "engine_internal.hpp"

namespace fdmc {

struct SobolLane {
    const uint32_t* rec;     // LDS records (LDS mode) or null
    const uint32_t* grec;    // global records of this replica
    uint32_t ga, gb, w;      // lane Gray bits 0..2, 3..5; wave index
    uint32_t hi_gray;        // gray(base + 64 w) (global mode)
    bool lds_mode;
};

// LDS mode only (callers whose dimension never exceeds SOBOL_LDS_MAX_D).
__device__ __forceinline__ uint32_t sobol_coord_lds(const SobolLane& s, int d) {
    const uint32_t* r = s.rec + d * 20;
    return xor3(r[s.ga], r[8 + s.gb], r[16 + s.w]);
}

// Global mode: the lane's Gray bits from the record, the higher bits of the
// tile's point index (hi_gray) one direction number each.
__device__ __forceinline__ uint32_t sobol_coord_glb(const SobolLane& s, int d) {
    const uint32_t* r = s.grec + (int64_t)d * SDIR_WORDS;
    uint32_t x = r[s.ga] ^ r[8 + s.gb];
    uint32_t g = s.hi_gray;
    for (; g & (g - 1u); g &= g - 1u) {          // two set bits at a time
        const uint32_t b0 = __builtin_ctz(g);
        g &= g - 1u;
        x = xor3(x, r[32 + b0], r[32 + __builtin_ctz(g)]);
    }
    if (g) x ^= r[32 + __builtin_ctz(g)];
    return x;
}

__device__ __forceinline__ uint32_t sobol_coord(const SobolLane& s, int d) {
    if (s.lds_mode) {
        const uint32_t* r = s.rec + d * 20;
        return xor3(r[s.ga], r[8 + s.gb], r[16 + s.w]);
    }
    return sobol_coord_glb(s, d);
}

// Words 16..19 of the LDS direction records for the 256-point tile n (points
// 256 n .. 256 n + 255, below 2^32): the wave offsets' direction sums XOR the
// sum of Lv[j] over the set bits of gray(256 n), bits 7..31.  `full`: form the
// sum (its loads are independent); otherwise the records hold tile n - 1, and
// gray(256 n) differs from gray(256 (n - 1)) in bits 7 and 8 + ctz(n) only, so
// two direction numbers are XORed in.  Threads d = tid, tid + 256, ... < D.
// LDS slot of bridge dimension d on a uniform bridge of G >= 1 blocks of 16
// dates: the skeleton's dimensions (d < G) keep their index, and the 15
// in-block dimensions of block i (G 2^lv + 2^lv i + q at in-block depth lv) sit
// together at slots G + 15 i + 2^lv - 1 + q, so one block's records are read
// from one base address with immediate offsets.
__device__ __forceinline__ int block_slot(int d, int G) {
    if (d < G) return d;
    const int lv = 31 - __builtin_clz((uint32_t)d / (uint32_t)G);
    const int r = d - (G << lv);
    return G + 15 * (r >> lv) + (1 << lv) - 1 + (r & ((1 << lv) - 1));
}

// The direction records of dimensions 0..D-1 into LDS, at block_slot(d, G)
// for G > 0, at d for G = 0.
__device__ __forceinline__ void rec_fill(uint32_t* rec, const uint32_t* grec, int D, int G, int tid) {
    for (int i = tid; i < D * 16; i += 256) {
        const int d = i >> 4, k = i & 15;
        rec[(G > 0 ? block_slot(d, G) : d) * 20 + k] = grec[(int64_t)d * SDIR_WORDS + k];
    }
}

// G > 0: records at block_slot(d, G) (rec_fill with the same G).
__device__ __forceinline__ void sobol_tile_words(uint32_t* rec, const uint32_t* grec, int D, uint32_t n, bool full,
                                                 int tid, int G = 0) {
    const uint32_t b32 = n << 8, gbase = b32 ^ (b32 >> 1);
    for (int d = tid; d < D; d += 256) {
        const uint32_t* gr = grec + (int64_t)d * SDIR_WORDS;
        uint4* rw = (uint4*)(rec + (G > 0 ? block_slot(d, G) : d) * 20 + 16);
        if (full) {
            uint32_t s = 0;
#pragma unroll
            for (int j = 7; j < 32; j++) s ^= ((gbase >> j) & 1u) ? gr[32 + j] : 0u;
            *rw = make_uint4(s ^ gr[16], s ^ gr[17], s ^ gr[18], s ^ gr[19]);
        } else {
            const uint32_t dl = gr[32 + 7] ^ gr[32 + 8 + __builtin_ctz(n)];
            const uint4 v = *rw;
            *rw = make_uint4(v.x ^ dl, v.y ^ dl, v.z ^ dl, v.w ^ dl);
        }
    }
}

// A Brownian-bridge midpoint W_M from the ends of its interval and the normal of
// coordinate j (direction records in LDS for bridges of at most SOBOL_LDS_MAX_D
// dates, LDS = true, else read from global memory): with bb_precompute's FP32
// weights (wl, wr, sd), or on a uniform date grid (UNI) with weights 1/2 and the
// sd of the depth, read from no table per coordinate.  The path and Greeks
// kernels both build their bridges from these, so the same points give the
// same paths.
template <bool LDS>
__device__ __forceinline__ uint32_t bb_coord(const SobolLane& sl, int j) {
    if constexpr (LDS) return sobol_coord_lds(sl, j);
    else return sobol_coord_glb(sl, j);
}

template <bool UNI, bool LDS>
__device__ __forceinline__ float bb_mid(const PathLaunch& a, const SobolLane& sl, const float4* itab, float WL,
                                        float WR, int j, int depth) {
    if constexpr (UNI) {
        const float sd = a.bb_sd[depth];
        return 0.5f * (WL + WR) - icdf_u32_tab_scaled_neg<true>(bb_coord<LDS>(sl, j), itab, sd);
    } else {
        const float4 c = ((const float4*)a.bbw)[j];
        return c.x * WL + c.y * WR + c.z * icdf_u32_tab<true>(bb_coord<LDS>(sl, j), itab);
    }
}

// W at the last date: sd sqrt(T) times coordinate 0's normal.
template <bool UNI, bool LDS>
__device__ __forceinline__ float bb_terminal(const PathLaunch& a, const SobolLane& sl, const float4* itab) {
    if constexpr (UNI) return -icdf_u32_tab_scaled_neg<true>(bb_coord<LDS>(sl, 0), itab, a.bb_sd[0]);
    else return ((const float4*)a.bbw)[0].z * icdf_u32_tab<true>(bb_coord<LDS>(sl, 0), itab);
}

// The 16 increments of block i of a uniform-grid bridge (weights 1/2, one sd
// per depth) from W at the block's ends: each interval's increment d splits at
// its midpoint into d/2 - y and d/2 + y (y = -sd z of the node), which is
// W_M = (W_L + W_R)/2 + sd z written for the increments (no W inside the block,
// no differences).  Node j0 + q at in-block depth l, j0 = (G + i) 2^l.
// The increments are kept 2^l times their size at depth l, so a child of d is
// d -+ 2^(l+1) sd z: one FMA, with no halving and no product sd z of its own;
// the 16 increments come out 16 times their size, and the steps take them with
// the vol parameters times 1/16 (vol_scaled: a power of two, so sigma dw is the
// same product).
// SLOT: the records are at block_slot(j, G) (LDS only), so the block's 15
// nodes are read from one base.
template <bool LDS, bool SLOT>
__device__ __forceinline__ void bb_block_uniform(const PathLaunch& a, const SobolLane& sl, const float4* itab, int G,
                                                 int lgG, int i, float WL, float WR, float (&dw)[16]) {
    static_assert(LDS || !SLOT, "slots are an LDS layout");
    dw[0] = WR - WL;
#pragma unroll
    for (int l = 0; l < 4; l++) {
        const int j0 = SLOT ? G + 15 * i + (1 << l) - 1 : (G << l) + (i << l);
        const float sds = a.bb_sd[lgG + l + 1] * (float)(2 << l);
#pragma unroll
        for (int q = (1 << l) - 1; q >= 0; q--) {
            const float y = icdf_u32_tab_scaled_neg<true>(bb_coord<LDS>(sl, j0 + q), itab, 1.0f);
            const float d = dw[q];
            dw[2 * q] = fmaf(-sds, y, d);
            dw[2 * q + 1] = fmaf(sds, y, d);
        }
    }
}

// The vol parameters of a leg times s (GBM sigma, Dupire sigma0, the quadratic
// local vol's a and b): the leg for increments that are 1/s times too large.
__device__ __forceinline__ LegParams vol_scaled(LegParams P, float s) {
    P.sigma *= s;
    P.sigma0 *= s;
    P.lv_a *= s;
    P.lv_b *= s;
    return P;
}

template <int MK, int PK>
struct Leg {
    float S, V, run;
    __device__ __forceinline__ void init(const LegParams& P) {
        S = P.S0;
        V = P.v0;
        run = run_init<PK>(P);
    }
    __device__ __forceinline__ void step1(const LegParams& P, float dw, float t, float h) {
        S = step_1f<MK>(P, S, dw, t, h);
        if constexpr (is_path_dep<PK>()) run_update<PK>(run, S);
    }
    __device__ __forceinline__ void stepH(const LegParams& P, float dw1, float dw2, float h, float em) {
        step_heston<MK == MK_HESTON_MIL>(P, S, V, dw1, dw2, h, em);
        if constexpr (is_path_dep<PK>()) run_update<PK, true>(run, S);
    }
    __device__ __forceinline__ float pay(const LegParams& P, int n, float bgk) const {
        return payoff<PK>(P, S, run, n, bgk);
    }
};

// Path state: main leg + optional control leg.
template <int MK, int PK, int CV, int CPK>
struct PathState {
    Leg<MK, PK> m;
    Leg<MK_GBM, CPK> c;      // CV_GBM: independent GBM leg; CV_SHARED: only .run is used
    __device__ __forceinline__ void init(const PathLaunch& a) {
        m.init(a.leg[0]);
        if constexpr (CV == CV_GBM) c.init(a.leg[1]);
        if constexpr (CV == CV_SHARED) c.run = run_init<CPK>(a.leg[1]);
    }
    __device__ __forceinline__ void step(const PathLaunch& a, float dw, int k) {
        const float t = (float)k * a.h;
        m.step1(a.leg[0], dw, t, a.h);
        if constexpr (CV == CV_GBM) c.step1(a.leg[1], dw, t, a.h);
        if constexpr (CV == CV_SHARED) {
            if constexpr (is_path_dep<CPK>()) run_update<CPK>(c.run, m.S);
        }
    }
    __device__ __forceinline__ void stepH(const PathLaunch& a, float dw1, float dw2) {
        m.stepH(a.leg[0], dw1, dw2, a.h, a.em);
    }
    __device__ __forceinline__ void finish(const PathLaunch& a, float& ym, float& yc) const {
        ym = m.pay(a.leg[0], a.n_steps, a.bgk[0]);
        if constexpr (CV == CV_GBM) yc = c.pay(a.leg[1], a.n_steps, a.bgk[1]);
        else if constexpr (CV == CV_SHARED) yc = payoff<CPK>(a.leg[1], m.S, c.run, a.n_steps, a.bgk[1]);
        else yc = 0.0f;
    }
};

template <int MK, int PK, int CV, int CPK, int NZ, int LB>
// Compiled for three waves per SIMD (168 VGPRs, under 100 B of spills): 23 %
// faster than the compiler's own allocation (190 VGPRs, two waves) on the
// 256-date bridge pass, measured on MI355X.
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3))) void k_path(PathLaunch a) {
    constexpr int NACC = (CV == CV_NONE) ? 2 : 7;
    extern __shared__ __align__(16) unsigned char smem[];
    const int tid = threadIdx.x, lane = tid & 63;
    const int r = blockIdx.y;
    const int D = a.D;
    constexpr bool BRIDGE = (NZ == NZ_SOBOL_BB || NZ == NZ_SOBOL_BB_GLOBAL);
    constexpr bool SOBOL = (NZ == NZ_SOBOL_RAW || BRIDGE);
    // inverse-CDF table (icdf_u32_tab), filled before the first barrier-separated use
    constexpr bool ITAB = SOBOL;
    __shared__ __align__(16) float4 itab[ITAB ? FDMC_ICDF_TAB_ROWS : 1];
    if constexpr (ITAB) {
        icdf_tab_load(itab, threadIdx.x, 256);
        __syncthreads();
    }
    const bool lds_mode = SOBOL && NZ != NZ_SOBOL_BB_GLOBAL && D <= SOBOL_LDS_MAX_D;
    uint32_t* rec = (uint32_t*)smem;
    const int rec_words = lds_mode ? D * 20 : 0;
    float* skel = (float*)(smem + (size_t)rec_words * 4);
    // the bridge skeleton (dates < 16) or the walk stack of the block boundaries
    const int skel_words = BRIDGE ? (LB == 0 ? a.G + 1 : 32 - __builtin_clz(a.G)) * 256 : 0;
    double* red = (double*)(smem + (((size_t)(rec_words + skel_words) * 4 + 15) & ~(size_t)15));

    const uint32_t* grec = SOBOL ? a.sdir + (int64_t)r * D * SDIR_WORDS : nullptr;
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
    const int gslot = (NZ == NZ_SOBOL_BB && LB == 4 && a.bb_uniform) ? a.G : 0;
    if (SOBOL && lds_mode) rec_fill(rec, grec, D, gslot, tid);

    double acc[NACC];
#pragma unroll
    for (int i = 0; i < NACC; i++) acc[i] = 0.0;
    // the legs for the 16-date bridge blocks' scaled increments (bb_block_uniform)
    PathLaunch a16 = a;
    if constexpr (BRIDGE && LB == 4) {
        a16.leg[0] = vol_scaled(a.leg[0], 0.0625f);
        a16.leg[1] = vol_scaled(a.leg[1], 0.0625f);
    }

    const int64_t last = a.first + a.count;
    const int64_t tile0 = (a.first >> 8) + (int64_t)blockIdx.x * a.tiles_per_wg;
    for (int it = 0; it < a.tiles_per_wg; it++) {
        const int64_t base = (tile0 + it) << 8;
        if (base >= last) break;
        if constexpr (SOBOL) {
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

        PathState<MK, PK, CV, CPK> ps;
        ps.init(a);

        if constexpr (NZ == NZ_ROWS) {
            // increments from the PCA GEMM: row r * count + (idx - first)
            const float* src = a.rows + ((int64_t)r * a.count + (valid ? idx - a.first : 0)) * a.ld;
            for (int k = 0; k < a.n_steps; k++) ps.step(a, src[k], k);
        } else if constexpr (NZ == NZ_PRNG) {
            // Philox normals, 4 per call; normal n = k * nf + c.
            const int nn = a.n_steps * a.nf;
            float z[4];
            if constexpr (is_heston(MK)) {
                for (int b = 0; 4 * b < nn; b++) {
                    prng_normals4(a.seed + (uint32_t)r, a.tag, (uint64_t)idx, (uint32_t)b, z);
#pragma unroll
                    for (int h2 = 0; h2 < 2; h2++) {
                        if (2 * b + h2 < a.n_steps) {
                            const float d1 = z[2 * h2] * a.sqrt_h, d2 = z[2 * h2 + 1] * a.sqrt_h;
                            ps.stepH(a, d1, a.leg[0].L10 * d1 + a.leg[0].L11 * d2);
                        }
                    }
                }
            } else {
                for (int b = 0; 4 * b < nn; b++) {
                    prng_normals4(a.seed + (uint32_t)r, a.tag, (uint64_t)idx, (uint32_t)b, z);
                    if (4 * b + 4 <= a.n_steps) {      // a full group: no per-step date test (a select per step)
#pragma unroll
                        for (int c = 0; c < 4; c++) ps.step(a, z[c] * a.sqrt_h, 4 * b + c);
                    } else {
#pragma unroll
                        for (int c = 0; c < 4; c++)
                            if (4 * b + c < a.n_steps) ps.step(a, z[c] * a.sqrt_h, 4 * b + c);
                    }
                }
            }
        } else if constexpr (NZ == NZ_SOBOL_RAW) {
            if constexpr (is_heston(MK)) {
                for (int k = 0; k < a.n_steps; k++) {
                    const float d1 = icdf_u32_tab<true>(sobol_coord(sl, 2 * k), itab) * a.sqrt_h;
                    const float d2 = icdf_u32_tab<true>(sobol_coord(sl, 2 * k + 1), itab) * a.sqrt_h;
                    ps.stepH(a, d1, a.leg[0].L10 * d1 + a.leg[0].L11 * d2);
                }
            } else {
                for (int k = 0; k < a.n_steps; k++)
                    ps.step(a, icdf_u32_tab<true>(sobol_coord(sl, k), itab) * a.sqrt_h, k);
            }
        } else {  // NZ_SOBOL_BB, NZ_SOBOL_BB_GLOBAL (one-factor models)
            // W_M from the ends of its interval and coordinate j's normal: with
            // bb_precompute's weights, or on a uniform date grid (weights 1/2, the
            // sd of the depth; no weight loads per coordinate).  Direction
            // records in LDS up to SOBOL_LDS_MAX_D dates, global beyond.
            const int G = a.G, lgG = 31 - __builtin_clz(G);
            auto bridge = [&](auto uniform_tag, auto lds_tag) {
                constexpr bool UNI = decltype(uniform_tag)::value;
                constexpr bool LDS = decltype(lds_tag)::value;
                auto mid = [&](float WL, float WR, int j, int depth) {
                    return bb_mid<UNI, LDS>(a, sl, itab, WL, WR, j, depth);
                };
                if constexpr (LB == 0) {
                    // fewer than 16 dates: the whole skeleton in LDS, one step per entry
                    float* sk = skel + tid;
                    sk[0] = 0.0f;
                    sk[G * 256] = bb_terminal<UNI, LDS>(a, sl, itab);
                    for (int j = 1; j < G; j++) {
                        const int lev = 31 - __builtin_clz(j);
                        const int q = j - (1 << lev);
                        const int st = G >> lev;
                        const int L = q * st, R = L + st, M = L + (st >> 1);
                        sk[M * 256] = mid(sk[L * 256], sk[R * 256], j, lev + 1);
                    }
                    float left = 0.0f;
                    for (int i = 0; i < G; i++) {
                        const float right = sk[(i + 1) * 256];
                        ps.step(a, right - left, i);
                        left = right;
                    }
                } else {
                    // G = 2^m blocks of B dates.  The skeleton (W at the block
                    // boundaries) is walked left to right instead of stored: rt[k]
                    // (a per-lane stack in LDS, k <= m) holds W at the right end of
                    // the aligned 2^k-block interval that contains the current
                    // boundary; moving to boundary i makes the intervals of size
                    // <= 2^ctz(i) start there, and their new right ends are bridge
                    // midpoints generated at that moment, from the same operands as
                    // the top-down skeleton.
                    constexpr int B = 1 << LB;
                    const int m = lgG;
                    float* rt = skel + tid;              // rt[k * 256]
                    // the midpoint of [L, L + 2^(k+1)) blocks: skeleton level m-1-k
                    auto node = [&](int k, int L, float WL, float WR) {
                        const int lev = m - 1 - k;
                        return mid(WL, WR, (1 << lev) + (L >> (k + 1)), lev + 1);
                    };
                    float Wp = 0.0f;
                    rt[m * 256] = bb_terminal<UNI, LDS>(a, sl, itab);
                    for (int k = m - 1; k >= 0; k--) rt[k * 256] = node(k, 0, Wp, rt[(k + 1) * 256]);
                    for (int i = 0; i < G; i++) {
                        if (i > 0) {
                            Wp = rt[0];
                            const int c = __builtin_ctz(i);
                            rt[c * 256] = rt[(c + 1) * 256];
                            for (int k = c - 1; k >= 0; k--) rt[k * 256] = node(k, i, Wp, rt[(k + 1) * 256]);
                        }
                        if constexpr (UNI && B == 16) {
                            float dw[16];
                            bb_block_uniform<LDS, LDS>(a, sl, itab, G, lgG, i, Wp, rt[0], dw);
#pragma unroll
                            for (int k = 0; k < B; k++) ps.step(a16, dw[k], i * B + k);
                        } else {
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
                                    W[Mp] = mid(W[Lp], W[Rp], j0 + q, lgG + l + 1);
                                }
                            }
#pragma unroll
                            for (int k = 0; k < B; k++) ps.step(a, W[k + 1] - W[k], i * B + k);
                        }
                    }
                }
            };
            using LdsRec = std::integral_constant<bool, NZ == NZ_SOBOL_BB>;
            if (a.bb_uniform) bridge(std::true_type{}, LdsRec{});
            else bridge(std::false_type{}, LdsRec{});
        }

        float ym, yc;
        ps.finish(a, ym, yc);
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
