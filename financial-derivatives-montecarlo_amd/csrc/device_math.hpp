// device_math.hpp -- per-lane math shared by the gfx950 path kernels.
//
//  * splitmix32 / Hong-Hickernell row products (methods_cuda.cu:753-795)
//  * inverse normal CDF: own least-squares fit (gen/fit_icdf.py) of -Phi^-1 on
//    the folded uniform as a table of piecewise cubics read from LDS
//  * Philox4x32-10 (Salmon et al., SC'11) + Box-Muller for pseudo-random paths
//  * Euler steppers and payoff accumulators (methods_cuda.cu:271-352)
//  * wave64 FP64 reductions
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "icdf_coeffs.h"
#include "kernel_params.hpp"

namespace fdmc {

__host__ __device__ __forceinline__ uint32_t splitmix32(uint32_t x) {
    x += 0x9E3779B9u;
    x = (x ^ (x >> 16)) * 0x21F0AAADu;
    x = (x ^ (x >> 15)) * 0x735A2D97u;
    return x ^ (x >> 15);
}

__device__ __forceinline__ float fast_log2(float x) { return __builtin_amdgcn_logf(x); }
__device__ __forceinline__ float fast_exp2(float x) { return __builtin_amdgcn_exp2f(x); }
__device__ __forceinline__ float fast_ln(float x) { return __builtin_amdgcn_logf(x) * 0.69314718055994531f; }
__device__ __forceinline__ float fast_exp(float x) { return __builtin_amdgcn_exp2f(x * 1.44269504088896341f); }
__device__ __forceinline__ float fast_sqrt(float x) { return __builtin_amdgcn_sqrtf(x); }

// Inverse standard normal CDF of u = (s + 1/2) 2^-32, never 0 or 1: with
// q = min(u, 1-u) folded to the lower half from the integer (both tails keep
// full relative precision), Phi^-1(u) = -+g(q), g = -Phi^-1 >= 0 on (0, 1/2].
// g is a table of cubics (own fit, gen/fit_icdf.py), 16 segments per binade of
// q read directly from q's FP32 encoding: the row is bits 19..31 of the
// encoding (biased exponent and top 4 mantissa bits) minus FDMC_ICDF_ROW0 (q in
// [2^-33, 1/2] spans biased exponents 94..126), the variable the remaining 19
// mantissa bits under the exponent of 1.0 (the float 1 + x 2^-23, one AND-OR).
// No logarithm and no branch: a 16-byte LDS read, three FMAs and the caller's
// scale.  Kernels copy the table into LDS with
// icdf_tab_load before use.
static __constant__ __align__(16) float g_icdf_tab[FDMC_ICDF_TAB_ROWS * 4] = FDMC_ICDF_TAB;

__device__ __forceinline__ void icdf_tab_load(float4* lds, int tid, int nthreads) {
    for (int i = tid; i < FDMC_ICDF_TAB_ROWS; i += nthreads) lds[i] = ((const float4*)g_icdf_tab)[i];
}

// The table row of encoding b: bits 19..31 of b minus FDMC_ICDF_ROW0 rows.  A
// kernel whose table sits at least FDMC_ICDF_ROW0 rows (24,064 bytes) into LDS
// gets a shift and a mask, the constant folding into the ds_read's offset
// field (BFE = false).  Below that address the constant is negative, which
// that unsigned field cannot hold, and the compiler adds it separately: a
// third instruction.  BFE = true costs two instructions wherever the table
// sits: a bit-field extract, written as the instruction itself (as a builtin
// the compiler turns it back into a shift and a mask), and one shift-and-add
// onto the table's base minus the constant.
template <bool BFE>
__device__ __forceinline__ float4 icdf_row(const float4* tab, uint32_t b) {
    if constexpr (BFE) {
        uint32_t row;
        asm("v_bfe_u32 %0, %1, 19, 12" : "=v"(row) : "v"(b));
        return *(const float4*)((const char*)(tab - FDMC_ICDF_ROW0) + (row << 4));
    } else {
        return tab[((b >> 19) & 0xFFFu) - FDMC_ICDF_ROW0];
    }
}

// -sd Phi^-1((s + 1/2) 2^-32) from the LDS table, for a scale sd > 0.  The
// fold to the lower half comes from reading s as a signed integer i:
// v = i 2^-32 + 2^-33 has |v| = q for both halves, and the row index and the
// variable are read from v's bits below the sign bit, so |v| is never formed.
// (Within 64 of u = 1/2, FP32 rounds q to exactly 1/2, whose row gives g = 0.)
// The result takes its sign bit from s with one OR, and callers fold the
// negation into the sum they add it to.
template <bool BFE = false>
__device__ __forceinline__ float icdf_u32_tab_scaled_neg(uint32_t s, const float4* tab, float sd) {
    const uint32_t b = __float_as_uint(fmaf((float)(int32_t)s, 0x1p-32f, 0x1p-33f));
    const float4 c = icdf_row<BFE>(tab, b);
    const float v = __uint_as_float((b & 0x7FFFFu) | 0x3F800000u);
    const float p = fmaf(fmaf(fmaf(c.w, v, c.z), v, c.y), v, c.x);
    const float y = sd * p;                                // sd g(q) >= 0
    return __uint_as_float(__float_as_uint(y) | (s & 0x80000000u));
}

template <bool BFE = false>
__device__ __forceinline__ float icdf_u32_tab(uint32_t s, const float4* tab) {
    return -icdf_u32_tab_scaled_neg<BFE>(s, tab, 1.0f);
}

// a ^ b ^ c in one VALU instruction: gfx950's three-input bitwise operation
// v_bitop3_b32 with truth table 0x96 (the parity of the three input bits).
__device__ __forceinline__ uint32_t xor3(uint32_t a, uint32_t b, uint32_t c) {
    return __builtin_amdgcn_bitop3_b32(a, b, c, 0x96);
}

// ---- Philox4x32-10 -------------------------------------------------------
struct u32x4 { uint32_t x, y, z, w; };

// The 32 x 32 -> 64-bit products are written as one 64-bit multiply so that
// each becomes a single v_mad_u64_u32 (both halves) instead of a
// v_mul_lo_u32 + v_mul_hi_u32 pair: half the quarter-rate multiplies per round.
__device__ __forceinline__ u32x4 philox4x32_10(u32x4 c, uint32_t k0, uint32_t k1) {
#pragma unroll
    for (int r = 0; r < 10; r++) {
        const uint64_t p0 = (uint64_t)0xD2511F53u * c.x, p1 = (uint64_t)0xCD9E8D57u * c.z;
        const uint32_t lo0 = (uint32_t)p0, hi0 = (uint32_t)(p0 >> 32);
        const uint32_t lo1 = (uint32_t)p1, hi1 = (uint32_t)(p1 >> 32);
        c = u32x4{xor3(hi1, c.y, k0), lo1, xor3(hi0, c.w, k1), lo0};
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    return c;
}

// Two normals by Box-Muller from Philox words (a, b):
// r = sqrt(-2 ln((a + 1/2) 2^-32)); (r cos 2 pi b 2^-32, r sin 2 pi b 2^-32).
// __builtin_amdgcn_{cos,sin}f take their argument in revolutions.
__device__ __forceinline__ void box_muller(uint32_t a, uint32_t b, float& z0, float& z1) {
    const float u1 = fmaf((float)a, 0x1p-32f, 0x1p-33f);
    const float r = fast_sqrt(-2.0f * 0.69314718055994531f * fast_log2(u1));
    const float th = (float)(b >> 8) * 0x1p-24f;
    z0 = r * __builtin_amdgcn_cosf(th);
    z1 = r * __builtin_amdgcn_sinf(th);
}

// Four normals of stream (seed, tag, path), block b (normals 4b .. 4b+3).
__device__ __forceinline__ void prng_normals4(uint32_t seed, uint32_t tag, uint64_t path,
                                              uint32_t b, float z[4]) {
    const u32x4 w = philox4x32_10(u32x4{b, (uint32_t)path, (uint32_t)(path >> 32), tag}, seed,
                                  0x46444D43u);
    box_muller(w.x, w.y, z[0], z[1]);
    box_muller(w.z, w.w, z[2], z[3]);
}

// ---- wave64 / workgroup FP64 reductions ------------------------------------
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// Sum NA per-thread values over a workgroup of NW waves; result valid in
// thread 0.  lds must hold NW * NA doubles.
template <int NA, int NW>
__device__ __forceinline__ void block_sum(double (&v)[NA], double* lds) {
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
    for (int i = 0; i < NA; i++) {
        const double s = wave_sum(v[i]);
        if (lane == 0) lds[wv * NA + i] = s;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
#pragma unroll
        for (int i = 0; i < NA; i++) {
            double s = 0.0;
#pragma unroll
            for (int w = 0; w < NW; w++) s += lds[w * NA + i];
            v[i] = s;
        }
    }
}

// ---- steppers (methods_cuda.cu:271-291, 388-406) ---------------------------
// One-factor Euler step at time t with step h and increment dw,
// S + mu S h + sigma_loc S dw (methods_cuda.cu:271-291) with mu h hoisted out
// of the loops (the term-by-term form takes four operations: without
// reassociation nothing of mu S h can be hoisted): local vols as
// S + S (mu h + sigma_loc dw), two FMAs; GBM as (sigma S) dw + (S + S mu h),
// three (its two-FMA chain with a constant sigma sends the compiler into a very
// long optimisation of the unrolled paths, and S (1 + mu h + sigma dw) would
// round mu h against 1 on every step: a drift bias of order 1e-3 relative).
template <int MK>
__device__ __forceinline__ float step_1f(const LegParams& P, float S, float dw, float t, float h) {
    float sl;
    if constexpr (MK == MK_GBM) {
        return fmaf(P.sigma * S, dw, fmaf(S, P.mu * h, S));
    } else if constexpr (MK == MK_DUPIRE) {
        // sigma0 exp(-alpha t) (S/S0)^(beta-1)
        sl = P.sigma0 * fast_exp(-P.alpha * t) * fast_exp2(P.beta_m1 * fast_log2(S * P.inv_S0));
    } else {  // MK_LVQ: a + b exp(-c t) ln(S/s_ref)^2
        const float l = fast_ln(S * P.inv_sref);
        sl = P.lv_a + P.lv_b * fast_exp(-P.lv_c * t) * l * l;
    }
    return fmaf(S, fmaf(sl, dw, P.mu * h), S);
}

// Full-truncation exponential-Euler Heston step (methods_cuda.cu:278-284) with
// e^{-kappa h} precomputed; MIL adds the Milstein term mil_q (dW2^2 - h),
// mil_q = xi^2/4, of the optional scheme.
// max(V, 0) as one v_max_f32, written as the instruction: fmaxf (or any
// builtin the compiler maps to it) costs two, the compiler first quieting a
// possible signalling NaN in V (v_max_f32 V, V).  The same value for every
// finite V, and V is finite on every path (each step forms it from finite
// terms).
__device__ __forceinline__ float max0(float x) {
    float r;
    asm("v_max_f32 %0, 0, %1" : "=v"(r) : "v"(x));
    return r;
}
template <bool MIL>
__device__ __forceinline__ void step_heston(const LegParams& P, float& S, float& V, float dw1,
                                            float dw2, float h, float em) {
    const float sq = fast_sqrt(max0(V));
    S = S + P.mu * S * h + sq * S * dw1;
    const float v = P.theta + em * (V - P.theta) + P.xi * sq * dw2;
    if constexpr (MIL) V = fmaf(P.mil_q, dw2 * dw2 - h, v);
    else V = v;
}

// ---- payoff accumulators (methods_cuda.cu:298-352) --------------------------
template <int PK>
__device__ __forceinline__ float run_init(const LegParams& P) {
    if constexpr (PK == PK_LOOKBACK) return P.S0;
    else return 0.0f;
}
// HW: the running minimum / maximum as one v_min_f32 / v_max_f32 written as
// the instruction, for loops that carry `run` from one iteration to the next
// one step at a time (fminf / fmaxf then cost two: the compiler quiets a
// possible signalling NaN in the loop-carried value first).  Unrolled blocks
// of steps keep fminf / fmaxf, which the compiler chains into v_min3 / v_max3.
// The two agree for every input the paths produce (quiet NaNs included).
template <int PK, bool HW = false>
__device__ __forceinline__ void run_update(float& run, float S) {
    if constexpr (PK == PK_ASIAN) run += S;
    else if constexpr (PK == PK_GEOM_ASIAN) run += fast_ln(S);
    else if constexpr (PK == PK_LOOKBACK) {
        if constexpr (HW) asm("v_min_f32 %0, %1, %2" : "=v"(run) : "v"(run), "v"(S));
        else run = fminf(run, S);
    } else if constexpr (PK == PK_BARRIER) {
        if constexpr (HW) asm("v_max_f32 %0, %1, %2" : "=v"(run) : "v"(run), "v"(S));
        else run = fmaxf(run, S);
    }
}
template <int PK>
__device__ __forceinline__ constexpr bool is_path_dep() {
    return PK == PK_ASIAN || PK == PK_GEOM_ASIAN || PK == PK_LOOKBACK || PK == PK_BARRIER;
}
// Terminal payoff; bgk is the precomputed exp(-beta sigma sqrt(h)) of the level.
template <int PK>
__device__ __forceinline__ float payoff(const LegParams& P, float S_T, float run, int n, float bgk) {
    if constexpr (PK == PK_EUROPEAN || PK == PK_BASKET) {
        return fmaxf(S_T - P.K, 0.0f) * P.discount;
    } else if constexpr (PK == PK_ASIAN) {
        return fmaxf(run / (float)n - P.K, 0.0f);
    } else if constexpr (PK == PK_GEOM_ASIAN) {
        return fmaxf(fast_exp(run / (float)n) - P.K, 0.0f);
    } else if constexpr (PK == PK_LOOKBACK) {
        // S_T - min * bgk (methods_cuda.cu:346-347), and 0 for a path whose
        // terminal value is not positive or NaN: it left the Euler scheme's
        // domain (S <= 0, where a local vol's log is NaN), and every payoff pays
        // 0 there, as the calls do through fmaxf.  For S_T > 0 the value is
        // unchanged: min <= S_T and bgk <= 1.
        return (S_T > 0.0f) ? S_T - run * bgk : 0.0f;
    } else {  // PK_BARRIER
        return (run / bgk >= P.B) ? 0.0f : fmaxf(S_T - P.K, 0.0f) * P.discount;
    }
}

}  // namespace fdmc
