// k_book.inl -- device code of the contract-book kernel (see k_book.hip).
#pragma once

#include "k_greeks.inl"

namespace fdmc {

// Workgroup-uniform tables (the local-vol time factors, the bridge sds, the
// entries and the contracts) are read through the constant address space:
// nothing writes them while the kernel runs, and the compiler then issues
// scalar loads into SGPRs instead of vector loads into VGPRs (it cannot prove
// generic global pointers unclobbered in a kernel that also stores).
template <class T>
using cst = const __attribute__((address_space(4))) T;
template <class T>
__device__ __forceinline__ cst<T>* as_const(const T* p) {
    return (cst<T>*)p;
}

struct BookGeom {
    int l, nf, nc, G;
    float hf, hc, muh_f, muh_c, bgk_f, bgk_c, T;
    float unit;   // price units per path unit: s_ref for the local vol (see book_body), else 1
    float vs_f, vs_c;   // vega of the fine / coarse leg per carried vol-level derivative (scaled blocks)
};

// The quadratic local vol's paths are stepped in units of s_ref (S0, K and B
// divided by s_ref per work item), so log2(S / s_ref) is one log with no
// multiply on every fine and coarse step.  Every payoff is linear in the path
// (and the Euler step homogeneous in S), so the price and its derivatives with
// respect to the vol level and the drift are s_ref times their values in path
// units, and delta (a ratio of two prices) is unchanged.
template <int MK>
constexpr bool book_scaled = MK == MK_LVQ;

// Sensitivities carried along a path: dS/dS0, dS/d(vol parameter), dS/dmu.
struct Tan3 {
    float s, v, m;
};

// The book's Euler step S' = S + mu S h + sigma_loc S dw, evaluated as
// S' = fma(S, fma(sigma_loc, dw, mu h), S) with mu h hoisted per work item (two
// VALU operations; the kernel is VALU-issue bound).  The local vol's time
// factor comes from a table (uniform across the workgroup: a work item is one
// (contract, level), so the step time is the same for every lane):
//   quadratic local vol: v = b e^{-c t} ln(2)^2 and, S in units of s_ref,
//       sigma_loc = a + v log2(S)^2  (= a + b e^{-c t} ln(S/s_ref)^2);
//   Dupire: v = e^{-alpha t}, sigma_loc = sigma0 v (S/S0)^(beta-1);  GBM: unused.
template <int MK>
__device__ __forceinline__ float book_sigma(const LegParams& P, float S, float v, float& aux) {
    if constexpr (MK == MK_GBM) {
        aux = 0.0f;
        return P.sigma;
    } else if constexpr (MK == MK_DUPIRE) {
        aux = fast_exp2(P.beta_m1 * fast_log2(S * P.inv_S0));   // (S/S0)^(beta-1)
        return P.sigma0 * v * aux;
    } else {
        aux = fast_log2(S);                                       // log2(S/s_ref): S in units of s_ref
        return fmaf(v * aux, aux, P.lv_a);                        // v aux is shared with book_partials
    }
}

template <int MK>
__device__ __forceinline__ float book_euler(const LegParams& P, float S, float dw, float v, float muh) {
    if constexpr (MK == MK_GBM) {
        // the reference's form: as one fma chain in S the unrolled GBM path sends the
        // compiler into a very long optimisation, and GBM has no local-vol work to save
        return S + muh * S + P.sigma * S * dw;
    } else {
        float aux;
        const float sl = book_sigma<MK>(P, S, v, aux);
        return fmaf(S, fmaf(sl, dw, muh), S);
    }
}

// Partial derivatives of book_euler: fS = dS'/dS, fvol = dS'/d(sigma, sigma0
// or a), fmu = dS'/dmu, fS0 = dS'/dS0 through the Dupire normalisation (the only
// model whose local vol depends on S0; book_s0_tangent).
template <int MK>
__device__ __forceinline__ void book_partials(const LegParams& P, float S, float dw, float v, float h, float muh,
                                              float& fS, float& fvol, float& fmu, float& fS0) {
    constexpr float TWO_OVER_LN2 = 2.8853900817779268f;   // 2 / ln 2
    float aux;
    const float sl = book_sigma<MK>(P, S, v, aux);
    fmu = S * h;
    fS0 = 0.0f;
    if constexpr (MK == MK_GBM) {
        fS = fmaf(sl, dw, 1.0f + muh);
        fvol = S * dw;
    } else if constexpr (MK == MK_DUPIRE) {
        fS = fmaf(dw * sl, P.beta_m1 + 1.0f, 1.0f + muh);
        fvol = v * aux * S * dw;                       // d/d sigma0
        fS0 = -sl * P.beta_m1 * P.inv_S0 * S * dw;     // local-vol normalisation by S0
    } else {
        // d sigma_loc/dS * S = 2 b e^{-c t} ln(S/s_ref) = (2 / ln 2) v log2(S/s_ref)
        fS = fmaf(dw, fmaf(v * aux, TWO_OVER_LN2, sl), 1.0f + muh);
        fvol = S * dw;                                 // d/d a
    }
}

// Models whose step depends on S0 other than through S (Dupire's (S/S0)^(beta-1)):
// only they carry the fS0 term in the tangents.
template <int MK>
constexpr bool book_s0_tangent = MK == MK_DUPIRE;

// One Euler step S -> S' and, with TAN, the tangent update
// t' = fS t + (f0, fv, fm) (step_partials are the step's partial derivatives).
template <int MK, bool TAN>
__device__ __forceinline__ void book_step(const LegParams& P, float& S, Tan3& t, float dw, float v, float h,
                                          float muh) {
    if constexpr (TAN) {
        float fS, fv, fm, f0;
        book_partials<MK>(P, S, dw, v, h, muh, fS, fv, fm, f0);
        t.s = book_s0_tangent<MK> ? fmaf(fS, t.s, f0) : fS * t.s;
        t.v = fS * t.v + fv;
        t.m = fS * t.m + fm;
    }
    S = book_euler<MK>(P, S, dw, v, muh);
}

// Levels 0..BOOK_SHORT_L (at most BOOK_SHORT_D dates) take the register-only
// path.
constexpr int BOOK_SHORT_L = 4;
constexpr int BOOK_SHORT_D = 1 << BOOK_SHORT_L;

// Records of a level of G >= 2 blocks of 16 dates sit at block_slot(d, G).

// Greeks carried as forward tangents along the path (GM = 2, payoffs with a
// pathwise derivative).
#define LIP_TAN(PK, GM) ((PK) != PK_BARRIER && (GM) == 2)

// Running payoff state of one leg (and, with TAN, the tangent sums the
// payoff's derivative needs: sum over dates for the Asian, the tangent at the
// running minimum for the lookback).
template <int MK, int PK, bool TAN>
struct BookLeg {
    float S, run;
    int kmin;            // index of the running minimum (lookback), 0 = S0
    Tan3 t, acc;         // current tangent; Asian sum / lookback tangent at the minimum
    __device__ __forceinline__ void init(const LegParams& P) {
        S = P.S0;
        run = run_init<PK>(P);
        kmin = 0;
        t = Tan3{1.0f, 0.0f, 0.0f};
        acc = (PK == PK_LOOKBACK) ? Tan3{1.0f, 0.0f, 0.0f} : Tan3{0.0f, 0.0f, 0.0f};
    }
    __device__ __forceinline__ void step(const LegParams& P, float dw, float v, float h, float muh, int idx) {
        book_step<MK, TAN>(P, S, t, dw, v, h, muh);
        if constexpr (PK == PK_LOOKBACK) {
            if (S < run) {
                run = S;
                kmin = idx;
                if constexpr (TAN) acc = t;
            }
        } else {
            run_update<PK>(run, S);
            if constexpr (TAN && PK == PK_ASIAN) {
                acc.s += t.s;
                acc.v += t.v;
                acc.m += t.m;
            }
        }
    }
};

// Direct sensitivities of a payoff: `each` at every S_k (k >= 1), `last` at
// S_n, `cmin` at the running minimum.
template <int PK>
__device__ __forceinline__ void book_direct(const LegParams& P, float S_T, float run, int n, float bgk, float& each,
                                            float& last, float& cmin) {
    each = 0.0f;
    last = 0.0f;
    cmin = 0.0f;
    if constexpr (PK == PK_ASIAN) each = (run / (float)n > P.K) ? 1.0f / (float)n : 0.0f;
    if constexpr (PK == PK_EUROPEAN) last = (S_T > P.K) ? P.discount : 0.0f;
    if constexpr (PK == PK_LOOKBACK) {   // the payoff is 0 unless S_T > 0
        last = (S_T > 0.0f) ? 1.0f : 0.0f;
        cmin = (S_T > 0.0f) ? -bgk : 0.0f;
    }
}

// Adjoint state of one leg during the reverse sweep over blocks: adj = dY/dS
// at the current block boundary through later dates only; v, m, s0 the
// accumulated parameter adjoints.
struct BookAdj {
    float adj, v, m, s0;
};

// One block of the reverse sweep: from the checkpoint S, a forward sweep of the
// block's n steps accumulates the block Jacobian dS_end/dS_start, the
// in-block parameter tangents and the direct payoff terms weighted by
// dS_k/dS_start; then
//   adj_start = sum_k d_k dS_k/dS_start + adj_end dS_end/dS_start,
//   theta_bar += sum_k d_k dS_k/dtheta + adj_end dS_end/dtheta.
// For a one-dimensional state this is the exact vector-Jacobian product of the
// block, with no per-step storage.
// PAIR: the coarse leg, whose increments dw[k] are those of the bridge's
// two-date intervals and whose step times are the even fine times (vb holds the
// block's fine-grid time factors, in registers).  The direct terms d_k by payoff: the Asian's `each` at every
// date, the lookback's `cmin` at its running minimum; the final date's `last`
// is the initial adjoint of the sweep (A.adj at S_n), so no step carries it.
template <int MK, int PK, int N, bool PAIR>
__device__ __forceinline__ void book_block_adj(const LegParams& P, float S, const float* dw, int nsteps, int k0,
                                               const float* vb, float h, float muh, float each, float cmin, int kmin,
                                               BookAdj& A) {
    constexpr bool DIRECT = PK == PK_ASIAN || PK == PK_LOOKBACK;
    float jS = 1.0f, jv = 0.0f, jm = 0.0f, j0 = 0.0f;
    float dS = 0.0f, dv = 0.0f, dm = 0.0f, d0 = 0.0f;
#pragma unroll
    for (int k = 0; k < N; k++) {
        if (k < nsteps) {
            const int gi = k0 + k;
            const float w = dw[k];
            const float v = vb[PAIR ? 2 * k : k];
            float fS, fv, fm, f0;
            book_partials<MK>(P, S, w, v, h, muh, fS, fv, fm, f0);
            S = book_euler<MK>(P, S, w, v, muh);
            jS = fS * jS;
            jv = fS * jv + fv;
            jm = fS * jm + fm;
            if constexpr (book_s0_tangent<MK>) j0 = fS * j0 + f0;
            if constexpr (DIRECT) {
                const float d = (PK == PK_ASIAN) ? each : (gi + 1 == kmin ? cmin : 0.0f);
                dS += d * jS;
                dv += d * jv;
                dm += d * jm;
                if constexpr (book_s0_tangent<MK>) d0 += d * j0;
            }
        }
    }
    A.v += dv + A.adj * jv;
    A.m += dm + A.adj * jm;
    A.s0 += d0 + A.adj * j0;
    A.adj = dS + A.adj * jS;
}

// Payoffs of both legs, their Greeks and the accumulation of the level
// difference.  `sweep(ef, mf, ec, mc, Af, Ac)` runs the adjoint reverse sweep of
// the fine and coarse legs given the payoffs' direct sensitivities (each date,
// running minimum; the final date's is already in Af.adj, Ac.adj).
template <int MK, int PK, int GM, class Sweep>
__device__ __forceinline__ void book_tail(const LegParams& P, const BookGeom& g,
                                          const BookLeg<MK, PK, LIP_TAN(PK, GM)>& F,
                                          const BookLeg<MK, PK, LIP_TAN(PK, GM)>& Cc, bool valid,
                                          double (&acc)[BOOK_NACC], Sweep&& sweep) {
    constexpr bool ADJ = PK != PK_BARRIER && GM == 1;
    constexpr bool TAN = LIP_TAN(PK, GM);
    const int nf = g.nf, nc = g.nc;
    const float pf = payoff<PK>(P, F.S, F.run, nf, g.bgk_f);
    const float pc = (nc > 0) ? payoff<PK>(P, Cc.S, Cc.run, nc, g.bgk_c) : 0.0f;

    // ---- Greeks of P_f and P_c: delta, vega, rho
    Tan3 gf{0.0f, 0.0f, 0.0f}, gc{0.0f, 0.0f, 0.0f};
    if constexpr (TAN) {
        auto grad = [&](const BookLeg<MK, PK, TAN>& Lg, int n, float bgk, float pay) {
            Tan3 r{0.0f, 0.0f, 0.0f};
            if constexpr (PK == PK_EUROPEAN) {
                if (Lg.S > P.K) r = Tan3{P.discount * Lg.t.s, P.discount * Lg.t.v, P.discount * Lg.t.m};
                r.m -= g.T * pay;    // d/dr of the discount e^{-rT}
            } else if constexpr (PK == PK_ASIAN) {
                if (Lg.run / (float)n > P.K) {
                    const float w = 1.0f / (float)n;
                    r = Tan3{Lg.acc.s * w, Lg.acc.v * w, Lg.acc.m * w};
                }
            } else if constexpr (PK == PK_LOOKBACK) {
                if (Lg.S > 0.0f) r = Tan3{Lg.t.s - bgk * Lg.acc.s, Lg.t.v - bgk * Lg.acc.v, Lg.t.m - bgk * Lg.acc.m};
            }
            return r;
        };
        gf = grad(F, nf, g.bgk_f, pf);
        if (nc > 0) gc = grad(Cc, nc, g.bgk_c, pc);
    }
    if constexpr (ADJ) {
        float ef, lf, mf, ec, lc, mc;
        book_direct<PK>(P, F.S, F.run, nf, g.bgk_f, ef, lf, mf);
        book_direct<PK>(P, Cc.S, Cc.run, nc > 0 ? nc : 1, g.bgk_c, ec, lc, mc);
        BookAdj Af{lf, 0.0f, 0.0f, 0.0f}, Ac{lc, 0.0f, 0.0f, 0.0f};   // dY/dS_n direct
        sweep(ef, mf, ec, mc, Af, Ac);
        if (F.kmin == 0) Af.adj += mf;     // S0 itself is the running minimum
        if (Cc.kmin == 0) Ac.adj += mc;
        gf = Tan3{Af.adj + Af.s0, Af.v, Af.m};
        gc = Tan3{Ac.adj + Ac.s0, Ac.v, Ac.m};
        if constexpr (PK == PK_EUROPEAN) {
            gf.m -= g.T * pf;       // d/dr of the discount e^{-rT}
            gc.m -= g.T * pc;
        }
        if (nc == 0) gc = Tan3{0.0f, 0.0f, 0.0f};
    }
    if constexpr (ADJ || TAN) {
        // a leg that left the Euler scheme's domain (S <= 0 under the local vol,
        // after which its states are NaN) pays 0 and contributes 0 to the
        // Greeks; the other leg of the pair counts as usual
        if (!(fabsf(gf.s) + fabsf(gf.v) + fabsf(gf.m) < INFINITY)) gf = Tan3{0.0f, 0.0f, 0.0f};
        if (!(fabsf(gc.s) + fabsf(gc.v) + fabsf(gc.m) < INFINITY)) gc = Tan3{0.0f, 0.0f, 0.0f};
    }
    if (valid) {
        acc[0] += (double)((pf - pc) * g.unit);
        if constexpr (ADJ || TAN) {     // no pathwise Greeks for the barrier: the sums stay 0
            acc[1] += (double)(gf.s - gc.s);
            acc[2] += (double)((gf.v * g.vs_f - gc.v * g.vs_c) * g.unit);
            acc[3] += (double)((gf.m - gc.m) * g.unit);
        }
    }
}

// Levels of NF <= 16 dates (NF = 1, 2, 4, 8, 16): the whole bridge in
// registers, unrolled, with no skeleton in memory; the same arithmetic, in the
// same order, as one block of the long path in its unscaled form (the long
// path's blocks of the quadratic local vol keep scaled increments instead,
// book_path_long; the increments differ from these by rounding only).
// dm[d]: XOR of the current tile's high Sobol words relative to the words in
// the LDS records (the work item's first tile), per dimension.
template <int MK, int PK, int GM, int NF>
__device__ __forceinline__ void book_path_short(const LegParams& P, const BookGeom& g, cst<float>* sdt,
                                                cst<float>* vt, const SobolLane& sl,
                                                const float4* itab, const uint32_t (&dm)[NF], bool valid,
                                                double (&acc)[BOOK_NACC]) {
    constexpr bool TAN = LIP_TAN(PK, GM);
    constexpr int NC = NF / 2;
    // the bridge in increments: dw[0] starts as W_T; level lev splits each
    // interval's increment d at its midpoint into d/2 - y and d/2 + y (y = -sd z
    // of node 2^lev + q), which is W_M = (W_L + W_R)/2 + sd z written for the
    // increments; the two-date intervals' increments are the coarse steps
    float dw[NF > 1 ? NF : 1], dc[NC > 0 ? NC : 1];
    // icdf_u32_tab_scaled_neg gives -sd z
    dw[0] = -icdf_u32_tab_scaled_neg(sobol_coord(sl, 0) ^ dm[0], itab, sdt[0]);
#pragma unroll
    for (int lev = 0; (1 << lev) < NF; lev++) {
        if ((2 << lev) == NF) {
#pragma unroll
            for (int q = 0; q < NC; q++) dc[q] = dw[q];
        }
        const float sd = sdt[lev + 1];
#pragma unroll
        for (int q = (1 << lev) - 1; q >= 0; q--) {
            const int j = (1 << lev) + q;
            const float y = icdf_u32_tab_scaled_neg(sobol_coord(sl, j) ^ dm[j], itab, sd);
            const float d = dw[q];
            dw[2 * q] = fmaf(d, 0.5f, -y);
            dw[2 * q + 1] = fmaf(d, 0.5f, y);
        }
    }
    BookLeg<MK, PK, TAN> F, Cc;
    F.init(P);
    Cc.init(P);
    float vb[NF];
#pragma unroll
    for (int k = 0; k < NF; k++) vb[k] = vt[k];
#pragma unroll
    for (int k = 0; k < NF; k++) {
        F.step(P, dw[k], vb[k], g.hf, g.muh_f, k + 1);
        if (k & 1) Cc.step(P, dc[k >> 1], vb[k - 1], g.hc, g.muh_c, (k >> 1) + 1);
    }
    book_tail<MK, PK, GM>(P, g, F, Cc, valid, acc,
                          [&](float ef, float mf, float ec, float mc, BookAdj& Af, BookAdj& Ac) {
                              if constexpr (NC > 0)
                                  book_block_adj<MK, PK, NC, true>(P, P.S0, dc, NC, 0, vb, g.hc, g.muh_c, ec, mc,
                                                                   Cc.kmin, Ac);
                              book_block_adj<MK, PK, NF, false>(P, P.S0, dw, NF, 0, vb, g.hf, g.muh_f, ef, mf,
                                                                F.kmin, Af);
                          });
}

// Levels of 32 or more dates (G = 2^m blocks of 16, 1 <= m <= 5).  The bridge
// skeleton (W at the block boundaries) is not stored: the forward pass walks
// the boundaries left to right keeping, per dyadic size 2^k, W at the right end
// of the aligned interval that holds the current boundary (rt[k]); moving to
// boundary i makes the intervals of size 2^k <= 2^ctz(i) start at i, the one of
// size 2^ctz(i) ends where its parent does, the smaller ones end at new
// midpoints.  Every skeleton point is generated once, from the same operands
// as the top-down bridge, so the values are identical.  The adjoint's reverse
// sweep walks right to left the same way (lt[k]: left ends).  Each block's 15
// inner points are generated in registers; the adjoint regenerates them block
// by block from the Sobol index, from checkpoints of S.
template <int MK, int PK, int GM>
__device__ __forceinline__ void book_path_long(const LegParams& P, const BookGeom& g, cst<float>* sdt,
                                               cst<float>* vt, const SobolLane& sl, const float4* itab,
                                               float* ckf, float* ckc, bool valid, double (&acc)[BOOK_NACC]) {
    constexpr bool ADJ = PK != PK_BARRIER && GM == 1;
    constexpr bool TAN = LIP_TAN(PK, GM);
    constexpr int MMAX = BOOK_NL - 1 - 4;        // log2 of the most blocks per path
    const int G = g.G;
    const int m = 31 - __builtin_clz(G);
    // W at the midpoint of [L, L + 2^(k+1)) from the ends: the skeleton point
    // of bridge level m-1-k, index L >> (k+1), coordinate 2^(m-1-k) + (L >> (k+1))
    auto node = [&](int k, int L, float WL, float WR) {
        const int lev = m - 1 - k;
        const float sd = sdt[lev + 1];
        return 0.5f * (WL + WR) -
               icdf_u32_tab_scaled_neg(sobol_coord(sl, (1 << lev) + (L >> (k + 1))), itab, sd);
    };
    // icdf_u32_tab_scaled_neg gives -sd z: the bridge subtracts it
    const float WG = -icdf_u32_tab_scaled_neg(sobol_coord(sl, 0), itab, sdt[0]);
    float sdb[4];
#pragma unroll
    for (int lv = 0; lv < 4; lv++) sdb[lv] = sdt[m + lv + 1];
    // Under the quadratic local vol the blocks are built in scaled increments: an
    // increment at in-block depth lv is kept 2^lv times its size, so the children
    // of d are d -+ 2^(lv+1) sd z with no halving and no product sd z of their own
    // (one FMA each, one VALU less per normal); the fine steps' increments come
    // out 16 and the coarse steps' 8 times their size, and the steps use the
    // vol's level and time factors times 1/16 and 1/8 (the voltab's planes 1 and
    // 2; powers of two, so sigma dw is the same product).  Of the Euler step's
    // partial derivatives only d/d(vol level) = S dw carries the scale: the
    // legs' vegas are multiplied by 1/16 and 1/8 at the end (BookGeom::vs_*).
    constexpr bool SC = MK == MK_LVQ;
    float sds[4];
#pragma unroll
    for (int lv = 0; lv < 4; lv++) sds[lv] = sdb[lv] * (float)(2 << lv);
    LegParams P16 = P, P8 = P;
    P16.lv_a = P.lv_a * 0.0625f;
    P8.lv_a = P.lv_a * 0.125f;
    // the block's increments from W at its ends, level by level in increments
    // (as the short path): each interval's increment d splits into d/2 - y and
    // d/2 + y; dc receives the two-date intervals' increments (coarse steps)
    auto gen_block = [&](int i, float WL, float WR, float (&dw)[16], float (&dc)[8]) {
        dw[0] = WR - WL;
        const int slot0 = G + 15 * i;       // the block's records (block_slot)
#pragma unroll
        for (int lv = 0; lv < 4; lv++) {
            if (lv == 3) {
#pragma unroll
                for (int q = 0; q < 8; q++) dc[q] = dw[q];
            }
#pragma unroll
            for (int q = (1 << lv) - 1; q >= 0; q--) {
                const float d = dw[q];
                if constexpr (SC) {
                    const float y = icdf_u32_tab_scaled_neg(sobol_coord(sl, slot0 + (1 << lv) - 1 + q), itab, 1.0f);
                    dw[2 * q] = fmaf(-sds[lv], y, d);
                    dw[2 * q + 1] = fmaf(sds[lv], y, d);
                } else {
                    const float y = icdf_u32_tab_scaled_neg(sobol_coord(sl, slot0 + (1 << lv) - 1 + q), itab, sdb[lv]);
                    dw[2 * q] = fmaf(d, 0.5f, -y);
                    dw[2 * q + 1] = fmaf(d, 0.5f, y);
                }
            }
        }
    };

    BookLeg<MK, PK, TAN> F, Cc;
    F.init(P);
    Cc.init(P);
    // the block's 16 local-vol time factors, loaded once into registers (the
    // coarse step at fine index 2 gc uses the even ones)
    auto load_vb = [&](int i, float (&vb)[16]) {
#pragma unroll
        for (int k = 0; k < 16; k++) vb[k] = vt[i * 16 + k];
    };
    float rt[MMAX + 1];
    float Wp = 0.0f;                              // W at the current boundary i
#pragma unroll
    for (int k = 0; k <= MMAX; k++)
        if (k == m) rt[k] = WG;
#pragma unroll
    for (int k = MMAX - 1; k >= 0; k--)
        if (k < m) rt[k] = node(k, 0, Wp, rt[k + 1]);
    for (int i = 0; i < G; i++) {
        if (i > 0) {
            Wp = rt[0];
            const int c = __builtin_ctz(i);       // < m
#pragma unroll
            for (int k = MMAX - 1; k >= 0; k--) {
                if (k == c) rt[k] = rt[k + 1];
                else if (k < c) rt[k] = node(k, i, Wp, rt[k + 1]);
            }
        }
        float dw[16], dc[8], vb[16];
        gen_block(i, Wp, rt[0], dw, dc);
        if constexpr (ADJ) {
            ckf[i * 256] = F.S;
            ckc[i * 256] = Cc.S;
        }
        if constexpr (SC) {
            cst<float>* v16 = vt + BOOK_VT_PLANE + i * 16;
            cst<float>* v8 = vt + 2 * BOOK_VT_PLANE + i * 16;
            float vc[8];
#pragma unroll
            for (int k = 0; k < 16; k++) vb[k] = v16[k];
#pragma unroll
            for (int k = 0; k < 8; k++) vc[k] = v8[2 * k];
#pragma unroll
            for (int k = 0; k < 16; k++) {
                const int gi = i * 16 + k;
                F.step(P16, dw[k], vb[k], g.hf, g.muh_f, gi + 1);
                if (k & 1) Cc.step(P8, dc[k >> 1], vc[k >> 1], g.hc, g.muh_c, (gi >> 1) + 1);
            }
        } else {
            load_vb(i, vb);
#pragma unroll
            for (int k = 0; k < 16; k++) {
                const int gi = i * 16 + k;
                F.step(P, dw[k], vb[k], g.hf, g.muh_f, gi + 1);
                if (k & 1) Cc.step(P, dc[k >> 1], vb[k - 1], g.hc, g.muh_c, (gi >> 1) + 1);
            }
        }
    }
    BookGeom gt = g;
    if constexpr (SC) {
        gt.vs_f = 0.0625f;
        gt.vs_c = 0.125f;
    }
    book_tail<MK, PK, GM>(P, gt, F, Cc, valid, acc,
                          [&](float ef, float mf, float ec, float mc, BookAdj& Af, BookAdj& Ac) {
                              float lt[MMAX + 1];
                              float WR = WG;        // W at the current block's right end i + 1
#pragma unroll
                              for (int k = 0; k <= MMAX; k++)
                                  if (k == m) lt[k] = 0.0f;
#pragma unroll
                              for (int k = MMAX - 1; k >= 0; k--)
                                  if (k < m) lt[k] = node(k, G - (2 << k), lt[k + 1], WR);
                              for (int i = G - 1; i >= 0; i--) {
                                  if (i < G - 1) {
                                      WR = lt[0];
                                      const int p = i + 1, c = __builtin_ctz(p);   // < m
#pragma unroll
                                      for (int k = MMAX - 1; k >= 0; k--) {
                                          if (k == c) lt[k] = lt[k + 1];
                                          else if (k < c) lt[k] = node(k, p - (2 << k), lt[k + 1], WR);
                                      }
                                  }
                                  float dw[16], dc[8], vb[16];
                                  gen_block(i, lt[0], WR, dw, dc);
                                  if constexpr (SC) {
                                      float v8[16];
                                      cst<float>* p16 = vt + BOOK_VT_PLANE + i * 16;
                                      cst<float>* p8 = vt + 2 * BOOK_VT_PLANE + i * 16;
#pragma unroll
                                      for (int k = 0; k < 16; k++) {
                                          vb[k] = p16[k];
                                          v8[k] = (k & 1) ? 0.0f : p8[k];    // the coarse steps read the even ones
                                      }
                                      book_block_adj<MK, PK, 8, true>(P8, ckc[i * 256], dc, 8, i * 8, v8, g.hc,
                                                                      g.muh_c, ec, mc, Cc.kmin, Ac);
                                      book_block_adj<MK, PK, 16, false>(P16, ckf[i * 256], dw, 16, i * 16, vb,
                                                                        g.hf, g.muh_f, ef, mf, F.kmin, Af);
                                  } else {
                                      load_vb(i, vb);
                                      book_block_adj<MK, PK, 8, true>(P, ckc[i * 256], dc, 8, i * 8, vb, g.hc,
                                                                      g.muh_c, ec, mc, Cc.kmin, Ac);
                                      book_block_adj<MK, PK, 16, false>(P, ckf[i * 256], dw, 16, i * 16, vb,
                                                                        g.hf, g.muh_f, ef, mf, F.kmin, Af);
                                  }
                              }
                          });
}

// Brownian bridge on the uniform grid of 2^l dates (bb_precompute,
// utils.cpp:54-101): every midpoint weight is 1/2 and the conditional sd
// depends only on the depth, so the bridge is W_M = (W_L + W_R) / 2 + sd[k] z
// with sdt[0] = sqrt(T) for the terminal value and sdt[k] for the midpoints
// of schedule level k - 1 (FP32 values of bb_precompute's weights).
//
// GM: 0 price only, 1 adjoint Greeks (reverse sweep over 16-step blocks with
// checkpoints, increments regenerated from the Sobol index), 2 forward
// tangents (carried along the forward pass; no second pass).
// The tiles [t0, t1) of a work item at a level of NF <= 16 dates.  The LDS
// records hold the high Sobol words of the item's first tile; each lane keeps
// their change since then per dimension in registers (dm), from the
// direction numbers of bits 7..31 in LDS (hl[d][25]): tile n changes
// gray(256 n) in bits 7 and 8 + ctz(n) only.  No barrier between tiles.
template <int MK, int GM, int NF>
__device__ __forceinline__ void book_tiles_short(const LegParams& P, const BookGeom& g, cst<float>* sdt,
                                                 cst<float>* vt, const SobolLane& sl,
                                                 const float4* itab, const uint32_t* hl, int kind, int64_t tfirst,
                                                 int t0, int t1, int64_t first, int64_t count, int tid,
                                                 double (&acc)[BOOK_NACC]) {
    uint32_t dm[NF];
#pragma unroll
    for (int d = 0; d < NF; d++) dm[d] = 0u;
    for (int t = t0; t < t1; t++) {
        const uint32_t n = (uint32_t)(tfirst + t);
        if (t > t0) {
            const int c = __builtin_ctz(n);
#pragma unroll
            for (int d = 0; d < NF; d++) dm[d] ^= hl[d * 25] ^ hl[d * 25 + 1 + c];
        }
        const int64_t idx = ((int64_t)n << 8) + tid;
        const bool valid = idx >= first && idx < first + count;
        switch (kind) {
        case PK_EUROPEAN: book_path_short<MK, PK_EUROPEAN, GM, NF>(P, g, sdt, vt, sl, itab, dm, valid, acc); break;
        case PK_ASIAN: book_path_short<MK, PK_ASIAN, GM, NF>(P, g, sdt, vt, sl, itab, dm, valid, acc); break;
        case PK_LOOKBACK: book_path_short<MK, PK_LOOKBACK, GM, NF>(P, g, sdt, vt, sl, itab, dm, valid, acc); break;
        default: book_path_short<MK, PK_BARRIER, GM, NF>(P, g, sdt, vt, sl, itab, dm, valid, acc); break;
        }
    }
}

// A workgroup-uniform value read from memory, moved into a scalar register.
__device__ __forceinline__ int sreg(int x) { return __builtin_amdgcn_readfirstlane(x); }
__device__ __forceinline__ float sreg(float x) { return __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(x))); }
__device__ __forceinline__ int64_t sreg(int64_t x) {
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)x);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)((uint64_t)x >> 32));
    return (int64_t)(((uint64_t)hi << 32) | lo);
}
__device__ __forceinline__ double sreg(double x) { return __longlong_as_double(sreg((int64_t)__double_as_longlong(x))); }

template <int MK, int GM>
__device__ __forceinline__ void book_body(const BookLaunch& a) {
    // static LDS: addresses fold into the instructions' offset fields instead
    // of adding a dynamic base to each
    __shared__ __align__(16) unsigned char smem[book_lds_static()];
    const int tid = threadIdx.x, lane = tid & 63;
    uint32_t* rec = (uint32_t*)smem;                                     // [BOOK_DMAX][20]
    constexpr size_t REC_BYTES = (size_t)BOOK_DMAX * 20 * 4;
    float4* itab = (float4*)(smem + REC_BYTES);                          // inverse-CDF table
    double* red = (double*)(smem + REC_BYTES + (size_t)FDMC_ICDF_TAB_ROWS * 16);
    int* s_item = (int*)(red + 4 * BOOK_NACC);
    // the adjoint's checkpoints of S (fine, coarse) at the 16-step block starts
    float* wscr = a.scratch + (size_t)blockIdx.x * book_scratch_floats();
    float* ckf = wscr + tid;
    float* ckc = ckf + BOOK_GMAX * 256;

    SobolLane sl;
    {
        const uint32_t gl = (uint32_t)(lane ^ (lane >> 1));
        sl.rec = rec;
        sl.grec = nullptr;
        sl.ga = gl & 7u;
        sl.gb = gl >> 3;
        sl.w = (uint32_t)(tid >> 6);
        sl.hi_gray = 0;
        sl.lds_mode = true;
    }
    icdf_tab_load(itab, tid, 256);        // visible after the loop's first barrier
    const int total = *a.total_items;     // written by k_book_scan
    int last_e = -1;
    for (;;) {
        __syncthreads();
        if (tid == 0) *s_item = (int)atomicAdd(a.queue, 1u);
        __syncthreads();
        // uniform over the workgroup: the entry, contract, tables and loop bounds
        // derived from it live in scalar registers and load through scalar loads
        const int item = __builtin_amdgcn_readfirstlane(*s_item);
        if (item >= total) break;         // every workgroup leaves once the queue is empty
        int lo = 0, hi = a.n_entries - 1; // last entry with tile_begin <= item
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (as_const(a.entries)[mid].tile_begin <= item) lo = mid;
            else hi = mid - 1;
        }
        // the entry and contract are uniform over the workgroup, but the
        // compiler cannot prove their tables read-only and loads them with
        // vector loads: readfirstlane moves every field into scalar registers
        // (fewer VGPRs live across the tile loops)
        cst<BookEntry>& Ev = as_const(a.entries)[lo];
        BookEntry E;
        E.c = Ev.c;
        E.l = Ev.l;
        E.r = Ev.r;
        E.tpi = Ev.tpi;
        E.tile_begin = Ev.tile_begin;
        E.ntiles = 0;
        E.first = Ev.first;
        E.count = Ev.count;
        cst<BookContract>& C = as_const(a.contracts)[E.c];
        const double CT = C.T;
        BookGeom g;
        g.l = E.l;
        g.nf = 1 << E.l;
        g.nc = g.nf >> 1;
        g.G = g.nf >= 16 ? g.nf >> 4 : g.nf;
        g.hf = (float)(CT / (double)g.nf);
        g.hc = g.nc > 0 ? (float)(CT / (double)g.nc) : 0.0f;
        g.muh_f = a.leg.mu * g.hf;
        g.muh_c = a.leg.mu * g.hc;
        g.bgk_f = C.bgk[E.l];
        g.bgk_c = E.l > 0 ? C.bgk[E.l - 1] : 1.0f;
        g.T = C.Tf;
        const int kind = C.kind;
        LegParams P = a.leg;
        P.K = C.K;
        P.B = C.B;
        P.discount = C.discount;
        g.unit = 1.0f;
        g.vs_f = 1.0f;
        g.vs_c = 1.0f;
        if constexpr (book_scaled<MK>) {
            P.S0 *= P.inv_sref;
            P.K *= P.inv_sref;
            P.B *= P.inv_sref;
            g.unit = 1.0f / P.inv_sref;
        }
        const int D = g.nf;
        const int slot = C.slot[E.l], t_idx = C.t_idx;
        const uint32_t* grec = a.sdir[E.l] + ((size_t)slot * a.R + E.r) * (size_t)D * SDIR_WORDS;
        cst<float>* sdt = as_const(a.bbtab) + ((size_t)t_idx * BOOK_NL + E.l) * BOOK_NL;
        cst<float>* vt = as_const(a.voltab) + (size_t)t_idx * BOOK_VT_STRIDE + (size_t)(D - 1);
        const int64_t tfirst = E.first >> 8;
        const int nt_all = (int)(((E.first + E.count - 1) >> 8) - tfirst + 1);
        const int t0 = (item - E.tile_begin) * E.tpi;
        const int t1 = min(t0 + E.tpi, nt_all);
        double acc[BOOK_NACC];
#pragma unroll
        for (int i = 0; i < BOOK_NACC; i++) acc[i] = 0.0;
        // LDS records of the entry (direction words of the lane and wave bits),
        // then words 16..19: the wave offsets' direction sums XOR the sum over
        // the set bits of gray(base) (bits 7..31; base is a multiple of 256
        // below 2^32) for the item's first tile, loads in parallel.  The
        // previous item ended with barriers, so the records are free.
        // Levels of at most 16 dates (dimensions) also keep the direction
        // numbers of bits 7..31 in LDS after the 16 records (hl[d][25]).
        uint32_t* hl = rec + BOOK_SHORT_D * 20;
        // records in LDS slots: identity for levels of at most 16 dates,
        // block_slot (block-contiguous) above
        const int Gs = E.l > BOOK_SHORT_L ? g.G : D;
        if (lo != last_e) {
            for (int i = tid; i < D * 16; i += 256) {
                const int d = i >> 4, k = i & 15;
                rec[block_slot(d, Gs) * 20 + k] = grec[(int64_t)d * SDIR_WORDS + k];
            }
            if (E.l <= BOOK_SHORT_L)
                for (int i = tid; i < D * 25; i += 256) hl[i] = grec[(int64_t)(i / 25) * SDIR_WORDS + 39 + i % 25];
            last_e = lo;
        }
        {
            const uint32_t b32 = (uint32_t)(tfirst + t0) << 8, gbase = b32 ^ (b32 >> 1);
            for (int d = tid; d < D; d += 256) {
                const uint32_t* gr = grec + (int64_t)d * SDIR_WORDS;
                uint32_t s = 0;
#pragma unroll
                for (int j = 7; j < 32; j++) s ^= ((gbase >> j) & 1u) ? gr[32 + j] : 0u;
                *(uint4*)(rec + block_slot(d, Gs) * 20 + 16) = make_uint4(s ^ gr[16], s ^ gr[17], s ^ gr[18], s ^ gr[19]);
            }
        }
        __syncthreads();
        // Each further tile n changes gray(256 n) in bits 7 and 8 + ctz(n) only:
        // two direction numbers per dimension.
        if (E.l <= BOOK_SHORT_L) {
            switch (E.l) {
            case 0: book_tiles_short<MK, GM, 1>(P, g, sdt, vt, sl, itab, hl, kind, tfirst, t0, t1, E.first, E.count, tid, acc); break;
            case 1: book_tiles_short<MK, GM, 2>(P, g, sdt, vt, sl, itab, hl, kind, tfirst, t0, t1, E.first, E.count, tid, acc); break;
            case 2: book_tiles_short<MK, GM, 4>(P, g, sdt, vt, sl, itab, hl, kind, tfirst, t0, t1, E.first, E.count, tid, acc); break;
            case 3: book_tiles_short<MK, GM, 8>(P, g, sdt, vt, sl, itab, hl, kind, tfirst, t0, t1, E.first, E.count, tid, acc); break;
            default: book_tiles_short<MK, GM, 16>(P, g, sdt, vt, sl, itab, hl, kind, tfirst, t0, t1, E.first, E.count, tid, acc); break;
            }
        } else {
            for (int t = t0; t < t1; t++) {
                const uint32_t n = (uint32_t)(tfirst + t);
                if (t > t0) {
                    __syncthreads();      // every lane is done with the previous tile's words
                    const int c = __builtin_ctz(n);
                    for (int d = tid; d < D; d += 256) {
                        const uint32_t* gr = grec + (int64_t)d * SDIR_WORDS;
                        const uint32_t dl = gr[32 + 7] ^ gr[32 + 8 + c];
                        uint4* r = (uint4*)(rec + block_slot(d, Gs) * 20 + 16);
                        const uint4 v = *r;
                        *r = make_uint4(v.x ^ dl, v.y ^ dl, v.z ^ dl, v.w ^ dl);
                    }
                    __syncthreads();
                }
                const int64_t idx = ((int64_t)n << 8) + tid;
                const bool valid = idx >= E.first && idx < E.first + E.count;
                switch (kind) {
                case PK_EUROPEAN: book_path_long<MK, PK_EUROPEAN, GM>(P, g, sdt, vt, sl, itab, ckf, ckc, valid, acc); break;
                case PK_ASIAN: book_path_long<MK, PK_ASIAN, GM>(P, g, sdt, vt, sl, itab, ckf, ckc, valid, acc); break;
                case PK_LOOKBACK: book_path_long<MK, PK_LOOKBACK, GM>(P, g, sdt, vt, sl, itab, ckf, ckc, valid, acc); break;
                default: book_path_long<MK, PK_BARRIER, GM>(P, g, sdt, vt, sl, itab, ckf, ckc, valid, acc); break;
                }
            }
        }
        block_sum<BOOK_NACC, 4>(acc, red);
        if (tid == 0) {
            double* out = a.partials + (int64_t)item * BOOK_NACC;
#pragma unroll
            for (int i = 0; i < BOOK_NACC; i++) out[i] = acc[i];
        }
    }
}

// Two compilations of the same body: the compiler's own register allocation,
// and one asked for three waves per SIMD (it may spill to reach it).  The
// launch-shape autotuner times both.
template <int MK, int GM>
__global__ __launch_bounds__(256) void k_book(BookLaunch a) {
    book_body<MK, GM>(a);
}

template <int MK, int GM>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3))) void k_book_o3(BookLaunch a) {
    book_body<MK, GM>(a);
}

// Kernel of a (model, Greeks mode) for one register-allocation choice.
template <template <int, int> class Sel, int MK>
const void* book_kernel_sel(int gm) {
    switch (gm) {
    case 0: return Sel<MK, 0>::fn();
    case 1: return Sel<MK, 1>::fn();
    case 2: return Sel<MK, 2>::fn();
    default: return nullptr;
    }
}

template <template <int, int> class Sel>
const void* book_kernel_any(int mk, int gm) {
    switch (mk) {
    case MK_GBM: return book_kernel_sel<Sel, MK_GBM>(gm);
    case MK_DUPIRE: return book_kernel_sel<Sel, MK_DUPIRE>(gm);
    case MK_LVQ: return book_kernel_sel<Sel, MK_LVQ>(gm);
    default: return nullptr;
    }
}

}  // namespace fdmc
