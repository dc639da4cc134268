// k_sobol.hip -- Hong-Hickernell scrambled Sobol on gfx950.
//
// The reference scrambles every coordinate of every point by 32 row products
// (32 splitmix32 hashes + 32 popcounts, methods_cuda.cu:781-795), inside a
// kernel that also walks curand's Sobol state (:799-825).  The scramble is
// F2-linear: scrambled(i) = L raw(i) ^ e and raw(i) = XOR_{j in gray(i)} v_j,
// hence scrambled(i) = e ^ XOR_{j in gray(i)} (L v_j).  k_build_sdir computes
// the 32 scrambled direction words L v_j once per (replica, dimension) with the
// reference's own row construction, plus small lookup tables; the path kernels
// then need 2-3 table reads and XORs per coordinate.  Results are bit-identical.
#include "device_math.hpp"
#include "engine_internal.hpp"

namespace fdmc {

// One thread per (slot, dimension).  salts[slot] is the replica salt, or
// salt_base + slot when salts is null.
__global__ __launch_bounds__(256) void k_build_sdir(const uint32_t* __restrict__ dirv,
                                                    const uint32_t* __restrict__ salts, uint32_t salt_base,
                                                    int n_slots, int D, uint32_t* __restrict__ out) {
    const int64_t gid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (gid >= (int64_t)n_slots * D) return;
    const int slot = (int)(gid / D), d = (int)(gid - (int64_t)slot * D);
    const uint32_t salt = salts ? salts[slot] : salt_base + (uint32_t)slot;
    const uint32_t base = splitmix32(salt) ^ (uint32_t)d;
    uint32_t row[32];
#pragma unroll
    for (int k = 0; k < 32; k++) {
        const uint32_t seed = splitmix32(base ^ (uint32_t)k);
        const uint32_t top = (k == 0) ? 0u : (0xFFFFFFFFu << (32 - k));
        row[k] = (1u << (31 - k)) | (seed & top);
    }
    const uint32_t e = splitmix32(base ^ 0xA5A5A5A5u);
    uint32_t* rec = out + gid * SDIR_WORDS;
    uint32_t Lv[32];
#pragma unroll 4
    for (int j = 0; j < 32; j++) {
        const uint32_t v = dirv[(int64_t)d * 32 + j];
        uint32_t x = 0;
#pragma unroll
        for (int k = 0; k < 32; k++) x |= ((uint32_t)__popc(row[k] & v) & 1u) << (31 - k);
        Lv[j] = x;
        rec[32 + j] = x;
    }
#pragma unroll
    for (int k = 0; k < 8; k++) {
        uint32_t a = e, b = 0;
#pragma unroll
        for (int j = 0; j < 3; j++) {
            if ((k >> j) & 1) { a ^= Lv[j]; b ^= Lv[j + 3]; }
        }
        rec[k] = a;
        rec[8 + k] = b;
    }
#pragma unroll
    for (int w = 0; w < 4; w++) {
        const uint32_t i = (uint32_t)w << 6, g = i ^ (i >> 1);
        uint32_t x = 0;
#pragma unroll
        for (int j = 5; j < 8; j++)
            if ((g >> j) & 1u) x ^= Lv[j];
        rec[16 + w] = x;
    }
#pragma unroll
    for (int k = 20; k < 32; k++) rec[k] = 0u;
}

// Generic scrambled point from a record (any index).
__device__ __forceinline__ uint32_t sobol_from_record(const uint32_t* __restrict__ rec, uint32_t i) {
    uint32_t g = i ^ (i >> 1);
    uint32_t x = rec[0];  // TA[0] = shift
    while (g) {
        x ^= rec[32 + __builtin_ctz(g)];
        g &= g - 1u;
    }
    return x;
}

// out[d * n + p] = scrambled point (offset + p) of dimension d (layout of
// kernel_gen_hh_scrambled_sobol_normal, methods_cuda.cu:799-825).
template <bool NORMAL>
__global__ __launch_bounds__(256) void k_sobol_gen(const uint32_t* __restrict__ sdir, int D, int64_t n,
                                                   uint32_t offset, void* __restrict__ out) {
    const int64_t gid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    // inverse-CDF table (icdf_u32_tab), filled before the first barrier-separated use
    constexpr bool ITAB = NORMAL;
    __shared__ __align__(16) float4 itab[ITAB ? FDMC_ICDF_TAB_ROWS : 1];
    if constexpr (ITAB) {
        icdf_tab_load(itab, threadIdx.x, 256);
        __syncthreads();
    }
    if (gid >= (int64_t)D * n) return;
    const int d = (int)(gid / n);
    const int64_t p = gid - (int64_t)d * n;
    const uint32_t x = sobol_from_record(sdir + (int64_t)d * SDIR_WORDS, offset + (uint32_t)p);
    if constexpr (NORMAL) ((float*)out)[gid] = icdf_u32_tab(x, itab);
    else ((uint32_t*)out)[gid] = x;
}

__global__ __launch_bounds__(256) void k_icdf(const uint32_t* __restrict__ in, int64_t n,
                                              float* __restrict__ out) {
    const int64_t gid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    // inverse-CDF table (icdf_u32_tab), filled before the first barrier-separated use
    constexpr bool ITAB = true;
    __shared__ __align__(16) float4 itab[ITAB ? FDMC_ICDF_TAB_ROWS : 1];
    if constexpr (ITAB) {
        icdf_tab_load(itab, threadIdx.x, 256);
        __syncthreads();
    }
    if (gid < n) out[gid] = icdf_u32_tab(in[gid], itab);
}

// Deterministic second-stage reduction: out[s][i] = sum_{w < nw} part[s][w][i],
// one workgroup per slot, fixed assignment of partials to threads.
__global__ __launch_bounds__(256) void k_reduce_partials(const double* __restrict__ part, int nw,
                                                         int nacc, double* __restrict__ out) {
    __shared__ double lds[4 * 16];
    const int s = blockIdx.x;
    const double* p = part + (int64_t)s * nw * nacc;
    for (int i = 0; i < nacc; i++) {
        double v[1] = {0.0};
        for (int w = threadIdx.x; w < nw; w += 256) v[0] += p[(int64_t)w * nacc + i];
        block_sum<1, 4>(v, lds);
        if (threadIdx.x == 0) out[(int64_t)s * nacc + i] = v[0];
        __syncthreads();
    }
}

// Segmented variant: out[e][i] = sum_{t < len_e} part[begin_e + t][i], entries
// taken from the multilevel entry table.
__global__ __launch_bounds__(256) void k_reduce_entries(const double* __restrict__ part,
                                                        const LevelEntry* __restrict__ ent, int nacc,
                                                        double* __restrict__ out) {
    __shared__ double lds[4 * 16];
    const LevelEntry E = ent[blockIdx.x];
    const double* p = part + (int64_t)E.tile_begin * nacc;
    for (int i = 0; i < nacc; i++) {
        double v[1] = {0.0};
        for (int w = threadIdx.x; w < E.ntiles; w += 256) v[0] += p[(int64_t)w * nacc + i];
        block_sum<1, 4>(v, lds);
        if (threadIdx.x == 0) out[(int64_t)blockIdx.x * nacc + i] = v[0];
        __syncthreads();
    }
}

__global__ void k_zero_u32(unsigned int* p, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0u;
}

// ---- host launchers ----------------------------------------------------------
hipError_t launch_build_sdir(const uint32_t* dirv, const uint32_t* d_salts, int n_slots, int D,
                             uint32_t* out, hipStream_t st, uint32_t salt_base) {
    const int64_t n = (int64_t)n_slots * D;
    const int blocks = (int)((n + 255) / 256);
    hipLaunchKernelGGL(k_build_sdir, dim3(blocks), dim3(256), 0, st, dirv, d_salts, salt_base, n_slots, D, out);
    return hipGetLastError();
}

hipError_t launch_sobol_gen(const uint32_t* sdir, int D, int64_t n, uint32_t offset, bool normal,
                            void* out, hipStream_t st) {
    const int64_t tot = (int64_t)D * n;
    const int64_t blocks = (tot + 255) / 256;
    if (blocks > 0x7FFFFFFF) return hipErrorInvalidValue;
    if (normal)
        hipLaunchKernelGGL(k_sobol_gen<true>, dim3((unsigned)blocks), dim3(256), 0, st, sdir, D, n, offset, out);
    else
        hipLaunchKernelGGL(k_sobol_gen<false>, dim3((unsigned)blocks), dim3(256), 0, st, sdir, D, n, offset, out);
    return hipGetLastError();
}

hipError_t launch_icdf(const uint32_t* in, int64_t n, float* out, hipStream_t st) {
    const int64_t blocks = (n + 255) / 256;
    if (blocks > 0x7FFFFFFF) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_icdf, dim3((unsigned)blocks), dim3(256), 0, st, in, n, out);
    return hipGetLastError();
}

hipError_t launch_reduce_partials(const double* part, int n_slots, int nw, int nacc, double* out,
                                  hipStream_t st) {
    hipLaunchKernelGGL(k_reduce_partials, dim3(n_slots), dim3(256), 0, st, part, nw, nacc, out);
    return hipGetLastError();
}

hipError_t launch_reduce_entries(const double* part, const LevelEntry* ent, int n_entries, int nacc,
                                 double* out, hipStream_t st) {
    hipLaunchKernelGGL(k_reduce_entries, dim3(n_entries), dim3(256), 0, st, part, ent, nacc, out);
    return hipGetLastError();
}

hipError_t launch_zero_u32(unsigned int* p, int n, hipStream_t st) {
    hipLaunchKernelGGL(k_zero_u32, dim3((n + 255) / 256), dim3(256), 0, st, p, n);
    return hipGetLastError();
}

}  // namespace fdmc
