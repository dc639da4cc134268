// engine_internal.hpp -- declarations shared by the engine's host code and the
// kernel translation units (not part of the public ABI).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "kernel_params.hpp"

namespace fdmc {

// k_sobol.hip
// d_salts null: replica salts salt_base + slot (no salt array in device memory)
hipError_t launch_build_sdir(const uint32_t* dirv, const uint32_t* d_salts, int n_slots, int D,
                             uint32_t* out, hipStream_t st, uint32_t salt_base = 0);
hipError_t launch_sobol_gen(const uint32_t* sdir, int D, int64_t n, uint32_t offset, bool normal,
                            void* out, hipStream_t st);
hipError_t launch_icdf(const uint32_t* in, int64_t n, float* out, hipStream_t st);
hipError_t launch_reduce_partials(const double* part, int n_slots, int nw, int nacc, double* out,
                                  hipStream_t st);

hipError_t launch_reduce_entries(const double* part, const LevelEntry* ent, int n_entries, int nacc,
                                 double* out, hipStream_t st);
hipError_t launch_zero_u32(unsigned int* p, int n, hipStream_t st);

// k_levels_{raw,bb,prng}.hip: multilevel kernel for (model, payoff), or null.
const void* levels_kernel_raw(int mk, int pk);
const void* levels_kernel_bb(int mk, int pk);
const void* levels_kernel_prng(int mk, int pk);
const void* levels_kernel_rows(int mk, int pk);

// k_greeks_{raw,bb,prng}.hip: price + adjoint Greeks kernel, or null.
const void* greeks_kernel_raw(int mk, int pk, int cv, int cpk, int lb);
const void* greeks_kernel_bb(int mk, int pk, int cv, int cpk, int lb);
const void* greeks_kernel_prng(int mk, int pk, int cv, int cpk, int lb);

// k_pca.hip
hipError_t launch_gen_z(const uint32_t* sdir, int D, int ld, int R, int64_t first, int64_t count, bool bf16,
                        void* out, hipStream_t st);
// epi: product epilogue (baskets of GBM assets) instead of storing C
hipError_t launch_gemm_nt(const void* A, const void* Bt, float* C, int Mrows, int N, int K, bool bf16,
                          hipStream_t st, const ProdEpi* epi = nullptr);
// k_is.hip
hipError_t launch_is_gbm(const ISLaunch& a, int nwg, hipStream_t st);
// k_basket.hip
const void* basket_kernel(int mk, int src, int na);
size_t basket_lds_bytes(int src, int D, int na);
const void* basket_wide_kernel(int mk, int src, bool corr, int n);
int basket_wide_block(bool corr);   // lanes per workgroup
int basket_wide_tile(bool corr);    // paths per tile
size_t basket_wide_lds_bytes(int n, bool corr);

// k_book.hip
// gm: 0 price, 1 adjoint Greeks, 2 tangent Greeks; occ 1: the three-waves-per-SIMD compilation
const void* book_kernel(int mk, int gm, int occ);
const void* book_kernel_o3(int mk, int gm);   // k_book_o3.hip
hipError_t launch_book_reset(const BookLaunch& a, hipStream_t s);
hipError_t launch_book_plan(const BookLaunch& a, hipStream_t s);
hipError_t launch_book_sweep(const void* fn, const BookLaunch& a, int grid, hipStream_t s,
                             hipEvent_t e0 = nullptr, hipEvent_t e1 = nullptr);
hipError_t launch_book_pass(const void* fn, const BookLaunch& a, int grid, hipStream_t s,
                            hipEvent_t e0 = nullptr, hipEvent_t e1 = nullptr);
hipError_t launch_book_status(const BookLaunch& a, hipStream_t s);
hipError_t launch_book_final(const BookLaunch& a, hipStream_t s);

// k_path_{raw,bb,prng}.hip: kernel for a (model, payoff, control) combination, or null.
const void* path_kernel_raw(int mk, int pk, int cv, int cpk, int lb);
const void* path_kernel_bb(int mk, int pk, int cv, int cpk, int lb, bool global_records);
const void* path_kernel_prng(int mk, int pk, int cv, int cpk, int lb);
const void* path_kernel_rows(int mk, int pk, int cv, int cpk, int lb);

}  // namespace fdmc
