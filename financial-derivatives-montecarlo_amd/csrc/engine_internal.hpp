// engine_internal.hpp -- declarations shared by the engine's host code and the
// kernel translation units (not part of the public ABI).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "kernel_params.hpp"

namespace fdmc {

// k_sobol.hip
hipError_t launch_build_sdir(const uint32_t* dirv, const uint32_t* d_salts, int n_slots, int D,
                             uint32_t* out, hipStream_t st);
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

// k_greeks_{raw,bb,prng}.hip: price + adjoint Greeks kernel, or null.
const void* greeks_kernel_raw(int mk, int pk, int cv, int cpk, int lb);
const void* greeks_kernel_bb(int mk, int pk, int cv, int cpk, int lb);
const void* greeks_kernel_prng(int mk, int pk, int cv, int cpk, int lb);

// k_path_{raw,bb,prng}.hip: kernel for a (model, payoff, control) combination, or null.
const void* path_kernel_raw(int mk, int pk, int cv, int cpk, int lb);
const void* path_kernel_bb(int mk, int pk, int cv, int cpk, int lb);
const void* path_kernel_prng(int mk, int pk, int cv, int cpk, int lb);

}  // namespace fdmc
