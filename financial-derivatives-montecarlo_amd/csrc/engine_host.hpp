// engine_host.hpp -- host-side internals shared by the engine's translation
// units (engine.cpp, book.cpp): error handling, device buffers, the context.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/fdmc.h"
#include "engine_internal.hpp"
#include "kernel_params.hpp"

namespace fdmc_host {

// Message of the last failed C ABI call of this thread (fdmc_last_error).
extern thread_local std::string g_last_error;

struct Fail {
    fdmc_status code;
    std::string msg;
};

[[noreturn]] inline void fail(fdmc_status c, const std::string& m) { throw Fail{c, m}; }

inline void hip_check(hipError_t e, const char* what) {
    if (e != hipSuccess) fail(FDMC_E_DEVICE, std::string(what) + ": " + hipGetErrorString(e));
}

// Runs f and turns every exception into a status code; nothing crosses the ABI.
template <class F>
fdmc_status guarded(F&& f) {
    try {
        f();
        g_last_error.clear();
        return FDMC_OK;
    } catch (const Fail& e) {
        g_last_error = e.msg;
        return e.code;
    } catch (const std::bad_alloc&) {
        g_last_error = "host allocation failed";
        return FDMC_E_ALLOC;
    } catch (const std::exception& e) {
        g_last_error = e.what();
        return FDMC_E_INTERNAL;
    }
}

inline double now_s() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

// Grow-only device buffer.
struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
    void* ensure(size_t n) {
        if (n <= bytes) return p;
        if (p) (void)hipFree(p);
        p = nullptr;
        bytes = 0;
        size_t want = std::max(n, (size_t)4096);
        if (hipMalloc(&p, want) != hipSuccess) {
            p = nullptr;
            fail(FDMC_E_ALLOC, "hipMalloc of " + std::to_string(want) + " bytes failed");
        }
        bytes = want;
        return p;
    }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        bytes = 0;
    }
};

}  // namespace fdmc_host

#define HIPC(call) fdmc_host::hip_check((call), #call)

struct fdmc_context {
    int device = 0;
    int n_cu = 256;
    hipStream_t stream = nullptr;
    uint32_t* d_dirv = nullptr;   // [FDMC_JK_NDIM][32]
    fdmc_host::DevBuf partials, sums, salts, sdir, gen_sdir, pca_z, pca_dw, basket_L;
    std::vector<uint32_t> sdir_key;  // salts + D of the cached sdir table
    int sdir_D = -1;
    double* h_sums = nullptr;     // pinned
    size_t h_sums_cap = 0;
    unsigned char* noalloc_scratch = nullptr;   // this device's address of the static scratch array
    bool timing = false;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    int64_t timed_launches = 0;
    double timed_ms = 0.0;
};

// Model/payoff flattening shared by the estimators and the book (engine.cpp).
fdmc::LegParams make_leg(const fdmc_model& m, const fdmc_payoff& p);
int model_kind_dev(const fdmc_model& m);
