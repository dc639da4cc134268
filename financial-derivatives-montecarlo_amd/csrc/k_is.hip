// k_is.hip -- importance-sampled GBM European call (reference kernel_is_gbm,
// methods_cuda.cu:685-707).  Each lane is one path: Philox normals z_k
// (stream (seed, tag), path index), increments (z_k + zs) sqrt(h), the GBM
// Euler step, and the payoff max(S_T - K, 0) e^{-rT} weighted by the
// likelihood ratio exp(-zs sum z_k - n zs^2 / 2) of the shifted measure.
// Sums {Y, Y^2} per workgroup; a workgroup takes tiles_per_wg 256-path tiles.
#include "k_path.inl"

namespace fdmc {

__global__ __launch_bounds__(256) void k_is_gbm(ISLaunch a) {
    __shared__ double red[4 * 2];
    const int tid = threadIdx.x;
    double acc[2] = {0.0, 0.0};
    const int64_t tile0 = (int64_t)blockIdx.x * a.tiles_per_wg;
    const int n = a.n_steps;
    for (int it = 0; it < a.tiles_per_wg; it++) {
        const int64_t base = (tile0 + it) << 8;
        if (base >= a.count) break;   // the same for every lane of the workgroup
        const int64_t idx = base + tid;
        float S = a.S0, zsum = 0.0f;
        for (int b = 0; 4 * b < n; b++) {
            float z[4];
            prng_normals4(a.seed, a.tag, (uint64_t)idx, (uint32_t)b, z);
#pragma unroll
            for (int c = 0; c < 4; c++)
                if (4 * b + c < n) {
                    zsum += z[c];
                    const float dw = (z[c] + a.zs) * a.sqrt_h;
                    S = S + a.mu * S * a.h + a.sigma * S * dw;
                }
        }
        const float lr = fast_exp(-a.zs * zsum - 0.5f * a.zs * a.zs * (float)n);
        const float y = fmaxf(S - a.K, 0.0f) * a.discount * lr;
        if (idx < a.count) {
            acc[0] += (double)y;
            acc[1] += (double)y * (double)y;
        }
    }
    block_sum<2, 4>(acc, red);
    if (tid == 0) {
        a.partials[(int64_t)blockIdx.x * 2] = acc[0];
        a.partials[(int64_t)blockIdx.x * 2 + 1] = acc[1];
    }
}

hipError_t launch_is_gbm(const ISLaunch& a, int nwg, hipStream_t st) {
    hipLaunchKernelGGL(k_is_gbm, dim3((unsigned)nwg), dim3(256), 0, st, a);
    return hipGetLastError();
}

}  // namespace fdmc
