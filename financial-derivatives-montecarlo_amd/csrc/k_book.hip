// k_book.hip -- a book of contracts priced (with pathwise-adjoint Greeks) by
// multilevel QMC, every contract, level and replica of a pass in one persistent
// launch, and the allocation loop itself on the device so that a sequence of
// passes can be captured in one hipGraph.
//
// Contract c (the book's one-factor model; payoff European, Asian, lookback or
// barrier; maturity T_c; n_c = 2^{L_c} dates) uses refinement M = 2 and levels
// l = 0 .. L_c.  Level l couples a fine path of 2^l Euler steps with a coarse
// path of 2^{l-1} steps on summed pairs of fine increments (kernel_mlmc,
// methods_cuda.cu:445-518); the noise is the Brownian bridge of the level's
// scrambled-Sobol points, salt_base + (c BOOK_NL + l) R + r.  Allocation and
// merges are run_mlqmc_cuda's (methods_cuda.cu:1680-2019), per contract:
//   S2_l   = var(replica means) * N_l,
//   N_l*   = ceil(2 / (R eps^2) sqrt(S2_l / 2^l) sum_k sqrt(S2_k 2^k)),
//   E_l   <- (E_l N + m n) / (N + n),   S2_l <- (S2_l N + v n^2) / (N + n),
// and the contract stops once sum_l S2_l / (N_l R) <= eps^2 / 2.
//
// One pass:
//   k_book_plan    per contract: points to add per level (pilot_n on pass 0)
//   k_book_scan    one workgroup: tiles and work items of every entry, prefix sum
//   k_book         persistent grid draining the work items through an atomic
//   k_book_reduce  one wave per entry: fixed-order sum of its items' partials
//   k_book_merge   per (contract, level): replica means -> E_l, S2_l
// k_book_status counts the contracts that need another pass; k_book_final
// writes the per-contract results.  Sums never depend on scheduling: an entry's
// split into work items depends only on its own point count.
//
// Greeks (Lipschitz payoffs only): reverse sweep over 16-step blocks of the fine
// path with the coarse path swept alongside; S_f and S_c are checkpointed at
// block starts in a per-workgroup global scratch, increments are regenerated
// from the Sobol index and the bridge skeleton.
#include "k_book.inl"

namespace fdmc {

// ---- the allocation loop ---------------------------------------------------------
// 0: the contract needs a pass (extra[] filled); 1: converged; 2: out of passes.
__device__ int book_decide(const BookLaunch& a, const BookContract& C, const BookState* s, const BookCtl& k,
                           int64_t (&extra)[BOOK_NL]) {
#pragma unroll
    for (int l = 0; l < BOOK_NL; l++) extra[l] = 0;
    if (k.iter >= a.max_iter) return 2;
    if (k.iter == 0) {
        for (int l = 0; l <= C.L; l++) extra[l] = a.pilot_n;
        return 0;
    }
    double var = 0.0;
    for (int l = 0; l <= C.L; l++) var += s[l].S2[0] / ((double)s[l].N * a.R);
    if (var <= 0.5 * a.eps * a.eps) return 1;
    double st = 0.0;
    for (int l = 0; l <= C.L; l++) st += sqrt(s[l].S2[0] * (double)(1 << l));
    bool any = false;
    for (int l = 0; l <= C.L; l++) {
        int64_t nopt = (int64_t)ceil(2.0 / ((double)a.R * a.eps * a.eps) * sqrt(s[l].S2[0] / (double)(1 << l)) * st);
        if (nopt < 100) nopt = 100;
        int64_t ex = nopt > s[l].N ? nopt - s[l].N : 0;
        // a top-up adds at least a tenth of the level's points: with R replicas the
        // variance estimate moves enough that exact Giles top-ups can crawl for
        // dozens of passes just above the target
        if (ex > 0) ex = max(ex, (s[l].N + 9) / 10);
        if (s[l].N + ex > 4294967296LL) ex = 4294967296LL - s[l].N;   // Sobol index stays in 32 bits
        extra[l] = ex;
        any = any || ex > 0;
    }
    return any ? 0 : 1;
}

__global__ void k_book_reset(BookLaunch a) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.n_contracts) return;
    const int c = a.c_begin + i;
    a.ctl[c] = BookCtl{0, 0, 0, 0};
    for (int l = 0; l < BOOK_NL; l++) a.state[(size_t)c * BOOK_NL + l] = BookState{};
}

__global__ void k_book_plan(BookLaunch a) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.n_contracts) return;
    const int c = a.c_begin + i;
    const BookContract& C = a.contracts[c];
    BookState* s = a.state + (size_t)c * BOOK_NL;
    BookCtl k = a.ctl[c];
    int64_t extra[BOOK_NL];
#pragma unroll
    for (int l = 0; l < BOOK_NL; l++) extra[l] = 0;
    if (!k.done) {
        const int d = book_decide(a, C, s, k, extra);
        if (d == 0) {
            k.iter++;
        } else {
            k.done = 1;
            k.conv = d == 1;
        }
        a.ctl[c] = k;
    }
    for (int l = 0; l <= C.L; l++) {
        s[l].extra = extra[l];
        for (int r = 0; r < a.R; r++) {
            BookEntry& E = a.entries[a.entry_pos[C.entry0 + l * a.R + r]];
            E.first = s[l].N;
            E.count = extra[l];
        }
    }
}

// One workgroup: tiles -> work items of every entry, exclusive prefix sum.
__global__ __launch_bounds__(1024) void k_book_scan(BookLaunch a) {
    __shared__ int part[1024];
    const int tid = threadIdx.x, n = a.n_entries;
    const int per = (n + 1023) / 1024, b = min(n, tid * per), e = min(n, b + per);
    int sum = 0;
    for (int i = b; i < e; i++) {
        BookEntry& E = a.entries[i];
        const int64_t nt = E.count > 0 ? ((E.first + E.count - 1) >> 8) - (E.first >> 8) + 1 : 0;
        int tpi = max(1, a.items_target >> E.l);
        tpi = (int)max<int64_t>(tpi, (nt + BOOK_MAX_ITEMS - 1) / BOOK_MAX_ITEMS);
        E.tpi = tpi;
        E.ntiles = (int)((nt + tpi - 1) / tpi);
        sum += E.ntiles;
    }
    part[tid] = sum;
    __syncthreads();
    for (int o = 1; o < 1024; o <<= 1) {
        const int v = tid >= o ? part[tid - o] : 0;
        __syncthreads();
        part[tid] += v;
        __syncthreads();
    }
    int run = tid > 0 ? part[tid - 1] : 0;
    for (int i = b; i < e; i++) {
        a.entries[i].tile_begin = run;
        run += a.entries[i].ntiles;
    }
    if (tid == 1023) *a.total_items = part[1023];
    if (tid == 0) *a.queue = 0u;
}

__global__ __launch_bounds__(256) void k_book_reduce(BookLaunch a) {
    const int e = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
    if (e >= a.n_entries) return;      // uniform per wave
    const BookEntry E = a.entries[e];
    double v[BOOK_NACC];
#pragma unroll
    for (int i = 0; i < BOOK_NACC; i++) v[i] = 0.0;
    for (int w = lane; w < E.ntiles; w += 64) {
        const double* p = a.partials + ((int64_t)E.tile_begin + w) * BOOK_NACC;
#pragma unroll
        for (int i = 0; i < BOOK_NACC; i++) v[i] += p[i];
    }
#pragma unroll
    for (int i = 0; i < BOOK_NACC; i++) v[i] = wave_sum(v[i]);
    if (lane == 0) {
#pragma unroll
        for (int i = 0; i < BOOK_NACC; i++) a.entry_sums[(int64_t)e * BOOK_NACC + i] = v[i];
    }
}

// run_mlqmc_cuda's merge (methods_cuda.cu:1975-1985) for the price difference
// and, alike, for each Greek.
__global__ void k_book_merge(BookLaunch a) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.n_contracts * BOOK_NL) return;
    const int c = a.c_begin + i / BOOK_NL, l = i % BOOK_NL;
    const BookContract& C = a.contracts[c];
    if (l > C.L) return;
    BookState& s = a.state[(size_t)c * BOOK_NL + l];
    const int64_t ex = s.extra;
    if (ex <= 0) return;
    const int R = a.R;
    const double wn = (double)ex, wo = (double)s.N;
    for (int q = 0; q < BOOK_NACC; q++) {
        double m = 0.0;
        for (int r = 0; r < R; r++) m += a.entry_sums[(int64_t)a.entry_pos[C.entry0 + l * R + r] * BOOK_NACC + q] / wn;
        m /= R;
        double v = 0.0;
        for (int r = 0; r < R; r++) {
            const double d = a.entry_sums[(int64_t)a.entry_pos[C.entry0 + l * R + r] * BOOK_NACC + q] / wn - m;
            v += d * d;
        }
        v = R > 1 ? v / (R - 1) : 0.0;
        if (s.N == 0) {
            s.E[q] = m;
            s.S2[q] = v * wn;
        } else {
            s.E[q] = (s.E[q] * wo + m * wn) / (wo + wn);
            s.S2[q] = (s.S2[q] * wo + v * wn * wn) / (wo + wn);
        }
    }
    s.N += ex;
}

// One workgroup: number of contracts of the launch that need another pass.
__global__ __launch_bounds__(256) void k_book_status(BookLaunch a) {
    __shared__ int cnt[256];
    int n = 0;
    for (int i = threadIdx.x; i < a.n_contracts; i += 256) {
        const int c = a.c_begin + i;
        const BookCtl k = a.ctl[c];
        if (k.done) continue;
        int64_t extra[BOOK_NL];
        if (book_decide(a, a.contracts[c], a.state + (size_t)c * BOOK_NL, k, extra) == 0) n++;
    }
    cnt[threadIdx.x] = n;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) cnt[threadIdx.x] += cnt[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) *a.active = cnt[0];
}

__global__ void k_book_final(BookLaunch a) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.n_contracts) return;
    const int c = a.c_begin + i;
    const BookContract& C = a.contracts[c];
    const BookState* s = a.state + (size_t)c * BOOK_NL;
    double val[BOOK_NACC] = {0.0, 0.0, 0.0, 0.0}, var[BOOK_NACC] = {0.0, 0.0, 0.0, 0.0};
    int64_t n = 0, ps = 0;
    for (int l = 0; l <= C.L; l++) {
        for (int q = 0; q < BOOK_NACC; q++) {
            val[q] += s[l].E[q];
            var[q] += s[l].S2[q] / ((double)s[l].N * a.R);
        }
        n += s[l].N * a.R;
        ps += s[l].N * a.R * ((int64_t)1 << l);
    }
    const double nan = __longlong_as_double(0x7ff8000000000000LL);
    const bool lip = C.kind != PK_BARRIER && a.greeks;
    BookOut o;
    o.v[0] = val[0];
    o.v[1] = sqrt(var[0]);
    for (int q = 1; q < BOOK_NACC; q++) {
        o.v[1 + q] = lip ? val[q] : nan;
        o.v[4 + q] = lip ? sqrt(var[q]) : nan;
    }
    o.n_samples = n;
    o.path_steps = ps;
    const BookCtl k = a.ctl[c];
    o.iterations = k.iter;
    // a contract whose last pass met the stopping rule is marked done only by
    // the next plan, which the host skips once no contract needs work
    bool conv = k.conv != 0;
    if (!k.done) {
        int64_t extra[BOOK_NL];
        conv = book_decide(a, C, s, k, extra) == 1;
    }
    // an estimate that is not finite never counts as converged
    for (int q = 0; q < BOOK_NACC; q++)
        if (!(fabs(val[q]) < INFINITY && var[q] < INFINITY)) conv = false;
    o.flags = conv ? 0 : 2;
    a.out[c] = o;
}

// ---- host launchers ---------------------------------------------------------------

template <int MK, int GM>
struct BookDefault {
    static const void* fn() { return (const void*)&k_book<MK, GM>; }
};

const void* book_kernel(int mk, int gm, int occ) {
    return occ ? book_kernel_o3(mk, gm) : book_kernel_any<BookDefault>(mk, gm);
}

hipError_t launch_book_reset(const BookLaunch& a, hipStream_t s) {
    hipLaunchKernelGGL(k_book_reset, dim3((a.n_contracts + 63) / 64), dim3(64), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_book_plan(const BookLaunch& a, hipStream_t s) {
    hipLaunchKernelGGL(k_book_plan, dim3((a.n_contracts + 63) / 64), dim3(64), 0, s, a);
    return hipGetLastError();
}

// scan -> persistent pass -> reduce (no planning or merging: level-sum hook and
// tuner).  With events e0/e1 the persistent kernel alone is bracketed by them.
hipError_t launch_book_sweep(const void* fn, const BookLaunch& a, int grid, hipStream_t s, hipEvent_t e0,
                             hipEvent_t e1) {
    hipLaunchKernelGGL(k_book_scan, dim3(1), dim3(1024), 0, s, a);
    hipError_t e = hipSuccess;
    if (e0 && (e = hipEventRecord(e0, s)) != hipSuccess) return e;
    void* args[] = {(void*)&a};
    e = hipLaunchKernel(fn, dim3(grid), dim3(256), args, 0, s);   // LDS is static
    if (e != hipSuccess) return e;
    if (e1 && (e = hipEventRecord(e1, s)) != hipSuccess) return e;
    hipLaunchKernelGGL(k_book_reduce, dim3((a.n_entries + 3) / 4), dim3(256), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_book_pass(const void* fn, const BookLaunch& a, int grid, hipStream_t s, hipEvent_t e0,
                            hipEvent_t e1) {
    hipLaunchKernelGGL(k_book_plan, dim3((a.n_contracts + 63) / 64), dim3(64), 0, s, a);
    hipError_t e = launch_book_sweep(fn, a, grid, s, e0, e1);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_book_merge, dim3((a.n_contracts * BOOK_NL + 63) / 64), dim3(64), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_book_status(const BookLaunch& a, hipStream_t s) {
    hipLaunchKernelGGL(k_book_status, dim3(1), dim3(256), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_book_final(const BookLaunch& a, hipStream_t s) {
    hipLaunchKernelGGL(k_book_final, dim3((a.n_contracts + 63) / 64), dim3(64), 0, s, a);
    return hipGetLastError();
}

}  // namespace fdmc
