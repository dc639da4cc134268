// book.cpp -- the contract book (fdmc_book_* in include/fdmc.h): device
// tables, the batched loop (device-side planning, passes captured in one
// hipGraph), the plain per-contract loop and the launch-shape autotuner.
//
// The reference prices one contract per run_mlqmc_cuda call
// (methods_cuda.cu:1680-2019): a host loop that launches every level of every
// replica separately and synchronises after each.  The plain mode keeps that
// shape (one contract at a time, one host round trip per pass) on this
// engine's kernels; the batched mode runs every contract of the book in the
// same passes with no host round trip between them.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "engine_host.hpp"

using namespace fdmc;
using namespace fdmc_host;

struct fdmc_book {
    fdmc_context* ctx = nullptr;
    int n = 0, R = 16, pilot_n = 400, max_iter = 50, greeks = 1, ipg = 4;
    uint32_t salt_base = 42;
    int mk = MK_LVQ;
    LegParams leg{};
    std::vector<BookContract> hc;
    std::vector<BookEntry> h_batched, h_plain;
    int n_entries = 0;
    DevBuf contracts, entries_b, entries_p, pos_b, pos_p, state, ctl, bbtab, voltab, partials, entry_sums, counters,
        out, scratch;
    DevBuf sdir[BOOK_NL];
    int* h_active = nullptr;          // pinned
    // launch shape (autotuned on the first run)
    bool tuned = false;
    int occ = 0, wg_per_cu = 2, items_target = 512, max_grid = 0;
    double tune_s = 0.0;
    // captured passes of the batched mode
    hipGraph_t graph = nullptr;
    hipGraphExec_t graph_exec = nullptr;
    double graph_eps = -1.0;
    int graph_occ = -1, graph_grid = -1;

    void drop_graph() {
        if (graph_exec) (void)hipGraphExecDestroy(graph_exec);
        if (graph) (void)hipGraphDestroy(graph);
        graph_exec = nullptr;
        graph = nullptr;
        graph_eps = -1.0;
    }
    ~fdmc_book() {
        drop_graph();
        if (h_active) (void)hipHostFree(h_active);
        DevBuf* all[] = {&contracts, &entries_b, &entries_p, &pos_b, &pos_p, &state, &ctl, &bbtab, &voltab,
                         &partials, &entry_sums, &counters, &out, &scratch};
        for (DevBuf* b : all) b->release();
        for (auto& s : sdir) s.release();
    }
};

namespace {

int book_payoff_kind(int32_t kind) {
    switch (kind) {
    case FDMC_PAYOFF_EUROPEAN: return PK_EUROPEAN;
    case FDMC_PAYOFF_ASIAN: return PK_ASIAN;
    case FDMC_PAYOFF_LOOKBACK: return PK_LOOKBACK;
    case FDMC_PAYOFF_BARRIER: return PK_BARRIER;
    default: return -1;
    }
}

// Conditional sds of bb_precompute's dyadic schedule (utils.cpp:54-101) for
// N = 2^l uniform steps over [0, T]: out[0] = sqrt(T) (terminal value), out[k]
// = sd of the midpoints of schedule level k - 1.  On a uniform grid every
// midpoint of one level has the same sd and weights 1/2 in FP32, which is
// what the book's bridge uses.
void bridge_sd(int N, double T, float* out) {
    std::vector<double> t((size_t)N + 1);
    for (int i = 0; i <= N; i++) t[i] = i * T / N;
    out[0] = (float)std::sqrt(T);
    for (int lev = 0; (1 << lev) < N; lev++) {
        const int stride = N >> lev, M = stride / 2;
        out[lev + 1] = (float)std::sqrt((t[M] - t[0]) * (t[stride] - t[M]) / (t[stride] - t[0]));
    }
}

BookLaunch make_launch(fdmc_book* b, bool plain, int c, double eps) {
    BookLaunch a{};
    a.leg = b->leg;
    a.R = b->R;
    a.pilot_n = b->pilot_n;
    a.max_iter = b->max_iter;
    a.greeks = b->greeks;
    a.items_target = b->items_target;
    a.eps = eps;
    a.contracts = (const BookContract*)b->contracts.p;
    a.state = (BookState*)b->state.p;
    a.ctl = (BookCtl*)b->ctl.p;
    for (int l = 0; l < BOOK_NL; l++) a.sdir[l] = (const uint32_t*)b->sdir[l].p;
    a.bbtab = (const float*)b->bbtab.p;
    a.voltab = (const float*)b->voltab.p;
    unsigned int* cnt = (unsigned int*)b->counters.p;
    a.queue = cnt;
    a.total_items = (int*)(cnt + 1);
    a.active = (int*)(cnt + 2);
    a.partials = (double*)b->partials.p;
    a.entry_sums = (double*)b->entry_sums.p;
    a.scratch = (float*)b->scratch.p;
    a.out = (BookOut*)b->out.p;
    if (!plain) {
        a.c_begin = 0;
        a.n_contracts = b->n;
        a.entries = (BookEntry*)b->entries_b.p;
        a.entry_pos = (const int*)b->pos_b.p;
        a.n_entries = b->n_entries;
    } else {
        const BookContract& C = b->hc[c];
        a.c_begin = c;
        a.n_contracts = 1;
        a.entries = (BookEntry*)b->entries_p.p + C.entry0;
        a.entry_pos = (const int*)b->pos_p.p;
        a.n_entries = (C.L + 1) * b->R;
    }
    return a;
}

int occupancy(const void* fn) {
    int nb = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, fn, 256, 0) != hipSuccess) nb = 1;   // static LDS
    return std::max(nb, 1);
}


// Launch-shape autotuner: times one sweep of every entry of the book at a fixed
// size (512 points per (contract, level, replica)) for each candidate
// (default or three-waves-per-SIMD compilation) x (persistent workgroups per
// CU) and keeps the fastest.  Only
// the launch shape is tuned: the split of entries into work items
// (items_target) is fixed, so results do not depend on the choice.
void autotune(fdmc_book* b) {
    const double t0 = now_s();
    fdmc_context* c = b->ctx;
    std::vector<BookEntry> e = b->h_batched;
    for (auto& x : e) {
        x.first = 0;
        x.count = 512;
    }
    HIPC(hipMemcpyAsync(b->entries_b.p, e.data(), e.size() * sizeof(BookEntry), hipMemcpyHostToDevice, c->stream));
    hipEvent_t ev0, ev1;
    HIPC(hipEventCreate(&ev0));
    HIPC(hipEventCreate(&ev1));
    float best = 1e30f;
    int best_occ = 0, best_wpc = 1;
    try {
        for (int occ = 0; occ < 2; occ++) {
            const void* fn = book_kernel(b->mk, b->greeks, occ);
            const int nb = std::min(occupancy(fn), 4);
            for (int wpc = 1; wpc <= nb; wpc++) {
                const int grid = std::min(c->n_cu * wpc, b->max_grid);
                BookLaunch a = make_launch(b, false, 0, 1.0);
                HIPC(launch_book_sweep(fn, a, grid, c->stream));   // warm-up
                HIPC(hipEventRecord(ev0, c->stream));
                HIPC(launch_book_sweep(fn, a, grid, c->stream));
                HIPC(hipEventRecord(ev1, c->stream));
                HIPC(hipEventSynchronize(ev1));
                float ms = 0.0f;
                HIPC(hipEventElapsedTime(&ms, ev0, ev1));
                if (ms < best) {
                    best = ms;
                    best_occ = occ;
                    best_wpc = wpc;
                }
            }
        }
    } catch (...) {
        (void)hipEventDestroy(ev0);
        (void)hipEventDestroy(ev1);
        throw;
    }
    (void)hipEventDestroy(ev0);
    (void)hipEventDestroy(ev1);
    b->occ = best_occ;
    b->wg_per_cu = best_wpc;
    b->tuned = true;
    b->tune_s = now_s() - t0;
}

int book_grid(fdmc_book* b) { return std::min(b->ctx->n_cu * b->wg_per_cu, b->max_grid); }

void read_results(fdmc_book* b, fdmc_book_result* res, int64_t* total_steps) {
    std::vector<BookOut> o((size_t)b->n);
    HIPC(hipMemcpyAsync(o.data(), b->out.p, o.size() * sizeof(BookOut), hipMemcpyDeviceToHost, b->ctx->stream));
    HIPC(hipStreamSynchronize(b->ctx->stream));
    int64_t tot = 0;
    for (int i = 0; i < b->n; i++) {
        fdmc_book_result& r = res[i];
        r.price = o[i].v[0];
        r.std_error = o[i].v[1];
        r.delta = o[i].v[2];
        r.vega = o[i].v[3];
        r.rho = o[i].v[4];
        r.delta_se = o[i].v[5];
        r.vega_se = o[i].v[6];
        r.rho_se = o[i].v[7];
        r.n_samples = o[i].n_samples;
        r.path_steps = o[i].path_steps;
        r.levels = b->hc[i].L + 1;
        r.iterations = o[i].iterations;
        r.flags = o[i].flags ? FDMC_FLAG_NOT_CONVERGED : 0u;
        r.pad = 0;
        tot += o[i].path_steps;
    }
    *total_steps = tot;
}

// With context timing on: the same passes launched one by one (no graph), the
// persistent kernel of each pass bracketed by HIP events.
void run_batched_timed(fdmc_book* b, double eps, fdmc_book_stats* st) {
    fdmc_context* c = b->ctx;
    const int grid = book_grid(b);
    const void* fn = book_kernel(b->mk, b->greeks, b->occ);
    BookLaunch a = make_launch(b, false, 0, eps);
    const double t0 = now_s();
    HIPC(launch_book_reset(a, c->stream));
    int passes = 0;
    for (int it = 0; it <= b->max_iter; it++) {
        HIPC(launch_book_pass(fn, a, grid, c->stream, c->ev0, c->ev1));
        HIPC(launch_book_status(a, c->stream));
        HIPC(hipMemcpyAsync(b->h_active, a.active, sizeof(int), hipMemcpyDeviceToHost, c->stream));
        HIPC(hipStreamSynchronize(c->stream));
        float ms = 0.0f;
        HIPC(hipEventElapsedTime(&ms, c->ev0, c->ev1));
        c->timed_ms += ms;
        c->timed_launches++;
        passes++;
        if (*b->h_active == 0) break;
    }
    HIPC(launch_book_final(a, c->stream));
    st->time_s = now_s() - t0;
    st->graph_replays = 0;
    st->passes = passes;
}

void run_batched(fdmc_book* b, double eps, fdmc_book_stats* st) {
    fdmc_context* c = b->ctx;
    if (c->timing) return run_batched_timed(b, eps, st);
    const int grid = book_grid(b);
    const void* fn = book_kernel(b->mk, b->greeks, b->occ);
    BookLaunch a = make_launch(b, false, 0, eps);
    if (!b->graph_exec || b->graph_eps != eps || b->graph_occ != b->occ ||
        b->graph_grid != grid) {
        b->drop_graph();
        HIPC(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
        hipError_t e = hipSuccess;
        for (int i = 0; i < b->ipg && e == hipSuccess; i++) e = launch_book_pass(fn, a, grid, c->stream);
        if (e == hipSuccess) e = launch_book_status(a, c->stream);
        hipGraph_t g = nullptr;
        const hipError_t e2 = hipStreamEndCapture(c->stream, &g);
        if (e != hipSuccess || e2 != hipSuccess) {
            if (g) (void)hipGraphDestroy(g);
            HIPC(e);
            HIPC(e2);
        }
        b->graph = g;
        HIPC(hipGraphInstantiate(&b->graph_exec, g, nullptr, nullptr, 0));
        b->graph_eps = eps;
        b->graph_occ = b->occ;
        b->graph_grid = grid;
    }
    const double t0 = now_s();
    HIPC(launch_book_reset(a, c->stream));
    int replays = 0;
    const int max_replays = (b->max_iter + b->ipg) / b->ipg + 1;
    for (;;) {
        HIPC(hipGraphLaunch(b->graph_exec, c->stream));
        HIPC(hipMemcpyAsync(b->h_active, a.active, sizeof(int), hipMemcpyDeviceToHost, c->stream));
        HIPC(hipStreamSynchronize(c->stream));
        replays++;
        if (*b->h_active == 0 || replays >= max_replays) break;
    }
    HIPC(launch_book_final(a, c->stream));
    st->time_s = now_s() - t0;     // read_results' copy is added by the caller
    st->graph_replays = replays;
    st->passes = replays * b->ipg;
}

void run_plain(fdmc_book* b, double eps, fdmc_book_stats* st) {
    fdmc_context* c = b->ctx;
    const int grid = book_grid(b);
    const void* fn = book_kernel(b->mk, b->greeks, b->occ);
    const double t0 = now_s();
    BookLaunch all = make_launch(b, false, 0, eps);
    HIPC(launch_book_reset(all, c->stream));
    int passes = 0;
    for (int k = 0; k < b->n; k++) {
        BookLaunch a = make_launch(b, true, k, eps);
        for (int it = 0; it <= b->max_iter; it++) {
            HIPC(launch_book_pass(fn, a, grid, c->stream, c->timing ? c->ev0 : nullptr,
                                  c->timing ? c->ev1 : nullptr));
            HIPC(launch_book_status(a, c->stream));
            HIPC(hipMemcpyAsync(b->h_active, a.active, sizeof(int), hipMemcpyDeviceToHost, c->stream));
            HIPC(hipStreamSynchronize(c->stream));
            if (c->timing) {
                float ms = 0.0f;
                HIPC(hipEventElapsedTime(&ms, c->ev0, c->ev1));
                c->timed_ms += ms;
                c->timed_launches++;
            }
            passes++;
            if (*b->h_active == 0) break;
        }
    }
    HIPC(launch_book_final(all, c->stream));
    st->time_s = now_s() - t0;
    st->graph_replays = 0;
    st->passes = passes;
}

}  // namespace

extern "C" {

fdmc_status fdmc_book_create(fdmc_context* ctx, const fdmc_model* model, const fdmc_book_contract* contracts,
                             int32_t n, const fdmc_book_config* cfg, fdmc_book** out) {
    return guarded([&] {
        if (!ctx || !model || !contracts || !cfg || !out) fail(FDMC_E_INVALID, "null argument");
        *out = nullptr;
        if (n <= 0) fail(FDMC_E_INVALID, "the book needs at least one contract");
        if (model->kind != FDMC_MODEL_GBM && model->kind != FDMC_MODEL_DUPIRE &&
            model->kind != FDMC_MODEL_LOCALVOL_QUAD)
            fail(FDMC_E_UNSUPPORTED, "book model must be GBM, Dupire or the quadratic local vol");
        if (!(model->S0 > 0.0)) fail(FDMC_E_INVALID, "S0 must be positive");
        if (cfg->R < 2 || cfg->R > 1024) fail(FDMC_E_INVALID, "R must be in [2, 1024]");
        if (cfg->pilot_n < 2) fail(FDMC_E_INVALID, "pilot_n must be >= 2");
        if (cfg->greeks < 0 || cfg->greeks > 2) fail(FDMC_E_INVALID, "greeks must be 0, 1 (adjoint) or 2 (tangent)");
        if (cfg->max_iter < 1) fail(FDMC_E_INVALID, "max_iter must be >= 1");
        if (cfg->iters_per_graph < 1 || cfg->iters_per_graph > 64)
            fail(FDMC_E_INVALID, "iters_per_graph must be in [1, 64]");
        if (cfg->variant < -1 || cfg->variant > 1) fail(FDMC_E_INVALID, "variant must be -1 (autotune), 0 or 1");
        if (cfg->variant >= 0 && (cfg->wg_per_cu < 1 || cfg->wg_per_cu > 4))
            fail(FDMC_E_INVALID, "wg_per_cu must be in [1, 4] with a fixed variant");
        std::vector<double> Ts;
        for (int i = 0; i < n; i++) {
            const fdmc_book_contract& k = contracts[i];
            const std::string at = " (contract " + std::to_string(i) + ")";
            if (book_payoff_kind(k.kind) < 0)
                fail(FDMC_E_UNSUPPORTED, "book payoffs are European, Asian, lookback or barrier" + at);
            if (k.n_steps < 1 || k.n_steps > BOOK_DMAX || (k.n_steps & (k.n_steps - 1)))
                fail(FDMC_E_INVALID, "n_steps must be a power of two in [1, 512]" + at);
            if (!(k.T > 0.0) || !std::isfinite(k.T)) fail(FDMC_E_INVALID, "T must be positive" + at);
            if (!std::isfinite(k.K) || !std::isfinite(k.B) || !std::isfinite(k.r) || !std::isfinite(k.sigma_bgk))
                fail(FDMC_E_INVALID, "non-finite contract parameter" + at);
            if (std::find(Ts.begin(), Ts.end(), k.T) == Ts.end()) Ts.push_back(k.T);
        }
        if ((int)Ts.size() > BOOK_MAX_T) fail(FDMC_E_UNSUPPORTED, "a book holds at most 8 distinct maturities");
        const int64_t n_entries64 = [&] {
            int64_t s = 0;
            for (int i = 0; i < n; i++) s += (int64_t)(31 - __builtin_clz((unsigned)contracts[i].n_steps) + 1) * cfg->R;
            return s;
        }();
        if (n_entries64 * BOOK_MAX_ITEMS > 0x7FFFFFFF) fail(FDMC_E_RANGE, "book too large for one pass");

        auto* b = new fdmc_book;
        try {
            b->ctx = ctx;
            b->n = n;
            b->R = cfg->R;
            b->pilot_n = cfg->pilot_n;
            b->max_iter = cfg->max_iter;
            b->greeks = cfg->greeks;
            b->ipg = cfg->iters_per_graph;
            b->salt_base = cfg->salt_base;
            b->mk = model_kind_dev(*model);
            if (cfg->variant >= 0) {
                b->occ = cfg->variant;
                b->wg_per_cu = cfg->wg_per_cu;
                b->tuned = true;
            }
            fdmc_payoff p{};
            p.kind = FDMC_PAYOFF_EUROPEAN;
            b->leg = make_leg(*model, p);
            b->n_entries = (int)n_entries64;
            HIPC(hipSetDevice(ctx->device));

            // contracts, entry ids and Sobol record slots
            b->hc.resize((size_t)n);
            int slots[BOOK_NL] = {0};
            int e0 = 0;
            for (int i = 0; i < n; i++) {
                const fdmc_book_contract& k = contracts[i];
                BookContract& C = b->hc[i];
                C = BookContract{};
                C.kind = book_payoff_kind(k.kind);
                C.L = 31 - __builtin_clz((unsigned)k.n_steps);
                C.t_idx = (int)(std::find(Ts.begin(), Ts.end(), k.T) - Ts.begin());
                C.entry0 = e0;
                e0 += (C.L + 1) * b->R;
                C.K = (float)k.K;
                C.B = (float)k.B;
                C.discount = (k.kind == FDMC_PAYOFF_EUROPEAN || k.kind == FDMC_PAYOFF_BARRIER)
                                 ? (float)std::exp(-k.r * k.T) : 1.0f;
                C.Tf = (float)k.T;
                C.T = k.T;
                for (int l = 0; l < BOOK_NL; l++) {
                    const float h = (float)(k.T / (double)(1 << l));
                    const bool bgk = k.kind == FDMC_PAYOFF_LOOKBACK || k.kind == FDMC_PAYOFF_BARRIER;
                    C.bgk[l] = bgk ? std::exp(-0.5826f * (float)k.sigma_bgk * std::sqrt(h)) : 1.0f;   // payoffs.hpp:8
                    C.slot[l] = l <= C.L ? slots[l]++ : -1;
                }
            }
            // entries: batched layout level-major (longest paths first), plain layout contract-major
            b->h_batched.reserve((size_t)b->n_entries);
            std::vector<int> pos_b((size_t)b->n_entries), pos_p((size_t)b->n_entries);
            for (int l = BOOK_NL - 1; l >= 0; l--)
                for (int i = 0; i < n; i++) {
                    const BookContract& C = b->hc[i];
                    if (l > C.L) continue;
                    for (int r = 0; r < b->R; r++) {
                        pos_b[(size_t)C.entry0 + (size_t)l * b->R + r] = (int)b->h_batched.size();
                        b->h_batched.push_back(BookEntry{i, l, r, 1, 0, 0, 0, 0});
                    }
                }
            b->h_plain.resize((size_t)b->n_entries);
            for (int i = 0; i < n; i++) {
                const BookContract& C = b->hc[i];
                for (int l = 0; l <= C.L; l++)
                    for (int r = 0; r < b->R; r++) {
                        const int local = (C.L - l) * b->R + r;
                        pos_p[(size_t)C.entry0 + (size_t)l * b->R + r] = local;
                        b->h_plain[(size_t)C.entry0 + local] = BookEntry{i, l, r, 1, 0, 0, 0, 0};
                    }
            }
            hipStream_t s = ctx->stream;
            auto upload = [&](DevBuf& d, const void* src, size_t bytes) {
                d.ensure(bytes);
                HIPC(hipMemcpyAsync(d.p, src, bytes, hipMemcpyHostToDevice, s));
            };
            upload(b->contracts, b->hc.data(), b->hc.size() * sizeof(BookContract));
            upload(b->entries_b, b->h_batched.data(), b->h_batched.size() * sizeof(BookEntry));
            upload(b->entries_p, b->h_plain.data(), b->h_plain.size() * sizeof(BookEntry));
            upload(b->pos_b, pos_b.data(), pos_b.size() * sizeof(int));
            upload(b->pos_p, pos_p.data(), pos_p.size() * sizeof(int));

            // scrambled direction records of every (contract, level, replica)
            DevBuf salts_d;
            for (int l = 0; l < BOOK_NL; l++) {
                if (slots[l] == 0) continue;
                std::vector<uint32_t> salts;
                salts.reserve((size_t)slots[l] * b->R);
                for (int i = 0; i < n; i++) {
                    if (l > b->hc[i].L) continue;
                    for (int r = 0; r < b->R; r++)
                        salts.push_back(b->salt_base + (uint32_t)((i * BOOK_NL + l) * b->R + r));
                }
                const int D = 1 << l;
                b->sdir[l].ensure(salts.size() * (size_t)D * SDIR_WORDS * sizeof(uint32_t));
                salts_d.ensure(salts.size() * sizeof(uint32_t));
                HIPC(hipMemcpyAsync(salts_d.p, salts.data(), salts.size() * sizeof(uint32_t), hipMemcpyHostToDevice,
                                    s));
                HIPC(launch_build_sdir(ctx->d_dirv, (const uint32_t*)salts_d.p, (int)salts.size(), D,
                                       (uint32_t*)b->sdir[l].p, s));
                HIPC(hipStreamSynchronize(s));   // salts_d is reused by the next level
            }
            salts_d.release();

            // bridge sds: every maturity, levels 0..9
            std::vector<float> bb((size_t)BOOK_MAX_T * BOOK_NL * BOOK_NL, 0.0f);
            for (size_t ti = 0; ti < Ts.size(); ti++)
                for (int l = 0; l < BOOK_NL; l++) bridge_sd(1 << l, Ts[ti], bb.data() + (ti * BOOK_NL + l) * BOOK_NL);
            upload(b->bbtab, bb.data(), bb.size() * sizeof(float));

            // local-vol time factor of every fine step: b e^{-c t} ln(2)^2 (quadratic
            // local vol, evaluated in log2), e^{-alpha t} (Dupire), t = k h in FP32 as
            // the steppers form it
            // Planes 1 and 2 hold the same factors times 1/16 and 1/8 (exact: powers
            // of two), for paths whose fine and coarse increments are built 16 and
            // 8 times their size (the scaled bridge blocks of k_book.inl).
            std::vector<float> vt((size_t)3 * BOOK_VT_PLANE, 0.0f);
            for (size_t ti = 0; ti < Ts.size(); ti++)
                for (int l = 0; l < BOOK_NL; l++) {
                    const int nf = 1 << l;
                    const float h = (float)(Ts[ti] / (double)nf);
                    float* row = vt.data() + ti * BOOK_VT_STRIDE + (size_t)(nf - 1);
                    for (int k = 0; k < nf; k++) {
                        const float t = (float)k * h;
                        if (b->mk == MK_LVQ) row[k] = b->leg.lv_b * std::exp(-b->leg.lv_c * t) * 0.480453014f;  // ln(2)^2
                        else if (b->mk == MK_DUPIRE) row[k] = std::exp(-b->leg.alpha * t);
                        row[BOOK_VT_PLANE + k] = row[k] * 0.0625f;
                        row[2 * BOOK_VT_PLANE + k] = row[k] * 0.125f;
                    }
                }
            upload(b->voltab, vt.data(), vt.size() * sizeof(float));

            // state, partials, scratch
            b->state.ensure((size_t)n * BOOK_NL * sizeof(BookState));
            b->ctl.ensure((size_t)n * sizeof(BookCtl));
            b->out.ensure((size_t)n * sizeof(BookOut));
            b->partials.ensure((size_t)b->n_entries * BOOK_MAX_ITEMS * BOOK_NACC * sizeof(double));
            b->entry_sums.ensure((size_t)b->n_entries * BOOK_NACC * sizeof(double));
            b->counters.ensure(64);
            b->max_grid = ctx->n_cu * 4;
            b->scratch.ensure((size_t)b->max_grid * book_scratch_floats() * sizeof(float));
            HIPC(hipHostMalloc((void**)&b->h_active, sizeof(int) * 4, hipHostMallocDefault));
            BookLaunch a = make_launch(b, false, 0, 1.0);
            HIPC(launch_book_reset(a, s));
            HIPC(hipStreamSynchronize(s));
        } catch (...) {
            delete b;
            throw;
        }
        *out = b;
    });
}

fdmc_status fdmc_book_run(fdmc_book* b, double eps, int32_t mode, fdmc_book_result* results,
                          fdmc_book_stats* stats) {
    return guarded([&] {
        if (!b || !results) fail(FDMC_E_INVALID, "null argument");
        if (!(eps > 0.0) || !std::isfinite(eps)) fail(FDMC_E_INVALID, "eps must be positive");
        if (mode != FDMC_BOOK_BATCHED && mode != FDMC_BOOK_PLAIN) fail(FDMC_E_INVALID, "unknown book mode");
        HIPC(hipSetDevice(b->ctx->device));
        fdmc_book_stats st{};
        if (!b->tuned) autotune(b);
        if (mode == FDMC_BOOK_BATCHED) run_batched(b, eps, &st);
        else run_plain(b, eps, &st);
        const double t0 = now_s();
        read_results(b, results, &st.path_steps);
        st.time_s += now_s() - t0;
        st.tune_s = b->tune_s;
        st.variant = b->occ;
        st.items_target = b->items_target;
        st.wg_per_cu = b->wg_per_cu;
        if (stats) *stats = st;
    });
}

fdmc_status fdmc_book_level_sums(fdmc_book* b, int32_t c, int32_t l, int64_t offset, int64_t n,
                                 double* sums_out) {
    return guarded([&] {
        if (!b || !sums_out) fail(FDMC_E_INVALID, "null argument");
        if (c < 0 || c >= b->n) fail(FDMC_E_INVALID, "contract index out of range");
        const BookContract& C = b->hc[c];
        if (l < 0 || l > C.L) fail(FDMC_E_INVALID, "level out of range for this contract");
        if (n <= 0 || offset < 0 || offset + n > 4294967296LL)
            fail(FDMC_E_RANGE, "points must lie in [0, 2^32)");
        HIPC(hipSetDevice(b->ctx->device));
        fdmc_context* ctx = b->ctx;
        if (!b->tuned) autotune(b);
        const int ne = (C.L + 1) * b->R;
        std::vector<BookEntry> e(b->h_plain.begin() + C.entry0, b->h_plain.begin() + C.entry0 + ne);
        for (auto& x : e) {
            x.first = offset;
            x.count = x.l == l ? n : 0;
        }
        BookLaunch a = make_launch(b, true, c, 1.0);
        HIPC(hipMemcpyAsync(a.entries, e.data(), e.size() * sizeof(BookEntry), hipMemcpyHostToDevice, ctx->stream));
        HIPC(launch_book_sweep(book_kernel(b->mk, b->greeks, b->occ), a, book_grid(b), ctx->stream));
        std::vector<double> s((size_t)ne * BOOK_NACC);
        HIPC(hipMemcpyAsync(s.data(), a.entry_sums, s.size() * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
        HIPC(hipStreamSynchronize(ctx->stream));
        for (int r = 0; r < b->R; r++) {
            const int local = (C.L - l) * b->R + r;
            for (int k = 0; k < BOOK_NACC; k++) sums_out[r * BOOK_NACC + k] = s[(size_t)local * BOOK_NACC + k];
        }
    });
}

fdmc_status fdmc_book_free(fdmc_book* b) {
    if (!b) return FDMC_OK;
    (void)hipSetDevice(b->ctx->device);
    (void)hipStreamSynchronize(b->ctx->stream);
    delete b;
    return FDMC_OK;
}

}  // extern "C"
