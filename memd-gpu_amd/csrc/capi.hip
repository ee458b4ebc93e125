// C-ABI layer of the MEMD engine: context, workspace arena, direction cache,
// plan construction and the sift-step loop (direct launches or one captured
// hipGraph replayed per step).
//
// Replaces the host driver memd<coord_t,real_t>() (memReducedEpochs_dev.cu:1074-1368),
// GenerateDirectionVector (:1029-1072) and the MEX gateway mexFunction
// (memdMatlabCuda.cu:996-1347).  Unlike the reference, the workspace persists
// across calls, knot counts never leave the device, and no host synchronisation
// happens inside the sifting loop.
#include "../../include/memd_amd.h"
#include "directions.h"
#include "engine.h"

#include <cstdarg>
#include <cstdint>
#include <cstdlib>
#include <cstdio>
#include <cstring>
#include <map>
#include <string>
#include <vector>

using namespace memd;

struct memd_ctx {
    int device = 0;
    hipStream_t user = nullptr;     // caller's stream (nullptr = legacy default stream)
    hipStream_t own = nullptr;      // owned stream used for graph capture / replay
    hipEvent_t ev_in = nullptr, ev_out = nullptr, t0 = nullptr, t1 = nullptr;
    std::string err = "no error";
    char* arena = nullptr;
    size_t arena_bytes = 0;
    std::map<std::string, double*> dircache;
    // last captured step graph and the plan it was captured for
    hipGraphExec_t graph_exec = nullptr;
    std::string graph_key;
    int* flag = nullptr;            // device word for the coordinate check
    hipEvent_t pev[kNumStepKernels + 1] = {};   // per-kernel timing events (profiling entry point)
};

namespace {

constexpr size_t kAlign = 256;
inline size_t align_up(size_t v) { return (v + kAlign - 1) / kAlign * kAlign; }

memd_status fail(memd_ctx* ctx, memd_status code, const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    if (ctx) ctx->err = buf;
    return code;
}

#define HIP_TRY(ctx, call)                                                                   \
    do {                                                                                     \
        hipError_t e_ = (call);                                                              \
        if (e_ != hipSuccess)                                                                \
            return fail((ctx), MEMD_ERR_HIP, "%s failed: %s", #call, hipGetErrorString(e_)); \
    } while (0)

struct EpochIn {
    int n, T, K, N_imf, S;
    const double* y;
    const double* V;
    double* imfs;
    int32_t* h_sift_counts;
    int32_t* h_n_imf_out;
};

// Byte layout of one epoch's workspace inside the arena.
struct EpochLayout {
    size_t running, mask, pre, E, X, fac, redsys, yknot, ysol, coef, sep, mean, tstat, red, sift, st, total;
};

EpochLayout layout_for(int n, int T, int K, int N_imf, bool records) {
    EpochLayout L{};
    size_t W = (size_t)(T + 63) / 64;
    size_t Ecap = (size_t)knot_capacity(T);
    size_t nb = (size_t)solver_block_capacity((int)Ecap);
    size_t off = 0;
    auto take = [&](size_t bytes) { size_t at = off; off += align_up(bytes); return at; };
    L.running = take((size_t)n * T * 8);
    L.mask = take((size_t)K * 2 * W * 8);
    L.pre = take((size_t)K * 2 * W * 4);
    L.E = take((size_t)K * 2 * 4);
    L.X = take((size_t)K * 2 * Ecap * 4);
    L.fac = take((size_t)K * 2 * 6 * Ecap * 8);
    L.redsys = take((size_t)K * 2 * 3 * nb * 8);
    L.yknot = take((size_t)K * 2 * n * Ecap * 8);
    L.ysol = take((size_t)K * 2 * n * Ecap * 8);
    L.coef = take(records ? (size_t)K * 2 * n * Ecap * 32 : 0);   // only the record-reading evaluations
    L.sep = take((size_t)K * 2 * n * nb * 8);
    L.mean = take((size_t)n * T * 8);
    L.tstat = take((size_t)T * 16);
    L.red = take(64);
    L.sift = take((size_t)(N_imf > 0 ? N_imf : 1) * 4);
    L.st = take(sizeof(EpochState));
    L.total = off;
    return L;
}

memd_status ensure_arena(memd_ctx* ctx, size_t bytes) {
    if (bytes <= ctx->arena_bytes) return MEMD_OK;
    HIP_TRY(ctx, hipDeviceSynchronize());
    if (ctx->graph_exec) {
        (void)hipGraphExecDestroy(ctx->graph_exec);
        ctx->graph_exec = nullptr;
        ctx->graph_key.clear();
    }
    if (ctx->arena) {
        HIP_TRY(ctx, hipFree(ctx->arena));
        ctx->arena = nullptr;
        ctx->arena_bytes = 0;
    }
    size_t want = bytes + bytes / 8;   // headroom so similar shapes reuse the arena
    hipError_t e = hipMalloc((void**)&ctx->arena, want);
    if (e != hipSuccess) {
        ctx->arena = nullptr;
        (void)hipGetLastError();
        return fail(ctx, MEMD_ERR_NOMEM, "workspace allocation of %zu bytes failed: %s", want,
                    hipGetErrorString(e));
    }
    ctx->arena_bytes = want;
    return MEMD_OK;
}

memd_status validate_shape(memd_ctx* ctx, size_t n, size_t T, size_t K, size_t N_imf, size_t S) {
    if (n < 1 || T < 4 || K < 1 || N_imf < 1 || S < 1)
        return fail(ctx, MEMD_ERR_ARG, "invalid shape n=%zu T=%zu K=%zu N_imf=%zu S=%zu (need n>=1, T>=4, K>=1, N_imf>=1, S>=1)",
                    n, T, K, N_imf, S);
    if (n > (size_t)kMaxChannels)
        return fail(ctx, MEMD_ERR_UNSUPPORTED, "n=%zu channels exceeds the engine limit of %d", n, kMaxChannels);
    if (T > (size_t)(1u << 30) || K > 65536 || N_imf > 4096 || S > (1u << 24))
        return fail(ctx, MEMD_ERR_UNSUPPORTED, "shape too large (T=%zu K=%zu N_imf=%zu S=%zu)", T, K, N_imf, S);
    return MEMD_OK;
}

hipStream_t work_stream(memd_ctx* ctx) { return ctx->user; }

// Evaluation-kernel tile variant: MEMD_EVAL_VARIANT (tuning runs) or the default.
int eval_variant_choice() {
    const char* v = std::getenv("MEMD_EVAL_VARIANT");
    if (v && *v) {
        int i = std::atoi(v);
        if (i >= 0 && i < kNumEvalVariants) return i;
    }
    return kEvalDefault;
}

// Core: run a set of epochs (all already FP64 and on the device) to completion.
memd_status run_epochs(memd_ctx* ctx, const std::vector<EpochIn>& in, const memd_opts* opts, double* elapsed,
                       memd_kernel_profile* prof = nullptr) {
    memd_opts o;
    if (opts) o = *opts; else memd_opts_default(&o);
    if (o.mode != MEMD_MODE_FIX && o.mode != MEMD_MODE_STOP)
        return fail(ctx, MEMD_ERR_ARG, "unknown mode %d", o.mode);
    if (o.window != MEMD_WINDOW_REFWARP32 && o.window != MEMD_WINDOW_FULL)
        return fail(ctx, MEMD_ERR_ARG, "unknown extrema window %d", o.window);
    const int count = (int)in.size();
    if (count == 0) return MEMD_OK;
    if (o.d_knot_counts && count != 1)
        return fail(ctx, MEMD_ERR_ARG, "knot-count tracing is only available for single-epoch calls");
    int nmax = 0;
    for (const EpochIn& I : in) nmax = I.n > nmax ? I.n : nmax;
    if (o.mode == MEMD_MODE_STOP && nmax > kStopMaxChannels)
        return fail(ctx, MEMD_ERR_UNSUPPORTED, "stop mode supports at most %d channels (got %d)", kStopMaxChannels, nmax);

    // ---- workspace plan
    const int variant = eval_variant_choice();
    const bool records = (o.mode == MEMD_MODE_STOP || variant != kEvalFused);
    std::vector<EpochLayout> lay(count);
    std::vector<size_t> base(count);
    size_t total = 0;
    for (int e = 0; e < count; ++e) {
        lay[e] = layout_for(in[e].n, in[e].T, in[e].K, in[e].N_imf, records);
        base[e] = total;
        total += lay[e].total;
    }
    size_t desc_at = total;
    total += align_up(sizeof(EpochDesc) * count);
    // block tables of the fixed-size launches and the solver work lists
    std::vector<int> nblk(kNumMaps, 0);
    std::vector<long> lcap(kNumLists, 0);
    {
        std::vector<EpochDesc> tmp(count);
        for (int e = 0; e < count; ++e) {
            tmp[e].n = in[e].n; tmp[e].T = in[e].T; tmp[e].K = in[e].K;
            tmp[e].Ecap = knot_capacity(in[e].T);
            tmp[e].nbmax = solver_block_capacity(tmp[e].Ecap);
            tmp[e].mode = o.mode;
            tmp[e].coef_records = records ? 1 : 0;
            for (int f = 0; f < kNumMaps; ++f) {
                long v = (long)nblk[f] + blocks_for(f, tmp[e], variant);
                if (v > 0x7fffffffL) return fail(ctx, MEMD_ERR_UNSUPPORTED, "batch too large for one launch");
                nblk[f] = (int)v;
            }
            for (int l = 0; l < kNumLists; ++l) lcap[l] += list_capacity(l, tmp[e]);
        }
    }
    // XCD affinity when the batch has enough epochs to keep every XCD busy
    const int groups = count >= kXcds ? kXcds : 1;
    std::vector<std::vector<long>> gcap(kNumLists, std::vector<long>(kXcds, 0));
    for (int e = 0; e < count; ++e) {
        EpochDesc tmp{};
        tmp.n = in[e].n; tmp.T = in[e].T; tmp.K = in[e].K;
        tmp.Ecap = knot_capacity(in[e].T);
        tmp.nbmax = solver_block_capacity(tmp.Ecap);
        tmp.mode = o.mode;
        tmp.coef_records = records ? 1 : 0;
        for (int l = 0; l < kNumLists; ++l) gcap[l][groups > 1 ? e % kXcds : 0] += list_capacity(l, tmp);
    }
    size_t tab_at[kNumMaps], list_at[kNumLists];
    for (int f = 0; f < kNumMaps; ++f) {
        tab_at[f] = total;
        total += align_up(sizeof(int2) * (size_t)nblk[f]);
    }
    for (int l = 0; l < kNumLists; ++l) {
        if (lcap[l] > 0x7fffffffL) return fail(ctx, MEMD_ERR_UNSUPPORTED, "batch too large for one launch");
        list_at[l] = total;
        total += align_up(sizeof(int4) * (size_t)lcap[l]);
    }
    size_t cnt_at = total;
    total += align_up(sizeof(int) * kNumLists * kXcds);
    size_t stats_at = total;
    total += align_up(sizeof(unsigned long long) * 4);
    size_t ndone_at = total;
    total += align_up(sizeof(int));
    memd_status rs = ensure_arena(ctx, total);
    if (rs != MEMD_OK) return rs;

    std::vector<EpochDesc> eds(count);
    for (int e = 0; e < count; ++e) {
        const EpochIn& I = in[e];
        const EpochLayout& L = lay[e];
        char* b = ctx->arena + base[e];
        EpochDesc& d = eds[e];
        d.n = I.n; d.T = I.T; d.K = I.K; d.W = (I.T + 63) / 64;
        d.Ecap = knot_capacity(I.T);
        d.nbmax = solver_block_capacity(d.Ecap);
        d.N_imf = I.N_imf; d.S = I.S; d.mode = o.mode; d.window = o.window;
        d.sd = o.sd; d.sd2 = o.sd2; d.tol = o.tol;
        d.y = I.y; d.V = I.V; d.imfs = I.imfs;
        d.running = (double*)(b + L.running);
        d.mask = (unsigned long long*)(b + L.mask);
        d.pre = (int*)(b + L.pre);
        d.E = (int*)(b + L.E);
        d.X = (int*)(b + L.X);
        d.fac = (double*)(b + L.fac);
        d.red = (double*)(b + L.redsys);
        d.yknot = (double*)(b + L.yknot);
        d.ysol = (double*)(b + L.ysol);
        d.coef = (double*)(b + L.coef);
        d.sep = (double*)(b + L.sep);
        d.mean = (double*)(b + L.mean);
        d.tstat = (double*)(b + L.tstat);
        d.stopred = (int*)(b + L.red);
        d.d_trace = o.d_knot_counts;
        d.max_passes = o.d_knot_counts ? o.max_passes : 0;
        d.sift_counts = (int*)(b + L.sift);
        d.st = (EpochState*)(b + L.st);
        d.stats = prof ? (unsigned long long*)(ctx->arena + stats_at) : nullptr;
        d.ndone = (int*)(ctx->arena + ndone_at);
        d.coef_records = records ? 1 : 0;
    }
    std::vector<int2> tabs;
    hipStream_t s = work_stream(ctx);
    int* ndone_dev = (int*)(ctx->arena + ndone_at);
    HIP_TRY(ctx, hipMemsetAsync(ndone_dev, 0, sizeof(int), s));
    HIP_TRY(ctx, hipMemcpyAsync(ctx->arena + desc_at, eds.data(), sizeof(EpochDesc) * count, hipMemcpyHostToDevice, s));
    for (int f = 0; f < kNumMaps; ++f) {
        // block order: block id b runs on XCD b mod 8; place epoch e's blocks on
        // XCD e mod 8 (falling back to any queue once its own is drained)
        std::vector<std::vector<int2>> q(groups);
        for (int e = 0; e < count; ++e) {
            int nb = blocks_for(f, eds[e], variant);
            for (int b = 0; b < nb; ++b) q[groups > 1 ? e % kXcds : 0].push_back(make_int2(e, b));
        }
        std::vector<size_t> head(groups, 0);
        for (int b = 0; b < nblk[f]; ++b) {
            int x = b % groups;
            if (head[x] == q[x].size()) {
                size_t best = 0;
                for (int y = 0; y < groups; ++y)
                    if (q[y].size() - head[y] > best) { best = q[y].size() - head[y]; x = y; }
            }
            tabs.push_back(q[x][head[x]++]);
        }
    }
    {
        size_t pos = 0;
        for (int f = 0; f < kNumMaps; ++f) {
            if (nblk[f] > 0)
                HIP_TRY(ctx, hipMemcpyAsync(ctx->arena + tab_at[f], tabs.data() + pos, sizeof(int2) * nblk[f],
                                            hipMemcpyHostToDevice, s));
            pos += nblk[f];
        }
    }

    StepPlan plan{};
    plan.eps = (const EpochDesc*)(ctx->arena + desc_at);
    plan.count = count;
    for (int f = 0; f < kNumMaps; ++f) {
        plan.tab[f] = (const int2*)(ctx->arena + tab_at[f]);
        plan.blocks[f] = nblk[f];
    }
    plan.wl.count = (int*)(ctx->arena + cnt_at);
    plan.wl.groups = groups;
    for (int l = 0; l < kNumLists; ++l) {
        plan.wl.item[l] = (int4*)(ctx->arena + list_at[l]);
        long at = 0;
        for (int x = 0; x < kXcds; ++x) {
            plan.wl.off[l][x] = (int)at;
            plan.wl.cap[l][x] = (int)gcap[l][x];
            at += gcap[l][x];
        }
        long g = lcap[l] < kListGrid ? lcap[l] : kListGrid;
        if (groups > 1) g = kListGrid;   // a multiple of kXcds, every XCD gets its share
        plan.grid[l] = (int)(g > 0 ? g : 1);
    }
    plan.mode = o.mode;
    plan.nmax = nmax;
    plan.eval_variant = variant;
    plan.coef_records = records ? 1 : 0;

    // Fix mode: exactly (N_imf-1)*S steps.  Stop mode: at most (N_imf-1)*(S+1)
    // passes (S subtractions plus the deciding pass per IMF); the host replays
    // steps in chunks and stops early once every epoch reports done.
    long steps = 0;
    for (int e = 0; e < count; ++e) {
        long st = (long)(in[e].N_imf - 1) * (o.mode == MEMD_MODE_STOP ? in[e].S + 1 : in[e].S);
        if (st > steps) steps = st;
    }

    if (prof) {
        std::memset(prof, 0, sizeof *prof);
        HIP_TRY(ctx, hipMemsetAsync(ctx->arena + stats_at, 0, sizeof(unsigned long long) * 4, s));
        HIP_TRY(ctx, hipEventRecord(ctx->pev[0], s));
        HIP_TRY(ctx, launch_init(plan, s));
        HIP_TRY(ctx, hipEventRecord(ctx->pev[1], s));
        HIP_TRY(ctx, hipEventSynchronize(ctx->pev[1]));
        float ms = 0.f;
        HIP_TRY(ctx, hipEventElapsedTime(&ms, ctx->pev[0], ctx->pev[1]));
        prof->ms_per_launch[kNumStepKernels] = ms;
        prof->launches[kNumStepKernels] = 1;
        double acc[kNumStepKernels] = {};
        HIP_TRY(ctx, hipEventRecord(ctx->t0, s));
        long ran = 0;
        for (long i = 0; i < steps; ++i) {
            HIP_TRY(ctx, launch_step_timed(plan, s, ctx->pev));
            HIP_TRY(ctx, hipEventSynchronize(ctx->pev[kNumStepKernels]));
            ++ran;
            for (int kk = 0; kk < kNumStepKernels; ++kk) {
                HIP_TRY(ctx, hipEventElapsedTime(&ms, ctx->pev[kk], ctx->pev[kk + 1]));
                acc[kk] += ms;
            }
            if (o.mode == MEMD_MODE_STOP) {
                int nd = 0;
                HIP_TRY(ctx, hipMemcpy(&nd, ndone_dev, sizeof(int), hipMemcpyDeviceToHost));
                if (nd >= count) break;
            }
        }
        steps = ran;
        HIP_TRY(ctx, hipEventRecord(ctx->t1, s));
        for (int kk = 0; kk < kNumStepKernels; ++kk) {
            prof->launches[kk] = steps;
            prof->ms_per_launch[kk] = steps ? acc[kk] / (double)steps : 0.0;
        }
        prof->steps = steps;
        unsigned long long st4[4];
        HIP_TRY(ctx, hipMemcpy(st4, ctx->arena + stats_at, sizeof st4, hipMemcpyDeviceToHost));
        prof->knots = (int64_t)st4[0];
        prof->knot_channels = (int64_t)st4[1];
        prof->eval_points = (int64_t)st4[2];
        prof->samples = (int64_t)st4[3];
    } else {
    HIP_TRY(ctx, launch_init(plan, s));
    HIP_TRY(ctx, hipEventRecord(ctx->t0, s));
    if (o.use_graph && steps > 1) {
        // capture one step on the owned stream; the plan lives in the arena, so the
        // key is the arena address plus the launch geometry
        std::string key = std::to_string((uintptr_t)ctx->arena) + "/" + std::to_string(count) + "/" +
                          std::to_string(plan.mode) + "/" + std::to_string(plan.nmax) + "/" +
                          std::to_string(plan.eval_variant);
        for (int f = 0; f < kNumMaps; ++f) key += "/" + std::to_string(plan.blocks[f]);
        key += "/g" + std::to_string(plan.wl.groups);
        for (int l = 0; l < kNumLists; ++l) {
            key += "/" + std::to_string(plan.grid[l]);
            for (int x = 0; x < kXcds; ++x) key += ":" + std::to_string(plan.wl.cap[l][x]);
        }
        if (!ctx->graph_exec || ctx->graph_key != key) {
            if (ctx->graph_exec) {
                HIP_TRY(ctx, hipGraphExecDestroy(ctx->graph_exec));
                ctx->graph_exec = nullptr;
            }
            hipGraph_t g = nullptr;
            HIP_TRY(ctx, hipStreamBeginCapture(ctx->own, hipStreamCaptureModeThreadLocal));
            hipError_t le = launch_step(plan, ctx->own);
            hipError_t ce = hipStreamEndCapture(ctx->own, &g);
            if (le != hipSuccess || ce != hipSuccess) {
                if (g) (void)hipGraphDestroy(g);
                return fail(ctx, MEMD_ERR_HIP, "graph capture failed: %s / %s", hipGetErrorString(le),
                            hipGetErrorString(ce));
            }
            hipError_t ie = hipGraphInstantiate(&ctx->graph_exec, g, nullptr, nullptr, 0);
            (void)hipGraphDestroy(g);
            if (ie != hipSuccess) {
                ctx->graph_exec = nullptr;
                return fail(ctx, MEMD_ERR_HIP, "graph instantiate failed: %s", hipGetErrorString(ie));
            }
            ctx->graph_key = key;
        }
    }
    {
        const bool graph = o.use_graph && steps > 1;
        auto launch_steps = [&](long nsteps) -> memd_status {
            if (graph) {
                HIP_TRY(ctx, hipEventRecord(ctx->ev_in, s));
                HIP_TRY(ctx, hipStreamWaitEvent(ctx->own, ctx->ev_in, 0));
                for (long i = 0; i < nsteps; ++i) HIP_TRY(ctx, hipGraphLaunch(ctx->graph_exec, ctx->own));
                HIP_TRY(ctx, hipEventRecord(ctx->ev_out, ctx->own));
                HIP_TRY(ctx, hipStreamWaitEvent(s, ctx->ev_out, 0));
            } else {
                for (long i = 0; i < nsteps; ++i) HIP_TRY(ctx, launch_step(plan, s));
            }
            return MEMD_OK;
        };
        if (o.mode == MEMD_MODE_STOP) {
            // epochs finish after data-dependent pass counts: replay in chunks and
            // stop once the device-side done counter covers the batch
            long left = steps;
            while (left > 0) {
                long chunk = left < 4 ? left : 4;
                memd_status r2 = launch_steps(chunk);
                if (r2 != MEMD_OK) return r2;
                left -= chunk;
                int nd = 0;
                HIP_TRY(ctx, hipMemcpyAsync(&nd, ndone_dev, sizeof(int), hipMemcpyDeviceToHost, s));
                HIP_TRY(ctx, hipStreamSynchronize(s));
                if (nd >= count) break;
            }
        } else {
            memd_status r2 = launch_steps(steps);
            if (r2 != MEMD_OK) return r2;
        }
    }
    HIP_TRY(ctx, hipEventRecord(ctx->t1, s));
    }

    // results: per-epoch state and sift counts
    std::vector<EpochState> sts(count);
    std::vector<std::vector<int>> sc(count);
    for (int e = 0; e < count; ++e) {
        HIP_TRY(ctx, hipMemcpyAsync(&sts[e], eds[e].st, sizeof(EpochState), hipMemcpyDeviceToHost, s));
        if (in[e].h_sift_counts) {
            sc[e].resize(in[e].N_imf);
            HIP_TRY(ctx, hipMemcpyAsync(sc[e].data(), eds[e].sift_counts, sizeof(int) * in[e].N_imf,
                                        hipMemcpyDeviceToHost, s));
        }
    }
    HIP_TRY(ctx, hipStreamSynchronize(s));
    if (elapsed) {
        float ms = 0.f;
        HIP_TRY(ctx, hipEventElapsedTime(&ms, ctx->t0, ctx->t1));
        *elapsed = ms * 1e-3;
    }
    for (int e = 0; e < count; ++e) {
        if (!sts[e].done)
            return fail(ctx, MEMD_ERR_HIP, "epoch %d did not finish (imf=%d nsift=%d)", e, sts[e].imf, sts[e].nsift);
        if (in[e].h_sift_counts)
            for (int i = 0; i < in[e].N_imf; ++i) in[e].h_sift_counts[i] = sc[e][i];
        if (in[e].h_n_imf_out) *in[e].h_n_imf_out = sts[e].n_imf_out;
    }
    return MEMD_OK;
}

memd_status check_coords(memd_ctx* ctx, const int32_t* d_x, size_t T) {
    if (!d_x) return MEMD_OK;
    int* flag = ctx->flag;
    hipStream_t s = work_stream(ctx);
    HIP_TRY(ctx, hipMemsetAsync(flag, 0, sizeof(int), s));
    HIP_TRY(ctx, launch_check_coords(d_x, (int)T, flag, s));
    int h = 0;
    HIP_TRY(ctx, hipMemcpyAsync(&h, flag, sizeof(int), hipMemcpyDeviceToHost, s));
    HIP_TRY(ctx, hipStreamSynchronize(s));
    if (h) return fail(ctx, MEMD_ERR_COORDS, "sample coordinates must be 0..T-1 (the reference's callers pass exactly that)");
    return MEMD_OK;
}

std::string dir_key(int n, int K, const int32_t* P) {
    std::string k = std::to_string(n) + ":" + std::to_string(K);
    for (int j = 0; j < n; ++j) k += "," + std::to_string(P[j]);
    return k;
}

memd_status cached_directions(memd_ctx* ctx, int n, int K, const int32_t* P, const double** out) {
    std::string key = dir_key(n, K, P);
    auto it = ctx->dircache.find(key);
    if (it != ctx->dircache.end()) { *out = it->second; return MEMD_OK; }
    std::vector<double> V((size_t)K * n);
    directions_host(n, K, P, V.data());
    double* d = nullptr;
    hipError_t e = hipMalloc((void**)&d, V.size() * sizeof(double));
    if (e != hipSuccess) return fail(ctx, MEMD_ERR_NOMEM, "direction cache allocation failed: %s", hipGetErrorString(e));
    HIP_TRY(ctx, hipMemcpy(d, V.data(), V.size() * sizeof(double), hipMemcpyHostToDevice));
    ctx->dircache[key] = d;
    *out = d;
    return MEMD_OK;
}

struct Scratch {  // FP64 staging for the f32 entry points (separate allocation)
    double* p = nullptr;
    ~Scratch() { if (p) (void)hipFree(p); }
};

}  // namespace

// ======================================================================= API

extern "C" {

const char* memd_version(void) { return "memd-amd 0.1 (gfx950, fp64)"; }

void memd_opts_default(memd_opts* o) {
    if (!o) return;
    std::memset(o, 0, sizeof *o);
    o->mode = MEMD_MODE_FIX;
    o->window = MEMD_WINDOW_REFWARP32;
    o->sd = 0.075;
    o->sd2 = 0.75;
    o->tol = 0.075;
    o->use_graph = 1;
}

size_t memd_workspace_bytes(size_t n, size_t T, size_t K) {
    if (n < 1 || T < 4 || K < 1) return 0;
    const bool records = eval_variant_choice() != kEvalFused;
    return layout_for((int)n, (int)T, (int)K, 1, records).total;
}

memd_status memd_ctx_create(int device, memd_ctx** out) {
    if (!out) return MEMD_ERR_ARG;
    *out = nullptr;
    int ndev = 0;
    hipError_t e = hipGetDeviceCount(&ndev);
    if (e != hipSuccess || ndev == 0) {
        (void)hipGetLastError();
        return MEMD_ERR_HIP;
    }
    memd_ctx* ctx = new memd_ctx();
    if (device < 0) {
        if (hipGetDevice(&ctx->device) != hipSuccess) { delete ctx; return MEMD_ERR_HIP; }
    } else {
        if (device >= ndev) { delete ctx; return MEMD_ERR_ARG; }
        ctx->device = device;
    }
    if (hipSetDevice(ctx->device) != hipSuccess ||
        hipStreamCreateWithFlags(&ctx->own, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&ctx->ev_in, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&ctx->ev_out, hipEventDisableTiming) != hipSuccess ||
        hipEventCreate(&ctx->t0) != hipSuccess || hipEventCreate(&ctx->t1) != hipSuccess ||
        hipMalloc((void**)&ctx->flag, sizeof(int)) != hipSuccess || configure_kernels() != hipSuccess) {
        memd_ctx_destroy(ctx);
        return MEMD_ERR_HIP;
    }
    for (int i = 0; i <= kNumStepKernels; ++i) {
        if (hipEventCreate(&ctx->pev[i]) == hipSuccess) continue;
        memd_ctx_destroy(ctx);
        return MEMD_ERR_HIP;
    }
    *out = ctx;
    return MEMD_OK;
}

memd_status memd_ctx_destroy(memd_ctx* ctx) {
    if (!ctx) return MEMD_OK;
    (void)hipSetDevice(ctx->device);
    (void)hipDeviceSynchronize();
    if (ctx->graph_exec) (void)hipGraphExecDestroy(ctx->graph_exec);
    for (auto& kv : ctx->dircache) (void)hipFree(kv.second);
    if (ctx->arena) (void)hipFree(ctx->arena);
    if (ctx->flag) (void)hipFree(ctx->flag);
    if (ctx->own) (void)hipStreamDestroy(ctx->own);
    if (ctx->ev_in) (void)hipEventDestroy(ctx->ev_in);
    if (ctx->ev_out) (void)hipEventDestroy(ctx->ev_out);
    if (ctx->t0) (void)hipEventDestroy(ctx->t0);
    if (ctx->t1) (void)hipEventDestroy(ctx->t1);
    for (int i = 0; i <= kNumStepKernels; ++i)
        if (ctx->pev[i]) (void)hipEventDestroy(ctx->pev[i]);
    delete ctx;
    return MEMD_OK;
}

memd_status memd_ctx_set_stream(memd_ctx* ctx, void* hip_stream) {
    if (!ctx) return MEMD_ERR_ARG;
    ctx->user = (hipStream_t)hip_stream;
    return MEMD_OK;
}

const char* memd_last_error(memd_ctx* ctx) { return ctx ? ctx->err.c_str() : "null context"; }

memd_status memd_directions_f64(memd_ctx* ctx, size_t n, size_t K, const int32_t* h_primes, double* d_V) {
    if (!ctx) return MEMD_ERR_ARG;
    if (n < 1 || K < 1 || !d_V) return fail(ctx, MEMD_ERR_ARG, "invalid direction arguments n=%zu K=%zu", n, K);
    if (n > (size_t)kMaxChannels || K > 65536) return fail(ctx, MEMD_ERR_UNSUPPORTED, "n or K too large");
    HIP_TRY(ctx, hipSetDevice(ctx->device));
    std::vector<int32_t> P(n);
    if (h_primes) std::memcpy(P.data(), h_primes, n * sizeof(int32_t));
    else reference_primes((int)n, (int)K, P.data());
    if (!primes_valid((int)n, P.data())) return fail(ctx, MEMD_ERR_ARG, "prime table needs P[0] != 0 and P[j>0] >= 2");
    const double* dV = nullptr;
    memd_status rs = cached_directions(ctx, (int)n, (int)K, P.data(), &dV);
    if (rs != MEMD_OK) return rs;
    HIP_TRY(ctx, hipMemcpyAsync(d_V, dV, n * K * sizeof(double), hipMemcpyDeviceToDevice, work_stream(ctx)));
    HIP_TRY(ctx, hipStreamSynchronize(work_stream(ctx)));
    return MEMD_OK;
}

memd_status memd_run_f64(memd_ctx* ctx, size_t n, size_t T, size_t K, size_t N_imf, size_t S, const int32_t* d_x,
                         const double* d_y, const double* d_V, double* d_IMFs, const memd_opts* opts, double* elapsed_s) {
    if (!ctx) return MEMD_ERR_ARG;
    memd_status rs = validate_shape(ctx, n, T, K, N_imf, S);
    if (rs != MEMD_OK) return rs;
    if (!d_y || !d_V || !d_IMFs) return fail(ctx, MEMD_ERR_ARG, "null device pointer");
    HIP_TRY(ctx, hipSetDevice(ctx->device));
    rs = check_coords(ctx, d_x, T);
    if (rs != MEMD_OK) return rs;
    std::vector<EpochIn> in{{(int)n, (int)T, (int)K, (int)N_imf, (int)S, d_y, d_V, d_IMFs,
                             opts ? opts->h_sift_counts : nullptr, opts ? opts->h_n_imf_out : nullptr}};
    return run_epochs(ctx, in, opts, elapsed_s);
}

memd_status memd_run_f32(memd_ctx* ctx, size_t n, size_t T, size_t K, size_t N_imf, size_t S, const int32_t* d_x,
                         const float* d_y, const float* d_V, float* d_IMFs, const memd_opts* opts, double* elapsed_s) {
    if (!ctx) return MEMD_ERR_ARG;
    memd_status rs = validate_shape(ctx, n, T, K, N_imf, S);
    if (rs != MEMD_OK) return rs;
    if (!d_y || !d_V || !d_IMFs) return fail(ctx, MEMD_ERR_ARG, "null device pointer");
    HIP_TRY(ctx, hipSetDevice(ctx->device));
    size_t ny = n * T, nv = K * n, no = N_imf * n * T;
    Scratch sc;
    hipError_t e = hipMalloc((void**)&sc.p, (ny + nv + no) * sizeof(double));
    if (e != hipSuccess) return fail(ctx, MEMD_ERR_NOMEM, "FP64 staging allocation failed: %s", hipGetErrorString(e));
    hipStream_t s = work_stream(ctx);
    HIP_TRY(ctx, launch_f32_to_f64(d_y, sc.p, ny, s));
    HIP_TRY(ctx, launch_f32_to_f64(d_V, sc.p + ny, nv, s));
    rs = memd_run_f64(ctx, n, T, K, N_imf, S, d_x, sc.p, sc.p + ny, sc.p + ny + nv, opts, elapsed_s);
    if (rs != MEMD_OK) return rs;
    HIP_TRY(ctx, launch_f64_to_f32(sc.p + ny + nv, d_IMFs, no, s));
    HIP_TRY(ctx, hipStreamSynchronize(s));
    return MEMD_OK;
}

static memd_status mex_common(memd_ctx* ctx, int32_t n, int32_t K, int32_t N_imf, int32_t S, size_t numel,
                              const int32_t* d_index, const int32_t* d_primes, size_t* T_out, const double** dV) {
    if (n < 1 || K < 1 || N_imf < 1 || S < 1 || numel % (size_t)n != 0)
        return fail(ctx, MEMD_ERR_ARG, "invalid MEX arguments n=%d K=%d N_imf=%d S=%d numel=%zu", n, K, N_imf, S, numel);
    size_t T = numel / (size_t)n;   // memdMatlabCuda.cu:1038-1039
    memd_status rs = validate_shape(ctx, (size_t)n, T, (size_t)K, (size_t)N_imf, (size_t)S);
    if (rs != MEMD_OK) return rs;
    if (!d_primes) return fail(ctx, MEMD_ERR_ARG, "null primes array");
    HIP_TRY(ctx, hipSetDevice(ctx->device));
    rs = check_coords(ctx, d_index, T);
    if (rs != MEMD_OK) return rs;
    std::vector<int32_t> P(n);
    HIP_TRY(ctx, hipMemcpy(P.data(), d_primes, sizeof(int32_t) * n, hipMemcpyDeviceToHost));
    if (!primes_valid(n, P.data())) return fail(ctx, MEMD_ERR_ARG, "prime table needs P[0] != 0 and P[j>0] >= 2");
    rs = cached_directions(ctx, n, K, P.data(), dV);
    if (rs != MEMD_OK) return rs;
    *T_out = T;
    return MEMD_OK;
}

memd_status memd_mex_compat_f64(memd_ctx* ctx, int32_t n, int32_t K, int32_t N_imf, int32_t S, size_t numel,
                                const double* d_data, const int32_t* d_index, const int32_t* d_primes, double* d_out) {
    if (!ctx) return MEMD_ERR_ARG;
    if (!d_data || !d_out) return fail(ctx, MEMD_ERR_ARG, "null device pointer");
    size_t T = 0;
    const double* dV = nullptr;
    memd_status rs = mex_common(ctx, n, K, N_imf, S, numel, d_index, d_primes, &T, &dV);
    if (rs != MEMD_OK) return rs;
    return memd_run_f64(ctx, n, T, K, N_imf, S, nullptr, d_data, dV, d_out, nullptr, nullptr);
}

memd_status memd_mex_compat_f32(memd_ctx* ctx, int32_t n, int32_t K, int32_t N_imf, int32_t S, size_t numel,
                                const float* d_data, const int32_t* d_index, const int32_t* d_primes, float* d_out) {
    if (!ctx) return MEMD_ERR_ARG;
    if (!d_data || !d_out) return fail(ctx, MEMD_ERR_ARG, "null device pointer");
    size_t T = 0;
    const double* dV = nullptr;
    memd_status rs = mex_common(ctx, n, K, N_imf, S, numel, d_index, d_primes, &T, &dV);
    if (rs != MEMD_OK) return rs;
    size_t ny = (size_t)n * T, no = (size_t)N_imf * n * T;
    Scratch sc;
    hipError_t e = hipMalloc((void**)&sc.p, (ny + no) * sizeof(double));
    if (e != hipSuccess) return fail(ctx, MEMD_ERR_NOMEM, "FP64 staging allocation failed: %s", hipGetErrorString(e));
    hipStream_t s = work_stream(ctx);
    HIP_TRY(ctx, launch_f32_to_f64(d_data, sc.p, ny, s));
    rs = memd_run_f64(ctx, n, T, K, N_imf, S, nullptr, sc.p, dV, sc.p + ny, nullptr, nullptr);
    if (rs != MEMD_OK) return rs;
    HIP_TRY(ctx, launch_f64_to_f32(sc.p + ny, d_out, no, s));
    HIP_TRY(ctx, hipStreamSynchronize(s));
    return MEMD_OK;
}

static memd_status batch_common(memd_ctx* ctx, const memd_epoch* epochs, size_t count, const memd_opts* opts,
                                double* elapsed_s, memd_kernel_profile* prof) {
    if (!ctx) return MEMD_ERR_ARG;
    if (count == 0) return MEMD_OK;
    if (!epochs) return fail(ctx, MEMD_ERR_ARG, "null epoch array");
    std::vector<EpochIn> in(count);
    for (size_t i = 0; i < count; ++i) {
        const memd_epoch& ep = epochs[i];
        memd_status rs = validate_shape(ctx, ep.n, ep.T, ep.K, ep.N_imf, ep.S);
        if (rs != MEMD_OK) return fail(ctx, rs, "epoch %zu: %s", i, ctx->err.c_str());
        if (!ep.d_y || !ep.d_V || !ep.d_IMFs) return fail(ctx, MEMD_ERR_ARG, "epoch %zu: null device pointer", i);
        in[i] = {(int)ep.n, (int)ep.T, (int)ep.K, (int)ep.N_imf, (int)ep.S, ep.d_y, ep.d_V, ep.d_IMFs,
                 ep.h_sift_counts, ep.h_n_imf_out};
    }
    HIP_TRY(ctx, hipSetDevice(ctx->device));
    return run_epochs(ctx, in, opts, elapsed_s, prof);
}

memd_status memd_run_batch_f64(memd_ctx* ctx, const memd_epoch* epochs, size_t count, const memd_opts* opts,
                               double* elapsed_s) {
    return batch_common(ctx, epochs, count, opts, elapsed_s, nullptr);
}

memd_status memd_profile_batch_f64(memd_ctx* ctx, const memd_epoch* epochs, size_t count, const memd_opts* opts,
                                   memd_kernel_profile* out) {
    if (!out) return fail(ctx, MEMD_ERR_ARG, "null profile output");
    return batch_common(ctx, epochs, count, opts, nullptr, out);
}

}  // extern "C"
