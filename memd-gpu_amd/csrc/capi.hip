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
#include "binio.h"
#include "directions.h"
#include "engine.h"

#include <cmath>
#include <cstdarg>
#include <cstdint>
#include <cstdlib>
#include <cstdio>
#include <algorithm>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>

using namespace memd;

// Workspace of the allocation-free entry point (memd_run_noalloc_f64): one
// array in every device's copy of the code object, held until the process
// ends.  Host code never passes `scratch` itself to a kernel or a copy (on the
// host the name is a stand-in, not a device address); it uses the address
// hipGetSymbolAddress returns for the current device, resolved once per device
// when the first context on it is created (scratch_address).
constexpr size_t kScratchBytes = (size_t)MEMD_NOALLOC_SCRATCH_BYTES;
__device__ __align__(256) char scratch[kScratchBytes];

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
    std::map<int, memd_tuning> tuning;          // tile shapes per channel count (memd_autotune / set_tuning)
    char* scratch = nullptr;        // this device's copy of `scratch` (allocation-free entry point)
};

namespace {

// Device address of `scratch` per device, resolved with that device current.
hipError_t scratch_address(int device, char** out) {
    static std::mutex m;
    static std::map<int, char*> table;
    std::lock_guard<std::mutex> lock(m);
    auto it = table.find(device);
    if (it != table.end()) {
        *out = it->second;
        return hipSuccess;
    }
    void* p = nullptr;
    hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    table[device] = (char*)p;
    *out = (char*)p;
    return hipSuccess;
}

constexpr size_t kAlign = 256;
inline size_t align_up(size_t v) { return (v + kAlign - 1) / kAlign * kAlign; }

memd_status fail(memd_ctx* ctx, memd_status code, const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    if (ctx) ctx->err = buf;
    else set_global_error(buf);
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
    size_t running, curt, mask, pre, E, mir, X, fac, redsys, ysol, coef, sep, sepin, mean, tstat, red, sift, st, total;
};

EpochLayout layout_for(int n, int T, int K, int N_imf, bool records) {
    EpochLayout L{};
    size_t W = (size_t)(T + 63) / 64;
    size_t Ecap = (size_t)knot_capacity(T);
    size_t nb = (size_t)solver_block_capacity((int)Ecap);
    size_t off = 0;
    auto take = [&](size_t bytes) { size_t at = off; off += align_up(bytes); return at; };
    L.running = take((size_t)n * T * 8);
    L.curt = take(((size_t)T + 4 * (size_t)K) * channel_stride(n) * 8);   // then yend
    L.mask = take((size_t)K * 2 * W * 8);
    L.pre = take((size_t)K * 2 * W * 4);
    L.E = take((size_t)K * 2 * 4 * 2);   // E, then Eraw
    L.mir = take((size_t)K * 2 * 16);
    L.X = take((size_t)K * 2 * Ecap * 4);
    L.fac = take((size_t)K * 2 * kFacPlanes * Ecap * 2 * 8);
    L.redsys = take((size_t)K * 2 * 3 * nb * 8);
    const size_t np = (size_t)channel_stride(n);
    L.ysol = take((size_t)K * 2 * np * Ecap * 8);
    L.coef = take(records ? (size_t)K * 2 * n * Ecap * 32 : 0);   // the record path's array only
    L.sep = take((size_t)K * 2 * np * (nb + 1) * 8);
    L.sepin = take((size_t)K * 2 * nb * 3 * n * 8);
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
    if (n * T > ((size_t)1 << 31))   // knot rows are addressed with 32-bit offsets
        return fail(ctx, MEMD_ERR_UNSUPPORTED, "n*T=%zu exceeds the engine limit of 2^31 samples per epoch", n * T);
    return MEMD_OK;
}

hipStream_t work_stream(memd_ctx* ctx) { return ctx->user; }

// Built-in tile shapes for n channels: k_block groups of blocks_per_item(n)
// solver blocks; evaluation tiles of 16 / TW channels close to n.
memd_tuning default_tuning(int n) {
    memd_tuning t{};
    t.n = n;
    t.block_groups = block_bpw_fit(n, blocks_per_item(n));
    t.eval_tile_words = n <= 4 ? 4 : (n <= 8 ? 2 : 1);
    return t;
}

// Tile shapes used for n channels: the context's installed choice, else the
// built-in one.
memd_tuning tuning_for(const memd_ctx* ctx, int n) {
    auto it = ctx->tuning.find(n);
    return it != ctx->tuning.end() ? it->second : default_tuning(n);
}

// Descriptor fields that depend only on the epoch's shape and the options.
EpochDesc desc_geometry(const memd_ctx* ctx, const EpochIn& I, const memd_opts& o, bool records) {
    EpochDesc d{};
    d.n = I.n; d.T = I.T; d.K = I.K; d.W = (I.T + 63) / 64;
    d.np = channel_stride(I.n);
    d.Ecap = knot_capacity(I.T);
    d.nbmax = solver_block_capacity(d.Ecap);
    d.N_imf = I.N_imf; d.S = I.S; d.mode = o.mode; d.window = o.window;
    d.sd = o.sd; d.sd2 = o.sd2; d.tol = o.tol;
    d.y = I.y; d.V = I.V; d.imfs = I.imfs;
    d.coef_records = records ? 1 : 0;
    d.spline = o.spline;
    d.boundary = o.boundary;
    const memd_tuning tu = tuning_for(ctx, I.n);
    d.bpw = tu.block_groups;
    d.etw = tu.eval_tile_words;
    // k_block items: at most 2K * groups; workgroups per epoch: the item count at a
    // typical knot density (one knot per 16 samples), 1..256
    d.bcap = 2 * I.K * ((d.nbmax + d.bpw - 1) / d.bpw);
    const long typical = 2L * I.K * (((long)I.T / 512 + 1 + d.bpw - 1) / d.bpw);
    d.bblocks = (int)std::min(256L, std::max(1L, typical));
    return d;
}

// Point a descriptor at its workspace (layout L starting at b).
void bind_workspace(EpochDesc& d, char* b, const EpochLayout& L) {
    d.running = (double*)(b + L.running);
    d.mask = (unsigned long long*)(b + L.mask);
    d.pre = (int*)(b + L.pre);
    d.E = (int*)(b + L.E);
    d.Eraw = d.E + 2 * (size_t)d.K;   // desc_geometry set K
    d.mir = (int4*)(b + L.mir);
    d.X = (int*)(b + L.X);
    d.fac = (double*)(b + L.fac);
    d.red = (double*)(b + L.redsys);
    d.curt = (double*)(b + L.curt);
    d.yend = d.curt + (size_t)d.T * d.np;   // the rows after the signal (desc_geometry set T and np)
    d.ysol = (double*)(b + L.ysol);
    d.coef = (double*)(b + L.coef);
    d.sep = (double*)(b + L.sep);
    d.sepin = (double*)(b + L.sepin);
    d.mean = (double*)(b + L.mean);
    d.tstat = (double*)(b + L.tstat);
    d.stopred = (int*)(b + L.red);
    d.sift_counts = (int*)(b + L.sift);
    d.st = (EpochState*)(b + L.st);
}

// Device region of a step plan: descriptors, block tables of the fixed-size
// launches, the solver work lists and their counters, statistics, the done
// counter, a table of epochs to initialise and a state snapshot.  Offsets are
// relative to the region's base.
struct PlanLayout {
    size_t desc_at = 0, tab_at[kNumMaps] = {}, list_at[kNumLists] = {};
    size_t cnt_at = 0, stats_at = 0, ndone_at = 0, fresh_at = 0, states_at = 0, total = 0;
    size_t bitem_at = 0, bcnt_at = 0;   // per-epoch k_block items (int4) and counters (one line each)
    int nblk[kNumMaps] = {};
    long lcap[kNumLists] = {};
    long gcap[kNumLists][kXcds] = {};
    int groups = 1;
};

// grp[e] in [0, groups) is the XCD group of epoch e (groups == 1: no affinity).
memd_status plan_layout(memd_ctx* ctx, const std::vector<EpochDesc>& eds, const std::vector<int>& grp, int groups,
                        size_t start, PlanLayout* P) {
    const int count = (int)eds.size();
    *P = PlanLayout{};
    P->groups = groups;
    for (int e = 0; e < count; ++e) {
        for (int f = 0; f < kNumMaps; ++f) {
            long v = (long)P->nblk[f] + blocks_for(f, eds[e]);
            if (v > 0x7fffffffL) return fail(ctx, MEMD_ERR_UNSUPPORTED, "batch too large for one launch");
            P->nblk[f] = (int)v;
        }
        for (int l = 0; l < kNumLists; ++l) {
            long c = list_capacity(l, eds[e]);
            P->lcap[l] += c;
            P->gcap[l][grp[e]] += c;
        }
    }
    size_t total = start;
    P->desc_at = total;
    total += align_up(sizeof(EpochDesc) * (size_t)count);
    for (int f = 0; f < kNumMaps; ++f) {
        P->tab_at[f] = total;
        total += align_up(sizeof(int2) * (size_t)P->nblk[f]);
    }
    for (int l = 0; l < kNumLists; ++l) {
        if (P->lcap[l] > 0x7fffffffL) return fail(ctx, MEMD_ERR_UNSUPPORTED, "batch too large for one launch");
        P->list_at[l] = total;
        total += align_up(sizeof(int4) * (size_t)P->lcap[l]);
    }
    P->cnt_at = total;
    total += align_up(kCountBytes);
    size_t bitems = 0;
    for (int e = 0; e < count; ++e) bitems += (size_t)eds[e].bcap;
    P->bitem_at = total;
    total += align_up(sizeof(int4) * bitems);
    P->bcnt_at = total;
    total += align_up(kBcountStride * (size_t)count);
    P->stats_at = total;
    total += align_up(sizeof(unsigned long long) * 4);
    P->ndone_at = total;
    total += align_up(sizeof(int));
    P->fresh_at = total;
    total += align_up(sizeof(int2) * (size_t)P->nblk[kMapInit]);
    P->states_at = total;
    total += align_up(sizeof(EpochState) * (size_t)count);
    P->total = total;
    return MEMD_OK;
}

// Upload descriptors and block tables into the region and fill the plan.  Block
// id b runs on XCD b mod 8: epoch e's blocks are placed at ids of XCD grp[e]
// (falling back to any queue once its own is drained).  tabs is host staging
// that must outlive the asynchronous copies (the caller synchronises later).
memd_status write_plan(memd_ctx* ctx, std::vector<EpochDesc>& eds, const std::vector<int>& grp,
                       const PlanLayout& P, char* region, int mode, bool records,
                       std::vector<int2>& tabs, hipStream_t s, StepPlan* out) {
    const int count = (int)eds.size();
    const int groups = P.groups;
    // the sub-list an epoch appends to is the one its capacity was planned in
    size_t bpos = 0;
    for (int e = 0; e < count; ++e) {
        eds[e].xcd = grp[e];
        eds[e].bitem = (int4*)(region + P.bitem_at) + bpos;
        eds[e].bcount = (int*)(region + P.bcnt_at + kBcountStride * (size_t)e);
        bpos += (size_t)eds[e].bcap;
    }
    HIP_TRY(ctx, hipMemcpyAsync(region + P.desc_at, eds.data(), sizeof(EpochDesc) * count, hipMemcpyHostToDevice, s));
    tabs.clear();
    for (int f = 0; f < kNumMaps; ++f) {
        std::vector<std::vector<int2>> q(groups);
        for (int e = 0; e < count; ++e) {
            int nb = blocks_for(f, eds[e]);
            for (int b = 0; b < nb; ++b) q[grp[e]].push_back(make_int2(e, b));
        }
        std::vector<size_t> head(groups, 0);
        for (int b = 0; b < P.nblk[f]; ++b) {
            int x = b % groups;
            if (head[x] == q[x].size()) {
                size_t best = 0;
                for (int y = 0; y < groups; ++y)
                    if (q[y].size() - head[y] > best) { best = q[y].size() - head[y]; x = y; }
            }
            tabs.push_back(q[x][head[x]++]);
        }
    }
    size_t pos = 0;
    for (int f = 0; f < kNumMaps; ++f) {
        if (P.nblk[f] > 0)
            HIP_TRY(ctx, hipMemcpyAsync(region + P.tab_at[f], tabs.data() + pos, sizeof(int2) * P.nblk[f],
                                        hipMemcpyHostToDevice, s));
        pos += P.nblk[f];
    }
    StepPlan plan{};
    plan.eps = (const EpochDesc*)(region + P.desc_at);
    plan.count = count;
    for (int f = 0; f < kNumMaps; ++f) {
        plan.tab[f] = (const int2*)(region + P.tab_at[f]);
        plan.blocks[f] = P.nblk[f];
    }
    plan.wl.count = (int*)(region + P.cnt_at);
    plan.wl.groups = groups;
    for (int l = 0; l < kNumLists; ++l) {
        plan.wl.item[l] = (int4*)(region + P.list_at[l]);
        long at = 0;
        for (int x = 0; x < kXcds; ++x) {
            plan.wl.off[l][x] = (int)at;
            plan.wl.cap[l][x] = (int)P.gcap[l][x];
            at += P.gcap[l][x];
        }
        // one-wave workgroups for the factor and reduced lists: a larger persistent grid
        const long gmax = (l == kListFactor || l == kListReduced) ? 4L * kListGrid : (long)kListGrid;
        long g = P.lcap[l] < gmax ? P.lcap[l] : gmax;
        if (groups > 1) g = gmax;   // a multiple of kXcds, every XCD gets its share
        plan.grid[l] = (int)(g > 0 ? g : 1);
    }
    int nmax = 0, nlds = 0;
    std::vector<int> ns(count), bpws(count);
    for (int e = 0; e < count; ++e) {
        nmax = eds[e].n > nmax ? eds[e].n : nmax;
        const int tile_ch = eds[e].n <= kFlagLdsMaxCh ? eds[e].n : kFlagChunk;   // chunked above the limit
        if (tile_ch > nlds) nlds = tile_ch;
        ns[e] = eds[e].n;
        bpws[e] = eds[e].bpw;
    }
    const BlockLds bl = block_lds(ns.data(), bpws.data(), count);
    plan.block_rmax = bl.rmax;
    plan.block_tile_off = bl.tile_off;
    plan.block_lds = bl.bytes;
    plan.mode = mode;
    plan.nmax = nmax;
    plan.nlds = nlds;
    plan.coef_records = records ? 1 : 0;
    *out = plan;
    return MEMD_OK;
}

// Packing of a batch onto the XCDs: every sift step runs all epochs, so a step
// lasts as long as the most loaded XCD.  Longest-processing-time first: epochs
// by decreasing n * T * K (the work of a step; ties keep batch order) go to the
// least loaded XCD.  Packing never changes an epoch's arithmetic.
void xcd_pack(const std::vector<EpochIn>& in, std::vector<int>& grp) {
    const int count = (int)in.size();
    std::vector<int> order(count);
    std::vector<long> cost(count);
    for (int e = 0; e < count; ++e) {
        order[e] = e;
        cost[e] = (long)in[e].n * (long)in[e].T * (long)in[e].K;
    }
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return cost[a] > cost[b]; });
    long load[kXcds] = {};
    for (int e : order) {
        int g = 0;
        for (int x = 1; x < kXcds; ++x)
            if (load[x] < load[g]) g = x;
        grp[e] = g;
        load[g] += cost[e];
    }
}

// Every field of a step plan, in order: the launch arguments of one captured step.
std::string plan_key(const StepPlan& p) {
    std::string k;
    auto add = [&](unsigned long long v) { k += std::to_string(v); k += ','; };
    add((uintptr_t)p.eps);
    add((unsigned long long)p.count);
    for (int f = 0; f < kNumMaps; ++f) {
        add((uintptr_t)p.tab[f]);
        add((unsigned long long)p.blocks[f]);
    }
    for (int l = 0; l < kNumLists; ++l) {
        add((uintptr_t)p.wl.item[l]);
        for (int x = 0; x < kXcds; ++x) {
            add((unsigned long long)p.wl.off[l][x]);
            add((unsigned long long)p.wl.cap[l][x]);
        }
        add((unsigned long long)p.grid[l]);
    }
    add((uintptr_t)p.wl.count);
    add((unsigned long long)p.wl.groups);
    add((unsigned long long)p.mode);
    add((unsigned long long)p.nmax);
    add((unsigned long long)p.nlds);
    add((unsigned long long)p.coef_records);
    add((unsigned long long)p.block_rmax);
    add((unsigned long long)p.block_tile_off);
    add((unsigned long long)p.block_lds);
    return k;
}

// Device memory a run works in: the context's arena (grown on demand) or a
// fixed buffer that is never reallocated (the allocation-free entry point; the
// step is then launched directly, since instantiating a graph allocates).
struct RunArena {
    char* fixed = nullptr;   // nullptr: the context's arena
    size_t fixed_bytes = 0;
};

// Option values both the run and the server entry points accept.
memd_status check_opts(memd_ctx* ctx, const memd_opts& o) {
    if (o.mode != MEMD_MODE_FIX && o.mode != MEMD_MODE_STOP)
        return fail(ctx, MEMD_ERR_ARG, "unknown mode %d", o.mode);
    if (o.window != MEMD_WINDOW_REFWARP32 && o.window != MEMD_WINDOW_FULL)
        return fail(ctx, MEMD_ERR_ARG, "unknown extrema window %d", o.window);
    if (o.spline != MEMD_SPLINE_NATURAL && o.spline != MEMD_SPLINE_NOT_A_KNOT)
        return fail(ctx, MEMD_ERR_ARG, "unknown spline end condition %d", o.spline);
    if (o.boundary != MEMD_BOUNDARY_CLAMP && o.boundary != MEMD_BOUNDARY_MIRROR)
        return fail(ctx, MEMD_ERR_ARG, "unknown end treatment %d", o.boundary);
    return MEMD_OK;
}

// The record-path kernels (spline records in HBM, one lane per sample walking the
// channels; at most kStopMaxChannels channels) serve stop mode, and both modes
// with mirrored ends; fix mode with the reference's ends uses the fused evaluation.
bool record_path(const memd_opts& o) { return o.mode == MEMD_MODE_STOP || o.boundary == MEMD_BOUNDARY_MIRROR; }

memd_status check_record_channels(memd_ctx* ctx, const memd_opts& o, int n) {
    if (record_path(o) && n > kStopMaxChannels)
        return fail(ctx, MEMD_ERR_UNSUPPORTED, "stop mode and mirrored ends support at most %d channels (got %d)",
                    kStopMaxChannels, n);
    return MEMD_OK;
}

// Core: run a set of epochs (all already FP64 and on the device) to completion.
memd_status run_epochs(memd_ctx* ctx, const std::vector<EpochIn>& in, const memd_opts* opts, double* elapsed,
                       memd_kernel_profile* prof = nullptr, RunArena where = RunArena{}) {
    memd_opts o;
    if (opts) o = *opts; else memd_opts_default(&o);
    memd_status rs = check_opts(ctx, o);
    if (rs != MEMD_OK) return rs;
    const int count = (int)in.size();
    if (count == 0) return MEMD_OK;
    if (o.d_knot_counts && count != 1)
        return fail(ctx, MEMD_ERR_ARG, "knot-count tracing is only available for single-epoch calls");
    int nmax = 0;
    for (const EpochIn& I : in) nmax = I.n > nmax ? I.n : nmax;
    if ((rs = check_record_channels(ctx, o, nmax)) != MEMD_OK) return rs;

    // ---- workspace plan: every epoch's workspace, then the plan region
    const bool records = record_path(o);
    std::vector<EpochLayout> lay(count);
    std::vector<size_t> base(count);
    std::vector<EpochDesc> eds(count);
    size_t total = 0;
    for (int e = 0; e < count; ++e) {
        lay[e] = layout_for(in[e].n, in[e].T, in[e].K, in[e].N_imf, records);
        base[e] = total;
        total += lay[e].total;
        eds[e] = desc_geometry(ctx, in[e], o, records);
    }
    // XCD affinity when the batch has enough epochs to keep every XCD busy
    const int groups = count >= kXcds ? kXcds : 1;
    std::vector<int> grp(count, 0);
    if (groups > 1) xcd_pack(in, grp);
    PlanLayout P;
    rs = plan_layout(ctx, eds, grp, groups, total, &P);
    if (rs != MEMD_OK) return rs;
    char* arena = nullptr;
    if (where.fixed) {
        if (P.total > where.fixed_bytes)
            return fail(ctx, MEMD_ERR_UNSUPPORTED,
                        "this call needs %zu bytes of workspace; the allocation-free entry point holds %zu",
                        P.total, where.fixed_bytes);
        arena = where.fixed;
    } else {
        rs = ensure_arena(ctx, P.total);
        if (rs != MEMD_OK) return rs;
        arena = ctx->arena;
    }
    const size_t stats_at = P.stats_at;
    int* ndone_dev = (int*)(arena + P.ndone_at);
    for (int e = 0; e < count; ++e) {
        EpochDesc& d = eds[e];
        bind_workspace(d, arena + base[e], lay[e]);
        d.d_trace = o.d_knot_counts;
        d.max_passes = o.d_knot_counts ? o.max_passes : 0;
        d.stats = prof ? (unsigned long long*)(arena + stats_at) : nullptr;
        d.ndone = ndone_dev;
    }
    hipStream_t s = work_stream(ctx);
    HIP_TRY(ctx, launch_zero_ints(ndone_dev, 1, s));   // a kernel, not a memset node: callers may capture
    StepPlan plan{};
    std::vector<int2> tabs;
    rs = write_plan(ctx, eds, grp, P, arena, records ? kModeStop : kModeFix, records, tabs, s, &plan);
    if (rs != MEMD_OK) return rs;

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
        HIP_TRY(ctx, hipMemsetAsync(arena + stats_at, 0, sizeof(unsigned long long) * 4, s));
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
        HIP_TRY(ctx, hipMemcpy(st4, arena + stats_at, sizeof st4, hipMemcpyDeviceToHost));
        prof->knots = (int64_t)st4[0];
        prof->knot_channels = (int64_t)st4[1];
        prof->eval_points = (int64_t)st4[2];
        prof->samples = (int64_t)st4[3];
    } else {
    HIP_TRY(ctx, launch_init(plan, s));
    HIP_TRY(ctx, hipEventRecord(ctx->t0, s));
    const bool graph = o.use_graph && steps > 1 && !where.fixed;
    if (graph) {
        // capture one step on the owned stream.  The graph bakes in every launch
        // argument, all of which come from the plan (descriptor array, block tables,
        // work lists and their counters, grids, LDS sizes): it is replayed only for
        // a plan equal to the captured one in every field.
        const std::string key = plan_key(plan);
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

struct Staging {  // FP64 staging for the f32 entry points and the autotuner (separate allocations)
    double* p = nullptr;
    ~Staging() { if (p) (void)hipFree(p); }
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
    return layout_for((int)n, (int)T, (int)K, 1, false).total;
}

size_t memd_workspace_bytes_opts(size_t n, size_t T, size_t K, size_t N_imf, const memd_opts* opts) {
    if (n < 1 || T < 4 || K < 1 || N_imf < 1) return 0;
    const bool records = opts && record_path(*opts);
    return layout_for((int)n, (int)T, (int)K, (int)N_imf, records).total;
}

memd_status memd_ctx_create(int device, memd_ctx** out) {
    if (!out) return MEMD_ERR_ARG;
    *out = nullptr;
    int ndev = 0;
    hipError_t e = hipGetDeviceCount(&ndev);
    if (e != hipSuccess || ndev == 0) {
        (void)hipGetLastError();
        return fail(nullptr, MEMD_ERR_HIP, "no HIP device available (%s)",
                    e != hipSuccess ? hipGetErrorString(e) : "device count is 0");
    }
    memd_ctx* ctx = new memd_ctx();
    if (device < 0) {
        if (hipGetDevice(&ctx->device) != hipSuccess) { delete ctx; return MEMD_ERR_HIP; }
    } else {
        if (device >= ndev) {
            delete ctx;
            return fail(nullptr, MEMD_ERR_ARG, "device %d does not exist (%d visible)", device, ndev);
        }
        ctx->device = device;
    }
    const char* step = "hipSetDevice";
    hipError_t he = hipSetDevice(ctx->device);
    if (he == hipSuccess) { step = "hipStreamCreateWithFlags"; he = hipStreamCreateWithFlags(&ctx->own, hipStreamNonBlocking); }
    if (he == hipSuccess) { step = "hipEventCreateWithFlags"; he = hipEventCreateWithFlags(&ctx->ev_in, hipEventDisableTiming); }
    if (he == hipSuccess) he = hipEventCreateWithFlags(&ctx->ev_out, hipEventDisableTiming);
    if (he == hipSuccess) { step = "hipEventCreate"; he = hipEventCreate(&ctx->t0); }
    if (he == hipSuccess) he = hipEventCreate(&ctx->t1);
    if (he == hipSuccess) { step = "hipMalloc"; he = hipMalloc((void**)&ctx->flag, sizeof(int)); }
    if (he == hipSuccess) { step = "hipFuncSetAttribute"; he = configure_kernels(); }
    if (he == hipSuccess) { step = "hipGetSymbolAddress(scratch)"; he = scratch_address(ctx->device, &ctx->scratch); }
    if (he != hipSuccess) {
        memd_ctx_destroy(ctx);
        return fail(nullptr, MEMD_ERR_HIP, "context set-up on device %d: %s failed: %s", device, step,
                    hipGetErrorString(he));
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
    if (!ctx) return fail(nullptr, MEMD_ERR_ARG, "null context");
    ctx->user = (hipStream_t)hip_stream;
    return MEMD_OK;
}

const char* memd_last_error(memd_ctx* ctx) { return ctx ? ctx->err.c_str() : global_error(); }

memd_status memd_directions_f64(memd_ctx* ctx, size_t n, size_t K, const int32_t* h_primes, double* d_V) {
    if (!ctx) return fail(nullptr, MEMD_ERR_ARG, "null context");
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

memd_status memd_directions_f32(memd_ctx* ctx, size_t n, size_t K, const int32_t* h_primes, float* d_V) {
    if (!ctx) return fail(nullptr, MEMD_ERR_ARG, "null context");
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
    HIP_TRY(ctx, launch_f64_to_f32(dV, d_V, n * K, work_stream(ctx)));
    HIP_TRY(ctx, hipStreamSynchronize(work_stream(ctx)));
    return MEMD_OK;
}

memd_status memd_run_f64(memd_ctx* ctx, size_t n, size_t T, size_t K, size_t N_imf, size_t S, const int32_t* d_x,
                         const double* d_y, const double* d_V, double* d_IMFs, const memd_opts* opts, double* elapsed_s) {
    if (!ctx) return fail(nullptr, MEMD_ERR_ARG, "null context");
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

memd_status memd_run_noalloc_f64(memd_ctx* ctx, size_t n, size_t T, size_t K, size_t N_imf, size_t S,
                                 const int32_t* d_x, const double* d_y, const double* d_V, double* d_IMFs,
                                 const memd_opts* opts, double* elapsed_s) {
    if (!ctx) return fail(nullptr, MEMD_ERR_ARG, "null context");
    memd_status rs = validate_shape(ctx, n, T, K, N_imf, S);
    if (rs != MEMD_OK) return rs;
    if (!d_y || !d_V || !d_IMFs) return fail(ctx, MEMD_ERR_ARG, "null device pointer");
    if (!ctx->scratch) return fail(ctx, MEMD_ERR_HIP, "the allocation-free workspace was not resolved on device %d",
                                   ctx->device);
    HIP_TRY(ctx, hipSetDevice(ctx->device));
    rs = check_coords(ctx, d_x, T);
    if (rs != MEMD_OK) return rs;
    std::vector<EpochIn> in{{(int)n, (int)T, (int)K, (int)N_imf, (int)S, d_y, d_V, d_IMFs,
                             opts ? opts->h_sift_counts : nullptr, opts ? opts->h_n_imf_out : nullptr}};
    RunArena where;
    where.fixed = ctx->scratch;
    where.fixed_bytes = kScratchBytes;
    return run_epochs(ctx, in, opts, elapsed_s, nullptr, where);
}

memd_status memd_run_f32(memd_ctx* ctx, size_t n, size_t T, size_t K, size_t N_imf, size_t S, const int32_t* d_x,
                         const float* d_y, const float* d_V, float* d_IMFs, const memd_opts* opts, double* elapsed_s) {
    if (!ctx) return fail(nullptr, MEMD_ERR_ARG, "null context");
    memd_status rs = validate_shape(ctx, n, T, K, N_imf, S);
    if (rs != MEMD_OK) return rs;
    if (!d_y || !d_V || !d_IMFs) return fail(ctx, MEMD_ERR_ARG, "null device pointer");
    HIP_TRY(ctx, hipSetDevice(ctx->device));
    size_t ny = n * T, nv = K * n, no = N_imf * n * T;
    Staging sc;
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
    if (!ctx) return fail(nullptr, MEMD_ERR_ARG, "null context");
    if (!d_data || !d_out) return fail(ctx, MEMD_ERR_ARG, "null device pointer");
    size_t T = 0;
    const double* dV = nullptr;
    memd_status rs = mex_common(ctx, n, K, N_imf, S, numel, d_index, d_primes, &T, &dV);
    if (rs != MEMD_OK) return rs;
    return memd_run_f64(ctx, n, T, K, N_imf, S, nullptr, d_data, dV, d_out, nullptr, nullptr);
}

memd_status memd_mex_compat_f32(memd_ctx* ctx, int32_t n, int32_t K, int32_t N_imf, int32_t S, size_t numel,
                                const float* d_data, const int32_t* d_index, const int32_t* d_primes, float* d_out) {
    if (!ctx) return fail(nullptr, MEMD_ERR_ARG, "null context");
    if (!d_data || !d_out) return fail(ctx, MEMD_ERR_ARG, "null device pointer");
    size_t T = 0;
    const double* dV = nullptr;
    memd_status rs = mex_common(ctx, n, K, N_imf, S, numel, d_index, d_primes, &T, &dV);
    if (rs != MEMD_OK) return rs;
    size_t ny = (size_t)n * T, no = (size_t)N_imf * n * T;
    Staging sc;
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
    if (!ctx) return fail(nullptr, MEMD_ERR_ARG, "null context");
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

memd_status memd_ctx_set_tuning(memd_ctx* ctx, const memd_tuning* t) {
    if (!ctx) return fail(nullptr, MEMD_ERR_ARG, "null context");
    if (!t) return fail(ctx, MEMD_ERR_ARG, "null tuning");
    if (t->n < 1 || t->n > kMaxChannels) return fail(ctx, MEMD_ERR_ARG, "tuning: n=%d out of range", t->n);
    if (t->block_groups < 1 || t->block_groups > kBlockBpwMax || block_bpw_fit(t->n, t->block_groups) != t->block_groups)
        return fail(ctx, MEMD_ERR_ARG, "tuning: block_groups=%d does not fit %d channels (1..%d)", t->block_groups, t->n,
                    block_bpw_fit(t->n, kBlockBpwMax));
    if (t->eval_tile_words != 1 && t->eval_tile_words != 2 && t->eval_tile_words != 4)
        return fail(ctx, MEMD_ERR_ARG, "tuning: eval_tile_words=%d (1, 2 or 4)", t->eval_tile_words);
    ctx->tuning[t->n] = *t;
    return MEMD_OK;
}

memd_status memd_ctx_get_tuning(memd_ctx* ctx, int32_t n, memd_tuning* out) {
    if (!ctx) return fail(nullptr, MEMD_ERR_ARG, "null context");
    if (!out) return fail(ctx, MEMD_ERR_ARG, "null output");
    if (n < 1 || n > kMaxChannels) return fail(ctx, MEMD_ERR_ARG, "n=%d out of range", n);
    *out = tuning_for(ctx, n);
    return MEMD_OK;
}

// Times every candidate shape on a synthetic batch of (n, T, K) epochs, sized to
// fill the device (about 5.4e8 n*T*K per step, at most 256 epochs), with one IMF
// of 4 sifts under the profiling path.  The block solve's time depends only on
// block_groups and the evaluation's only on eval_tile_words, so the candidates
// are run in pairs and each knob keeps its fastest value.
memd_status memd_autotune(memd_ctx* ctx, size_t n, size_t T, size_t K, memd_tuning* out) {
    if (!ctx) return fail(nullptr, MEMD_ERR_ARG, "null context");
    memd_status rs = validate_shape(ctx, n, T, K, 2, 4);
    if (rs != MEMD_OK) return rs;
    HIP_TRY(ctx, hipSetDevice(ctx->device));
    const size_t work = n * T * K;
    size_t count = (size_t)5.4e8 / (work ? work : 1);
    count = count < 1 ? 1 : (count > 256 ? 256 : count);
    // synthetic channels: three sinusoids per channel plus uniform noise (fixed LCG)
    std::vector<double> hy(n * T);
    uint64_t lcg = 0x9e3779b97f4a7c15ull;
    for (size_t c = 0; c < n; ++c)
        for (size_t t = 0; t < T; ++t) {
            lcg = lcg * 6364136223846793005ull + 1442695040888963407ull;
            const double u = (double)(lcg >> 11) * (1.0 / 9007199254740992.0) - 0.5;
            hy[c * T + t] = std::sin(0.031 * (double)t + 0.7 * (double)c) + 0.5 * std::sin(0.173 * (double)t + (double)c) +
                            0.25 * std::sin(0.61 * (double)t) + 0.2 * u;
        }
    Staging y, V, imfs;
    hipError_t e = hipMalloc((void**)&y.p, n * T * sizeof(double));
    if (e == hipSuccess) e = hipMalloc((void**)&V.p, K * n * sizeof(double));
    if (e == hipSuccess) e = hipMalloc((void**)&imfs.p, count * 2 * n * T * sizeof(double));
    if (e != hipSuccess) return fail(ctx, MEMD_ERR_NOMEM, "autotune buffers: %s", hipGetErrorString(e));
    HIP_TRY(ctx, hipMemcpy(y.p, hy.data(), n * T * sizeof(double), hipMemcpyHostToDevice));
    rs = memd_directions_f64(ctx, n, K, nullptr, V.p);
    if (rs != MEMD_OK) return rs;
    std::vector<EpochIn> in(count);
    for (size_t i = 0; i < count; ++i)
        in[i] = {(int)n, (int)T, (int)K, 2, 4, y.p, V.p, imfs.p + i * 2 * n * T, nullptr, nullptr};
    std::vector<int> bg, tw = {1, 2, 4};
    for (int b : {1, 2, 4, 8})
        if (block_bpw_fit((int)n, b) == b) bg.push_back(b);
    const bool had = ctx->tuning.count((int)n) != 0;
    const memd_tuning saved = tuning_for(ctx, (int)n);
    const memd_tuning def = default_tuning((int)n);
    std::vector<double> bms(bg.size(), 1e300), ems(tw.size(), 1e300);
    const size_t rounds = bg.size() > tw.size() ? bg.size() : tw.size();
    memd_opts o;
    memd_opts_default(&o);
    for (size_t r = 0; r <= rounds; ++r) {   // timing pass 0 warms up (plan region, caches)
        const size_t i = r == 0 ? 0 : r - 1;
        const size_t ib = i < bg.size() ? i : bg.size() - 1, it = i < tw.size() ? i : tw.size() - 1;
        memd_tuning cand = def;
        cand.block_groups = bg[ib];
        cand.eval_tile_words = tw[it];
        ctx->tuning[(int)n] = cand;
        memd_kernel_profile prof;
        rs = run_epochs(ctx, in, &o, nullptr, &prof);
        if (rs != MEMD_OK) break;
        if (r == 0) continue;
        bms[ib] = std::min(bms[ib], prof.ms_per_launch[3]);
        ems[it] = std::min(ems[it], prof.ms_per_launch[6]);
    }
    if (had) ctx->tuning[(int)n] = saved;
    else ctx->tuning.erase((int)n);
    if (rs != MEMD_OK) return rs;
    // fastest value of each knob; the built-in one unless another is faster by
    // more than 3% (timing noise does not flip the choice)
    auto pick = [](const std::vector<int>& vals, const std::vector<double>& t, int dflt, double* tbest) {
        size_t b = 0, d = 0;
        for (size_t i = 0; i < vals.size(); ++i) {
            if (t[i] < t[b]) b = i;
            if (vals[i] == dflt) d = i;
        }
        if (vals[d] == dflt && t[b] > 0.97 * t[d]) b = d;
        *tbest = t[b];
        return vals[b];
    };
    memd_tuning best = def;
    best.block_groups = pick(bg, bms, def.block_groups, &best.block_ms);
    best.eval_tile_words = pick(tw, ems, def.eval_tile_words, &best.eval_ms);
    ctx->tuning[(int)n] = best;
    if (out) *out = best;
    return MEMD_OK;
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

// ======================================================================= server
//
// Served continuous batching.  Workspaces live in one pool allocated at
// creation; the plan region (descriptors, tables, lists) is rebuilt whenever
// the active set changes.  Continuing epochs keep their workspace (and thus
// their state) across rebuilds; joining epochs are initialised through a
// table that covers only them.

struct memd_server {
    struct Req {
        EpochIn in{};
        EpochLayout lay{};
        int state = MEMD_REQ_QUEUED;
        size_t off = 0;
        int grp = 0;
        long weight = 0;      // n * T * K, the XCD load estimate
        long skipped = 0;     // admission rounds this request was passed over
    };
    memd_ctx* ctx = nullptr;
    memd_opts opts{};
    bool records = false;
    char* pool = nullptr;
    size_t pool_bytes = 0, pool_used = 0;
    std::map<size_t, size_t> free_blocks;   // offset -> bytes, coalesced
    char* region = nullptr;
    size_t region_bytes = 0;
    int max_active = 0;
    std::map<int64_t, Req> reqs;
    std::vector<int64_t> queue;             // FIFO
    std::vector<int64_t> active;
    std::vector<int64_t> finished;          // completion order, popped by memd_server_completed
    size_t finished_head = 0;
    int64_t next_id = 1;
    long load[kXcds] = {};
    // current plan and its host staging (kept alive across asynchronous copies)
    std::vector<EpochDesc> eds;
    std::vector<int> grp;
    std::vector<int2> tabs, fresh;
    std::vector<EpochState> states;
    PlanLayout P;
    StepPlan plan{};
};

namespace {

// A queued request passed over this many admission rounds blocks backfill
// behind it until it fits.
constexpr long kMaxSkip = 8;

bool pool_alloc(memd_server* S, size_t bytes, size_t* off) {
    for (auto it = S->free_blocks.begin(); it != S->free_blocks.end(); ++it) {
        if (it->second < bytes) continue;
        *off = it->first;
        size_t rem = it->second - bytes, at = it->first + bytes;
        S->free_blocks.erase(it);
        if (rem) S->free_blocks[at] = rem;
        S->pool_used += bytes;
        return true;
    }
    return false;
}

void pool_free(memd_server* S, size_t off, size_t bytes) {
    S->pool_used -= bytes;
    auto it = S->free_blocks.emplace(off, bytes).first;
    auto nx = std::next(it);
    if (nx != S->free_blocks.end() && it->first + it->second == nx->first) {
        it->second += nx->second;
        S->free_blocks.erase(nx);
    }
    if (it != S->free_blocks.begin()) {
        auto pv = std::prev(it);
        if (pv->first + pv->second == it->first) {
            pv->second += it->second;
            S->free_blocks.erase(it);
        }
    }
}

// Admit queued requests that fit, in FIFO order; returns the admitted ids.
void server_admit(memd_server* S, std::vector<int64_t>& joined) {
    std::vector<int64_t> keep;
    bool blocked = false;
    for (int64_t id : S->queue) {
        memd_server::Req& r = S->reqs[id];
        size_t off = 0;
        bool room = !(S->max_active > 0 && (int)S->active.size() >= S->max_active);
        if (!blocked && room && pool_alloc(S, r.lay.total, &off)) {
            r.off = off;
            r.state = MEMD_REQ_RUNNING;
            int g = 0;
            for (int x = 1; x < kXcds; ++x)
                if (S->load[x] < S->load[g]) g = x;
            r.grp = g;
            S->load[g] += r.weight;
            S->active.push_back(id);
            joined.push_back(id);
            continue;
        }
        if (!blocked && room && ++r.skipped > kMaxSkip) blocked = true;
        keep.push_back(id);
    }
    S->queue.swap(keep);
}

// Rebuild the plan over the active set and initialise the joining epochs.
memd_status server_rebuild(memd_server* S, const std::vector<int64_t>& joined, hipStream_t s) {
    memd_ctx* ctx = S->ctx;
    const int count = (int)S->active.size();
    S->eds.resize(count);
    S->grp.resize(count);
    const int groups = count >= kXcds ? kXcds : 1;
    std::map<int64_t, int> index;
    for (int e = 0; e < count; ++e) {
        memd_server::Req& r = S->reqs[S->active[e]];
        S->eds[e] = desc_geometry(ctx, r.in, S->opts, S->records);
        bind_workspace(S->eds[e], S->pool + r.off, r.lay);
        S->grp[e] = groups > 1 ? r.grp : 0;
        index[S->active[e]] = e;
    }
    memd_status rs = plan_layout(ctx, S->eds, S->grp, groups, 0, &S->P);
    if (rs != MEMD_OK) return rs;
    if (S->P.total > S->region_bytes) {
        HIP_TRY(ctx, hipStreamSynchronize(s));
        if (S->region) HIP_TRY(ctx, hipFree(S->region));
        S->region = nullptr;
        S->region_bytes = 0;
        size_t want = S->P.total + S->P.total / 4;
        hipError_t e = hipMalloc((void**)&S->region, want);
        if (e != hipSuccess) {
            S->region = nullptr;
            (void)hipGetLastError();
            return fail(ctx, MEMD_ERR_NOMEM, "plan region allocation of %zu bytes failed: %s", want,
                        hipGetErrorString(e));
        }
        S->region_bytes = want;
    }
    int* ndone = (int*)(S->region + S->P.ndone_at);
    for (EpochDesc& d : S->eds) d.ndone = ndone;
    rs = write_plan(ctx, S->eds, S->grp, S->P, S->region, S->records ? kModeStop : kModeFix, S->records, S->tabs, s,
                    &S->plan);
    if (rs != MEMD_OK) return rs;
    // counters are zero between steps; a new region starts from zero too
    HIP_TRY(ctx, hipMemsetAsync(S->region + S->P.cnt_at, 0, kCountBytes, s));
    HIP_TRY(ctx, hipMemsetAsync(S->region + S->P.bcnt_at, 0, kBcountStride * S->eds.size(), s));
    HIP_TRY(ctx, hipMemsetAsync(ndone, 0, sizeof(int), s));
    S->fresh.clear();
    for (int64_t id : joined) {
        int e = index[id];
        int nb = blocks_for(kMapInit, S->eds[e]);
        for (int b = 0; b < nb; ++b) S->fresh.push_back(make_int2(e, b));
    }
    if (!S->fresh.empty()) {
        int2* ft = (int2*)(S->region + S->P.fresh_at);
        HIP_TRY(ctx, hipMemcpyAsync(ft, S->fresh.data(), sizeof(int2) * S->fresh.size(), hipMemcpyHostToDevice, s));
        HIP_TRY(ctx, launch_init_table(S->plan, ft, (int)S->fresh.size(), s));
    }
    return MEMD_OK;
}

// Snapshot the states and retire every finished epoch.
memd_status server_retire(memd_server* S, hipStream_t s) {
    memd_ctx* ctx = S->ctx;
    const int count = (int)S->active.size();
    S->states.resize(count);
    EpochState* dst = (EpochState*)(S->region + S->P.states_at);
    HIP_TRY(ctx, launch_gather_states(S->plan, dst, s));
    HIP_TRY(ctx, hipMemcpyAsync(S->states.data(), dst, sizeof(EpochState) * count, hipMemcpyDeviceToHost, s));
    HIP_TRY(ctx, hipStreamSynchronize(s));
    std::vector<int64_t> still;
    for (int e = 0; e < count; ++e) {
        int64_t id = S->active[e];
        memd_server::Req& r = S->reqs[id];
        if (!S->states[e].done) {
            still.push_back(id);
            continue;
        }
        if (r.in.h_sift_counts)
            HIP_TRY(ctx, hipMemcpy(r.in.h_sift_counts, S->eds[e].sift_counts, sizeof(int32_t) * r.in.N_imf,
                                   hipMemcpyDeviceToHost));
        if (r.in.h_n_imf_out) *r.in.h_n_imf_out = S->states[e].n_imf_out;
        r.state = MEMD_REQ_DONE;
        pool_free(S, r.off, r.lay.total);
        S->load[r.grp] -= r.weight;
        S->finished.push_back(id);
    }
    S->active.swap(still);
    return MEMD_OK;
}

}  // namespace

extern "C" {

memd_status memd_server_create(memd_ctx* ctx, size_t hbm_pool_bytes, int32_t max_active, const memd_opts* opts,
                               memd_server** out) {
    if (!ctx) return fail(nullptr, MEMD_ERR_ARG, "null context");
    if (!out) return fail(ctx, MEMD_ERR_ARG, "null output pointer");
    *out = nullptr;
    memd_opts o;
    if (opts) o = *opts; else memd_opts_default(&o);
    memd_status rs = check_opts(ctx, o);
    if (rs != MEMD_OK) return rs;
    if (o.d_knot_counts || o.h_sift_counts || o.h_n_imf_out)
        return fail(ctx, MEMD_ERR_ARG, "per-call diagnostics are not available on a server (use the request's outputs)");
    if (hbm_pool_bytes < kAlign || max_active < 0) return fail(ctx, MEMD_ERR_ARG, "invalid pool size or active cap");
    HIP_TRY(ctx, hipSetDevice(ctx->device));
    memd_server* S = new memd_server();
    S->ctx = ctx;
    S->opts = o;
    S->records = record_path(o);
    S->max_active = max_active;
    S->pool_bytes = hbm_pool_bytes / kAlign * kAlign;
    hipError_t e = hipMalloc((void**)&S->pool, S->pool_bytes);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        delete S;
        return fail(ctx, MEMD_ERR_NOMEM, "server pool allocation of %zu bytes failed: %s", hbm_pool_bytes,
                    hipGetErrorString(e));
    }
    S->free_blocks[0] = S->pool_bytes;
    *out = S;
    return MEMD_OK;
}

memd_status memd_server_destroy(memd_server* S) {
    if (!S) return MEMD_OK;
    (void)hipSetDevice(S->ctx->device);
    (void)hipStreamSynchronize(work_stream(S->ctx));
    if (S->pool) (void)hipFree(S->pool);
    if (S->region) (void)hipFree(S->region);
    delete S;
    return MEMD_OK;
}

memd_status memd_server_submit(memd_server* S, const memd_epoch* ep, int64_t* id) {
    if (!S) return fail(nullptr, MEMD_ERR_ARG, "null server");
    memd_ctx* ctx = S->ctx;
    if (!ep || !id) return fail(ctx, MEMD_ERR_ARG, "null epoch or id pointer");
    memd_status rs = validate_shape(ctx, ep->n, ep->T, ep->K, ep->N_imf, ep->S);
    if (rs != MEMD_OK) return rs;
    if (!ep->d_y || !ep->d_V || !ep->d_IMFs) return fail(ctx, MEMD_ERR_ARG, "null device pointer");
    if ((rs = check_record_channels(ctx, S->opts, (int)ep->n)) != MEMD_OK) return rs;
    memd_server::Req r;
    r.in = {(int)ep->n, (int)ep->T, (int)ep->K, (int)ep->N_imf, (int)ep->S, ep->d_y, ep->d_V, ep->d_IMFs,
            ep->h_sift_counts, ep->h_n_imf_out};
    r.lay = layout_for(r.in.n, r.in.T, r.in.K, r.in.N_imf, S->records);
    if (r.lay.total > S->pool_bytes)
        return fail(ctx, MEMD_ERR_ADMISSION, "epoch needs %zu workspace bytes, the pool holds %zu", r.lay.total,
                    S->pool_bytes);
    r.weight = (long)r.in.n * r.in.T * r.in.K;
    *id = S->next_id++;
    S->reqs[*id] = r;
    S->queue.push_back(*id);
    return MEMD_OK;
}

memd_status memd_server_step(memd_server* S, int32_t steps, int32_t* n_active, int32_t* n_queued) {
    if (!S) return fail(nullptr, MEMD_ERR_ARG, "null server");
    memd_ctx* ctx = S->ctx;
    if (steps < 0) return fail(ctx, MEMD_ERR_ARG, "negative step count");
    HIP_TRY(ctx, hipSetDevice(ctx->device));
    hipStream_t s = work_stream(ctx);
    std::vector<int64_t> joined;
    const size_t before = S->active.size();
    server_admit(S, joined);
    // the plan is rebuilt when the set changed: epochs joined now, or retired in
    // the previous call (the active list then shrank below the plan's count)
    if (!joined.empty() || before != (size_t)S->plan.count) {
        memd_status rs = server_rebuild(S, joined, s);
        if (rs != MEMD_OK) return rs;
    }
    if (!S->active.empty()) {
        for (int i = 0; i < steps; ++i) HIP_TRY(ctx, launch_step(S->plan, s));
        memd_status rs = server_retire(S, s);
        if (rs != MEMD_OK) return rs;
    }
    if (n_active) *n_active = (int32_t)S->active.size();
    if (n_queued) *n_queued = (int32_t)S->queue.size();
    return MEMD_OK;
}

memd_status memd_server_completed(memd_server* S, int64_t* ids, int32_t cap, int32_t* n) {
    if (!S) return fail(nullptr, MEMD_ERR_ARG, "null server");
    if (!n || (cap > 0 && !ids)) return fail(S->ctx, MEMD_ERR_ARG, "null output pointer");
    int32_t k = 0;
    while (k < cap && S->finished_head < S->finished.size()) ids[k++] = S->finished[S->finished_head++];
    if (S->finished_head == S->finished.size()) {
        S->finished.clear();
        S->finished_head = 0;
    }
    *n = k;
    return MEMD_OK;
}

memd_status memd_server_poll(memd_server* S, int64_t id, int32_t* state) {
    if (!S) return fail(nullptr, MEMD_ERR_ARG, "null server");
    if (!state) return fail(S->ctx, MEMD_ERR_ARG, "null state pointer");
    auto it = S->reqs.find(id);
    if (it == S->reqs.end()) return fail(S->ctx, MEMD_ERR_NOT_FOUND, "unknown request id %lld", (long long)id);
    *state = it->second.state;
    return MEMD_OK;
}

memd_status memd_server_release(memd_server* S, int64_t id) {
    if (!S) return fail(nullptr, MEMD_ERR_ARG, "null server");
    auto it = S->reqs.find(id);
    if (it == S->reqs.end()) return fail(S->ctx, MEMD_ERR_NOT_FOUND, "unknown request id %lld", (long long)id);
    if (it->second.state != MEMD_REQ_DONE)
        return fail(S->ctx, MEMD_ERR_ARG, "request %lld is not finished", (long long)id);
    S->reqs.erase(it);
    return MEMD_OK;
}

size_t memd_server_pool_used(memd_server* S) { return S ? S->pool_used : 0; }

}  // extern "C"
