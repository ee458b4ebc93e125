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
#include <chrono>
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
    // direction sets by (n, K, primes), least recently used evicted beyond dircache_cap bytes
    struct DirEntry {
        double* p = nullptr;
        size_t bytes = 0;
        unsigned long long used = 0;
    };
    std::map<std::string, DirEntry> dircache;
    size_t dircache_bytes = 0, dircache_cap = (size_t)256 << 20;
    unsigned long long dircache_clock = 0;
    int fail_allocs = 0;            // testing hook: the next fail_allocs device allocations fail
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

// Every device allocation the library makes for a context goes through here, so
// that memd_debug_fail_allocations can make the next ones fail as out of memory.
hipError_t dev_alloc(memd_ctx* ctx, void** p, size_t bytes) {
    if (ctx && ctx->fail_allocs > 0) {
        --ctx->fail_allocs;
        *p = nullptr;
        return hipErrorOutOfMemory;
    }
    return hipMalloc(p, bytes);
}

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
    size_t running, vt, curt, mask, pre, E, mir, X, fac, redsys, ysol, coef, sep, sepin, mean, ampb, tstat, red, sift, st,
        total;
};

// Kernel path of a call: stop mode and mirrored ends run the stop-mode kernels
// (decision and apply); mirrored ends evaluate from the spline record array (the
// record path), the reference's ends with the fused evaluation, which in stop
// mode leaves per-block amplitude sums for k_stop_amp.
struct PathKind {
    bool stop_kernels = false;
    bool records = false;
    bool ampb = false;
};
PathKind path_for(const memd_opts& o) {
    PathKind k;
    k.records = o.boundary == MEMD_BOUNDARY_MIRROR;
    k.stop_kernels = o.mode == MEMD_MODE_STOP || k.records;
    k.ampb = o.mode == MEMD_MODE_STOP && !k.records;
    return k;
}

EpochLayout layout_for(int n, int T, int K, int N_imf, const PathKind& pk) {
    const bool records = pk.records;
    EpochLayout L{};
    size_t W = (size_t)(T + 63) / 64;
    size_t Ecap = (size_t)knot_capacity(T);
    size_t nb = (size_t)solver_block_capacity((int)Ecap);
    size_t off = 0;
    auto take = [&](size_t bytes) { size_t at = off; off += align_up(bytes); return at; };
    L.running = take((size_t)n * T * 8);
    L.vt = take((size_t)n * dir_pad(K) * 8);
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
    L.sepin = take((size_t)K * 2 * nb * np * 8);
    L.mean = take((size_t)n * T * 8);
    L.ampb = take(pk.ampb ? (size_t)K * ((n + 15) / 16) * T * 8 : 0);
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
    hipError_t e = dev_alloc(ctx, (void**)&ctx->arena, want);
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

// k_block workgroups an epoch's items are drained by: enough that the whole
// launch holds about kBlockWgTarget workgroups (4 per CU) when the batch is small,
// at most 128 per epoch in larger batches (an XCD then works through one epoch's
// items at a time, so its signal stays in that XCD's L2, and a workgroup drains
// several items, each loaded while the previous one is solved; 256 per epoch left
// half of them without an item at the sparse IMFs: 277 -> 266 us per launch at
// 32 ch x 4096, K=64, 64 epochs).
constexpr int kBlockWgTarget = 1024;

// Descriptor fields that depend only on the epoch's shape, the options and the
// number of epochs in the launch.
EpochDesc desc_geometry(const memd_ctx* ctx, const EpochIn& I, const memd_opts& o, const PathKind& pk, int batch) {
    const bool records = pk.records;
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
    d.etw = pk.ampb ? 1 : tu.eval_tile_words;   // stop mode's block sums take 16-channel tiles
    // k_block items: at most 2K * groups; workgroups per epoch: the item count at a
    // typical knot density (one knot per 16 samples), at most the cap above
    d.bcap = 2 * I.K * ((d.nbmax + d.bpw - 1) / d.bpw);
    const long typical = 2L * I.K * (((long)I.T / 512 + 1 + d.bpw - 1) / d.bpw);
    const long cap = std::max(128L, (long)kBlockWgTarget / std::max(1, batch));
    d.bblocks = (int)std::min(cap, std::max(1L, typical));
    return d;
}

// Point a descriptor at its workspace (layout L starting at b).
void bind_workspace(EpochDesc& d, char* b, const EpochLayout& L) {
    d.running = (double*)(b + L.running);
    d.vt = (double*)(b + L.vt);
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
    d.ampb = (double*)(b + L.ampb);
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
    size_t sift_at = 0, siftoff_at = 0;   // server probes: every epoch's sift counts and their offsets
    int sift_total = 0;
    size_t bitem_at = 0, bcnt_at = 0;   // per-epoch k_block items (int4) and counters (one line each)
    int nblk[kNumMaps] = {};
    long lcap[kNumLists] = {};
    long gcap[kNumLists][kXcds] = {};
    int groups = 1;
};

// grp[e] in [0, groups) is the XCD group of epoch e (groups == 1: no affinity), -1
// an epoch spread over all XCDs.
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
            if (grp[e] >= 0) {
                P->gcap[l][grp[e]] += c;
                P->lcap[l] += c;
            } else {   // spread: any XCD's sub-list may receive its items, and the list holds them all
                for (int x = 0; x < groups; ++x) P->gcap[l][x] += c;
                P->lcap[l] += c * groups;
            }
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
    total += align_up(sizeof(unsigned long long) * 4 * kStatSlots);
    P->ndone_at = total;
    total += align_up(sizeof(int));
    P->fresh_at = total;
    total += align_up(sizeof(int2) * (size_t)P->nblk[kMapInit]);
    P->states_at = total;
    total += align_up(sizeof(EpochState) * (size_t)count);
    for (int e = 0; e < count; ++e) P->sift_total += eds[e].N_imf;
    P->sift_at = total;
    total += align_up(sizeof(int) * (size_t)P->sift_total);
    P->siftoff_at = total;
    total += align_up(sizeof(int) * (size_t)count);
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
            for (int b = 0; b < nb; ++b) q[grp[e] >= 0 ? grp[e] : b % groups].push_back(make_int2(e, b));
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
        if (at != P.lcap[l])   // the sub-lists must tile the list's storage exactly
            return fail(ctx, MEMD_ERR_HIP, "work list %d: sub-lists hold %ld items, storage %ld", l, at, P.lcap[l]);
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
    plan.factor_wide = P.lcap[kListFactor] <= kFactorWideItems ? 1 : 0;
    *out = plan;
    return MEMD_OK;
}

// Packing of a batch onto the XCDs: every sift step runs all epochs, so a step
// lasts as long as the most loaded XCD.  An epoch with more than an XCD's fair
// share of the batch's work (n * T * K) is spread over all XCDs (grp -1: its
// blocks and work items go to every XCD), since pinned it alone would set the
// step time.  The rest, longest-processing-time first (by decreasing n * T * K,
// ties in batch order), go to the least loaded XCD.  Placement never changes an
// epoch's arithmetic.
void xcd_pack(const std::vector<EpochIn>& in, std::vector<int>& grp) {
    const int count = (int)in.size();
    std::vector<int> order(count);
    std::vector<long> cost(count);
    long total = 0;
    for (int e = 0; e < count; ++e) {
        order[e] = e;
        cost[e] = (long)in[e].n * (long)in[e].T * (long)in[e].K;
        total += cost[e];
    }
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return cost[a] > cost[b]; });
    long load[kXcds] = {};
    for (int e : order) {
        if (cost[e] * kXcds > total) {
            grp[e] = -1;
            for (int x = 0; x < kXcds; ++x) load[x] += cost[e] / kXcds;
            continue;
        }
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
    add((unsigned long long)p.factor_wide);
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
// channels; at most kStopMaxChannels channels) serve mirrored ends in both modes.
memd_status check_record_channels(memd_ctx* ctx, const memd_opts& o, int n) {
    if (path_for(o).records && n > kStopMaxChannels)
        return fail(ctx, MEMD_ERR_UNSUPPORTED, "mirrored ends support at most %d channels (got %d)", kStopMaxChannels, n);
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
    const PathKind pk = path_for(o);
    const bool records = pk.records;
    std::vector<EpochLayout> lay(count);
    std::vector<size_t> base(count);
    std::vector<EpochDesc> eds(count);
    size_t total = 0;
    for (int e = 0; e < count; ++e) {
        lay[e] = layout_for(in[e].n, in[e].T, in[e].K, in[e].N_imf, pk);
        base[e] = total;
        total += lay[e].total;
        eds[e] = desc_geometry(ctx, in[e], o, pk, count);
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
    // each epoch's first entry in the gathered sift counts (read back after the steps)
    std::vector<int> sift_off(count);
    for (int e = 0, at = 0; e < count; ++e) {
        sift_off[e] = at;
        at += in[e].N_imf;
    }
    HIP_TRY(ctx, hipMemcpyAsync(arena + P.siftoff_at, sift_off.data(), sizeof(int) * count, hipMemcpyHostToDevice, s));
    StepPlan plan{};
    std::vector<int2> tabs;
    rs = write_plan(ctx, eds, grp, P, arena, pk.stop_kernels ? kModeStop : kModeFix, records, tabs, s, &plan);
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
        HIP_TRY(ctx, hipMemsetAsync(arena + stats_at, 0, sizeof(unsigned long long) * 4 * kStatSlots, s));
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
        std::vector<unsigned long long> slots(4 * (size_t)kStatSlots);
        HIP_TRY(ctx, hipMemcpy(slots.data(), arena + stats_at, sizeof(unsigned long long) * slots.size(),
                               hipMemcpyDeviceToHost));
        unsigned long long st4[4] = {0, 0, 0, 0};
        for (int sl = 0; sl < kStatSlots; ++sl)
            for (int i = 0; i < 4; ++i) st4[i] += slots[4 * (size_t)sl + i];
        prof->knots = (int64_t)st4[0];
        prof->knot_channels = (int64_t)st4[1];
        prof->eval_points = (int64_t)st4[2];
        prof->samples = (int64_t)st4[3];
    } else {
    HIP_TRY(ctx, launch_init(plan, s));
    HIP_TRY(ctx, hipEventRecord(ctx->t0, s));
    const bool graph = o.use_graph && steps > 1 && !where.fixed;
    // fix mode: the graph holds G consecutive steps (the largest divisor of the step
    // count up to 16), so replays, and the gaps between them, are G times fewer;
    // stop mode polls between chunks and captures one step
    int G = 1;
    if (o.mode != MEMD_MODE_STOP)
        for (int d = 16; d > 1; --d)
            if (steps % d == 0) {
                G = d;
                break;
            }
    if (graph) {
        // capture G steps on the owned stream.  The graph bakes in every launch
        // argument, all of which come from the plan (descriptor array, block tables,
        // work lists and their counters, grids, LDS sizes): it is replayed only for
        // a plan equal to the captured one in every field, and the same G.
        const std::string key = plan_key(plan) + "G" + std::to_string(G);
        if (!ctx->graph_exec || ctx->graph_key != key) {
            if (ctx->graph_exec) {
                HIP_TRY(ctx, hipGraphExecDestroy(ctx->graph_exec));
                ctx->graph_exec = nullptr;
            }
            hipGraph_t g = nullptr;
            HIP_TRY(ctx, hipStreamBeginCapture(ctx->own, hipStreamCaptureModeThreadLocal));
            hipError_t le = hipSuccess;
            for (int i = 0; i < G && le == hipSuccess; ++i) le = launch_step(plan, ctx->own);
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
                for (long i = 0; i < nsteps; i += G) HIP_TRY(ctx, hipGraphLaunch(ctx->graph_exec, ctx->own));
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

    // results: every epoch's state and sift counts gathered into the plan region by
    // one kernel, then two copies (not two per epoch)
    std::vector<EpochState> sts(count);
    std::vector<int> sift_all((size_t)P.sift_total);
    {
        EpochState* dstates = (EpochState*)(arena + P.states_at);
        int* dsift = (int*)(arena + P.sift_at);
        HIP_TRY(ctx, launch_gather_probe(plan, dstates, dsift, (const int*)(arena + P.siftoff_at), s));
        HIP_TRY(ctx, hipMemcpyAsync(sts.data(), dstates, sizeof(EpochState) * count, hipMemcpyDeviceToHost, s));
        if (P.sift_total > 0)
            HIP_TRY(ctx, hipMemcpyAsync(sift_all.data(), dsift, sizeof(int) * (size_t)P.sift_total,
                                        hipMemcpyDeviceToHost, s));
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
            for (int i = 0; i < in[e].N_imf; ++i) in[e].h_sift_counts[i] = sift_all[(size_t)sift_off[e] + i];
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

// Evict least recently used direction sets until `incoming` more bytes fit the
// cap (an entry larger than the cap alone is still cached).  Work queued on the
// context's stream may read an entry, so the stream drains first.
memd_status dircache_trim(memd_ctx* ctx, size_t incoming) {
    bool synced = false;
    while (!ctx->dircache.empty() && ctx->dircache_bytes + incoming > ctx->dircache_cap) {
        auto lru = ctx->dircache.begin();
        for (auto it = ctx->dircache.begin(); it != ctx->dircache.end(); ++it)
            if (it->second.used < lru->second.used) lru = it;
        if (!synced) {
            HIP_TRY(ctx, hipStreamSynchronize(ctx->user));
            synced = true;
        }
        HIP_TRY(ctx, hipFree(lru->second.p));
        ctx->dircache_bytes -= lru->second.bytes;
        ctx->dircache.erase(lru);
    }
    return MEMD_OK;
}

memd_status cached_directions(memd_ctx* ctx, int n, int K, const int32_t* P, const double** out) {
    std::string key = dir_key(n, K, P);
    auto it = ctx->dircache.find(key);
    if (it != ctx->dircache.end()) {
        it->second.used = ++ctx->dircache_clock;
        *out = it->second.p;
        return MEMD_OK;
    }
    std::vector<double> V((size_t)K * n);
    directions_host(n, K, P, V.data());
    const size_t bytes = V.size() * sizeof(double);
    memd_status rs = dircache_trim(ctx, bytes);
    if (rs != MEMD_OK) return rs;
    double* d = nullptr;
    hipError_t e = dev_alloc(ctx, (void**)&d, bytes);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        return fail(ctx, MEMD_ERR_NOMEM, "direction cache allocation failed: %s", hipGetErrorString(e));
    }
    e = hipMemcpy(d, V.data(), bytes, hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        (void)hipFree(d);
        return fail(ctx, MEMD_ERR_HIP, "direction upload failed: %s", hipGetErrorString(e));
    }
    memd_ctx::DirEntry& en = ctx->dircache[key];
    en.p = d;
    en.bytes = bytes;
    en.used = ++ctx->dircache_clock;
    ctx->dircache_bytes += bytes;
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
    return layout_for((int)n, (int)T, (int)K, 1, PathKind{}).total;
}

size_t memd_workspace_bytes_opts(size_t n, size_t T, size_t K, size_t N_imf, const memd_opts* opts) {
    if (n < 1 || T < 4 || K < 1 || N_imf < 1) return 0;
    return layout_for((int)n, (int)T, (int)K, (int)N_imf, opts ? path_for(*opts) : PathKind{}).total;
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
    for (auto& kv : ctx->dircache) (void)hipFree(kv.second.p);
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

memd_status memd_ctx_set_direction_cache(memd_ctx* ctx, size_t max_bytes) {
    if (!ctx) return fail(nullptr, MEMD_ERR_ARG, "null context");
    HIP_TRY(ctx, hipSetDevice(ctx->device));
    ctx->dircache_cap = max_bytes;
    return dircache_trim(ctx, 0);
}

size_t memd_ctx_direction_cache_bytes(memd_ctx* ctx) { return ctx ? ctx->dircache_bytes : 0; }

memd_status memd_debug_fail_allocations(memd_ctx* ctx, int32_t count) {
    if (!ctx) return fail(nullptr, MEMD_ERR_ARG, "null context");
    if (count < 0) return fail(ctx, MEMD_ERR_ARG, "negative count");
    ctx->fail_allocs = count;
    return MEMD_OK;
}

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
    hipError_t e = dev_alloc(ctx, (void**)&sc.p, (ny + nv + no) * sizeof(double));
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
    hipError_t e = dev_alloc(ctx, (void**)&sc.p, (ny + no) * sizeof(double));
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
// fill the device (about 5.4e8 n*T*K per step, at most 256 epochs), with one sift
// of each of the first 5 IMFs under the profiling path: the knot densities of a
// whole decomposition, from the first IMF's dense knots (where larger solver
// groups win) to the sparse later ones (where a small batch needs more, smaller
// work items).  The block solve's time depends only on block_groups and the
// evaluation's only on eval_tile_words, so the candidates are run in pairs and
// each knob keeps its fastest value.
memd_status memd_autotune(memd_ctx* ctx, size_t n, size_t T, size_t K, memd_tuning* out) {
    if (!ctx) return fail(nullptr, MEMD_ERR_ARG, "null context");
    constexpr int kTuneSlots = 6, kTuneSifts = 1;   // 5 IMFs of one sift, and the residual
    memd_status rs = validate_shape(ctx, n, T, K, kTuneSlots, kTuneSifts);
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
    hipError_t e = dev_alloc(ctx, (void**)&y.p, n * T * sizeof(double));
    if (e == hipSuccess) e = dev_alloc(ctx, (void**)&V.p, K * n * sizeof(double));
    if (e == hipSuccess) e = dev_alloc(ctx, (void**)&imfs.p, count * kTuneSlots * n * T * sizeof(double));
    if (e != hipSuccess) return fail(ctx, MEMD_ERR_NOMEM, "autotune buffers: %s", hipGetErrorString(e));
    HIP_TRY(ctx, hipMemcpy(y.p, hy.data(), n * T * sizeof(double), hipMemcpyHostToDevice));
    rs = memd_directions_f64(ctx, n, K, nullptr, V.p);
    if (rs != MEMD_OK) return rs;
    std::vector<EpochIn> in(count);
    for (size_t i = 0; i < count; ++i)
        in[i] = {(int)n, (int)T, (int)K, kTuneSlots, kTuneSifts, y.p, V.p, imfs.p + i * kTuneSlots * n * T,
                 nullptr, nullptr};
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
// creation.  The plan region (descriptors, tables, lists) is rebuilt when epochs
// join (or are withdrawn), and one sift step of the plan is captured as a graph
// that every later call replays.  A call does not wait for its own steps: it
// queues them, takes a probe (every epoch's state and sift counts, copied to
// pinned host memory, and an event) and then waits only for the previous
// call's probe, so the device always has the next steps queued while the host
// retires what the older probe shows finished.  Finished epochs leave the
// active set at once (their pool block can be reused by the next join) and stay
// in the plan, where every kernel skips them, until the next rebuild.
// Continuing epochs keep their workspace (and thus their state) across
// rebuilds; joining epochs are initialised through a table that covers only them.

struct memd_server {
    struct Req {
        EpochIn in{};
        EpochLayout lay{};
        int state = MEMD_REQ_QUEUED;
        size_t off = 0;
        int grp = 0;
        long weight = 0;      // n * T * K, the XCD load estimate
        long skipped = 0;     // admission rounds this request was passed over
        double deadline = 0;  // host clock (seconds), 0 = none
    };
    struct Probe {
        hipEvent_t ev = nullptr;
        EpochState* h_states = nullptr;   // pinned
        int* h_sift = nullptr;            // pinned
        size_t cap_states = 0, cap_sift = 0;
        std::vector<int64_t> ids;         // the plan's epochs, in plan order, when the probe was taken
        std::vector<int> sift_off;
        long seq = 0;
        bool pending = false;
    };
    memd_ctx* ctx = nullptr;
    memd_opts opts{};
    PathKind path;
    char* pool = nullptr;
    size_t pool_bytes = 0, pool_used = 0;
    std::map<size_t, size_t> free_blocks;   // offset -> bytes, coalesced
    char* region = nullptr;
    size_t region_bytes = 0;
    int max_active = 0;
    std::map<int64_t, Req> reqs;
    std::vector<int64_t> queue;             // FIFO
    std::vector<int64_t> active;            // running epochs (a subset of plan_ids)
    std::vector<int64_t> finished;          // completion order, popped by memd_server_completed
    size_t finished_head = 0;
    int64_t next_id = 1;
    long load[kXcds] = {};
    // current plan, its graph and its host staging (kept alive across asynchronous copies)
    std::vector<EpochDesc> eds;
    std::vector<int> grp;
    std::vector<int2> tabs, fresh;
    std::vector<int64_t> plan_ids;
    std::vector<int> sift_off;
    PlanLayout P;
    StepPlan plan{};
    bool dirty = false;                     // rebuild before the next launch
    hipGraphExec_t graph = nullptr;
    std::vector<std::pair<long, hipGraphExec_t>> old_graphs;   // destroyed once probe seq completes
    hipEvent_t staged = nullptr;            // the last rebuild's host-to-device copies
    hipStream_t cap = nullptr;              // graph capture (replays run on the context's own stream)
    Probe probe[2];
    int probe_slot = 0;
    long probe_seq = 0, probe_done = 0;
    memd_server_info info{};
};

namespace {

// A queued request passed over this many admission rounds blocks backfill
// behind it until it fits.
constexpr long kMaxSkip = 8;

double host_seconds() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

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

// Epochs of at least this much work per step (n * T * K) are spread over all
// XCDs by the server: pinned to one, a long epoch would set every step's length
// for the short ones beside it.
constexpr long kSpreadWeight = 1L << 26;

// XCD placement of an admitted request (least loaded XCD, or all of them), and
// its removal (sign -1).
void server_place(memd_server* S, memd_server::Req& r, int sign) {
    if (sign > 0) {
        if (r.weight >= kSpreadWeight) {
            r.grp = -1;
        } else {
            int g = 0;
            for (int x = 1; x < kXcds; ++x)
                if (S->load[x] < S->load[g]) g = x;
            r.grp = g;
        }
    }
    if (r.grp < 0) {
        for (int x = 0; x < kXcds; ++x) S->load[x] += sign * (r.weight / kXcds);
    } else {
        S->load[r.grp] += sign * r.weight;
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
            server_place(S, r, +1);
            S->active.push_back(id);
            joined.push_back(id);
            continue;
        }
        if (!blocked && room && ++r.skipped > kMaxSkip) blocked = true;
        keep.push_back(id);
    }
    S->queue.swap(keep);
}

// Undo an admission whose plan could not be built: the requests go back to the
// front of the queue, their pool blocks are returned.
void server_unadmit(memd_server* S, const std::vector<int64_t>& joined) {
    std::vector<int64_t> q(joined.begin(), joined.end());
    for (int64_t id : joined) {
        memd_server::Req& r = S->reqs[id];
        pool_free(S, r.off, r.lay.total);
        server_place(S, r, -1);
        r.state = MEMD_REQ_QUEUED;
        S->active.erase(std::find(S->active.begin(), S->active.end(), id));
    }
    q.insert(q.end(), S->queue.begin(), S->queue.end());
    S->queue.swap(q);
}

// Destroy replaced graphs whose last launch precedes a completed probe.
void server_drop_graphs(memd_server* S, bool all) {
    std::vector<std::pair<long, hipGraphExec_t>> keep;
    for (auto& g : S->old_graphs) {
        if (all || g.first <= S->probe_done) (void)hipGraphExecDestroy(g.second);
        else keep.push_back(g);
    }
    S->old_graphs.swap(keep);
}

// Rebuild the plan over the active set, initialise the joining epochs and
// capture the step graph.
memd_status server_rebuild(memd_server* S, const std::vector<int64_t>& joined, hipStream_t s) {
    memd_ctx* ctx = S->ctx;
    const double t0 = host_seconds();
    const int count = (int)S->active.size();
    if (S->staged) HIP_TRY(ctx, hipEventSynchronize(S->staged));   // the previous plan's staging is free
    S->eds.resize(count);
    S->grp.resize(count);
    S->plan_ids = S->active;
    const int groups = count >= kXcds ? kXcds : 1;
    std::map<int64_t, int> index;
    for (int e = 0; e < count; ++e) {
        memd_server::Req& r = S->reqs[S->active[e]];
        S->eds[e] = desc_geometry(ctx, r.in, S->opts, S->path, count);
        bind_workspace(S->eds[e], S->pool + r.off, r.lay);
        S->grp[e] = groups > 1 ? r.grp : 0;
        index[S->active[e]] = e;
    }
    memd_status rs = plan_layout(ctx, S->eds, S->grp, groups, 0, &S->P);
    if (rs != MEMD_OK) return rs;
    if (S->P.total > S->region_bytes) {
        // the region is still read by queued steps and probes: drain, then replace it
        HIP_TRY(ctx, hipStreamSynchronize(s));
        S->probe_done = S->probe_seq;
        server_drop_graphs(S, true);
        if (S->graph) {
            (void)hipGraphExecDestroy(S->graph);
            S->graph = nullptr;
        }
        if (S->region) HIP_TRY(ctx, hipFree(S->region));
        S->region = nullptr;
        S->region_bytes = 0;
        size_t want = S->P.total + S->P.total / 4;
        hipError_t e = dev_alloc(ctx, (void**)&S->region, want);
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
    rs = write_plan(ctx, S->eds, S->grp, S->P, S->region, S->path.stop_kernels ? kModeStop : kModeFix, S->path.records,
                    S->tabs, s, &S->plan);
    if (rs != MEMD_OK) return rs;
    S->sift_off.resize(count);
    for (int e = 0, at = 0; e < count; ++e) {
        S->sift_off[e] = at;
        at += S->eds[e].N_imf;
    }
    if (count)
        HIP_TRY(ctx, hipMemcpyAsync(S->region + S->P.siftoff_at, S->sift_off.data(), sizeof(int) * count,
                                    hipMemcpyHostToDevice, s));
    // counters are zero between steps; a new region starts from zero too (kernels, not
    // memset nodes: the step graph is replayed in the same stream order)
    HIP_TRY(ctx, launch_zero_ints((int*)(S->region + S->P.cnt_at), (int)(kCountBytes / sizeof(int)), s));
    HIP_TRY(ctx, launch_zero_ints((int*)(S->region + S->P.bcnt_at), (int)(kBcountStride * count / sizeof(int)), s));
    HIP_TRY(ctx, launch_zero_ints(ndone, 1, s));
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
    HIP_TRY(ctx, hipEventRecord(S->staged, s));
    // one sift step of this plan as a graph (kernels only); the replaced graph lives
    // until the probe after its last launch has completed
    if (S->graph) S->old_graphs.push_back({S->probe_seq, S->graph});
    S->graph = nullptr;
    if (count) {
        hipGraph_t g = nullptr;
        HIP_TRY(ctx, hipStreamBeginCapture(S->cap, hipStreamCaptureModeThreadLocal));
        hipError_t le = launch_step(S->plan, S->cap);
        hipError_t ce = hipStreamEndCapture(S->cap, &g);
        if (le != hipSuccess || ce != hipSuccess) {
            if (g) (void)hipGraphDestroy(g);
            return fail(ctx, MEMD_ERR_HIP, "graph capture failed: %s / %s", hipGetErrorString(le),
                        hipGetErrorString(ce));
        }
        hipError_t ie = hipGraphInstantiate(&S->graph, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        if (ie != hipSuccess) {
            S->graph = nullptr;
            return fail(ctx, MEMD_ERR_HIP, "graph instantiate failed: %s", hipGetErrorString(ie));
        }
    }
    S->dirty = false;
    S->info.rebuilds += 1;
    S->info.rebuild_s += host_seconds() - t0;
    return MEMD_OK;
}

void server_consume(memd_server* S, memd_server::Probe& pr);

// Queue `steps` replays of the step graph and a probe behind them.
memd_status server_launch(memd_server* S, int steps, hipStream_t s) {
    memd_ctx* ctx = S->ctx;
    HIP_TRY(ctx, hipEventRecord(ctx->ev_in, s));
    HIP_TRY(ctx, hipStreamWaitEvent(ctx->own, ctx->ev_in, 0));
    for (int i = 0; i < steps; ++i) HIP_TRY(ctx, hipGraphLaunch(S->graph, ctx->own));
    HIP_TRY(ctx, hipEventRecord(ctx->ev_out, ctx->own));
    HIP_TRY(ctx, hipStreamWaitEvent(s, ctx->ev_out, 0));
    S->info.steps += steps;
    memd_server::Probe& pr = S->probe[S->probe_slot];
    S->probe_slot ^= 1;
    if (pr.pending) {   // normally consumed by the previous call already
        HIP_TRY(ctx, hipEventSynchronize(pr.ev));
        server_consume(S, pr);
    }
    const int count = S->plan.count;
    const size_t nsift = (size_t)S->P.sift_total;
    if (pr.cap_states < (size_t)count || pr.cap_sift < nsift) {
        if (pr.h_states) HIP_TRY(ctx, hipHostFree(pr.h_states));
        if (pr.h_sift) HIP_TRY(ctx, hipHostFree(pr.h_sift));
        pr.h_states = nullptr;
        pr.h_sift = nullptr;
        pr.cap_states = pr.cap_sift = 0;
        const size_t cs = (size_t)count * 2 + 16, cf = nsift * 2 + 64;
        HIP_TRY(ctx, hipHostMalloc((void**)&pr.h_states, sizeof(EpochState) * cs, hipHostMallocDefault));
        HIP_TRY(ctx, hipHostMalloc((void**)&pr.h_sift, sizeof(int) * cf, hipHostMallocDefault));
        pr.cap_states = cs;
        pr.cap_sift = cf;
    }
    EpochState* dst = (EpochState*)(S->region + S->P.states_at);
    int* dsift = (int*)(S->region + S->P.sift_at);
    HIP_TRY(ctx, launch_gather_probe(S->plan, dst, dsift, (const int*)(S->region + S->P.siftoff_at), s));
    HIP_TRY(ctx, hipMemcpyAsync(pr.h_states, dst, sizeof(EpochState) * count, hipMemcpyDeviceToHost, s));
    if (nsift) HIP_TRY(ctx, hipMemcpyAsync(pr.h_sift, dsift, sizeof(int) * nsift, hipMemcpyDeviceToHost, s));
    HIP_TRY(ctx, hipEventRecord(pr.ev, s));
    pr.ids = S->plan_ids;
    pr.sift_off = S->sift_off;
    pr.seq = ++S->probe_seq;
    pr.pending = true;
    S->info.probes += 1;
    return MEMD_OK;
}

// Retire every running epoch a completed probe shows finished.
void server_consume(memd_server* S, memd_server::Probe& pr) {
    pr.pending = false;
    if (pr.seq > S->probe_done) S->probe_done = pr.seq;
    bool retired = false;
    for (size_t e = 0; e < pr.ids.size(); ++e) {
        const EpochState& st = pr.h_states[e];
        if (!st.done) continue;
        auto it = S->reqs.find(pr.ids[e]);
        if (it == S->reqs.end() || it->second.state != MEMD_REQ_RUNNING) continue;
        memd_server::Req& r = it->second;
        if (r.in.h_sift_counts)
            std::memcpy(r.in.h_sift_counts, pr.h_sift + pr.sift_off[e], sizeof(int32_t) * r.in.N_imf);
        if (r.in.h_n_imf_out) *r.in.h_n_imf_out = st.n_imf_out;
        r.state = MEMD_REQ_DONE;
        // the block stays intact until a rebuild drops this epoch from the plan: a
        // new epoch placed there is initialised in stream order after that rebuild
        pool_free(S, r.off, r.lay.total);
        server_place(S, r, -1);
        S->active.erase(std::find(S->active.begin(), S->active.end(), pr.ids[e]));
        S->finished.push_back(pr.ids[e]);
        S->info.retired += 1;
        retired = true;
    }
    (void)retired;
}

// Wait for (blocking) or check (non-blocking) the probes older than `keep_seq`.
memd_status server_poll_probes(memd_server* S, bool block, long keep_seq) {
    memd_ctx* ctx = S->ctx;
    for (int pass = 0; pass < 2; ++pass) {
        // the older pending probe first
        memd_server::Probe* p = nullptr;
        for (auto& pr : S->probe)
            if (pr.pending && pr.seq < keep_seq && (!p || pr.seq < p->seq)) p = &pr;
        if (!p) break;
        if (block) {
            HIP_TRY(ctx, hipEventSynchronize(p->ev));
        } else {
            hipError_t q = hipEventQuery(p->ev);
            if (q == hipErrorNotReady) break;
            if (q != hipSuccess) HIP_TRY(ctx, q);
        }
        server_consume(S, *p);
    }
    return MEMD_OK;
}

// Withdraw a queued or running request (cancel or deadline).
memd_status server_withdraw(memd_server* S, int64_t id, int state) {
    memd_ctx* ctx = S->ctx;
    memd_server::Req& r = S->reqs[id];
    if (r.state == MEMD_REQ_QUEUED) {
        S->queue.erase(std::find(S->queue.begin(), S->queue.end(), id));
    } else {
        // mark it done on the device behind everything queued, and wait: afterwards no
        // kernel reads or writes its buffers
        hipStream_t s = work_stream(ctx);
        HIP_TRY(ctx, launch_mark_done((EpochState*)(S->pool + r.off + r.lay.st), s));
        HIP_TRY(ctx, hipStreamSynchronize(s));
        pool_free(S, r.off, r.lay.total);
        server_place(S, r, -1);
        S->active.erase(std::find(S->active.begin(), S->active.end(), id));
    }
    r.state = state;
    S->finished.push_back(id);
    S->info.withdrawn += 1;
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
    S->path = path_for(o);
    S->max_active = max_active;
    S->pool_bytes = hbm_pool_bytes / kAlign * kAlign;
    hipError_t e = dev_alloc(ctx, (void**)&S->pool, S->pool_bytes);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&S->staged, hipEventDisableTiming);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&S->cap, hipStreamNonBlocking);
    for (int i = 0; i < 2 && e == hipSuccess; ++i) e = hipEventCreateWithFlags(&S->probe[i].ev, hipEventDisableTiming);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        memd_server_destroy(S);
        return fail(ctx, e == hipErrorOutOfMemory ? MEMD_ERR_NOMEM : MEMD_ERR_HIP,
                    "server set-up (pool of %zu bytes) failed: %s", hbm_pool_bytes, hipGetErrorString(e));
    }
    S->free_blocks[0] = S->pool_bytes;
    *out = S;
    return MEMD_OK;
}

memd_status memd_server_destroy(memd_server* S) {
    if (!S) return MEMD_OK;
    (void)hipSetDevice(S->ctx->device);
    (void)hipStreamSynchronize(work_stream(S->ctx));
    server_drop_graphs(S, true);
    if (S->graph) (void)hipGraphExecDestroy(S->graph);
    for (auto& pr : S->probe) {
        if (pr.h_states) (void)hipHostFree(pr.h_states);
        if (pr.h_sift) (void)hipHostFree(pr.h_sift);
        if (pr.ev) (void)hipEventDestroy(pr.ev);
    }
    if (S->staged) (void)hipEventDestroy(S->staged);
    if (S->cap) (void)hipStreamDestroy(S->cap);
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
    r.lay = layout_for(r.in.n, r.in.T, r.in.K, r.in.N_imf, S->path);
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
    memd_status rs;
    // 1. deadlines
    if (S->info.deadlines) {
        const double now = host_seconds();
        std::vector<int64_t> late;
        for (auto& kv : S->reqs)
            if (kv.second.deadline > 0 && now >= kv.second.deadline &&
                (kv.second.state == MEMD_REQ_QUEUED || kv.second.state == MEMD_REQ_RUNNING))
                late.push_back(kv.first);
        for (int64_t id : late)
            if ((rs = server_withdraw(S, id, MEMD_REQ_TIMED_OUT)) != MEMD_OK) return rs;
    }
    // 2. what the device has already reported
    if ((rs = server_poll_probes(S, false, S->probe_seq + 1)) != MEMD_OK) return rs;
    // 3. admissions; the plan is rebuilt when epochs join, after a failed rebuild, or
    //    when finished epochs outnumber the running ones in it
    std::vector<int64_t> joined;
    server_admit(S, joined);
    const bool stale = S->plan.count > 2 * (int)S->active.size() + 8;
    if (!joined.empty() || S->dirty || stale || (S->plan.count == 0 && !S->active.empty())) {
        rs = server_rebuild(S, joined, s);
        if (rs != MEMD_OK) {
            server_unadmit(S, joined);
            S->dirty = true;
            return rs;
        }
    }
    // 4. queue the steps and a probe behind them
    long keep = S->probe_seq + 1;
    if (steps > 0 && !S->active.empty() && S->graph) {
        if ((rs = server_launch(S, steps, s)) != MEMD_OK) return rs;
        keep = S->probe_seq;   // this call's probe stays in flight
    }
    // 5. wait for the older probes and retire what they show
    if ((rs = server_poll_probes(S, true, keep)) != MEMD_OK) return rs;
    server_drop_graphs(S, false);
    if (n_active) *n_active = (int32_t)S->active.size();
    if (n_queued) *n_queued = (int32_t)S->queue.size();
    return MEMD_OK;
}

memd_status memd_server_cancel(memd_server* S, int64_t id) {
    if (!S) return fail(nullptr, MEMD_ERR_ARG, "null server");
    auto it = S->reqs.find(id);
    if (it == S->reqs.end()) return fail(S->ctx, MEMD_ERR_NOT_FOUND, "unknown request id %lld", (long long)id);
    if (it->second.state != MEMD_REQ_QUEUED && it->second.state != MEMD_REQ_RUNNING)
        return fail(S->ctx, MEMD_ERR_ARG, "request %lld has already ended (state %d)", (long long)id, it->second.state);
    HIP_TRY(S->ctx, hipSetDevice(S->ctx->device));
    return server_withdraw(S, id, MEMD_REQ_CANCELLED);
}

memd_status memd_server_set_deadline(memd_server* S, int64_t id, double seconds) {
    if (!S) return fail(nullptr, MEMD_ERR_ARG, "null server");
    auto it = S->reqs.find(id);
    if (it == S->reqs.end()) return fail(S->ctx, MEMD_ERR_NOT_FOUND, "unknown request id %lld", (long long)id);
    if (it->second.state != MEMD_REQ_QUEUED && it->second.state != MEMD_REQ_RUNNING)
        return fail(S->ctx, MEMD_ERR_ARG, "request %lld has already ended (state %d)", (long long)id, it->second.state);
    it->second.deadline = seconds > 0 ? host_seconds() + seconds : 0;
    if (seconds > 0) S->info.deadlines += 1;
    return MEMD_OK;
}

memd_status memd_server_get_info(memd_server* S, memd_server_info* out) {
    if (!S) return fail(nullptr, MEMD_ERR_ARG, "null server");
    if (!out) return fail(S->ctx, MEMD_ERR_ARG, "null output pointer");
    *out = S->info;
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
    if (it->second.state == MEMD_REQ_QUEUED || it->second.state == MEMD_REQ_RUNNING)
        return fail(S->ctx, MEMD_ERR_ARG, "request %lld has not ended", (long long)id);
    S->reqs.erase(it);
    return MEMD_OK;
}

size_t memd_server_pool_used(memd_server* S) { return S ? S->pool_used : 0; }

}  // extern "C"
