// gfx950 kernels of the MEMD engine.
//
// One sift step of every active epoch is a fixed sequence of launches, each
// covering all epochs of the batch:
//
//   flags    projection of the current signal onto every direction, fused with
//            strict-extremum tests and wave64 ballots -> one 64-bit flag word per
//            64 samples per (direction, envelope)
//            [reference: cublasSgemm :1182-1184, find_extrema_shfl_max/min :77-141]
//   compact  popcount prefix over the flag words, expansion into knot lists, and
//            the exact work items of the solver kernels for valid directions
//            [prescan_* :143-281, select_extrema_* :384-445]
//   factor   blocked tridiagonal, shared part: block factors, interval
//            reciprocals, spike vectors, reduced-system factorisation
//   knots    knot values (end knots clamped) and right-hand sides, coalesced
//   block    phase 1: per-channel block solves (fma chains in registers)
//   reduced  phase 2: separator values, channels as lanes
//   coef     phase 3 (second derivatives M) and the spline record (a,b,c,d) of
//            every interval of every channel, coalesced
//            [setBoundary* :447-584, tridiagonal_setup :704-736,
//             cusparseSgtsv2_nopivot :1227-1241, spline_coefficients :775-793]
//   eval     evaluation of both envelopes for every direction, mean over valid
//            directions and the sift update, fused; envelopes never leave
//            registers [interpolate :795-830, averageUppperLower :832-844,
//            averageDirection :846-864, updateSignal :866-874]
//   advance  per-epoch state machine (IMF / sift counters)
//
// Line numbers refer to memReducedEpochs_dev.cu.  All arithmetic is FP64 with
// contraction disabled (-ffp-contract=off); fma() appears exactly where the
// oracle's contract puts it, so results are bit-identical to the oracle.
//
// Fixed-size launches (init, flags, compact, eval, stop kernels) find their
// epoch through a per-block table; the solver kernels are persistent grids that
// drain work lists filled by k_compact with the exact number of items, so no
// workgroup is spent on knots that do not exist.
#include "engine.h"

namespace memd {

namespace {

__device__ __forceinline__ int window_limit(int T, int window) {
    if (window == kWindowFull) return T - 2;
    // the reference launches ceil(T/242) blocks of 8 warps; warp w compares
    // samples 30w+1 .. 30w+30 with their shuffle neighbours
    long G = ((long)T + 241) / 242;
    long lim = 240 * G;
    return (int)(lim < (long)(T - 2) ? lim : (long)(T - 2));
}

__device__ __forceinline__ double* cur_slot(const EpochDesc& ed, int imf) {
    return ed.imfs + (size_t)imf * ed.n * ed.T;
}

__device__ __forceinline__ bool dir_valid(const int* E, int k) {
    // averageUppperLower / averageDirection (:840, :857): both envelopes need > 3 knots
    return E[2 * k] > 3 && E[2 * k + 1] > 3;
}

// Knot values with the end-knot clamp (:447-486 max, :545-584 min), E >= 6 only.
struct KnotRow {
    const int* X;
    const double* cur;   // this channel's row of the current signal
    int E;
    int env;             // 0 = max envelope, 1 = min envelope

    __device__ __forceinline__ double raw(int u) const { return cur[X[u]]; }
    __device__ double Y(int u) const {
        if (E >= 6 && u == 0) {
            double y0 = raw(0), y1 = raw(1), y2 = raw(2);
            double s = (y2 - y1) / (double)(X[2] - X[1]);
            double tv = y1 - s * (double)(X[1] - X[0]);
            bool take = (env == 0) ? (tv > y0) : (tv < y0);
            return take ? tv : y0;
        }
        if (E >= 6 && u == E - 1) {
            int m = E - 1;
            double ym = raw(m), ym1 = raw(m - 1), ym2 = raw(m - 2);
            double s = (ym1 - ym2) / (double)(X[m - 1] - X[m - 2]);
            double tv = ym1 + s * (double)(X[m] - X[m - 1]);
            bool take = (env == 0) ? (tv > ym) : (tv < ym);
            return take ? tv : ym;
        }
        return raw(u);
    }
};

// Knot system geometry: rows u = 1..N (N = E-2), separators at u = jL
// (j = 1..J, J = (N-1)/L), block b = rows bL+1 .. (b<J ? (b+1)L-1 : N).
struct BlockGeom {
    int N, J, p, q;
};

__device__ __forceinline__ BlockGeom block_geom(int E, int b) {
    BlockGeom g;
    g.N = E - 2;
    g.J = (g.N - 1) / kSolverBlock;
    g.p = b * kSolverBlock + 1;
    g.q = (b < g.J) ? (b + 1) * kSolverBlock - 1 : g.N;
    return g;
}

// Pointers in the global address space: loads through them compile to
// global_load with a scalar base and a 32-bit per-lane offset instead of flat
// loads with 64-bit address arithmetic.
template <class T>
using GPtr = const __attribute__((address_space(1))) T*;
template <class T>
__device__ __forceinline__ GPtr<T> gptr(const T* p) {
    return (GPtr<T>)p;
}
// Read-only data every lane of a wave reads at the same address (direction
// vectors): the constant address space lets uniform loads use the scalar cache.
template <class T>
using CPtr = const __attribute__((address_space(4))) T*;
template <class T>
__device__ __forceinline__ CPtr<T> cptr(const T* p) {
    return (CPtr<T>)p;
}
template <class T>
using GMut = __attribute__((address_space(1))) T*;
// Four doubles loaded together from a 32-byte aligned address.
typedef double v4d __attribute__((ext_vector_type(4)));
template <class T>
__device__ __forceinline__ GMut<T> gmut(T* p) {
    return (GMut<T>)p;
}

// Row u (1..N, N = E-2) of a knot system from the three knot positions X[u-1],
// X[u], X[u+1].  Natural ends: the reference's rows
// (memReducedEpochs_dev.cu:704-736).  Not-a-knot: rows 1 and N with M_0 and
// M_{E-1} eliminated (integer products, exact in FP64).
__device__ __forceinline__ void sys_row3(int xm, int x0, int xp, int u, int E, int spline, double& lo, double& di,
                                         double& up) {
    lo = (double)(x0 - xm);
    up = (double)(xp - x0);
    di = (double)(2 * ((x0 - xm) + (xp - x0)));
    if (spline && E >= 4) {
        if (u == 1) {
            long long h0 = x0 - xm, h1 = xp - x0;
            lo = 0.0;
            di = (double)((h0 + h1) * (h0 + 2 * h1));
            up = (double)(h1 * h1 - h0 * h0);
        } else if (u == E - 2) {
            long long a = x0 - xm, b = xp - x0;
            lo = (double)(a * a - b * b);
            di = (double)((a + b) * (2 * a + b));
            up = 0.0;
        }
    }
}

// Plane `plane` (kFacAB, kFacIH, kFacRC) of the per-knot pairs of (direction,
// envelope) ke: the pair of row u at [2 u], [2 u + 1].
__device__ __forceinline__ double* fac_plane(const EpochDesc& ed, int ke, int plane) {
    return ed.fac + ((size_t)ke * kFacPlanes + plane) * ed.Ecap * 2;
}
typedef double v2d __attribute__((ext_vector_type(2)));

// Z of (direction, envelope) ke: Z[j] at [j * np], j = 0 .. J + 1.
__device__ __forceinline__ double* zsep_base(const EpochDesc& ed, int ke) {
    return ed.sep + (size_t)ke * (ed.nbmax + 1) * ed.np;
}

// Knot rows outside the blocks of the blocked solve: the end rows 0 and E-1 and
// the separators u = 32 j, 1 <= j <= J.  Their M is y^ itself (i0 = i1 = 0).
__device__ __forceinline__ bool outside_blocks(int u, int E, int J) {
    return u == 0 || u == E - 1 || (u >= kSolverBlock && (u & (kSolverBlock - 1)) == 0 && (u >> 5) <= J);
}

// Z indices of knot row u: (b, b + 1) for a row of block b, (0, 0) outside blocks.
__device__ __forceinline__ int2 zsep_index(int u, int E, int J) {
    if (outside_blocks(u, E, J)) return make_int2(0, 0);
    const int b = (u - 1) >> 5;
    return make_int2(b, b + 1);
}

__device__ __forceinline__ void zero_counters(const WorkLists& wl) {
    for (int l = 0; l < kNumLists; ++l)
        for (int g = 0; g < kXcds; ++g) wl.count[count_slot(l, g)] = 0;
}

// Persistent-grid iteration over one list: with XCD affinity, workgroup w
// drains the sub-list of XCD w mod 8.
struct ListCursor {
    const int4* base;
    int start, stride, cnt;
};

__device__ __forceinline__ ListCursor list_cursor(const WorkLists& wl, int list) {
    ListCursor cu;
    int g = wl.groups > 1 ? (int)(blockIdx.x % kXcds) : 0;
    cu.base = wl.item[list] + wl.off[list][g];
    cu.start = wl.groups > 1 ? (int)(blockIdx.x / kXcds) : (int)blockIdx.x;
    cu.stride = wl.groups > 1 ? (int)(gridDim.x / kXcds) : (int)gridDim.x;
    cu.cnt = min(wl.count[count_slot(list, g)], wl.cap[list][g]);
    return cu;
}

// ---------------------------------------------------------------- init

__global__ __launch_bounds__(256) void k_init(const EpochDesc* __restrict__ eps, const int2* __restrict__ tab,
                                              WorkLists wl) {
    const int2 br = tab[blockIdx.x];
    const EpochDesc& ed = eps[br.x];
    size_t nT = (size_t)ed.n * ed.T;
    size_t i = (size_t)br.y * 256 + threadIdx.x;
    if (i < nT) {
        double v = ed.y[i];
        ed.running[i] = v;
        ed.imfs[i] = v;   // slot 0 holds the signal being sifted (or the residual if N_imf == 1)
    }
    {   // the directions channel-major (the epoch's blocks stride over n * Kp entries)
        const int Kp = dir_pad(ed.K);
        const size_t nvt = (size_t)ed.n * Kp, stride = (nT + 255) / 256 * 256;
        for (size_t j = i; j < nvt; j += stride) {
            const int c = (int)(j / Kp), k = (int)(j - (size_t)c * Kp);
            ed.vt[j] = k < ed.K ? ed.V[(size_t)k * ed.n + c] : 0.0;
        }
    }
    if (br.y == 0 && threadIdx.x == 0) {
        EpochState* st = ed.st;
        st->imf = 0;
        st->nsift = 0;
        st->pass = 0;
        st->action = kActNone;
        st->act_imf = 0;
        st->done = (ed.N_imf <= 1) ? 1 : 0;
        st->n_imf_out = (ed.N_imf <= 1) ? 1 : 0;
        if (ed.N_imf <= 1) atomicAdd(ed.ndone, 1);
        for (int i2 = 0; i2 < ed.N_imf; ++i2) ed.sift_counts[i2] = 0;
        *ed.bcount = 0;
        if (br.x == 0) zero_counters(wl);
    }
}

// ------------------------------------------------------ projection + flags

// Neighbouring lanes' values by DPP wavefront shifts (no LDS round trip); lane 0
// (from_prev) and lane 63 (from_next) get no value and are filled by the caller.
__device__ __forceinline__ double wave_from_prev(double v) {   // lane i <- lane i-1
    const int lo = __builtin_amdgcn_mov_dpp(__double2loint(v), 0x138, 0xF, 0xF, false);   // wave_shr:1
    const int hi = __builtin_amdgcn_mov_dpp(__double2hiint(v), 0x138, 0xF, 0xF, false);
    return __hiloint2double(hi, lo);
}
__device__ __forceinline__ double wave_from_next(double v) {   // lane i <- lane i+1
    const int lo = __builtin_amdgcn_mov_dpp(__double2loint(v), 0x130, 0xF, 0xF, false);   // wave_shl:1
    const int hi = __builtin_amdgcn_mov_dpp(__double2hiint(v), 0x130, 0xF, 0xF, false);
    return __hiloint2double(hi, lo);
}

// The projections beside each wave of a k_flags tile, per direction d of a pass:
// Px[d][w + 1] = wave w's first and last, Px[d][0][1] = the projection at t0 - 1
// and Px[d][kFlagWaves + 1][0] the one at t0 + kFlagTile, so wave w's lane 0 finds its
// left neighbour at Px[d][w][1] and lane 63 its right one at Px[d][w + 2][0].
constexpr int kFlagWaves = kFlagTile / 64;
typedef double FlagEdges[kFlagWaves + 2][2];

// Lane 0 and 63 of every wave store their projections, threads 2d + side below 2 kd
// the halo projection `hacc` of direction d (side 0: t0 - 1, side 1: t0 + kFlagTile).
template <int DG>
__device__ __forceinline__ void put_edges(const double (&acc)[DG], double hacc, int kd, int tid, FlagEdges* Px) {
    const int lane = tid & 63, wv = tid >> 6;
    if (tid < 2 * kd) Px[tid >> 1][(tid & 1) ? kFlagWaves + 1 : 0][(tid & 1) ? 0 : 1] = hacc;
    if (lane == 0 || lane == 63) {
#pragma unroll
        for (int d = 0; d < DG; ++d) Px[d][wv + 1][lane == 63 ? 1 : 0] = acc[d];
    }
}

// Flag words of directions k0 .. k0 + kd - 1 from the tile's projections (acc[d]
// at sample t0 + tid) and their edge values: end samples are both extrema, the
// tested interior samples (within the window) strict extrema against both
// neighbours; two ballots per (direction, 64 samples), stored by lane 0.
template <int DG>
__device__ __forceinline__ void store_flags(const EpochDesc& ed, const double (&acc)[DG], const FlagEdges* Px, int k0,
                                            int kd, int t0, int tid) {
    const int T = ed.T, W = ed.W;
    const int lane = tid & 63, wv = tid >> 6;
    const int t = t0 + tid;
    const int w = (t0 >> 6) + wv;
    const int lim = window_limit(T, ed.window);
    // the lanes' classes as wave masks (per direction only the comparisons are
    // balloted, and the flag words are formed from masks in scalar registers)
    const unsigned long long endk = __ballot(t == 0 || t == T - 1);
    const unsigned long long test = __ballot(t > 0 && t < T - 1 && t <= lim);
    // global-address-space stores: a flat store would also hold up the LDS waits
    const GMut<unsigned long long> mk = gmut(ed.mask);
#pragma unroll
    for (int d = 0; d < DG; ++d) {
        if (d < kd) {   // uniform
            const int k = k0 + d;
            // neighbours: adjacent lanes; across waves and tiles the edge values (every
            // lane reads the wave's two: broadcast reads).  The lane shifts run with
            // every lane active, before the selects: under a lane condition the lane
            // beside a disabled one would read no value
            const double p = acc[d];
            const double sa = wave_from_prev(p), sb = wave_from_next(p);
            const double ea = Px[d][wv][1], eb = Px[d][wv + 2][0];
            const double a = lane == 0 ? ea : sa;
            const double b = lane == 63 ? eb : sb;
            const unsigned long long mmax = endk | (test & __ballot(p > a) & __ballot(p > b));
            const unsigned long long mmin = endk | (test & __ballot(p < a) & __ballot(p < b));
            if (lane == 0 && w < W) {
                mk[((size_t)k * 2 + 0) * W + w] = mmax;
                mk[((size_t)k * 2 + 1) * W + w] = mmin;
            }
        }
    }
}

// Stage channels c0 .. c0 + cn - 1 (cn <= kFlagChunk) of the signal at samples
// t0 - 1 .. t0 + kFlagTile into LDS rows of kFlagTile + 2 doubles (samples outside
// 0 .. T-1 are 0): thread i loads sample t0 + i of every channel in one round of
// loads, threads below 2 cn the two halo columns.
static_assert(kFlagLdsMaxCh <= kFlagChunk, "the whole-tile path stages its channels in one call");
__device__ __forceinline__ void stage_columns(const GPtr<double>& cur, int T, int c0, int cn, int t0,
                                              double* curT, int tid) {
    constexpr int TW = kFlagTile + 2;
    const int t = t0 + tid;
    const unsigned tc = (unsigned)min(t, T - 1);
    double v[kFlagChunk];
#pragma unroll
    for (int c = 0; c < kFlagChunk; ++c) v[c] = c < cn ? cur[(unsigned)(c0 + c) * (unsigned)T + tc] : 0.0;
    double hv = 0.0;
    const int hc = tid >> 1, th = (tid & 1) ? t0 + kFlagTile : t0 - 1;
    const bool hin = tid < 2 * cn && th >= 0 && th < T;
    if (hin) hv = cur[(unsigned)(c0 + hc) * (unsigned)T + (unsigned)th];
#pragma unroll
    for (int c = 0; c < kFlagChunk; ++c)
        if (c < cn) curT[c * TW + tid + 1] = t < T ? v[c] : 0.0;
    if (tid < 2 * cn) curT[hc * TW + ((tid & 1) ? TW - 1 : 0)] = hv;
}

// Sample-major copy of channels c0 .. c0 + cn - 1 of the staged tile for the
// knot-value gathers: lanes over (sample, channel), channel fastest.
__device__ __forceinline__ void copy_sample_major(const EpochDesc& ed, const double* curT, int c0, int cn, int t0,
                                                  int tid) {
    constexpr int TW = kFlagTile + 2;
    const int tn = min(kFlagTile, ed.T - t0);
    const unsigned long long cmagic = (0x100000000ull + (unsigned)cn - 1) / (unsigned)cn;   // idx / cn (as k_block)
    const GMut<double> ct = gmut(ed.curt + (size_t)t0 * ed.np + c0);
#pragma unroll 4
    for (int idx = tid; idx < cn * tn; idx += kFlagTile) {
        const int tl = (int)(((unsigned long long)idx * cmagic) >> 32);
        const int c = idx - tl * cn;
        ct[(unsigned)tl * (unsigned)ed.np + (unsigned)c] = curT[c * TW + tl + 1];
    }
}

// Projections of the tile for DG directions from k0 (the chains of directions
// past K run on the zero coefficients of the channel-major copy and are not
// used): acc[d] += V[k0 + d][c0 + c] * x(c) for c in 0 .. cn-1, in ascending c.  The
// DG coefficients of a channel are adjacent words: one base, immediate offsets.
template <int DG>
__device__ __forceinline__ void project(const EpochDesc& ed, const double* curT, int c0, int cn, int k0, int tid,
                                        double (&acc)[DG]) {
    constexpr int TW = kFlagTile + 2;
    const int Kp = dir_pad(ed.K);
    const CPtr<double> VT = cptr(ed.vt) + k0;
    for (int c = 0; c < cn; ++c) {
        const double x = curT[c * TW + tid + 1];
        const CPtr<double> v = VT + (unsigned)(c0 + c) * (unsigned)Kp;
#pragma unroll
        for (int d = 0; d < DG; ++d) acc[d] = fma((double)v[d], x, acc[d]);
    }
}

// Channel-chunked path of k_flags (n > kFlagLdsMaxCh): one workgroup per (group
// of kFlagGroup directions, tile).  The tile's channels pass through the LDS tile
// kFlagChunk at a time, and every direction's chain (and each halo chain) stays in
// registers across the chunks in ascending channel order: the same fma chain as
// with the whole tile staged.  The signal is read once per direction group.
__device__ void flags_chunked(const EpochDesc& ed, const EpochState* st, int by, double* curT, FlagEdges* Px) {
    constexpr int TW = kFlagTile + 2;
    constexpr int DG = kFlagGroup;
    constexpr int CH = kFlagChunk;
    const int n = ed.n, T = ed.T, K = ed.K;
    const int tiles = (T + kFlagTile - 1) / kFlagTile;
    const int grp = by / tiles;
    const int tile = by - grp * tiles;
    const int kg0 = grp * DG;
    const int kd = min(DG, K - kg0);
    const int t0 = tile * kFlagTile;
    const GPtr<double> cur = gptr(cur_slot(ed, st->imf));
    const CPtr<double> V = cptr(ed.V);
    const int tid = threadIdx.x;
    double acc[DG];
#pragma unroll
    for (int d = 0; d < DG; ++d) acc[d] = 0.0;
    // halo chains: thread 2d + side, for direction d and column 0 or TW - 1
    const unsigned hrow = (unsigned)min(kg0 + (tid >> 1), K - 1) * (unsigned)n;
    const int hcol = (tid & 1) ? TW - 1 : 0;
    double hacc = 0.0;
    for (int c0 = 0; c0 < n; c0 += CH) {
        const int cn = min(CH, n - c0);
        __syncthreads();   // the previous chunk has been read
        stage_columns(cur, T, c0, cn, t0, curT, tid);
        __syncthreads();
        if (grp == 0) copy_sample_major(ed, curT, c0, cn, t0, tid);
        if (tid < 2 * kd) {
            for (int c = 0; c < cn; ++c) hacc = fma((double)V[hrow + c0 + c], curT[c * TW + hcol], hacc);
        }
        project(ed, curT, c0, cn, kg0, tid, acc);
    }
    put_edges(acc, hacc, kd, tid, Px);
    __syncthreads();
    store_flags(ed, acc, Px, kg0, kd, t0, tid);
}

// One workgroup per (epoch, 256-sample tile) and all directions (n <= kFlagLdsMaxCh):
// the tile of the current signal (plus one sample each side) is staged in LDS
// once, then every pass of kFlagGroup directions runs its projections, extremum
// tests and ballots from LDS.
__global__ __launch_bounds__(kFlagTile) void k_flags(const EpochDesc* __restrict__ eps,
                                                     const int2* __restrict__ tab) {
    constexpr int TW = kFlagTile + 2;
    constexpr int DG = kFlagGroup;   // directions per pass: independent fma chains per thread
    extern __shared__ double curT[];   // [n][TW]: samples t0-1 .. t0+kFlagTile
    __shared__ FlagEdges Px[DG];
    const int2 br = tab[blockIdx.x];
    const EpochDesc& ed = eps[br.x];
    const EpochState* st = ed.st;
    if (st->done) return;
    if (ed.n > kFlagLdsMaxCh) {   // its blocks are (direction group, tile) pairs (blocks_for)
        flags_chunked(ed, st, br.y, curT, Px);
        return;
    }
    const int n = ed.n, T = ed.T, K = ed.K;
    const int t0 = br.y * kFlagTile;
    const GPtr<double> cur = gptr(cur_slot(ed, st->imf));
    const CPtr<double> V = cptr(ed.V);
    const int tid = threadIdx.x;
    stage_columns(cur, T, 0, n, t0, curT, tid);
    __syncthreads();
    copy_sample_major(ed, curT, 0, n, t0, tid);
    const int hcol = (tid & 1) ? TW - 1 : 0;
    for (int k0 = 0; k0 < K; k0 += DG) {
        const int kd = min(DG, K - k0);
        double acc[DG];
#pragma unroll
        for (int d = 0; d < DG; ++d) acc[d] = 0.0;
        project(ed, curT, 0, n, k0, tid, acc);
        // halo projections of the pass's directions, one chain per (direction, side)
        double hacc = 0.0;
        if (tid < 2 * kd) {
            const unsigned hrow = (unsigned)(k0 + (tid >> 1)) * (unsigned)n;
            for (int c = 0; c < n; ++c) hacc = fma((double)V[hrow + c], curT[c * TW + hcol], hacc);
        }
        __syncthreads();   // the previous pass has read its edge values
        put_edges(acc, hacc, kd, tid, Px);
        __syncthreads();
        store_flags(ed, acc, Px, k0, kd, t0, tid);
    }
}

// ------------------------------------------------------------ compaction

// k_block work items of one direction-envelope with E knots (groups of bpw
// solver blocks).
__host__ __device__ inline int block_items(int E, int bpw) {
    const int J = (E - 3) / kSolverBlock;
    return (J + 1 + bpw - 1) / bpw;
}

// Work items one direction-envelope with E knots contributes to each solver list.
__device__ __forceinline__ void list_sizes(const EpochDesc& ed, int E, int* m) {
    const int J = (E - 3) / kSolverBlock;
    m[kListFactor] = 1;
    m[kListReduced] = (J > 0 || ed.spline) ? 1 : 0;   // not-a-knot: end values
    m[kListCoef] = ed.coef_records ? (int)(((long)ed.n * E + 256 * kCoefItems - 1) / (256 * kCoefItems)) : 0;
}

// ------------------------------------------------------------ mirrored ends
//
// memd.m's boundary_conditions with nbsym = 2 (Rehman & Mandic's MEMD toolbox,
// after Rilling, Flandrin & Goncalves), in 0-based samples, on the interior
// extrema of one direction: per end, up to two extrema of each list nearest the
// end are reflected about the end extremum (when the end sample lies beyond the
// first extremum of the other kind) or about the end sample, which then joins
// one list; a reflection about an extremum that falls short of the end sample is
// redone about the end sample.  Same decisions as the oracle's mirror_ends.

// The interior extrema a list's end rules can reach: the first three and the
// last three (back[2] is the last), a >= 2 of them in all.
struct EndExtrema {
    int front[3], back[3], a;
    __device__ int at(int i) const { return i < 3 ? front[i] : back[i - (a - 3)]; }
};

// out[] = interior elements hi-1 down to max(lo, 0); returns the count.
__device__ __forceinline__ int take_reversed(const EndExtrema& v, int lo, int hi, int* out) {
    int m = 0;
    for (int i = hi - 1; i >= max(lo, 0); --i) out[m++] = v.at(i);
    return m;
}

struct MirrorEnds {
    int nl[2], nr[2];        // knots added left / right (0 max list, 1 min list)
    int ls[2][3], rs[2][3];  // their source samples, in position order
    int lsym, rsym;
};

// x0 / xl: projection at samples 0 and T-1; xa0, xi0: at the first maximum and
// minimum; xal, xil: at the last ones.
__device__ void mirror_ends(int last, double x0, double xl, double xa0, double xi0, double xal, double xil,
                            const EndExtrema& A, const EndExtrema& I, MirrorEnds& r) {
    const int a = A.a, b = I.a;
    const int imax0 = A.front[0], imin0 = I.front[0], imaxl = A.back[2], iminl = I.back[2];
    if (imax0 < imin0) {
        if (x0 > xi0) {
            r.nl[0] = take_reversed(A, 1, min(a, 3), r.ls[0]);
            r.nl[1] = take_reversed(I, 0, min(b, 2), r.ls[1]);
            r.lsym = imax0;
        } else {
            r.nl[0] = take_reversed(A, 0, min(a, 2), r.ls[0]);
            r.nl[1] = take_reversed(I, 0, min(b, 1), r.ls[1]);
            r.ls[1][r.nl[1]++] = 0;
            r.lsym = 0;
        }
    } else {
        if (x0 < xa0) {
            r.nl[0] = take_reversed(A, 0, min(a, 2), r.ls[0]);
            r.nl[1] = take_reversed(I, 1, min(b, 3), r.ls[1]);
            r.lsym = imin0;
        } else {
            r.nl[0] = take_reversed(A, 0, min(a, 1), r.ls[0]);
            r.ls[0][r.nl[0]++] = 0;
            r.nl[1] = take_reversed(I, 0, min(b, 2), r.ls[1]);
            r.lsym = 0;
        }
    }
    if (2 * r.lsym - r.ls[1][0] > 0 || 2 * r.lsym - r.ls[0][0] > 0) {
        if (r.lsym == imax0) r.nl[0] = take_reversed(A, 0, min(a, 2), r.ls[0]);
        else r.nl[1] = take_reversed(I, 0, min(b, 2), r.ls[1]);
        r.lsym = 0;
    }
    if (imaxl < iminl) {
        if (xl < xal) {
            r.nr[0] = take_reversed(A, a - 2, a, r.rs[0]);
            r.nr[1] = take_reversed(I, b - 3, b - 1, r.rs[1]);
            r.rsym = iminl;
        } else {
            r.rs[0][0] = last;
            r.nr[0] = 1 + take_reversed(A, a - 1, a, r.rs[0] + 1);
            r.nr[1] = take_reversed(I, b - 2, b, r.rs[1]);
            r.rsym = last;
        }
    } else {
        if (xl > xil) {
            r.nr[0] = take_reversed(A, a - 3, a - 1, r.rs[0]);
            r.nr[1] = take_reversed(I, b - 2, b, r.rs[1]);
            r.rsym = imaxl;
        } else {
            r.nr[0] = take_reversed(A, a - 2, a, r.rs[0]);
            r.rs[1][0] = last;
            r.nr[1] = 1 + take_reversed(I, b - 1, b, r.rs[1] + 1);
            r.rsym = last;
        }
    }
    if (2 * r.rsym - r.rs[1][r.nr[1] - 1] < last || 2 * r.rsym - r.rs[0][r.nr[0] - 1] < last) {
        if (r.rsym == imaxl) r.nr[0] = take_reversed(A, a - 2, a, r.rs[0]);
        else r.nr[1] = take_reversed(I, b - 2, b, r.rs[1]);
        r.rsym = last;
    }
}

// Sample whose value knot j (at position p) of a mirrored list carries.
__device__ __forceinline__ int knot_source(const int4& m, int j, int E, int p) {
    return j < m.x ? 2 * m.z - p : (j >= E - m.y ? 2 * m.w - p : p);
}

// One workgroup per (epoch, direction), one wave per envelope: the knot list
// (exclusive prefix of the flag words' popcounts, a wave scan), then the work
// items of a valid direction appended to the solver lists.  With mirrored ends
// the lists are compacted one slot in, and a valid direction's lists are then
// rebuilt around their interior extrema (mirror_ends) before the work is sized:
// the interior moves left by one when one knot is added at the left end.
__global__ __launch_bounds__(128) void k_compact(const EpochDesc* __restrict__ eps, const int2* __restrict__ tab,
                                                 WorkLists wl) {
    const int2 br = tab[blockIdx.x];
    const EpochDesc& ed = eps[br.x];
    const EpochState* st = ed.st;
    if (st->done) return;
    const int k = br.y;
    const int W = ed.W;
    const int tid = threadIdx.x;
    const int env = tid >> 6, lane = tid & 63;   // wave e compacts envelope e
    const bool mirror = ed.boundary != 0;
    __shared__ int sE[2];
    __shared__ int sEnds[2][6];   // mirrored ends: raw slots 1..3 and E-4..E-2 of each list
    {
        const int ke = 2 * k + env;
        const unsigned long long* m = ed.mask + (size_t)ke * W;
        int* pre = ed.pre + (size_t)ke * W;
        int* X = ed.X + (size_t)ke * ed.Ecap + (mirror ? 1 : 0);
        const int per = (W + 63) / 64;   // consecutive words per lane
        const int w0 = lane * per;
        const int w1 = min(w0 + per, W);
        int mine = 0;
        for (int w = w0; w < w1; ++w) mine += __popcll(m[w]);
        int incl = mine;   // inclusive scan over the wave
#pragma unroll
        for (int d = 1; d < 64; d <<= 1) {
            const int v = __shfl_up(incl, d);
            if (lane >= d) incl += v;
        }
        const int Etot = __shfl(incl, 63);
        int base = incl - mine;
        if (per <= 2 && Etot > 8 * W) {
            // dense envelope (more than 8 knots per flag word) of at most 2 words per lane:
            // the wave takes the words one at a time, lane l storing the position of bit
            // l, so consecutive knots come from consecutive lanes (coalesced stores instead
            // of one scattered store per knot and lane).  Longer epochs keep the per-lane
            // walk: a wave then runs W word steps, which cost more than the scattered
            // stores save when few epochs share the device
            unsigned long long mw[2];
            int bs[2];
            int run = base;
#pragma unroll
            for (int sl = 0; sl < 2; ++sl) {
                const int w = w0 + sl;
                const bool own = sl < per && w < W;
                mw[sl] = own ? m[w] : 0ull;
                bs[sl] = run;
                if (own) pre[w] = run;
                run += (int)__popcll(mw[sl]);
            }
#pragma unroll
            for (int sl = 0; sl < 2; ++sl) {
                if (sl < per) {   // uniform
                    for (int src = 0; src < 64; ++src) {
                        const int w = src * per + sl;
                        if (w >= W) break;   // uniform; later lanes' words are past W too
                        const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)mw[sl], src);
                        const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(mw[sl] >> 32), src);
                        const int bw = __builtin_amdgcn_readlane(bs[sl], src);
                        const unsigned long long bits = ((unsigned long long)hi << 32) | lo;
                        if ((bits >> lane) & 1ull) {
                            const int at = bw + (int)__builtin_amdgcn_mbcnt_hi(hi, __builtin_amdgcn_mbcnt_lo(lo, 0u));
                            const int pos = (w << 6) + lane;
                            if (mirror) {
                                if (at >= 1 && at <= 3) sEnds[env][at - 1] = pos;
                                if (at >= Etot - 4 && at <= Etot - 2) sEnds[env][3 + at - (Etot - 4)] = pos;
                            }
                            X[at] = pos;
                        }
                    }
                }
            }
        } else {
            for (int w = w0; w < w1; ++w) {
                unsigned long long bits = m[w];
                pre[w] = base;
                while (bits) {
                    const int b = __ffsll((long long)bits) - 1;
                    const int pos = (w << 6) + b;
                    if (mirror) {
                        if (base >= 1 && base <= 3) sEnds[env][base - 1] = pos;
                        if (base >= Etot - 4 && base <= Etot - 2) sEnds[env][3 + base - (Etot - 4)] = pos;
                    }
                    X[base++] = pos;
                    bits &= bits - 1;
                }
            }
        }
        if (lane == 63) {
            sE[env] = Etot;
            ed.E[ke] = Etot;
            ed.Eraw[ke] = Etot;
            if (ed.d_trace && st->pass < ed.max_passes)
                ed.d_trace[(size_t)st->pass * 2 * ed.K + ke] = Etot;
            if (ed.stats) {
                unsigned long long* sl = ed.stats + 4 * (blockIdx.x % kStatSlots);
                atomicAdd(&sl[0], (unsigned long long)Etot);
                atomicAdd(&sl[1], (unsigned long long)Etot * (unsigned long long)ed.n);
            }
        }
    }
    __syncthreads();
    if (!(sE[0] > 3 && sE[1] > 3)) return;
    if (mirror) {
        __shared__ double sx[6];
        // the end rules' lists live in LDS: thread 0 indexes them at run time, which
        // private arrays would turn into scratch
        __shared__ EndExtrema sA, sI;
        __shared__ MirrorEnds sR;
        if (tid < 6) {   // the projection (as k_flags: fma chain over ascending channels) at the
                         // samples the end rules compare
            const int T = ed.T;
            const int t = tid == 0 ? 0 : tid == 1 ? T - 1 : tid == 2 ? sEnds[0][0] : tid == 3 ? sEnds[1][0]
                        : tid == 4 ? sEnds[0][5] : sEnds[1][5];
            const double* Vk = ed.V + (size_t)k * ed.n;
            const double* row = ed.curt + (size_t)t * ed.np;
            double acc = 0.0;
            for (int c = 0; c < ed.n; ++c) acc = fma(Vk[c], row[c], acc);
            sx[tid] = acc;
        }
        __syncthreads();
        if (tid == 0) {
            for (int i = 0; i < 3; ++i) {
                sA.front[i] = sEnds[0][i]; sA.back[i] = sEnds[0][3 + i];
                sI.front[i] = sEnds[1][i]; sI.back[i] = sEnds[1][3 + i];
            }
            sA.a = sE[0] - 2;
            sI.a = sE[1] - 2;
            mirror_ends(ed.T - 1, sx[0], sx[1], sx[2], sx[3], sx[4], sx[5], sA, sI, sR);
        }
        __syncthreads();
        {
            const int ke = 2 * k + env;
            int* X = ed.X + (size_t)ke * ed.Ecap;   // raw list at X[1 .. Eraw]
            const int Eraw = sE[env], nin = Eraw - 2;
            const int nl = sR.nl[env], nr = sR.nr[env], lsym = sR.lsym, rsym = sR.rsym;
            if (nl == 1) {   // interior X[2 .. Eraw-1] -> X[1 .. Eraw-2]; each pass reads only
                             // slots above the ones it writes, which later passes never read
                for (int i0 = 0; i0 < nin; i0 += 64) {
                    const int i = i0 + lane;
                    const int v = i < nin ? X[2 + i] : 0;
                    if (i < nin) X[1 + i] = v;
                }
            }
            const int Em = nl + nin + nr;
            if (lane == 0) {   // after the interior has moved (the right knots may land on its
                               // old last slot)
                for (int i = 0; i < nl; ++i) X[i] = 2 * lsym - sR.ls[env][i];
                for (int i = 0; i < nr; ++i) X[nl + nin + i] = 2 * rsym - sR.rs[env][i];
                ed.E[ke] = Em;
                ed.mir[ke] = make_int4(nl, nr, lsym, rsym);
            }
        }
        __syncthreads();
        if (tid == 0 || tid == 64) sE[tid >> 6] = sR.nl[tid >> 6] + (sE[tid >> 6] - 2) + sR.nr[tid >> 6];
        __syncthreads();
    }
    __shared__ int s_base[2][kNumLists], s_m[2][kNumLists];
    // the epoch's XCD sub-list; a spread epoch (xcd < 0) deals its directions over all
    const int g = wl.groups > 1 ? (ed.xcd >= 0 ? ed.xcd : (k & (kXcds - 1))) : 0;
    if (tid == 0) {   // one counter update per list for both envelopes
        int m0[kNumLists], m1[kNumLists];
        list_sizes(ed, sE[0], m0);
        list_sizes(ed, sE[1], m1);
        for (int l = 0; l < kNumLists; ++l) {
            const int tot = m0[l] + m1[l];
            const int base = tot ? atomicAdd(&wl.count[count_slot(l, g)], tot) : 0;
            s_m[0][l] = m0[l];
            s_m[1][l] = m1[l];
            s_base[0][l] = base;
            s_base[1][l] = base + m0[l];
        }
    }
    __shared__ int s_bbase, s_bm[2];
    if (tid == 64) {   // the epoch's own k_block items, both envelopes together
        const int b0 = block_items(sE[0], ed.bpw), b1 = block_items(sE[1], ed.bpw);
        s_bm[0] = b0;
        s_bm[1] = b1;
        s_bbase = atomicAdd(ed.bcount, b0 + b1);
    }
    __syncthreads();
    for (int e2 = 0; e2 < 2; ++e2) {
        for (int l = 0; l < kNumLists; ++l)
            for (int i = tid; i < s_m[e2][l]; i += 128) {
                int at = s_base[e2][l] + i;
                if (at < wl.cap[l][g]) wl.item[l][wl.off[l][g] + at] = make_int4(br.x, 2 * k + e2, i, 0);
            }
        for (int i = tid; i < s_bm[e2]; i += 128) {
            const int at = s_bbase + (e2 ? s_bm[0] : 0) + i;
            if (at < ed.bcap) ed.bitem[at] = make_int4(2 * k + e2, i, sE[e2], 0);
        }
    }
}

// ------------------------------------------------- blocked tridiagonal solver
//
// Knot system of one (direction, envelope): rows u = 1..N with
// lo_u = h_{u-1}, di_u = 2(X_{u+1}-X_{u-1}), up_u = h_u and, per channel,
// r_u = 3 (sl_u - sl_{u-1}), sl_u = (Y_{u+1}-Y_u) * ih_u; M_0 = M_{E-1} = 0.
// Same arithmetic as the oracle's struct Blocked.

// Shared factors of one (direction, envelope): per block the forward factors,
// the reciprocals of the intervals it owns and its two spike vectors; then the
// reduced system's factorisation (one lane, J steps).  NT threads per item: 64
// when many items share the device, 256 when few do (a single epoch has 2K
// items), which cuts the rounds of the row and block phases by 4.  The
// arithmetic of every row is the same either way.
template <int NT>
__global__ __launch_bounds__(NT) void k_factor(const EpochDesc* __restrict__ eps, WorkLists wl) {
    constexpr int L = kSolverBlock;
    __shared__ double s_rlo[NT], s_rdi[NT], s_rup[NT];
    // the first NT blocks' boundary rows for the reduced system: the positions X[p-1],
    // X[p] and X[q] and the (al, be) of rows p and q
    __shared__ int s_bx[NT][3];
    __shared__ v2d s_bab[NT][2];
    const ListCursor cu = list_cursor(wl, kListFactor);
    const int lane = threadIdx.x;
    // A wave's loads and stores share one in-order completion counter, so a load
    // issued after stores waits for them too: each phase's loads are issued before
    // the previous phase's stores (the row phase one round ahead, the first block's
    // positions before the row phase), and the reduced system takes the blocks'
    // boundary rows from LDS instead of reloading what the block phase stored.
    for (int w = cu.start; w < cu.cnt; w += cu.stride) {
        const int4 it = cu.base[w];
        const EpochDesc& ed = eps[it.x];
        const int ke = it.y;
        const int E = ed.E[ke];
        const int spline = ed.spline;
        const GPtr<int> X = gptr(ed.X + (size_t)ke * ed.Ecap);
        const GMut<v2d> AB = gmut(reinterpret_cast<v2d*>(fac_plane(ed, ke, kFacAB)));   // (al, be) of row u
        const GMut<v2d> IH = gmut(reinterpret_cast<v2d*>(fac_plane(ed, ke, kFacIH)));   // (ih, i3h)
        const GMut<v2d> RC = gmut(reinterpret_cast<v2d*>(fac_plane(ed, ke, kFacRC)));   // (rd, cp)
        const int J = block_geom(E, 0).J;
        // the lane's first block's positions X[p-1 .. q+1] (lanes past the last block
        // load the last block's and do not use them)
        int xs[L + 2];
        {
            const BlockGeom g = block_geom(E, min(lane, J));
            const int len = g.q - g.p + 1;
#pragma unroll
            for (int i = 0; i < L + 2; ++i) xs[i] = X[g.p - 1 + min(i, len + 1)];
        }
        // every row: the reciprocals of interval u (u < E-1) and, at the rows outside
        // the blocks, al = be = 0; all lanes, 4 rows per lane per round, the next
        // round's positions loaded (clamped in range) before this round's stores
        {
            int xa[4], xb[4];
#pragma unroll
            for (int g = 0; g < 4; ++g) {
                const int u = min(lane + NT * g, E - 2);
                xa[g] = X[u];
                xb[g] = X[u + 1];
            }
            for (int u0 = lane; u0 < E; u0 += NT * 4) {
                int na[4], nb[4];
#pragma unroll
                for (int g = 0; g < 4; ++g) {
                    const int u = min(u0 + NT * (4 + g), E - 2);
                    na[g] = X[u];
                    nb[g] = X[u + 1];
                }
#pragma unroll
                for (int g = 0; g < 4; ++g) {
                    const int u = u0 + NT * g;
                    if (u < E) {
                        if (u < E - 1) {
                            const double h = (double)(xb[g] - xa[g]);
                            v2d v;
                            v.x = 1.0 / h;
                            v.y = 1.0 / (3.0 * h);
                            IH[u] = v;
                        }
                        if (outside_blocks(u, E, J)) AB[u] = v2d{0.0, 0.0};
                    }
                }
#pragma unroll
                for (int g = 0; g < 4; ++g) {
                    xa[g] = na[g];
                    xb[g] = nb[g];
                }
            }
        }
        // one lane per block: the block's rows live in registers, memory sees stores only
        for (int b = lane; b <= J; b += NT) {
            const BlockGeom g = block_geom(E, b);
            const int p = g.p, len = g.q - g.p + 1;
            if (b != lane) {   // blocks past the first NT (long epochs)
#pragma unroll
                for (int i = 0; i < L + 2; ++i) xs[i] = X[p - 1 + min(i, len + 1)];
            }
            const bool left = b >= 1, right = b < J;
            double cpr[L], zl[L];
            double cprev = 0.0, z = 0.0, zr = 0.0;
            int xq = 0;   // X[q]
#pragma unroll
            for (int i = 0; i < L; ++i) {
                if (i < len) {
                    double lo, di, up;
                    sys_row3(xs[i], xs[i + 1], xs[i + 2], p + i, E, spline, lo, di, up);
                    double den = (i == 0) ? di : fma(-lo, cprev, di);
                    double r = 1.0 / den;
                    cprev = up * r;
                    RC[p + i] = v2d{r, cprev};
                    cpr[i] = cprev;
                    // left spike A_b^{-1}(lo_p e_first), forward sweep
                    z = (i == 0) ? lo * r : fma(-lo, z, 0.0) * r;
                    zl[i] = z;
                    // right spike A_b^{-1}(up_q e_last): the forward sweep is +0 above the
                    // last row (every pivot is positive) and up_q * rd_q there
                    if (i == len - 1) {
                        zr = up * r;
                        xq = xs[i + 1];
                    }
                }
            }
            // phase-3 coefficients: the spikes, 0 where the block has none (the
            // first block has no left spike, the last no right spike)
            double x = 0.0, y = 0.0;
            v2d abq = v2d{0.0, 0.0}, abp = v2d{0.0, 0.0};   // (al, be) of rows q and p
#pragma unroll
            for (int i = L - 1; i >= 0; --i) {
                if (i < len) {
                    x = (i == len - 1) ? zl[i] : fma(-cpr[i], x, zl[i]);
                    y = (i == len - 1) ? zr : fma(-cpr[i], y, 0.0);
                    const v2d ab = v2d{left ? x : 0.0, right ? y : 0.0};
                    AB[p + i] = ab;
                    if (i == len - 1) abq = ab;
                    abp = ab;
                }
            }
            if (b < NT) {
                s_bx[b][0] = xs[0];
                s_bx[b][1] = xs[1];
                s_bx[b][2] = xq;
                s_bab[b][0] = abp;
                s_bab[b][1] = abq;
            }
        }
        __syncthreads();
        if (J > 0) {
            // reduced system: coefficients of every separator in parallel (LDS), then one
            // lane runs the factorisation chain
            // global-address-space stores (flat stores also count against the LDS wait
            // counter that the chain's LDS reads wait on)
            const GMut<double> Rlo = gmut(ed.red + ((size_t)ke * 3 + 0) * ed.nbmax);
            const GMut<double> Rrd = gmut(ed.red + ((size_t)ke * 3 + 1) * ed.nbmax);
            const GMut<double> Rcp = gmut(ed.red + ((size_t)ke * 3 + 2) * ed.nbmax);
            double cprev = 0.0;
            for (int j0 = 1; j0 <= J; j0 += NT) {
                const int j = j0 + lane;
                if (j <= J) {
                    // separator j = row sp: the last row of block j-1 and the first of block j
                    const int sp = j * L;
                    int xm, x0, xp;
                    v2d abm, abp;
                    if (j < NT) {
                        xm = s_bx[j - 1][2];
                        abm = s_bab[j - 1][1];
                        x0 = s_bx[j][0];
                        xp = s_bx[j][1];
                        abp = s_bab[j][0];
                    } else {
                        xm = X[sp - 1];
                        x0 = X[sp];
                        xp = X[sp + 1];
                        abm = AB[sp - 1];
                        abp = AB[sp + 1];
                    }
                    double lo = (double)(x0 - xm), up = (double)(xp - x0);
                    double di = (double)(2 * ((x0 - xm) + (xp - x0)));
                    s_rlo[lane] = (j >= 2) ? -(lo * abm.x) : 0.0;
                    s_rdi[lane] = fma(-up, abp.x, fma(-lo, abm.y, di));
                    s_rup[lane] = (j <= J - 1) ? -(up * abp.y) : 0.0;
                }
                __syncthreads();
                if (lane == 0) {
                    const int jn = min(NT, J - j0 + 1);
                    for (int jj = 0; jj < jn; ++jj) {
                        const int jg = j0 + jj;
                        double rlo = s_rlo[jj];
                        double den = (jg == 1) ? s_rdi[jj] : fma(-rlo, cprev, s_rdi[jj]);
                        double r = 1.0 / den;
                        cprev = s_rup[jj] * r;
                        Rlo[jg] = rlo;
                        Rrd[jg] = r;
                        Rcp[jg] = cprev;
                    }
                }
                __syncthreads();
            }
        }
    }
}

// End-knot clamp of memReducedEpochs_dev.cu:451-485 (max) / :549-583 (min),
// from the raw values at the end knot (ye), its neighbour (y1) and the next one
// (y2), and the positions x_e, x_1, x_2 (same operations as KnotRow::Y).
__device__ __forceinline__ double clamp_end(double ye, double y1, double y2, int xe, int x1, int x2, int env,
                                            bool left) {
    double tv;
    if (left) {
        double sl = (y2 - y1) / (double)(x2 - x1);
        tv = y1 - sl * (double)(x1 - xe);
    } else {
        double sl = (y1 - y2) / (double)(x1 - x2);
        tv = y1 + sl * (double)(xe - x1);
    }
    bool take = (env == 0) ? (tv > ye) : (tv < ye);
    return take ? tv : ye;
}

// Global inputs of tile row r of a k_block work item (phase 1a): the sample the
// knot carries, the factor row's positions and factors, the interval reciprocal.
struct BlockRowIn {
    int xs, xm, x0, xp;
    v2d rc;
    double ih;
};

// Geometry of a k_block work item (ke, group, E): its first row p0 and last row
// qlast of the solve, and the tile rows p0-1 .. qlast+2.
struct BlockItemGeom {
    int J, b0, nb, p0, qlast, R, RW;
};
__device__ __forceinline__ BlockItemGeom block_item_geom(const int4& it, int bpw) {
    BlockItemGeom g;
    g.J = block_geom(it.z, 0).J;
    g.b0 = it.y * bpw;
    g.nb = min(bpw, g.J - g.b0 + 1);
    g.p0 = g.b0 * kSolverBlock + 1;
    g.qlast = block_geom(it.z, g.b0 + g.nb - 1).q;
    g.R = g.qlast - g.p0 + 1;
    g.RW = g.R + 3;
    return g;
}

__device__ __forceinline__ BlockRowIn block_row_in(const EpochDesc& ed, const int4& it, const BlockItemGeom& g, int r) {
    const int ke = it.x, E = it.z;
    const GPtr<int> X = gptr(ed.X + (size_t)ke * ed.Ecap);
    const GPtr<v2d> RCp = gptr(reinterpret_cast<const v2d*>(fac_plane(ed, ke, kFacRC)));
    const GPtr<double> IHp = gptr(fac_plane(ed, ke, kFacIH));
    const int u = min(g.p0 + r, g.qlast);   // factor row (used when r < R)
    const int jr = min(g.p0 - 1 + r, E - 1);
    BlockRowIn v;
    v.xs = X[jr];
    if (ed.boundary != 0) v.xs = knot_source(ed.mir[ke], jr, E, v.xs);   // the sample the knot carries
    v.rc = RCp[u];
    v.xm = X[u - 1];
    v.x0 = X[u];
    v.xp = X[u + 1];
    v.ih = IHp[2 * (unsigned)min(g.p0 - 1 + r, E - 2)];
    return v;
}

// Knot values, right-hand sides and phase 1, per (direction, envelope) and group
// of EpochDesc::bpw solver blocks (rows p0..qlast, at most 32 * bpw).
//   1. The group's knot positions, shared factors and reciprocals are staged in
//      LDS.  The knot values Y of rows p0-1 .. qlast+2 of every channel are then
//      copied from the sample-major signal (one knot's channels are adjacent words)
//      into a row-major LDS tile by 16-byte global-to-LDS loads, which hold no
//      registers, so the whole tile is in flight at once.  The end knots are
//      clamped in the tile (memReducedEpochs_dev.cu:447-486, 545-584); the clamped
//      end values go to yend (the other knot values are the signal itself, which
//      the readers take from curt).
//   2. Lanes (block, channel): the right-hand side of the separator after the
//      block (it reads the next block's first row) is formed first; after a
//      barrier, r_u = 3 (sl_u - sl_{u-1}) with sl_u = (Y_{u+1} - Y_u) * ih_u and the
//      block's Thomas sweeps (y = A_b^{-1} r) run in place in the tile, a rolling
//      window of Y values in registers keeping every read ahead of the writes.
//   3. The backward sweep stores each block row's y as it is formed (lanes of a
//      wave are consecutive channels of one row, so the stores are coalesced).
__global__ __launch_bounds__(kBlockThreads, 4) void k_block(const EpochDesc* __restrict__ eps,
                                                         const int2* __restrict__ tab, int rmax, int tile_off) {
    extern __shared__ double lds[];
    // per tile row r (rmax + 3 rows; the entries past a row's use are written, never read)
    const int rw = rmax + 3;
    double* s_rd = lds;                   // factor rows p0 + r (r < R)
    double* s_cp = lds + rw;
    double* s_lo = lds + 2 * rw;
    double* s_ih = lds + 3 * rw;          // reciprocal of interval p0 - 1 + r (r <= R + 1)
    int* sX = reinterpret_cast<int*>(lds + 4 * rw);   // position of knot p0 - 1 + r
    // tile [RW][np]: row r = knot p0 - 1 + r (clamped to E - 1), channels innermost;
    // 16-byte aligned (tile_off)
    double* sY = reinterpret_cast<double*>(reinterpret_cast<char*>(lds) + tile_off);
    const int2 br = tab[blockIdx.x];   // (epoch, workgroup of the epoch)
    const EpochDesc& ed = eps[br.x];
    const int tid = threadIdx.x;
    const int cnt = min(*ed.bcount, ed.bcap);
    int4 it = br.y < cnt ? ed.bitem[br.y] : make_int4(0, 0, 0, 0);
    // phase-1a inputs of row tid of the current item: loaded one item ahead, together
    // with the previous item's tile, so an item costs one round of global loads
    BlockRowIn pre{};
    if (br.y < cnt) {   // a workgroup without items loads nothing
        const BlockItemGeom g0 = block_item_geom(it, ed.bpw);
        pre = block_row_in(ed, it, g0, min(tid, g0.RW - 1));
    }
    for (int w = br.y; w < cnt; w += ed.bblocks) {
        // the next item's record is loaded while this one is solved
        const int4 nit = w + ed.bblocks < cnt ? ed.bitem[w + ed.bblocks] : it;
        const int ke = it.x, env = ke & 1;
        const int n = ed.n, np = ed.np;
        const int E = it.z;
        const int bpw = ed.bpw;
        const BlockItemGeom geo = block_item_geom(it, bpw);
        const int J = geo.J, b0 = geo.b0, nb = geo.nb, p0 = geo.p0, qlast = geo.qlast, RW = geo.RW;
        // idx / np as (idx * ceil(2^32 / np)) >> 32 in 64 bits: exact for idx * (np - 1) <
        // 2^32 (the tile bounds idx < np * RW by 160 KB / 8 and np <= 256 here)
        const unsigned long long pmagic = (0x100000000ull + (unsigned)np - 1) / (unsigned)np;
        const GPtr<int> X = gptr(ed.X + (size_t)ke * ed.Ecap);
        const double* curt = ed.curt;   // [T][np] current signal, sample-major
        const GMut<double> ye = gmut(ed.yend + (size_t)ke * 2 * np);   // [2][np]: Y_0, Y_{E-1} clamped
        const GMut<double> ys = gmut(ed.ysol + (size_t)ke * ed.Ecap * np);
        const GMut<double> sip = gmut(ed.sepin + (size_t)ke * ed.nbmax * np);   // [separator][np]: right-hand side
        __syncthreads();   // the previous item's tile is no longer read
        // 1a. knot positions of tile rows 0..RW-1, factors of rows p0..qlast, reciprocals of
        //     intervals p0-1..qlast+1 (the last one: the next separator's right-hand side;
        //     clamped where the group ends the system); row tid's inputs were prefetched,
        //     rows past the workgroup (groups of more than 8 blocks) are loaded here
        const int spline = ed.spline;
        const bool mirror = ed.boundary != 0;
        auto put_row = [&](int r, const BlockRowIn& v) {
            const int u = min(p0 + r, qlast);
            double lo, di, up;
            sys_row3(v.xm, v.x0, v.xp, u, E, spline, lo, di, up);
            sX[r] = v.xs;
            s_ih[r] = v.ih;
            s_rd[r] = v.rc.x;
            s_cp[r] = v.rc.y;
            s_lo[r] = lo;
        };
        if (tid < RW) put_row(tid, pre);
        for (int r = tid + kBlockThreads; r < RW; r += kBlockThreads) put_row(r, block_row_in(ed, it, geo, r));
        const int X0 = X[0], X1 = X[min(1, E - 1)], X2 = X[min(2, E - 1)];
        const int XE1 = X[E - 1], XE2 = X[max(E - 2, 0)], XE3 = X[max(E - 3, 0)];
        __syncthreads();
        // the next item's row inputs, in flight with this item's tile
        {
            const BlockItemGeom gn = block_item_geom(nit, bpw);   // nit is a valid item (the next, or this one)
            pre = block_row_in(ed, nit, gn, min(tid, gn.RW - 1));
        }
        // 1b. the tile, 16 bytes (two channels of one row) per lane: lane l of a wave
        //     fills bytes 16 l .. 16 l + 15 from the instruction's LDS base
        {
            const int pairs = RW * np / 2;
            const int lane = tid & 63;
            for (int q0 = tid & ~63; q0 < pairs; q0 += kBlockThreads) {
                const int q = q0 + lane;
                if (q < pairs) {
                    const int f = 2 * q;
                    const int r = (int)(((unsigned long long)f * pmagic) >> 32);
                    const int c = f - r * np;
                    __builtin_amdgcn_global_load_lds(curt + (size_t)sX[r] * np + c,
                                                     (__attribute__((address_space(3))) void*)(sY + 2 * q0), 16, 0, 0);
                }
            }
            __builtin_amdgcn_s_waitcnt(0);   // the tile has landed (vmcnt) before the barrier
        }
        __syncthreads();
        // end knots of E >= 6 clamped in the tile (row 0 is knot 0 in the first
        // group; rows rE - 2 .. rE are knots E-3 .. E-1 in the last, and the rows past
        // rE repeat knot E-1)
        const int rE = E - 1 - (p0 - 1);
        const bool has_left = p0 == 1, has_right = rE <= RW - 1;
        {   // y^ = 0 at the end rows (natural ends; not-a-knot end values replace them
            // in k_reduced) and the zero padding Z[0], Z[J+1] of the separator values
            const GMut<double> Z = gmut(zsep_base(ed, ke));
            for (int c = tid; c < np; c += kBlockThreads) {
                if (has_left) {
                    ys[c] = 0.0;
                    Z[c] = 0.0;
                }
                if (b0 + nb - 1 == J) {
                    ys[(unsigned)(E - 1) * (unsigned)np + (unsigned)c] = 0.0;
                    Z[(unsigned)(J + 1) * (unsigned)np + (unsigned)c] = 0.0;
                }
            }
        }
        if (!mirror && E >= 6 && (has_left || has_right)) {
            for (int c = tid; c < n; c += kBlockThreads) {
                if (has_left) {
                    const double y = clamp_end(sY[c], sY[np + c], sY[2 * np + c], X0, X1, X2, env, true);
                    sY[c] = y;
                    ye[c] = y;
                }
                if (has_right) {
                    const double y = clamp_end(sY[rE * np + c], sY[(rE - 1) * np + c], sY[(rE - 2) * np + c], XE1,
                                               XE2, XE3, env, false);
                    for (int r = rE; r < RW; ++r) sY[r * np + c] = y;
                    ye[np + c] = y;
                }
            }
            __syncthreads();
        }
        // 2. per (block, channel) lane; n * nb may exceed the workgroup
        for (int lane = tid; lane < n * nb; lane += kBlockThreads) {   // separator rows q+1 keep their right-hand side
            const int slot = lane / n, c = lane - slot * n;
            if (b0 + slot >= J) continue;
            const BlockGeom g = block_geom(E, b0 + slot);
            const int q = g.q, off = g.p - p0, len = q - g.p + 1;
            const double* Yc = sY + off * np + c;   // Yc[i * np] = Y(p - 1 + i)
            const double sl = (Yc[(len + 2) * np] - Yc[(len + 1) * np]) * s_ih[off + len + 1];
            const double slm = (Yc[(len + 1) * np] - Yc[len * np]) * s_ih[off + len];
            sip[(size_t)(b0 + slot + 1) * np + c] = 3.0 * (sl - slm);   // r of separator j = b + 1
        }
        __syncthreads();
        for (int lane = tid; lane < n * nb; lane += kBlockThreads) {
            const int slot = lane / n, c = lane - slot * n;
            const BlockGeom g = block_geom(E, b0 + slot);
            const int p = g.p, q = g.q, off = p - p0, len = q - p + 1;
            double* Yc = sY + off * np + c;   // Yc[i * np] = Y(p - 1 + i)
            // forward sweep, z_i written over Y(p+i) once no read needs it.  The first
            // and last rows are peeled (the pivot of row 0, the not-a-knot rows u = 1 and
            // u = N = E-2, which are only ever a block's first or last row), and each
            // row's LDS inputs are read one row ahead: the tile, the reciprocals and
            // the factors share one LDS array, so the compiler would not move a read
            // above the previous row's store
            const bool nak = ed.spline && E >= 4;
            auto nak_scale = [&](int u, double r) {
                if (nak) {
                    if (u == 1) r = r * (double)(X2 - X1);                 // h_1
                    else if (u == E - 2) r = r * (double)(XE2 - XE3);      // h_{N-1}
                }
                return r;
            };
            double ym = Yc[0], y0 = Yc[np], yp = Yc[2 * np];
            double zp;
            {
                const double sl = (yp - y0) * s_ih[off + 1];
                const double slm = (y0 - ym) * s_ih[off];
                zp = nak_scale(p, 3.0 * (sl - slm)) * s_rd[off];
                Yc[np] = zp;
            }
            if (len > 1) {
                // row i's inputs: yq = Y(p+i+1), ihq = ih of interval p+i, loq, rdq
                double yq = Yc[3 * np], ihq = s_ih[off + 2], loq = s_lo[off + 1], rdq = s_rd[off + 1];
                double ihp = s_ih[off + 1];
#pragma unroll 4
                for (int i = 1; i < len - 1; ++i) {
                    const double yn = Yc[(i + 3) * np], ihn = s_ih[off + i + 2];
                    const double lon = s_lo[off + i + 1], rdn = s_rd[off + i + 1];
                    ym = y0;
                    y0 = yp;
                    yp = yq;
                    const double sl = (yp - y0) * ihq;
                    const double slm = (y0 - ym) * ihp;
                    const double r = 3.0 * (sl - slm);
                    zp = fma(-loq, zp, r) * rdq;
                    Yc[(i + 1) * np] = zp;
                    yq = yn;
                    ihp = ihq;
                    ihq = ihn;
                    loq = lon;
                    rdq = rdn;
                }
                {   // the last row
                    const int i = len - 1;
                    ym = y0;
                    y0 = yp;
                    yp = yq;
                    const double sl = (yp - y0) * ihq;
                    const double slm = (y0 - ym) * ihp;
                    zp = fma(-loq, zp, nak_scale(p + i, 3.0 * (sl - slm))) * rdq;
                    Yc[(i + 1) * np] = zp;
                }
            }
            // backward sweep: each row's y goes straight to ysol (a wave's lanes are
            // consecutive channels of one row: coalesced; k_reduced reads the rows beside
            // the separators from there); the next row's z and factor are read one row
            // ahead
            double x = zp;   // the last row: y = z
            ys[(unsigned)(p + len - 1) * (unsigned)np + (unsigned)c] = x;
            if (len > 1) {
                double zq = Yc[(len - 1) * np], cpq = s_cp[off + len - 2];
#pragma unroll 4
                for (int i = len - 2; i >= 0; --i) {
                    const double zn = Yc[max(i, 0) * np], cpn = s_cp[off + max(i - 1, 0)];
                    x = fma(-cpq, x, zq);
                    ys[(unsigned)(p + i) * (unsigned)np + (unsigned)c] = x;
                    zq = zn;
                    cpq = cpn;
                }
            }
        }
        it = nit;
    }
}

// Second derivative M_u of knot row u of one channel after phase 2: phase-3
// combination inside blocks, the separator value at separator rows, and the end
// values (0 for natural ends, me[0] / me[1] for not-a-knot).
// row[u * rst]: the channel's block solution at row u; sv[j * sst]: its separator
// value j.
__device__ __forceinline__ double knot_M(const double* row, int rst, const double* sv, int sst, const double* AB,
                                         int u, int E, int J) {
    constexpr int L = kSolverBlock;
    if (u == 0 || u == E - 1) return 0.0;
    if (u % L == 0 && u / L <= J) return sv[(u / L) * sst];
    int b = (u - 1) / L;
    double m = row[(size_t)u * rst];
    if (b >= 1) m = fma(-sv[b * sst], AB[2 * (size_t)u], m);
    if (b < J) m = fma(-sv[(b + 1) * sst], AB[2 * (size_t)u + 1], m);
    return m;
}

// Phase 2: separator values, one lane per channel, one wave per (direction,
// envelope) (no division: the reduced factorisation comes from k_factor).  The
// chain over the J separators is serial per channel, so its inputs are staged in
// LDS a chunk of separators at a time by 16-byte global-to-LDS loads (they hold
// no registers: a whole chunk is one round trip), and the chain reads LDS only;
// the back substitution stages the forward values the same way.  Arithmetic per
// channel and separator as the oracle's phase 2.
// Staging doubles per chunk (separators x 3 x np forward, x np back): 48 np, within
// [768, 3072], for the widest epoch of the launch.  A small stage keeps more of
// these latency-bound waves resident; a wide epoch still gets 16 separators per
// chunk up to 64 channels.
int reduced_stage(int nmax) {
    const int v = 48 * channel_stride(nmax);
    return v < 768 ? 768 : (v > 3072 ? 3072 : v);
}

__global__ __launch_bounds__(64, 2) void k_reduced(const EpochDesc* __restrict__ eps, WorkLists wl, int kRedStage) {
    constexpr int L = kSolverBlock;
    constexpr int kCh = (kMaxChannels + 63) / 64;   // channels per lane
    extern __shared__ __align__(16) double sIn[];   // kRedStage doubles
    __shared__ double sA[64], sB[64];
    __shared__ int sLo[64], sUp[64];
    const ListCursor cu = list_cursor(wl, kListReduced);
    const int lane = threadIdx.x;
    for (int w = cu.start; w < cu.cnt; w += cu.stride) {
        const int4 it = cu.base[w];
        const EpochDesc& ed = eps[it.x];
        const int ke = it.y;
        const int n = ed.n, np = ed.np;
        const int E = ed.E[ke];
        const int J = (E - 3) / L;
        const int* X = ed.X + (size_t)ke * ed.Ecap;
        const double* Rlo = ed.red + ((size_t)ke * 3 + 0) * ed.nbmax;
        const double* Rrd = ed.red + ((size_t)ke * 3 + 1) * ed.nbmax;
        const double* Rcp = ed.red + ((size_t)ke * 3 + 2) * ed.nbmax;
        const double* sip = ed.sepin + (size_t)ke * ed.nbmax * np;   // [j][np]: right-hand side of separator j
        double* sv = zsep_base(ed, ke);                                    // [j][np]: Z[j]
        double* row = ed.ysol + (size_t)ke * ed.Ecap * np;                 // [u][np]: y^ (separator rows)
        const GPtr<int> Xg = gptr(X);
        double z[kCh];
#pragma unroll
        for (int cc = 0; cc < kCh; ++cc) z[cc] = 0.0;
        // forward sweep
        const int CH = min(64, kRedStage / (3 * np));
        for (int j0 = 1; j0 <= J; j0 += CH) {
            const int cnt = min(CH, J - j0 + 1);
            __syncthreads();   // the previous chunk is no longer read
            {   // per separator j the three rows [y^ of row 32 j - 1, right-hand side, y^ of
                // row 32 j + 1] (the block solution beside the separator, from ysol), 16
                // bytes (two channels of one row) per lane; np is a multiple of 4
                const int rowd = 3 * np;
                const unsigned long long rmagic = (0x100000000ull + (unsigned)rowd - 1) / (unsigned)rowd;
                const int pieces = cnt * rowd / 2;
                for (int q0 = 0; q0 < pieces; q0 += 64) {
                    const int q = q0 + lane;
                    if (q < pieces) {
                        const int f = 2 * q;
                        const int jj = (int)(((unsigned long long)f * rmagic) >> 32);
                        const int rem = f - jj * rowd;
                        const int comp = rem >= 2 * np ? 2 : (rem >= np ? 1 : 0);
                        const int c2 = rem - comp * np;
                        const int j = j0 + jj;
                        const double* src = comp == 1 ? sip + (size_t)j * np + c2
                                                      : row + (size_t)(j * L + comp - 1) * np + c2;
                        __builtin_amdgcn_global_load_lds(src, (__attribute__((address_space(3))) void*)(sIn + 2 * q0), 16, 0, 0);
                    }
                }
            }
            if (lane < cnt) {
                const int j = j0 + lane, s0 = j * L;
                const int xm = Xg[s0 - 1], x0 = Xg[s0], xp = Xg[s0 + 1];
                sLo[lane] = x0 - xm;
                sUp[lane] = xp - x0;
                sA[lane] = Rlo[j];
                sB[lane] = Rrd[j];
            }
            __builtin_amdgcn_s_waitcnt(0);   // the staged chunk has landed (vmcnt) before the barrier
            __syncthreads();
#pragma unroll
            for (int cc = 0; cc < kCh; ++cc) {
                const int c = lane + 64 * cc;
                if (c < n) {
                    double zc = z[cc];
                    for (int jj = 0; jj < cnt; ++jj) {
                        const int j = j0 + jj;
                        const double lo = (double)sLo[jj], up = (double)sUp[jj];
                        const double* y3 = sIn + (size_t)3 * jj * np + c;
                        const double D = fma(-up, y3[2 * np], fma(-lo, y3[0], y3[np]));
                        zc = (j == 1) ? D * sB[jj] : fma(-sA[jj], zc, D) * sB[jj];
                        sv[(size_t)j * np + c] = zc;
                    }
                    z[cc] = zc;
                }
            }
        }
        // back substitution: x_J = z_J, x_j = fma(-cp_j, x_{j+1}, z_j); every separator
        // value also becomes y^ of its knot row
#pragma unroll
        for (int cc = 0; cc < kCh; ++cc) {
            const int c = lane + 64 * cc;
            if (c < n && J > 0) row[(size_t)(J * L) * np + c] = z[cc];
        }
        const int CB = min(64, kRedStage / np);
        for (int j1 = J - 1; j1 >= 1; j1 -= CB) {
            const int cnt = min(CB, j1), jlo = j1 - cnt + 1;
            __syncthreads();   // the forward values are stored (vmcnt) and the previous chunk read
            {
                const double* src = sv + (size_t)jlo * np;
                const int pieces = cnt * np / 2;
                for (int q0 = 0; q0 < pieces; q0 += 64) {
                    const int q = q0 + lane;
                    if (q < pieces)
                        __builtin_amdgcn_global_load_lds(src + 2 * q, (__attribute__((address_space(3))) void*)(sIn + 2 * q0),
                                                         16, 0, 0);
                }
            }
            if (lane < cnt) sA[lane] = Rcp[jlo + lane];
            __builtin_amdgcn_s_waitcnt(0);
            __syncthreads();
#pragma unroll
            for (int cc = 0; cc < kCh; ++cc) {
                const int c = lane + 64 * cc;
                if (c < n) {
                    double xc = z[cc];
                    for (int j = j1; j >= jlo; --j) {
                        xc = fma(-sA[j - jlo], xc, sIn[(size_t)(j - jlo) * np + c]);
                        sv[(size_t)j * np + c] = xc;
                        row[(size_t)(j * L) * np + c] = xc;
                    }
                    z[cc] = xc;
                }
            }
        }
        if (ed.spline) {
            // not-a-knot end values from M_1, M_2 and M_{N-1}, M_N (N = E-2): y^ of the
            // end rows (k_block wrote 0 there); each lane reads back its own channel's
            // stores
            const double* AB = fac_plane(ed, ke, kFacAB);
            const double* IH = fac_plane(ed, ke, kFacIH);
            for (int c = lane; c < n; c += 64) {
                const double* rc = row + c;
                const double* svc = sv + c;
                double M1 = knot_M(rc, np, svc, np, AB, 1, E, J);
                double M2 = knot_M(rc, np, svc, np, AB, 2, E, J);
                double MN1 = knot_M(rc, np, svc, np, AB, E - 3, E, J);
                double MN = knot_M(rc, np, svc, np, AB, E - 2, E, J);
                double h0 = (double)(X[1] - X[0]), h1 = (double)(X[2] - X[1]);
                double a = (double)(X[E - 2] - X[E - 3]), b = (double)(X[E - 1] - X[E - 2]);
                row[c] = ((h0 + h1) * M1 - h0 * M2) * IH[2 * 1];
                row[(size_t)(E - 1) * np + c] = ((a + b) * MN - b * MN1) * IH[2 * (size_t)(E - 3)];
            }
        }
        __syncthreads();   // the next item reuses the staging arrays
    }
}

// Phase 3 + coefficients for the record reader (stop mode's evaluation): the
// spline record (a, b, c, d) of every interval of every
// channel (memReducedEpochs_dev.cu:775-793) with the shared reciprocals, from M of
// its two knots (knot_m).  A work item is 256 * kCoefItems (channel,
// interval) pairs of one (direction, envelope); lanes run over consecutive
// intervals, and the loads of G pairs are issued together.  The fused fix-mode
// evaluation forms M the same way inside its record build.

// M of a knot row after phase 2 from its y^, its phase-3 coefficients al, be
// and the separator values Z[i0], Z[i1] (zsep_index): the blocked solve's
// x_u = fma(-X_{b+1}, be_u, fma(-X_b, al_u, y_u)) at block rows, where a missing
// spike has al or be = 0 and Z[0] = Z[J+1] = +0 (so the term adds -0: x + (-0) = x),
// and y^ itself at the end and separator rows (both Z = +0, al = be = 0).
__device__ __forceinline__ double knot_m(double yh, double al, double be, double z0, double z1) {
    return fma(-z1, be, fma(-z0, al, yh));
}

__global__ __launch_bounds__(256) void k_coef(const EpochDesc* __restrict__ eps, WorkLists wl) {
    constexpr int L = kSolverBlock;
    constexpr int G = 4;   // pairs whose loads are issued together
    const ListCursor cu = list_cursor(wl, kListCoef);
    for (int w = cu.start; w < cu.cnt; w += cu.stride) {
        const int4 it = cu.base[w];
        const EpochDesc& ed = eps[it.x];
        const int ke = it.y;
        const int n = ed.n, np = ed.np;
        const int E = ed.E[ke];
        const int J = (E - 3) / L;
        const int base = it.z * 256 * kCoefItems;
        const GPtr<v2d> AB = gptr(reinterpret_cast<const v2d*>(fac_plane(ed, ke, kFacAB)));
        const GPtr<v2d> IH = gptr(reinterpret_cast<const v2d*>(fac_plane(ed, ke, kFacIH)));
        const size_t kerow = (size_t)ke * n;
        // idx / D as (idx * ceil(2^40 / D)) >> 40: exact for idx < n * D when n * D^2 < 2^40
        // (the error term idx * (magic * D - 2^40) stays below 2^40); otherwise divide
        const int D = E - 1;
        const unsigned long long dmagic = ((1ull << 40) + (unsigned long long)D - 1) / (unsigned long long)D;
        const bool use_magic = (unsigned long long)n * (unsigned long long)D * (unsigned long long)D < (1ull << 40);
        auto divD = [&](int idx) -> int {
            return use_magic ? (int)(((unsigned long long)idx * dmagic) >> 40) : idx / D;
        };
        // records (a, b, c, d) per interval (memReducedEpochs_dev.cu:775-793) with the
        // shared reciprocals
        const int nI = E - 1;
        const int items = n * nI;
        const GPtr<int> X = gptr(ed.X + (size_t)ke * ed.Ecap);
        const GPtr<double> Yh = gptr(ed.ysol + (size_t)ke * ed.Ecap * np);
        const GPtr<double> Z = gptr(zsep_base(ed, ke));
        const bool mirror = ed.boundary != 0;
        const int4 mir = mirror ? ed.mir[ke] : make_int4(0, 0, 0, 0);
        double4* out = reinterpret_cast<double4*>(ed.coef);
        for (int r0 = 0; r0 < kCoefItems; r0 += G) {
            // y^, al, be and Z[i0], Z[i1] of both knots; knot values, positions, reciprocals
            double h0[G], hk[G], a0[G], e0[G], a1[G], e1[G], z00[G], z01[G], z10[G], z11[G];
            double y0[G], y1[G], ih[G], i3h[G];
            int cs[G], js[G], x0[G], x1[G];
#pragma unroll
            for (int a = 0; a < G; ++a) {
                const int idx = min(base + (r0 + a) * 256 + (int)threadIdx.x, items - 1);
                const int c = divD(idx);
                const int j = idx - c * nI;
                cs[a] = c;
                js[a] = j;
                const int2 zj = zsep_index(j, E, J), zk = zsep_index(j + 1, E, J);
                h0[a] = Yh[(unsigned)j * (unsigned)np + (unsigned)c];
                hk[a] = Yh[(unsigned)(j + 1) * (unsigned)np + (unsigned)c];
                const v2d ab0 = AB[j], ab1 = AB[j + 1];
                a0[a] = ab0.x;
                e0[a] = ab0.y;
                a1[a] = ab1.x;
                e1[a] = ab1.y;
                z00[a] = Z[(unsigned)zj.x * (unsigned)np + (unsigned)c];
                z01[a] = Z[(unsigned)zj.y * (unsigned)np + (unsigned)c];
                z10[a] = Z[(unsigned)zk.x * (unsigned)np + (unsigned)c];
                z11[a] = Z[(unsigned)zk.y * (unsigned)np + (unsigned)c];
                // knot values: the signal at the knot, the clamped end values (rows
                // T + 2 ke and T + 2 ke + 1 of the same array), or with mirrored ends the
                // signal at the knot's source sample
                int r0w, r1w;
                if (mirror) {
                    r0w = knot_source(mir, j, E, X[j]);
                    r1w = knot_source(mir, j + 1, E, X[j + 1]);
                } else {
                    r0w = (E >= 6 && j == 0) ? ed.T + 2 * ke : X[j];
                    r1w = (E >= 6 && j + 1 == E - 1) ? ed.T + 2 * ke + 1 : X[j + 1];
                }
                y0[a] = ed.curt[(size_t)r0w * np + c];
                y1[a] = ed.curt[(size_t)r1w * np + c];
                x0[a] = X[j];
                x1[a] = X[j + 1];
                const v2d ihp = IH[j];
                ih[a] = ihp.x;
                i3h[a] = ihp.y;
            }
#pragma unroll
            for (int a = 0; a < G; ++a) {
                const int idx = base + (r0 + a) * 256 + (int)threadIdx.x;
                if (idx >= items) continue;
                const double m0 = knot_m(h0[a], a0[a], e0[a], z00[a], z01[a]);
                const double m1 = knot_m(hk[a], a1[a], e1[a], z10[a], z11[a]);
                const double h = (double)(x1[a] - x0[a]);
                const double sl = (y1[a] - y0[a]) * ih[a];
                double4 r;
                r.x = y0[a];
                r.y = sl - (h * (2.0 * m0 + m1)) * (1.0 / 3.0);
                r.z = m0;
                r.w = (m1 - m0) * i3h[a];
                out[(kerow + cs[a]) * ed.Ecap + js[a]] = r;
            }
        }
    }
}

// --------------------------------------------- fused evaluate + mean + update

__device__ __forceinline__ int knot_interval(const EpochDesc& ed, int ke, int t, int E) {
    // j = #knots strictly before t, minus one, clamped to [0, E-2]
    int w = t >> 6, bit = t & 63;
    unsigned long long m = ed.mask[(size_t)ke * ed.W + w];
    int before = ed.pre[(size_t)ke * ed.W + w] + __popcll(m & ((1ull << bit) - 1ull));
    int j = before - 1;
    j = j < 0 ? 0 : j;
    j = j > E - 2 ? E - 2 : j;
    return j;
}

// The same interval with mirrored ends: the knots left of t are the added ones
// (positions X[0 .. nl-1] and X[E-nr .. E-1]) and the interior extrema, which are
// the flag bits before t less the flag of sample 0.
__device__ __forceinline__ int knot_interval_mirror(const EpochDesc& ed, int ke, int t, int E) {
    const int* X = ed.X + (size_t)ke * ed.Ecap;
    const int4 m = ed.mir[ke];
    int w = t >> 6, bit = t & 63;
    unsigned long long mk = ed.mask[(size_t)ke * ed.W + w];
    int before = ed.pre[(size_t)ke * ed.W + w] + __popcll(mk & ((1ull << bit) - 1ull)) - (t > 0 ? 1 : 0);
    for (int i = 0; i < m.x; ++i) before += X[i] < t ? 1 : 0;
    for (int i = E - m.y; i < E; ++i) before += X[i] < t ? 1 : 0;
    int j = before - 1;
    j = j < 0 ? 0 : j;
    j = j > E - 2 ? E - 2 : j;
    return j;
}

__device__ __forceinline__ double horner4(const double4& r, double tau) {
    return fma(fma(fma(r.w, tau, r.z), tau, r.y), tau, r.x);
}

__device__ __forceinline__ int uniform(int v) { return __builtin_amdgcn_readfirstlane(v); }

// Inclusive prefix sum within each row of 16 lanes (DPP row shifts; lanes whose
// source would cross the row add 0).
__device__ __forceinline__ void row_scan16(int& x) {
    x += __builtin_amdgcn_update_dpp(0, x, 0x111, 0xF, 0xF, false);   // row_shr:1
    x += __builtin_amdgcn_update_dpp(0, x, 0x112, 0xF, 0xF, false);   // row_shr:2
    x += __builtin_amdgcn_update_dpp(0, x, 0x114, 0xF, 0xF, false);   // row_shr:4
    x += __builtin_amdgcn_update_dpp(0, x, 0x118, 0xF, 0xF, false);   // row_shr:8
}

// Inclusive prefix sum over lanes 0 .. 31: within each 16-lane row, then lanes
// 16 .. 31 add the first row's total (lanes above 31 are not used).
__device__ __forceinline__ void seg_scan(int& x) {
    row_scan16(x);
    const int r0 = __builtin_amdgcn_readlane(x, 15);
    if (threadIdx.x & 16) x += r0;
}
static_assert(2 * kEvalBatch <= 32, "the formation's prefix sums cover 32 segments");

// The value of lane + K of the same 16-lane row (DPP row shift; lanes whose
// source would leave the row get 0, and their results are not used).
template <int K>
__device__ __forceinline__ double row_from(double v) {
    const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(v), 0x100 + K, 0xF, 0xF, false);   // row_shl:K
    const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(v), 0x100 + K, 0xF, 0xF, false);
    return __hiloint2double(hi, lo);
}

// Phase B of one direction for a lane's 4 consecutive samples: both envelopes
// (upper u, lower l) at each sample, the mean of the two added to the
// accumulator.  Each sample uses the record of X_j < t <= X_{j+1}.  On entry ru /
// rl point at the records of the first sample's intervals and xu / xl at their
// left knots' positions (all four loads are issued together).  The knots are exactly the set flag bits,
// so sample t enters the next interval, whose left knot is t - 1 (offset 1), when
// bit t - 1 is set: bit s - 1 of `cu` / `cl` is the flag of sample ts + s - 1 (the
// caller clears the flag of sample 0, the left knot of the first interval).
// Offsets are integers below 2^53, so the running offset tau + 1 is the same
// double as (double)(t - X_j).
__device__ __forceinline__ double4 rec_load(const v2d* ab, const v2d* cd) {
    const v2d x = ab[0], y = cd[0];
    double4 r;
    r.x = x.x; r.y = x.y; r.z = y.x; r.w = y.y;
    return r;
}

// DIFF: also dd[s] = e+ - e- at each sample (stop mode's amplitude).
template <bool DIFF>
__device__ __forceinline__ void eval_pair(const v2d* __restrict__ rua, const v2d* __restrict__ ruc,
                                          const int* __restrict__ xu, unsigned cu, const v2d* __restrict__ rla,
                                          const v2d* __restrict__ rlc, const int* __restrict__ xl, unsigned cl,
                                          int ts, int tlast, double (&acc)[4], double (&dd)[4]) {
    const int ju = xu[0], jl = xl[0];
    double4 qu = rec_load(rua, ruc), ql = rec_load(rla, rlc);
    double tu = (double)(ts - ju), tl = (double)(ts - jl);
    {
        const double eu = horner4(qu, tu), el = horner4(ql, tl);
        acc[0] = acc[0] + (eu + el) * 0.5;
        if (DIFF) dd[0] = eu - el;
    }
#pragma unroll
    for (int s = 1; s < 4; ++s) {
        if (ts + s <= tlast) {
            tu = tu + 1.0;
            tl = tl + 1.0;
            if ((cu >> (s - 1)) & 1u) {
                ++rua;
                ++ruc;
                qu = rec_load(rua, ruc);
                tu = 1.0;
            }
            if ((cl >> (s - 1)) & 1u) {
                ++rla;
                ++rlc;
                ql = rec_load(rla, rlc);
                tl = 1.0;
            }
            const double eu = horner4(qu, tu), el = horner4(ql, tl);
            acc[s] = acc[s] + (eu + el) * 0.5;
            if (DIFF) dd[s] = eu - el;
        }
    }
}

// Fused records + evaluation (fix mode).  One workgroup = one tile of 64 * TW
// samples (TW = EpochDesc::etw flag words) x 16 / TW channels; lane = (channel,
// quad of 4 consecutive samples).  Directions are taken in batches of at most
// kEvalBatch whose records fit the LDS budget:
//   formation (wave 0): per (direction, envelope) the knot count, the tile's flag
//     words, the knots before each word and the tile's interval window
//     [jlo, jlo + nint); prefix sums over the directions' record and knot counts
//     close the batch;
//   phase A: every (direction, envelope, interval, quad of <= 4 channels) item of the batch
//     builds the spline record (a,b,c,d) of its interval in LDS from M at its two
//     knots (knot_m: y^ and the phase-3 coefficients of the knot row, the separator
//     values), the knot values and the shared reciprocals -- the same operations
//     as the record path, so the records are bit-identical;
//   phase B: each lane evaluates both envelopes of every valid direction of the
//     batch at its 4 samples and adds their mean in ascending direction order.
// STOP (stop mode, TW = 1, the host guarantees it): the mean goes to ed.mean
// instead of updating the signal, and per valid direction each wave sums the
// squared envelope differences of its 4 channels in channel order (shuffles), wave
// 0 adds the 4 waves' sums in order and stores the block's sum for k_stop_amp:
// the amplitude's channel order of groups of 4 and blocks of 16.
template <bool STOP>
__global__ __launch_bounds__(256, 4) void k_eval_fused(const EpochDesc* __restrict__ eps,
                                                   const int2* __restrict__ tab) {
    constexpr int NS = 2 * kEvalBatch;   // segments (direction, envelope) per batch
    constexpr int NT = 256;              // threads per workgroup
    // spline records, segment s: [channel][stride] from its record base, as two
    // arrays of 16-byte halves (a, b) and (c, d): a wave's record reads and writes
    // are then 16 bytes apart, not every other 16 bytes of 32-byte records
    __shared__ v2d recAB[kEvalCap];
    __shared__ v2d recCD[kEvalCap];
    __shared__ int xk[kEvalXkCap];       // segment s: knot positions jlo .. jlo + nint
    __shared__ int4 sSegA[NS];           // item prefix, -, nint, jlo
    __shared__ int4 sSegB[NS];           // record base, row stride, E, knot-position base
    // segment s: its rows of ysol, zsep and fac, so phase A addresses are a base plus a
    // 32-bit offset (one (direction, envelope)'s rows stay far below 4 GB)
    __shared__ const double* sSegP[NS][3];
    __shared__ int sIp[NS + 1];          // item prefix; entries past the batch hold the total
    // segment s, tile word w at [s * TW + w]: the knots before the end of the word,
    // less jlo + 1, and the flag word (sample 0's flag cleared)
    __shared__ int sCum[kEvalSegWords];
    __shared__ unsigned long long sMsk[kEvalSegWords];
    __shared__ int sNb;
    __shared__ unsigned sValid;   // bit s: segment s of the batch is valid (both envelopes of its direction)
    __shared__ int cE[2 * kEvalChunk], cPr[2 * kEvalChunk];   // staged formation inputs
    __shared__ unsigned long long cM[kEvalChunkWords];          // entry i, word w at [i * TW + w]
    const int2 br = tab[blockIdx.x];
    const EpochDesc& ed = eps[br.x];
    const EpochState* st = ed.st;
    if (st->done) return;
    const int n = ed.n, T = ed.T, K = ed.K, W = ed.W;
    const double* curt = ed.curt;   // [T + 4K][np]: the signal sample-major, then the clamped end values
    // descriptor fields of the record build, read once: the descriptor is reached
    // through a generic pointer, so every LDS store would otherwise force a reload
    const int Ecap = ed.Ecap, np = ed.np, nbmax = ed.nbmax;
    const double* const ysol = ed.ysol;
    const double* const zsep = ed.sep;
    const double* const fac = ed.fac;
    const int* const Xall = ed.X;
    const int TW = ed.etw;
    const int lgq = 4 + (TW >> 1);       // log2 quads per channel: 4, 5, 6 for TW = 1, 2, 4
    const int CG = 16 / TW;
    const int nsmax = min(NS, kEvalSegWords / TW);                 // segments per batch
    const int chunk = min(kEvalChunk, kEvalChunkWords / (2 * TW));  // directions per staged chunk
    const int groups = (n + CG - 1) / CG;
    const int tile = br.y / groups;
    const int cg = br.y - tile * groups;
    const int cgn = min(CG, n - cg * CG);   // channels of this group
    const int w0 = tile * TW;
    const int t0 = 64 * w0;
    const int tlast = min(t0 + 64 * TW - 1, T - 1);
    const int tid = threadIdx.x;
    const int wbase = uniform(tid & ~63);
    // lane -> (channel cl, quad q), channels fastest: a wave's lanes share 64 / CG
    // quads' crossing patterns and read CG channel rows of one interval (distinct
    // banks: the row stride is odd); in stop mode (CG = 16) a quad's channels are one
    // 16-lane row for the amplitude sums
    const int cl = tid & (CG - 1);
    const int q = tid >> (8 - lgq);
    const int c = cg * CG + cl;
    const int ts = t0 + 4 * q;
    const int wi = q >> 4;               // tile word of the quad
    const int qsh = 4 * (q & 15);        // bit of sample ts in its flag word
    const bool lane_on = cl < cgn && ts <= tlast;
    const int nq = (cgn + 3) >> 2;              // channel quads of the group
    const int qmag = (65536 + nq - 1) / nq;     // rem / nq as (rem * qmag) >> 16: exact for rem < 2^15
    double acc[4] = {0.0, 0.0, 0.0, 0.0};
    double nv = 0.0;
    int kc1 = 0;   // end of the staged chunk of formation inputs
    const int G16 = (n + 15) >> 4;
    for (int kb = 0; kb < K;) {
        if (tid < 64) {   // formation
            if (kb >= kc1) {   // stage the next chunk of directions: lane (direction, envelope)
                const int kk = kb + (tid >> 1), ke = 2 * kk + (tid & 1);
                int E = 0, pr = 0;
                unsigned long long m[4] = {0, 0, 0, 0};
                if (tid < 2 * chunk && kk < K) {
                    E = gptr(ed.E)[ke];
                    pr = gptr(ed.pre)[(size_t)ke * W + w0];
                    const GPtr<unsigned long long> mk = gptr(ed.mask);
#pragma unroll
                    for (int w = 0; w < 4; ++w)
                        if (w < TW && w0 + w < W) m[w] = mk[(size_t)ke * W + w0 + w];
                }
                if (tid < 2 * chunk) {
                    cE[tid] = E;
                    cPr[tid] = pr;
#pragma unroll
                    for (int w = 0; w < 4; ++w)
                        if (w < TW) cM[tid * TW + w] = m[w];
                }
                kc1 = kb + chunk;
            }
            const int k = kb + (tid >> 1), env = tid & 1;
            // a batch stays inside the staged chunk
            const bool have = tid < nsmax && k < K && k < kc1;
            const int ci = 2 * (k - (kc1 - chunk)) + env;   // chunk entry
            int E = 0, pr = 0, inside = 0;
            int cum[4];
            unsigned long long msk[4];
            if (have) {
                E = cE[ci];
                pr = cPr[ci];
            }
            int run = pr;
#pragma unroll
            for (int w = 0; w < 4; ++w) {
                msk[w] = 0;
                if (have && w < TW && w0 + w < W) {
                    const unsigned long long m = cM[ci * TW + w];
                    const int lim = tlast - 64 * (w0 + w);   // bits below tlast in this word
                    msk[w] = m;
                    inside += lim >= 64 ? (int)__popcll(m) : (lim > 0 ? (int)__popcll(m & ((1ull << lim) - 1ull)) : 0);
                    run += (int)__popcll(m);
                }
                cum[w] = run;
            }
            // sample 0 is the left knot of interval 0, not a crossing into it: without
            // its flag, the knots below t = 0 count 1 like those below t = 1
            if (w0 == 0) msk[0] &= ~1ull;
            const int lo = min(max(pr - 1, 0), E - 2);
            const int hi = min(max(pr + inside - 1, 0), E - 2);
            const int ni = hi - lo + 1;
            const int Ep = __builtin_amdgcn_mov_dpp(E, 0xB1, 0xF, 0xF, false);   // quad_perm [1,0,3,2]: the other envelope
            const bool valid = have && E > 3 && Ep > 3;
            const int stride = ni | 1;   // odd: a wave's channels read disjoint banks
            const int need = valid ? cgn * stride : 0;
            const int nxk = valid ? ni + 1 : 0;
            const int nit = valid ? nq * ni : 0;
            int pneed = need, pxk = nxk, pit = nit;
            // inclusive prefix sums over the NS segments (lanes 0 .. NS - 1)
            seg_scan(pneed);
            seg_scan(pxk);
            seg_scan(pit);
            // leading directions whose records and knot positions fit (the first always does)
            const unsigned long long fit =
                __ballot(have && (tid & 1) && pneed <= kEvalCap && pxk <= kEvalXkCap);
            const int nb = max((int)__popcll(fit), 1);
            const int total = __builtin_amdgcn_readlane(pit, 2 * nb - 1);
            if (tid < 2 * nb) {
                sSegA[tid] = make_int4(pit - nit, 0, ni, lo);
                sSegB[tid] = make_int4(pneed - need, stride, E, pxk - nxk);
                const size_t kes = (size_t)(2 * kb + tid);   // ke of the segment
                sSegP[tid][0] = ysol + kes * Ecap * np;
                sSegP[tid][1] = zsep + kes * (nbmax + 1) * np;
                sSegP[tid][2] = fac + kes * kFacPlanes * Ecap * 2;
#pragma unroll
                for (int w = 0; w < 4; ++w) {
                    if (w < TW) {
                        sCum[tid * TW + w] = cum[w] - lo - 1;
                        sMsk[tid * TW + w] = msk[w];
                    }
                }
            }
            if (tid <= NS) sIp[tid] = tid < 2 * nb ? pit - nit : total;
            const unsigned long long vbits = __ballot(valid);
            if (tid == 0) {
                sNb = nb;
                sValid = (unsigned)vbits;
            }
        }
        __syncthreads();
        const int nb = uniform(sNb);
        const int total = uniform(sIp[NS]);
        const int lwv = tid & 63;
        const int ipl = sIp[min(lwv, NS)];   // lane s: the first item of segment s
        // phase A: items (interval, quad of <= 4 channels), quad fastest.  The
        // interval's positions, per-row records and reciprocals are loaded once for
        // its channels; the channel rows of the knot-major arrays are adjacent words.
        // Lanes past the last item redo it and store nothing, so every lane's loads
        // are in range and no register needs a default.
        for (int a0 = 0; wbase + NT * a0 < total; ++a0) {   // wave-uniform bound
            const int fw = wbase + NT * a0;                  // the wave's first item
            const int fr = tid + NT * a0;
            const bool on = fr < total;
            const int f = min(fr, total - 1);
            // the item's segment (the last s with ip[s] <= f): lane s of the wave holds
            // boundary ip[s]; those at or below the wave's first item are counted by a
            // ballot, the few inside the wave's items per lane
            const bool bin = lwv >= 1 && lwv < 2 * nb;
            int seg = (int)__popcll(__ballot(bin && ipl <= fw));
            for (unsigned long long in = __ballot(bin && ipl > fw && ipl <= fw + 63); in; in &= in - 1) {
                const int bnd = __builtin_amdgcn_readlane(ipl, (int)__builtin_ctzll(in));
                seg += (f >= bnd) ? 1 : 0;
            }
            const int4 A = sSegA[seg];
            const int4 B = sSegB[seg];
            const int E = B.z;
            const int ke = 2 * (kb + (seg >> 1)) + (seg & 1);
            const int rem = f - A.x;
            const int ii = (rem * qmag) >> 16;
            const int qd = rem - ii * nq;
            const int j = A.w + ii;
            const int c0 = cg * CG + 4 * qd;
            const int nch = min(4, cgn - 4 * qd);
            const int J = (E - 3) / kSolverBlock;
            // 4 channels from c0 (a multiple of 4) in rows of np (a multiple of 4)
            // words: 32-byte aligned; channels past the group are padding or the
            // next group's and are not used
            const double* const yb = sSegP[seg][0];
            const double* const zb0 = sSegP[seg][1];
            const double* const Fb = sSegP[seg][2];   // planes of (direction, envelope) ke
            const GPtr<int> Xr = gptr(Xall + (size_t)ke * Ecap);
            // the knot positions first: the knot-value loads below depend on them,
            // and the other loads are issued while they are in flight
            const int x0 = Xr[j];
            const int x1 = Xr[j + 1];
            const GPtr<v2d> ABr = gptr(reinterpret_cast<const v2d*>(Fb + (unsigned)(kFacAB * Ecap * 2 + 2 * j)));
            const GPtr<v2d> IHr = gptr(reinterpret_cast<const v2d*>(Fb + (unsigned)(kFacIH * Ecap * 2 + 2 * j)));
            const v2d ab0 = ABr[0], ab1 = ABr[1], ihp = IHr[0];
            const int nr = np >> 2;   // double4 per row
            const GPtr<v4d> Yr = gptr(reinterpret_cast<const v4d*>(yb + (unsigned)(j * np + c0)));
            const v4d yh0 = Yr[0], yh1 = Yr[nr];
            // separator values Z_b, Z_b+1 of the interval's block b: the pair of every
            // block row among its two knots; a knot outside the blocks (an end row or a
            // separator) has al = be = 0, so its M is its y^ whatever the pair
            const int zb = min(j >> 5, J);
            const GPtr<v4d> Zr = gptr(reinterpret_cast<const v4d*>(zb0 + (unsigned)(zb * np + c0)));
            const v4d za = Zr[0], zc = Zr[nr];
            __builtin_amdgcn_sched_barrier(0);   // keep the independent loads above
            // knot values: the signal at the two knots (sample-major rows), or the
            // clamped end values (E >= 6), rows T + 2 ke and T + 2 ke + 1 of the same
            // array
            const int r0 = (E >= 6 && j == 0) ? T + 2 * ke : x0;
            const int r1 = (E >= 6 && j + 1 == E - 1) ? T + 2 * ke + 1 : x1;
            const v4d yv0 = *gptr(reinterpret_cast<const v4d*>(curt + (unsigned)(r0 * np + c0)));
            const v4d yv1 = *gptr(reinterpret_cast<const v4d*>(curt + (unsigned)(r1 * np + c0)));
            if (on) {
                const double v0[4] = {yv0.x, yv0.y, yv0.z, yv0.w}, v1[4] = {yv1.x, yv1.y, yv1.z, yv1.w};
                const double y0[4] = {yh0.x, yh0.y, yh0.z, yh0.w}, y1[4] = {yh1.x, yh1.y, yh1.z, yh1.w};
                const double z0[4] = {za.x, za.y, za.z, za.w}, z1[4] = {zc.x, zc.y, zc.z, zc.w};
                const int i = j - A.w;
                if (qd == 0) {   // first quad of the group: knot positions
                    xk[B.w + i] = x0;
                    if (i == A.z - 1) xk[B.w + A.z] = x1;
                }
                const double h = (double)(x1 - x0);
#pragma unroll
                for (int q2 = 0; q2 < 4; ++q2) {
                    if (q2 < nch) {
                        const double mj = knot_m(y0[q2], ab0.x, ab0.y, z0[q2], z1[q2]);
                        const double mk = knot_m(y1[q2], ab1.x, ab1.y, z0[q2], z1[q2]);
                        const double slp = (v1[q2] - v0[q2]) * ihp.x;
                        const int at = B.x + (4 * qd + q2) * B.y + i;
                        recAB[at] = v2d{v0[q2], slp - (h * (2.0 * mj + mk)) * (1.0 / 3.0)};
                        recCD[at] = v2d{mj, (mk - mj) * ihp.y};
                    }
                }
            }
        }
        __syncthreads();
        // phase B.  Samples 0 and T-1 are knots of every direction, so the knots
        // below a sample t in 1..T-1 number 1..E-1 (and sample 0's cleared flag counts
        // 1 for t = 0): the interval index knots below - 1 lies in 0..E-2.  The flag
        // word shifted to the quad's first sample gives the knots at or after it in
        // the word and, in its low 3 bits, the quad's interval crossings.
        const unsigned vmask = (unsigned)uniform((int)sValid);
        for (int d = 0; d < nb; ++d) {
            if (!((vmask >> (2 * d)) & 1u)) continue;
            nv = nv + 1.0;
            double dd[4] = {0.0, 0.0, 0.0, 0.0};
            if (lane_on) {
                // both envelopes' flag words, knot counts and record offsets in one
                // round of LDS reads; then the first interval (in the tile's window)
                // and the crossing flags
                const unsigned long long m0 = sMsk[2 * d * TW + wi], m1 = sMsk[(2 * d + 1) * TW + wi];
                const int c0 = sCum[2 * d * TW + wi], c1 = sCum[(2 * d + 1) * TW + wi];
                const int4 B0 = sSegB[2 * d], B1 = sSegB[2 * d + 1];
                __builtin_amdgcn_sched_barrier(0);
                const unsigned long long h0 = m0 >> qsh, h1 = m1 >> qsh;
                const int i0 = c0 - (int)__popcll(h0), i1 = c1 - (int)__popcll(h1);
                const unsigned x0 = (unsigned)h0 & 7u, x1 = (unsigned)h1 & 7u;
                const int r0 = B0.x + __umul24(cl, B0.y) + i0, r1 = B1.x + __umul24(cl, B1.y) + i1;
                eval_pair<STOP>(recAB + r0, recCD + r0, xk + B0.w + i0, x0, recAB + r1, recCD + r1, xk + B1.w + i1,
                                x1, ts, tlast, acc, dd);
            }
            if (STOP) {
                // the block's amplitude sum per sample: the quad's 16 channels are one
                // 16-lane row (lane = channel), summed in the fixed order: channels in
                // groups of 4, ((c0 + c1) + c2) + c3, then the groups, ((g0 + g1) + g2) +
                // g3 (row shifts; the sums of lanes not at a group start are unused).
                // Samples past the tile and absent channels carry 0, which adds exactly.
#pragma unroll
                for (int s2 = 0; s2 < 4; ++s2) {
                    const double x = dd[s2] * dd[s2];
                    const double g = ((x + row_from<1>(x)) + row_from<2>(x)) + row_from<3>(x);
                    const double b = ((g + row_from<4>(g)) + row_from<8>(g)) + row_from<12>(g);
                    if (cl == 0 && ts + s2 <= tlast) ed.ampb[((size_t)(kb + d) * G16 + cg) * T + ts + s2] = b;
                }
            }
        }
        __syncthreads();
        kb += nb;
    }
    if (ed.stats && br.y == 0 && tid == 0) {
        unsigned long long* sl = ed.stats + 4 * (blockIdx.x % kStatSlots);
        atomicAdd(&sl[2], (unsigned long long)nv * (unsigned long long)n * (unsigned long long)T);
        atomicAdd(&sl[3], (unsigned long long)n * (unsigned long long)T);
    }
    if (!(cl < cgn)) return;
    if (STOP) {   // mean = sum / nvalid (0/0 = NaN with no valid direction: the decision then stops)
#pragma unroll
        for (int s2 = 0; s2 < 4; ++s2)
            if (ts + s2 <= tlast) ed.mean[(size_t)c * T + ts + s2] = acc[s2] / nv;
        return;
    }
    const GMut<double> cur = gmut(cur_slot(ed, st->imf));
    const bool last = (ed.mode == kModeFix) && (st->nsift == ed.S - 1);
    const GMut<double> nxt = gmut(cur_slot(ed, st->imf + 1));
    const GMut<double> run = gmut(ed.running);
    // the lane's samples (and residuals) are loaded together, from in-range indices
    // (global address space: flat accesses would make every store wait for all loads)
    double cv[4], rv[4];
#pragma unroll
    for (int s = 0; s < 4; ++s) cv[s] = cur[(size_t)c * T + min(ts + s, tlast)];
    if (last) {
#pragma unroll
        for (int s = 0; s < 4; ++s) rv[s] = run[(size_t)c * T + min(ts + s, tlast)];
    }
#pragma unroll
    for (int s = 0; s < 4; ++s) {
        const int t = ts + s;
        if (t <= tlast) {
            const size_t idx = (size_t)c * T + t;
            const double v = cv[s] - acc[s] / nv;
            cur[idx] = v;
            if (last) {
                const double rr = rv[s] - v;
                run[idx] = rr;
                nxt[idx] = rr;
            }
        }
    }
}

// Fix-mode state advance: one thread per epoch; also clears the work lists
// for the next step.
__global__ void k_advance(const EpochDesc* __restrict__ eps, int count, WorkLists wl) {
    int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e == 0) zero_counters(wl);
    if (e >= count) return;
    const EpochDesc& ed = eps[e];
    *ed.bcount = 0;
    EpochState* st = ed.st;
    if (st->done) return;
    st->pass += 1;
    st->nsift += 1;
    if (st->nsift == ed.S) {
        ed.sift_counts[st->imf] = ed.S;
        st->nsift = 0;
        st->imf += 1;
        if (st->imf == ed.N_imf - 1) {
            st->done = 1;
            st->n_imf_out = ed.N_imf;
            atomicAdd(ed.ndone, 1);
        }
    }
}

// ------------------------------------------------------------ stop mode
//
// memd.m-style criterion (not in the reference, which only runs fixed sift
// counts; see the oracle's header).  Per sample t: mean(c,t) = sum_valid
// ((e+ + e-)/2) / nvalid, amp(t) = sum_valid sqrt(sum_c (e+ - e-)^2)/2 / nvalid,
// q(t) = sum_c mean(c,t)^2, all sums in ascending order.

// One wavefront per 64 samples; every lane owns one sample and walks the
// channels in order, so the channel sums match the oracle bit for bit.
__global__ __launch_bounds__(64) void k_eval_stop(const EpochDesc* __restrict__ eps, const int2* __restrict__ tab) {
    extern __shared__ double sums[];   // [n][64]
    const int2 br = tab[blockIdx.x];
    const EpochDesc& ed = eps[br.x];
    const EpochState* st = ed.st;
    if (st->done) return;
    const int n = ed.n, T = ed.T, K = ed.K;
    const int lane = threadIdx.x;
    const int t = br.y * 64 + lane;
    const int tq = min(t, T - 1);
    const double4* coef = reinterpret_cast<const double4*>(ed.coef);
    for (int c = 0; c < n; ++c) sums[c * 64 + lane] = 0.0;
    double amps = 0.0, nv = 0.0;
    for (int k = 0; k < K; ++k) {
        if (!dir_valid(ed.E, k)) continue;
        nv = nv + 1.0;
        int ke0 = 2 * k, ke1 = 2 * k + 1;
        int j0, j1;
        if (ed.boundary != 0) {
            j0 = knot_interval_mirror(ed, ke0, tq, ed.E[ke0]);
            j1 = knot_interval_mirror(ed, ke1, tq, ed.E[ke1]);
        } else {
            j0 = knot_interval(ed, ke0, tq, ed.E[ke0]);
            j1 = knot_interval(ed, ke1, tq, ed.E[ke1]);
        }
        double tau0 = (double)(tq - ed.X[(size_t)ke0 * ed.Ecap + j0]);
        double tau1 = (double)(tq - ed.X[(size_t)ke1 * ed.Ecap + j1]);
        // sum_c (e+ - e-)^2: sequential within groups of 4 channels, 4 groups to a
        // block of 16, blocks in order (the fused evaluation's order)
        double s2 = 0.0, bsum = 0.0, gsum = 0.0;
        for (int c = 0; c < n; ++c) {
            double ep = horner4(coef[((size_t)ke0 * n + c) * ed.Ecap + j0], tau0);
            double em = horner4(coef[((size_t)ke1 * n + c) * ed.Ecap + j1], tau1);
            double d = ep - em;
            const double x = d * d;
            gsum = (c & 3) == 0 ? x : gsum + x;
            if ((c & 3) == 3 || c == n - 1) {
                const int g = (c >> 2) & 3;
                bsum = g == 0 ? gsum : bsum + gsum;
                if (g == 3 || c == n - 1) s2 = (c >> 4) == 0 ? bsum : s2 + bsum;
            }
            sums[c * 64 + lane] = sums[c * 64 + lane] + (ep + em) * 0.5;
        }
        amps = amps + sqrt(s2) * 0.5;
    }
    if (ed.stats && br.y == 0 && lane == 0) {
        unsigned long long* sl = ed.stats + 4 * (blockIdx.x % kStatSlots);
        atomicAdd(&sl[2], (unsigned long long)nv * (unsigned long long)n * (unsigned long long)T);
        atomicAdd(&sl[3], (unsigned long long)n * (unsigned long long)T);
    }
    if (t >= T) return;
    double q = 0.0;
    for (int c = 0; c < n; ++c) {
        double m = sums[c * 64 + lane] / nv;   // 0/0 = NaN with no valid direction, as fix mode
        ed.mean[(size_t)c * T + t] = m;
        q = q + m * m;
    }
    ed.tstat[2 * (size_t)t + 0] = q;
    ed.tstat[2 * (size_t)t + 1] = (nv > 0.0) ? amps / nv : 0.0;
    if (t == 0) ed.stopred[0] = (int)nv;
}

// Stop mode after the fused evaluation: per sample, the amplitude from the block
// sums (blocks of 16 channels in order, per valid direction in order) and
// |mean(t)|^2 (channels in order); the same outputs as k_eval_stop's epilogue.
__global__ __launch_bounds__(64) void k_stop_amp(const EpochDesc* __restrict__ eps, const int2* __restrict__ tab) {
    const int2 br = tab[blockIdx.x];
    const EpochDesc& ed = eps[br.x];
    if (ed.st->done) return;
    const int n = ed.n, T = ed.T, K = ed.K, G16 = (n + 15) >> 4;
    const int t = br.y * 64 + (int)threadIdx.x;
    if (t >= T) return;
    double amps = 0.0, nv = 0.0;
    for (int k = 0; k < K; ++k) {
        if (!dir_valid(ed.E, k)) continue;
        nv = nv + 1.0;
        const double* pb = ed.ampb + (size_t)k * G16 * T + t;
        double s2 = pb[0];
        for (int g = 1; g < G16; ++g) s2 = s2 + pb[(size_t)g * T];
        amps = amps + sqrt(s2) * 0.5;
    }
    double q = 0.0;
    for (int c = 0; c < n; ++c) {
        const double m = ed.mean[(size_t)c * T + t];
        q = q + m * m;
    }
    ed.tstat[2 * (size_t)t + 0] = q;
    ed.tstat[2 * (size_t)t + 1] = (nv > 0.0) ? amps / nv : 0.0;
    if (t == 0) ed.stopred[0] = (int)nv;
}

// One workgroup per epoch: the stop decision and the state transition.  Epochs in
// fix mode on these kernels (mirrored ends) sift exactly S times per IMF, as
// k_advance does.
__global__ __launch_bounds__(256) void k_stop_decide(const EpochDesc* __restrict__ eps, int count) {
    int e = blockIdx.x;
    if (e >= count) return;
    const EpochDesc& ed = eps[e];
    EpochState* st = ed.st;
    __shared__ int s_done;
    __shared__ int s_zero, s_cnt, s_any2;
    if (threadIdx.x == 0) {
        s_done = st->done;
        s_zero = 0;
        s_cnt = 0;
        s_any2 = 0;
    }
    __syncthreads();
    if (s_done) return;
    if (ed.mode == kModeFix) {
        if (threadIdx.x != 0) return;
        const int imf = st->imf;
        st->act_imf = imf;
        st->pass += 1;
        st->action = kActContinue;
        st->nsift += 1;
        if (st->nsift == ed.S) {
            ed.sift_counts[imf] = ed.S;
            st->nsift = 0;
            st->imf = imf + 1;
            if (imf + 1 == ed.N_imf - 1) {
                st->done = 1;
                st->n_imf_out = ed.N_imf;
                atomicAdd(ed.ndone, 1);
            }
        }
        return;
    }
    const int T = ed.T, K = ed.K;
    const bool nvzero = ed.stopred[0] == 0;
    // every amp(t) nonzero? (memd.m divides by amp only then)
    int zero = 0;
    for (int t = threadIdx.x; t < T; t += blockDim.x)
        if (!(ed.tstat[2 * (size_t)t + 1] != 0.0)) zero = 1;
    if (zero) atomicOr(&s_zero, 1);
    __syncthreads();
    const bool amp_nonzero = (s_zero == 0);
    int cnt = 0, any2 = 0;
    for (int t = threadIdx.x; t < T; t += blockDim.x) {
        double sx = sqrt(ed.tstat[2 * (size_t)t + 0]);
        if (amp_nonzero) sx = sx / ed.tstat[2 * (size_t)t + 1];
        if (sx > ed.sd) ++cnt;
        if (sx > ed.sd2) any2 = 1;
    }
    atomicAdd(&s_cnt, cnt);
    if (any2) atomicOr(&s_any2, 1);
    __syncthreads();
    if (threadIdx.x != 0) return;
    bool all_small = true, anynem = false;
    for (int k = 0; k < K; ++k) {
        int nem = (ed.Eraw[2 * k] - 2) + (ed.Eraw[2 * k + 1] - 2);
        if (nem >= 3) all_small = false;
        if (nem > 2) anynem = true;
    }
    const int imf = st->imf;
    st->act_imf = imf;
    st->pass += 1;
    if (st->nsift == 0 && all_small) {   // memd.m stop_emd: no IMF left to extract
        st->action = kActEnd;
        st->done = 1;
        st->n_imf_out = imf + 1;
        atomicAdd(ed.ndone, 1);
        return;
    }
    bool stop;
    if (nvzero) {
        stop = true;
    } else {
        stop = !(((((double)s_cnt / (double)T) > ed.tol) || s_any2) && anynem);
    }
    bool finish;
    if (stop) {
        st->action = kActFinishImf;
        finish = true;
    } else {
        st->action = kActContinue;
        st->nsift += 1;
        finish = (st->nsift == ed.S);
    }
    if (finish) {
        ed.sift_counts[imf] = st->nsift;
        st->nsift = 0;
        st->imf = imf + 1;
        if (imf + 1 == ed.N_imf - 1) {
            st->done = 1;
            st->n_imf_out = ed.N_imf;
            atomicAdd(ed.ndone, 1);
        }
    }
    // k_stop_apply reads "IMF finished" as imf != act_imf
}

// Elementwise application of the decision (grid = the init table: n*T per epoch).
__global__ __launch_bounds__(256) void k_stop_apply(const EpochDesc* __restrict__ eps, const int2* __restrict__ tab) {
    const int2 br = tab[blockIdx.x];
    const EpochDesc& ed = eps[br.x];
    const EpochState* st = ed.st;
    int act = st->action;
    if (act == kActEnd || act == kActNone) return;
    size_t nT = (size_t)ed.n * ed.T;
    size_t i = (size_t)br.y * 256 + threadIdx.x;
    if (i >= nT) return;
    const int imf = st->act_imf;
    const bool finish = st->imf != imf;
    double* cur = cur_slot(ed, imf);
    double v = cur[i];
    if (act == kActContinue) {
        v = v - ed.mean[i];
        cur[i] = v;
    }
    if (finish) {
        double r = ed.running[i] - v;
        ed.running[i] = r;
        cur_slot(ed, imf + 1)[i] = r;
    }
}

// End of a stop-mode step: epochs already done decide nothing in the next step,
// and the work lists are cleared.
__global__ void k_stop_reset(const EpochDesc* __restrict__ eps, int count, WorkLists wl) {
    int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e == 0) zero_counters(wl);
    if (e < count) *eps[e].bcount = 0;
    if (e < count && eps[e].st->done) eps[e].st->action = kActNone;
}

// ------------------------------------------------------------ misc kernels


__global__ void k_gather_probe(const EpochDesc* __restrict__ eps, int count, EpochState* __restrict__ states,
                               int* __restrict__ sift, const int* __restrict__ sift_off) {
    const int e = blockIdx.x;
    if (e >= count) return;
    const EpochDesc& ed = eps[e];
    if (threadIdx.x == 0) states[e] = *ed.st;
    const int at = sift_off[e];
    for (int i = threadIdx.x; i < ed.N_imf; i += blockDim.x) sift[at + i] = ed.sift_counts[i];
}

__global__ void k_mark_done(EpochState* st) {
    if (threadIdx.x == 0) {
        st->done = 1;
        st->action = kActNone;
    }
}

__global__ void k_zero_ints(int* __restrict__ p, int count) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < count) p[i] = 0;
}

__global__ void k_check_coords(const int32_t* __restrict__ x, int T, int* flag) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < T && x[t] != t) atomicOr(flag, 1);
}

__global__ void k_f32_to_f64(const float* __restrict__ in, double* __restrict__ out, size_t count) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < count) out[i] = (double)in[i];
}

__global__ void k_f64_to_f32(const double* __restrict__ in, float* __restrict__ out, size_t count) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < count) out[i] = (float)in[i];
}

}  // namespace

// ------------------------------------------------------------ host helpers

int knot_capacity(int T) { return T / 2 + 4; }
int solver_block_capacity(int Ecap) {
    int N = Ecap - 2;
    return (N >= 1) ? (N - 1) / kSolverBlock + 1 : 1;
}

int blocks_for(int map, const EpochDesc& e) {
    switch (map) {
    case kMapInit: return (int)(((size_t)e.n * e.T + 255) / 256);
    case kMapFlags:
        if (e.n <= kFlagLdsMaxCh) return (e.T + kFlagTile - 1) / kFlagTile;
        return ((e.K + kFlagGroup - 1) / kFlagGroup) * ((e.T + kFlagTile - 1) / kFlagTile);
    case kMapCompact: return e.K;
    case kMapBlock: return e.bblocks;
    case kMapAmp: return (e.mode == kModeStop && !e.coef_records) ? (e.T + 63) / 64 : 0;
    case kMapEval:
        if (e.coef_records) return (e.T + 63) / 64;
        return ((e.W + e.etw - 1) / e.etw) * ((e.n + 16 / e.etw - 1) / (16 / e.etw));
    }
    return 0;
}

int block_bpw_fit(int n, int want) {
    int b = want < 1 ? 1 : (want > kBlockBpwMax ? kBlockBpwMax : want);
    for (; b > 1; --b) {
        const BlockLds L = block_lds(&n, &b, 1);
        if (L.bytes <= (size_t)kBlockLdsMax) break;
    }
    return b;
}

BlockLds block_lds(const int* n, const int* bpw, int count) {
    int rmax = kSolverBlock;
    size_t tile = 0;
    for (int e = 0; e < count; ++e) {
        const int r = kSolverBlock * bpw[e];
        rmax = r > rmax ? r : rmax;
        const size_t v = (size_t)channel_stride(n[e]) * (size_t)(r + 3);
        tile = v > tile ? v : tile;
    }
    BlockLds L;
    L.rmax = rmax;
    // rd, cp, lo, ih: rmax + 3 doubles each; X: rmax + 3 ints; the tile [rows][np] starts
    // 16-byte aligned (global-to-LDS loads of 16 bytes per lane)
    L.tile_off = (sizeof(double) * 4 * ((size_t)rmax + 3) + ((size_t)rmax + 3) * sizeof(int) + 15) / 16 * 16;
    L.bytes = L.tile_off + tile * sizeof(double);
    return L;
}

long list_capacity(int list, const EpochDesc& e) {
    const long K2 = 2L * e.K;
    switch (list) {
    case kListFactor: return K2;
    case kListReduced: return K2;
    case kListCoef: return e.coef_records ? K2 * (((long)e.n * e.Ecap + 256L * kCoefItems - 1) / (256L * kCoefItems)) : 0;
    }
    return 0;
}

#define MEMD_TABLE_LAUNCH(kern, map, threads, lds) \
    hipLaunchKernelGGL(kern, dim3(p.blocks[map]), dim3(threads), lds, s, p.eps, p.tab[map])
#define MEMD_LIST_LAUNCH(kern, list, threads) \
    hipLaunchKernelGGL(kern, dim3(p.grid[list]), dim3(threads), 0, s, p.eps, p.wl)

hipError_t launch_init(const StepPlan& p, hipStream_t s) {
    hipLaunchKernelGGL(k_init, dim3(p.blocks[kMapInit]), dim3(256), 0, s, p.eps, p.tab[kMapInit], p.wl);
    return hipGetLastError();
}

hipError_t launch_init_table(const StepPlan& p, const int2* tab, int blocks, hipStream_t s) {
    if (blocks <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_init, dim3(blocks), dim3(256), 0, s, p.eps, tab, p.wl);
    return hipGetLastError();
}

hipError_t launch_gather_probe(const StepPlan& p, EpochState* states, int* sift, const int* sift_off, hipStream_t s) {
    if (p.count <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_gather_probe, dim3(p.count), dim3(64), 0, s, p.eps, p.count, states, sift, sift_off);
    return hipGetLastError();
}

hipError_t launch_mark_done(EpochState* st, hipStream_t s) {
    hipLaunchKernelGGL(k_mark_done, dim3(1), dim3(64), 0, s, st);
    return hipGetLastError();
}

static void launch_kernel(int which, const StepPlan& p, hipStream_t s) {
    switch (which) {
    case 0:   // per epoch: LDS-staged tiles (n <= kFlagLdsMaxCh) or (direction group, tile) blocks
        MEMD_TABLE_LAUNCH(k_flags, kMapFlags, kFlagTile, (size_t)p.nlds * (kFlagTile + 2) * sizeof(double));
        break;
    case 1:
        hipLaunchKernelGGL(k_compact, dim3(p.blocks[kMapCompact]), dim3(128), 0, s, p.eps, p.tab[kMapCompact], p.wl);
        break;
    case 2:
        if (p.factor_wide) MEMD_LIST_LAUNCH(k_factor<256>, kListFactor, 256);
        else MEMD_LIST_LAUNCH(k_factor<64>, kListFactor, 64);
        break;
    case 3:
        hipLaunchKernelGGL(k_block, dim3(p.blocks[kMapBlock]), dim3(kBlockThreads), p.block_lds, s, p.eps,
                           p.tab[kMapBlock], p.block_rmax, (int)p.block_tile_off);
        break;
    case 4: {
        const int stage = reduced_stage(p.nmax);
        hipLaunchKernelGGL(k_reduced, dim3(p.grid[kListReduced]), dim3(64), (size_t)stage * sizeof(double), s, p.eps,
                           p.wl, stage);
        break;
    }
    case 5:   // the fused evaluation forms M itself; records only for stop mode's evaluation
        if (p.coef_records) MEMD_LIST_LAUNCH(k_coef, kListCoef, 256);
        break;
    case 6:
        if (p.mode == kModeStop && p.coef_records)   // mirrored ends: the record path
            MEMD_TABLE_LAUNCH(k_eval_stop, kMapEval, 64, (size_t)p.nmax * 64 * sizeof(double));
        else if (p.mode == kModeStop)
            MEMD_TABLE_LAUNCH(k_eval_fused<true>, kMapEval, 256, 0);
        else
            MEMD_TABLE_LAUNCH(k_eval_fused<false>, kMapEval, 256, 0);
        break;
    case 7:
        if (p.mode == kModeStop) {
            if (!p.coef_records) MEMD_TABLE_LAUNCH(k_stop_amp, kMapAmp, 64, 0);
            hipLaunchKernelGGL(k_stop_decide, dim3(p.count), dim3(256), 0, s, p.eps, p.count);
        } else
            hipLaunchKernelGGL(k_advance, dim3((p.count + 63) / 64), dim3(64), 0, s, p.eps, p.count, p.wl);
        break;
    default:
        if (p.mode == kModeStop) {
            MEMD_TABLE_LAUNCH(k_stop_apply, kMapInit, 256, 0);
            hipLaunchKernelGGL(k_stop_reset, dim3((p.count + 63) / 64), dim3(64), 0, s, p.eps, p.count, p.wl);
        }
        break;
    }
}

hipError_t launch_step(const StepPlan& p, hipStream_t s) {
    for (int i = 0; i < kNumStepKernels; ++i) launch_kernel(i, p, s);
    return hipGetLastError();
}

hipError_t launch_step_timed(const StepPlan& p, hipStream_t s, hipEvent_t* ev) {
    hipError_t e;
    if ((e = hipEventRecord(ev[0], s)) != hipSuccess) return e;
    for (int i = 0; i < kNumStepKernels; ++i) {
        launch_kernel(i, p, s);
        if ((e = hipEventRecord(ev[i + 1], s)) != hipSuccess) return e;
    }
    return hipGetLastError();
}

hipError_t configure_kernels() {
    // the projection kernel's LDS tile holds up to max(kFlagLdsMaxCh, kFlagChunk) x 258
    // doubles (33 KB)
    hipError_t e0 = hipFuncSetAttribute(reinterpret_cast<const void*>(&k_block),
                                        hipFuncAttributeMaxDynamicSharedMemorySize, kBlockLdsMax);
    if (e0 != hipSuccess) return e0;
    hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(&k_flags),
                                       hipFuncAttributeMaxDynamicSharedMemorySize,
                                       (kFlagLdsMaxCh > kFlagChunk ? kFlagLdsMaxCh : kFlagChunk) * (kFlagTile + 2) *
                                           (int)sizeof(double));
    if (e != hipSuccess) return e;
    return hipFuncSetAttribute(reinterpret_cast<const void*>(&k_eval_stop),
                               hipFuncAttributeMaxDynamicSharedMemorySize,
                               kStopMaxChannels * 64 * (int)sizeof(double));
}

hipError_t launch_check_coords(const int32_t* x, int T, int* flag, hipStream_t s) {
    hipLaunchKernelGGL(k_check_coords, dim3((T + 255) / 256), dim3(256), 0, s, x, T, flag);
    return hipGetLastError();
}

hipError_t launch_zero_ints(int* p, int count, hipStream_t s) {
    if (count <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_zero_ints, dim3((count + 255) / 256), dim3(256), 0, s, p, count);
    return hipGetLastError();
}

hipError_t launch_f32_to_f64(const float* in, double* out, size_t count, hipStream_t s) {
    if (count == 0) return hipSuccess;
    hipLaunchKernelGGL(k_f32_to_f64, dim3((unsigned)((count + 255) / 256)), dim3(256), 0, s, in, out, count);
    return hipGetLastError();
}

hipError_t launch_f64_to_f32(const double* in, float* out, size_t count, hipStream_t s) {
    if (count == 0) return hipSuccess;
    hipLaunchKernelGGL(k_f64_to_f32, dim3((unsigned)((count + 255) / 256)), dim3(256), 0, s, in, out, count);
    return hipGetLastError();
}

}  // namespace memd
