// gfx950 kernels of the MEMD engine.
//
// One sift step of every active epoch is a fixed sequence of launches, each
// covering all epochs of the batch (block -> epoch through a prefix map):
//
//   flags    projection of the current signal onto every direction, fused with
//            strict-extremum tests and wave64 ballots -> one 64-bit flag word per
//            64 samples per (direction, envelope)
//            [reference: cublasSgemm :1182-1184, find_extrema_shfl_max/min :77-141]
//   compact  popcount prefix over the flag words and expansion into knot lists
//            [prescan_* :143-281, select_extrema_* :384-445]
//   solve1   blocked tridiagonal, phase 1: block ends and spikes per (block, channel)
//   reduced  phase 2: separator system per (direction, envelope), channels as lanes
//   solve3   phase 3: block re-solve + spline coefficients per interval
//            [setBoundary* :447-584, tridiagonal_setup :704-736,
//             cusparseSgtsv2_nopivot :1227-1241, spline_coefficients :775-793]
//   eval     spline evaluation of both envelopes for every direction, mean over
//            valid directions and the sift update, fused; envelopes stay in
//            registers [interpolate :795-830, averageUppperLower :832-844,
//            averageDirection :846-864, updateSignal :866-874]
//   advance  per-epoch state machine (IMF / sift counters)
//
// Line numbers refer to memReducedEpochs_dev.cu.  All arithmetic is FP64 with
// contraction disabled (-ffp-contract=off); fma() appears exactly where the
// oracle's contract puts it, so results are bit-identical to the oracle.
#include "engine.h"

namespace memd {

namespace {

__device__ __forceinline__ int find_epoch(const int* off, int count, int block) {
    // largest e with off[e] <= block (off is ascending, off[0] = 0)
    int lo = 0, hi = count - 1;
    while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        if (off[mid] <= block) lo = mid; else hi = mid - 1;
    }
    return lo;
}

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
    __device__ __forceinline__ double h(int u) const { return (double)(X[u + 1] - X[u]); }
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

// ---------------------------------------------------------------- init

__global__ __launch_bounds__(256) void k_init(const EpochDesc* __restrict__ eps, const int* __restrict__ off,
                                              int count) {
    int e = find_epoch(off, count, blockIdx.x);
    const EpochDesc& ed = eps[e];
    int local = blockIdx.x - off[e];
    size_t nT = (size_t)ed.n * ed.T;
    size_t i = (size_t)local * 256 + threadIdx.x;
    if (i < nT) {
        double v = ed.y[i];
        ed.running[i] = v;
        ed.imfs[i] = v;   // slot 0 holds the signal being sifted (or the residual if N_imf == 1)
    }
    if (local == 0 && threadIdx.x == 0) {
        EpochState* st = ed.st;
        st->imf = 0;
        st->nsift = 0;
        st->pass = 0;
        st->action = 0;
        st->done = (ed.N_imf <= 1) ? 1 : 0;
        st->n_imf_out = (ed.N_imf <= 1) ? 1 : 0;
        for (int i2 = 0; i2 < ed.N_imf; ++i2) ed.sift_counts[i2] = 0;
    }
}

// ------------------------------------------------------ projection + flags

__device__ __forceinline__ double project_at(const double* __restrict__ Vk, const double* __restrict__ cur,
                                             int n, int T, int t) {
    double acc = 0.0;
    for (int c = 0; c < n; ++c) acc = fma(Vk[c], cur[(size_t)c * T + t], acc);
    return acc;
}

__global__ __launch_bounds__(kFlagTile) void k_flags(const EpochDesc* __restrict__ eps,
                                                     const int* __restrict__ off, int count) {
    int e = find_epoch(off, count, blockIdx.x);
    const EpochDesc& ed = eps[e];
    const EpochState* st = ed.st;
    if (st->done) return;
    int local = blockIdx.x - off[e];
    const int n = ed.n, T = ed.T;
    int tiles = (T + kFlagTile - 1) / kFlagTile;
    int k = local / tiles;
    int tile = local - k * tiles;
    const double* cur = cur_slot(ed, st->imf);
    const double* Vk = ed.V + (size_t)k * n;
    __shared__ double Ps[kFlagTile + 2];
    int tid = threadIdx.x;
    int t0 = tile * kFlagTile;
    int t = t0 + tid;
    if (t < T) Ps[tid + 1] = project_at(Vk, cur, n, T, t);
    if (tid == 0 && t0 > 0) Ps[0] = project_at(Vk, cur, n, T, t0 - 1);
    if (tid == kFlagTile - 1 && t0 + kFlagTile < T) Ps[kFlagTile + 1] = project_at(Vk, cur, n, T, t0 + kFlagTile);
    __syncthreads();
    int lim = window_limit(T, ed.window);
    bool fmax = false, fmin = false;
    if (t < T) {
        if (t == 0 || t == T - 1) {
            fmax = fmin = true;
        } else if (t <= lim) {
            double p = Ps[tid + 1], a = Ps[tid], b = Ps[tid + 2];
            fmax = (p > a) && (p > b);
            fmin = (p < a) && (p < b);
        }
    }
    unsigned long long mmax = __ballot(fmax);
    unsigned long long mmin = __ballot(fmin);
    int lane = tid & 63;
    int w = (t0 >> 6) + (tid >> 6);
    if (lane == 0 && w < ed.W) {
        ed.mask[((size_t)k * 2 + 0) * ed.W + w] = mmax;
        ed.mask[((size_t)k * 2 + 1) * ed.W + w] = mmin;
    }
}

// ------------------------------------------------------------ compaction

__global__ __launch_bounds__(256) void k_compact(const EpochDesc* __restrict__ eps,
                                                 const int* __restrict__ off, int count) {
    int e = find_epoch(off, count, blockIdx.x);
    const EpochDesc& ed = eps[e];
    const EpochState* st = ed.st;
    if (st->done) return;
    int local = blockIdx.x - off[e];   // = k*2 + env
    const int W = ed.W;
    const unsigned long long* m = ed.mask + (size_t)local * W;
    int* pre = ed.pre + (size_t)local * W;
    int* X = ed.X + (size_t)local * ed.Ecap;
    int tid = threadIdx.x;
    int per = (W + 255) / 256;
    int w0 = tid * per;
    int w1 = min(w0 + per, W);
    int mine = 0;
    for (int w = w0; w < w1; ++w) mine += __popcll(m[w]);
    __shared__ int sc[256];
    sc[tid] = mine;
    __syncthreads();
    for (int d = 1; d < 256; d <<= 1) {   // inclusive Hillis-Steele scan of 256 partials
        int v = (tid >= d) ? sc[tid - d] : 0;
        __syncthreads();
        sc[tid] += v;
        __syncthreads();
    }
    int base = sc[tid] - mine;
    for (int w = w0; w < w1; ++w) {
        unsigned long long bits = m[w];
        pre[w] = base;
        while (bits) {
            int b = __ffsll((long long)bits) - 1;
            X[base++] = (w << 6) + b;
            bits &= bits - 1;
        }
    }
    if (tid == 255) {
        int Etot = sc[255];
        ed.E[local] = Etot;
        if (ed.stats) {
            atomicAdd(&ed.stats[0], (unsigned long long)Etot);
            atomicAdd(&ed.stats[1], (unsigned long long)Etot * (unsigned long long)ed.n);
        }
        if (ed.d_trace && st->pass < ed.max_passes)
            ed.d_trace[(size_t)st->pass * 2 * ed.K + local] = Etot;
    }
}

// ------------------------------------------------- blocked tridiagonal solver
//
// Knot system of one (direction, envelope): rows u = 1..N (N = E-2) with
// lo_u = h_{u-1}, di_u = 2(X_{u+1}-X_{u-1}), up_u = h_u and, per channel,
// r_u = 3 (sl_u - sl_{u-1}), sl_u = (Y_{u+1}-Y_u)/h_u; M_0 = M_{E-1} = 0.
// Blocks of L = kSolverBlock rows: separators at u = jL (j = 1..J, J = (N-1)/L),
// block b = rows bL+1 .. (b<J ? (b+1)L-1 : N).

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

__device__ __forceinline__ void solver_lane(const EpochDesc& ed, int local, int& k, int& env, int& b, int& c,
                                            bool& active) {
    int bpw = 256 / ed.n;
    int wgs = (ed.nbmax + bpw - 1) / bpw;
    int ke = local / wgs;          // k*2 + env
    int wg = local - ke * wgs;
    k = ke >> 1;
    env = ke & 1;
    int slot = threadIdx.x / ed.n;
    c = threadIdx.x - slot * ed.n;
    b = wg * bpw + slot;
    active = slot < bpw;
}

// Phase 1: per (block, channel) the first/last entries of A_b^{-1} r; per
// block also the spike ends (written by channel 0).  Forward sweep gives the
// last entries, reverse sweep the first ones.
__global__ __launch_bounds__(256) void k_solve1(const EpochDesc* __restrict__ eps,
                                                const int* __restrict__ off, int count) {
    int e = find_epoch(off, count, blockIdx.x);
    const EpochDesc& ed = eps[e];
    const EpochState* st = ed.st;
    if (st->done) return;
    int k, env, b, c;
    bool active;
    solver_lane(ed, blockIdx.x - off[e], k, env, b, c, active);
    if (!active || !dir_valid(ed.E, k)) return;
    int ke = 2 * k + env;
    int E = ed.E[ke];
    BlockGeom g = block_geom(E, b);
    if (g.J == 0 || b > g.J) return;
    KnotRow kr{ed.X + (size_t)ke * ed.Ecap, cur_slot(ed, st->imf) + (size_t)c * ed.T, E, env};
    const int p = g.p, q = g.q;
    const bool left = b >= 1, right = b < g.J;

    // forward sweep u = p..q
    double cp = 0.0, zy = 0.0, za = 0.0, zb = 0.0;
    double yprev = kr.Y(p - 1), ycur = kr.Y(p);
    double slprev = (ycur - yprev) / kr.h(p - 1);
    for (int u = p; u <= q; ++u) {
        double ynext = kr.Y(u + 1);
        double hu = kr.h(u), hl = kr.h(u - 1);
        double sl = (ynext - ycur) / hu;
        double r = 3.0 * (sl - slprev);
        double lo = hl, up = hu;
        double di = (double)(2 * ((kr.X[u] - kr.X[u - 1]) + (kr.X[u + 1] - kr.X[u])));
        double ga = (u == p) ? lo : 0.0;
        double gb = (u == q) ? up : 0.0;
        double den = (u == p) ? di : fma(-lo, cp, di);
        double rd = 1.0 / den;
        cp = up * rd;
        if (u == p) {
            zy = r * rd; za = ga * rd; zb = gb * rd;
        } else {
            zy = fma(-lo, zy, r) * rd;
            za = fma(-lo, za, ga) * rd;
            zb = fma(-lo, zb, gb) * rd;
        }
        ycur = ynext;
        slprev = sl;
    }
    double yq = zy, aq = za, bq = zb;

    // reverse sweep u = q..p
    double cl = 0.0;
    double ry = 0.0, ra = 0.0, rb = 0.0;
    double ynext = kr.Y(q + 1);
    ycur = kr.Y(q);
    double slnext = (ynext - ycur) / kr.h(q);
    for (int u = q; u >= p; --u) {
        double yp = kr.Y(u - 1);
        double hl = kr.h(u - 1), hu = kr.h(u);
        double sl = (ycur - yp) / hl;
        double r = 3.0 * (slnext - sl);
        double lo = hl, up = hu;
        double di = (double)(2 * ((kr.X[u] - kr.X[u - 1]) + (kr.X[u + 1] - kr.X[u])));
        double ga = (u == p) ? lo : 0.0;
        double gb = (u == q) ? up : 0.0;
        double den = (u == q) ? di : fma(-up, cl, di);
        double rd = 1.0 / den;
        cl = lo * rd;
        if (u == q) {
            ry = r * rd; ra = ga * rd; rb = gb * rd;
        } else {
            ry = fma(-up, ry, r) * rd;
            ra = fma(-up, ra, ga) * rd;
            rb = fma(-up, rb, gb) * rd;
        }
        ycur = yp;
        slnext = sl;
    }
    size_t bi = (size_t)ke * ed.nbmax + b;
    double* be = ed.bend + (bi * ed.n + c) * 2;
    be[0] = ry;   // first entry
    be[1] = yq;   // last entry
    if (c == 0) {
        double* sp = ed.bspk + bi * 4;
        sp[0] = left ? ra : 0.0;
        sp[1] = left ? aq : 0.0;
        sp[2] = right ? rb : 0.0;
        sp[3] = right ? bq : 0.0;
    }
}

// Phase 2: reduced system of the J separators, one lane per channel.
__global__ __launch_bounds__(256) void k_reduced(const EpochDesc* __restrict__ eps,
                                                 const int* __restrict__ off, int count) {
    int e = find_epoch(off, count, blockIdx.x);
    const EpochDesc& ed = eps[e];
    const EpochState* st = ed.st;
    if (st->done) return;
    int ke = blockIdx.x - off[e];
    int k = ke >> 1, env = ke & 1;
    int c = threadIdx.x;
    if (c >= ed.n || !dir_valid(ed.E, k)) return;
    int E = ed.E[ke];
    int N = E - 2;
    int J = (N - 1) / kSolverBlock;
    if (J == 0) return;
    KnotRow kr{ed.X + (size_t)ke * ed.Ecap, cur_slot(ed, st->imf) + (size_t)c * ed.T, E, env};
    const double* sp = ed.bspk + (size_t)ke * ed.nbmax * 4;
    const double* be = ed.bend + (size_t)ke * ed.nbmax * ed.n * 2;
    double* sv = ed.sep + (size_t)ke * ed.nbmax * ed.n * 2;
    double cp = 0.0, z = 0.0;
    for (int j = 1; j <= J; ++j) {
        int s = j * kSolverBlock;
        double ym = kr.Y(s - 1), y0 = kr.Y(s), yp = kr.Y(s + 1);
        double hl = kr.h(s - 1), hu = kr.h(s);
        double r = 3.0 * ((yp - y0) / hu - (y0 - ym) / hl);
        double lo = hl, up = hu;
        double di = (double)(2 * ((kr.X[s] - kr.X[s - 1]) + (kr.X[s + 1] - kr.X[s])));
        const double* spL = sp + (size_t)(j - 1) * 4;   // block j-1
        const double* spR = sp + (size_t)j * 4;         // block j
        double Rlo = (j >= 2) ? -(lo * spL[1]) : 0.0;
        double Rdi = fma(-up, spR[0], fma(-lo, spL[3], di));
        double Rup = (j <= J - 1) ? -(up * spR[2]) : 0.0;
        double yqL = be[((size_t)(j - 1) * ed.n + c) * 2 + 1];
        double ypR = be[((size_t)j * ed.n + c) * 2 + 0];
        double D = fma(-up, ypR, fma(-lo, yqL, r));
        double den = (j == 1) ? Rdi : fma(-Rlo, cp, Rdi);
        double rd = 1.0 / den;
        cp = Rup * rd;
        z = (j == 1) ? D * rd : fma(-Rlo, z, D) * rd;
        double* slot = sv + ((size_t)j * ed.n + c) * 2;
        slot[0] = z;
        slot[1] = cp;
    }
    double x = z;   // x_J = z_J
    for (int j = J - 1; j >= 1; --j) {
        double* slot = sv + ((size_t)j * ed.n + c) * 2;
        x = fma(-slot[1], x, slot[0]);
        slot[0] = x;
    }
}

// Phase 3: re-solve every block with its separator values and emit the spline
// coefficients of the intervals it owns (bL .. q).  With J == 0 this is the
// whole (single-block) solve.
__global__ __launch_bounds__(256) void k_solve3(const EpochDesc* __restrict__ eps,
                                                const int* __restrict__ off, int count) {
    constexpr int L = kSolverBlock;
    int e = find_epoch(off, count, blockIdx.x);
    const EpochDesc& ed = eps[e];
    const EpochState* st = ed.st;
    if (st->done) return;
    int k, env, b, c;
    bool active;
    solver_lane(ed, blockIdx.x - off[e], k, env, b, c, active);
    if (!active || !dir_valid(ed.E, k)) return;
    int ke = 2 * k + env;
    int E = ed.E[ke];
    BlockGeom g = block_geom(E, b);
    if (b > g.J) return;
    KnotRow kr{ed.X + (size_t)ke * ed.Ecap, cur_slot(ed, st->imf) + (size_t)c * ed.T, E, env};
    const int p = g.p, q = g.q;
    const double* sv = ed.sep + (size_t)ke * ed.nbmax * ed.n * 2;
    double XL = (b >= 1) ? sv[((size_t)b * ed.n + c) * 2] : 0.0;        // M at row p-1
    double XR = (b < g.J) ? sv[((size_t)(b + 1) * ed.n + c) * 2] : 0.0; // M at row q+1

    double zarr[L], cparr[L];
    double cp = 0.0, z = 0.0;
    double yprev = kr.Y(p - 1), ycur = kr.Y(p);
    double slprev = (ycur - yprev) / kr.h(p - 1);
#pragma unroll
    for (int i = 0; i < L; ++i) {
        int u = p + i;
        if (u <= q) {
            double ynext = kr.Y(u + 1);
            double hu = kr.h(u), hl = kr.h(u - 1);
            double sl = (ynext - ycur) / hu;
            double gv = 3.0 * (sl - slprev);
            double lo = hl, up = hu;
            double di = (double)(2 * ((kr.X[u] - kr.X[u - 1]) + (kr.X[u + 1] - kr.X[u])));
            if (u == p && b >= 1) gv = fma(-lo, XL, gv);
            if (u == q && b < g.J) gv = fma(-up, XR, gv);
            double den = (u == p) ? di : fma(-lo, cp, di);
            double rd = 1.0 / den;
            cp = up * rd;
            z = (u == p) ? gv * rd : fma(-lo, z, gv) * rd;
            zarr[i] = z;
            cparr[i] = cp;
            ycur = ynext;
            slprev = sl;
        }
    }
    // back substitution, emitting interval u's coefficients as soon as M_u is known
    double* cf = ed.coef + ((size_t)ke * ed.n + c) * (size_t)ed.Ecap * 4;
    double xnext = XR;   // M_{u+1}
    double ynext = kr.Y(q + 1);
#pragma unroll
    for (int i = L - 1; i >= 0; --i) {
        int u = p + i;
        if (u <= q) {
            double x = (u == q) ? zarr[i] : fma(-cparr[i], xnext, zarr[i]);
            double yu = kr.Y(u);
            double hu = kr.h(u);
            double sl = (ynext - yu) / hu;
            double* o = cf + (size_t)u * 4;
            o[0] = yu;
            o[1] = sl - (hu * (2.0 * x + xnext)) / 3.0;
            o[2] = x;
            o[3] = (xnext - x) / (3.0 * hu);
            xnext = x;
            ynext = yu;
        }
    }
    {   // interval p-1 (from separator row bL, or from knot 0)
        int u = p - 1;
        double x = XL;
        double yu = kr.Y(u);
        double hu = kr.h(u);
        double sl = (ynext - yu) / hu;
        double* o = cf + (size_t)u * 4;
        o[0] = yu;
        o[1] = sl - (hu * (2.0 * x + xnext)) / 3.0;
        o[2] = x;
        o[3] = (xnext - x) / (3.0 * hu);
    }
}

// --------------------------------------------- fused evaluate + mean + update

__device__ __forceinline__ double horner(const double* __restrict__ cf, double tau) {
    double2 ab = *reinterpret_cast<const double2*>(cf);
    double2 cd = *reinterpret_cast<const double2*>(cf + 2);
    return fma(fma(fma(cd.y, tau, cd.x), tau, ab.y), tau, ab.x);
}

__device__ __forceinline__ int knot_interval(const EpochDesc& ed, int ke, int w, int lane, int E) {
    // j = #knots strictly before t, minus one, clamped to [0, E-2]
    unsigned long long m = ed.mask[(size_t)ke * ed.W + w];
    int before = ed.pre[(size_t)ke * ed.W + w] + __popcll(m & ((1ull << lane) - 1ull));
    int j = before - 1;
    j = j < 0 ? 0 : j;
    j = j > E - 2 ? E - 2 : j;
    return j;
}

// Fix mode: cur -= sum_valid((e+ + e-)/2) / nvalid, and on the S-th sift of an
// IMF also running -= cur and the next slot <- running.
__global__ __launch_bounds__(64) void k_eval(const EpochDesc* __restrict__ eps,
                                             const int* __restrict__ off, int count) {
    int e = find_epoch(off, count, blockIdx.x);
    const EpochDesc& ed = eps[e];
    const EpochState* st = ed.st;
    if (st->done) return;
    const int n = ed.n, T = ed.T, K = ed.K;
    int local = blockIdx.x - off[e];
    int groups = (n + kEvalCPL - 1) / kEvalCPL;
    int tile = local / groups;
    int cg = local - tile * groups;
    int lane = threadIdx.x;
    int t = tile * 64 + lane;
    int c0 = cg * kEvalCPL;
    double acc[kEvalCPL];
#pragma unroll
    for (int i = 0; i < kEvalCPL; ++i) acc[i] = 0.0;
    double nv = 0.0;
    for (int k = 0; k < K; ++k) {
        if (!dir_valid(ed.E, k)) continue;
        nv = nv + 1.0;
        int E0 = ed.E[2 * k], E1 = ed.E[2 * k + 1];
        int j0 = knot_interval(ed, 2 * k, tile, lane, E0);
        int j1 = knot_interval(ed, 2 * k + 1, tile, lane, E1);
        double tau0 = (double)(t - ed.X[(size_t)(2 * k) * ed.Ecap + j0]);
        double tau1 = (double)(t - ed.X[(size_t)(2 * k + 1) * ed.Ecap + j1]);
        const double* cf0 = ed.coef + ((size_t)(2 * k) * n) * (size_t)ed.Ecap * 4 + (size_t)j0 * 4;
        const double* cf1 = ed.coef + ((size_t)(2 * k + 1) * n) * (size_t)ed.Ecap * 4 + (size_t)j1 * 4;
#pragma unroll
        for (int i = 0; i < kEvalCPL; ++i) {
            int c = c0 + i;
            if (c < n) {
                double ep = horner(cf0 + (size_t)c * ed.Ecap * 4, tau0);
                double em = horner(cf1 + (size_t)c * ed.Ecap * 4, tau1);
                acc[i] = acc[i] + (ep + em) * 0.5;
            }
        }
    }
    if (ed.stats && local == 0 && lane == 0) {
        atomicAdd(&ed.stats[2], (unsigned long long)nv * (unsigned long long)n * (unsigned long long)T);
        atomicAdd(&ed.stats[3], (unsigned long long)n * (unsigned long long)T);
    }
    if (t >= T) return;
    double* cur = cur_slot(ed, st->imf);
    bool last = (ed.mode == kModeFix) && (st->nsift == ed.S - 1);
    double* nxt = cur_slot(ed, st->imf + 1);
#pragma unroll
    for (int i = 0; i < kEvalCPL; ++i) {
        int c = c0 + i;
        if (c < n) {
            size_t idx = (size_t)c * T + t;
            double v = cur[idx] - acc[i] / nv;
            cur[idx] = v;
            if (last) {
                double r = ed.running[idx] - v;
                ed.running[idx] = r;
                nxt[idx] = r;
            }
        }
    }
}

// Fix-mode state advance: one thread per epoch.
__global__ void k_advance(const EpochDesc* __restrict__ eps, int count) {
    int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= count) return;
    const EpochDesc& ed = eps[e];
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
        }
    }
}

// ------------------------------------------------------------ misc kernels

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

int knot_capacity(int T) { return T / 2 + 2; }
int solver_block_capacity(int Ecap) {
    int N = Ecap - 2;
    return (N >= 1) ? (N - 1) / kSolverBlock + 1 : 1;
}
int solver_blocks_per_wg(int n) { return 256 / n; }

int blocks_init(const EpochDesc& e) { return (int)(((size_t)e.n * e.T + 255) / 256); }
int blocks_flags(const EpochDesc& e) { return e.K * ((e.T + kFlagTile - 1) / kFlagTile); }
int blocks_compact(const EpochDesc& e) { return e.K * 2; }
int blocks_solve(const EpochDesc& e) {
    int bpw = solver_blocks_per_wg(e.n);
    return e.K * 2 * ((e.nbmax + bpw - 1) / bpw);
}
int blocks_reduced(const EpochDesc& e) { return e.K * 2; }
int blocks_eval(const EpochDesc& e) { return ((e.T + 63) / 64) * ((e.n + kEvalCPL - 1) / kEvalCPL); }

hipError_t launch_init(const StepPlan& p, hipStream_t s) {
    hipLaunchKernelGGL(k_init, dim3(p.blocks_init), dim3(256), 0, s, p.eps, p.off_init, p.count);
    return hipGetLastError();
}

hipError_t launch_step(const StepPlan& p, hipStream_t s) {
    hipLaunchKernelGGL(k_flags, dim3(p.blocks_flags), dim3(kFlagTile), 0, s, p.eps, p.off_flags, p.count);
    hipLaunchKernelGGL(k_compact, dim3(p.blocks_compact), dim3(256), 0, s, p.eps, p.off_compact, p.count);
    hipLaunchKernelGGL(k_solve1, dim3(p.blocks_solve), dim3(256), 0, s, p.eps, p.off_solve, p.count);
    hipLaunchKernelGGL(k_reduced, dim3(p.blocks_reduced), dim3(256), 0, s, p.eps, p.off_reduced, p.count);
    hipLaunchKernelGGL(k_solve3, dim3(p.blocks_solve), dim3(256), 0, s, p.eps, p.off_solve, p.count);
    hipLaunchKernelGGL(k_eval, dim3(p.blocks_eval), dim3(64), 0, s, p.eps, p.off_eval, p.count);
    hipLaunchKernelGGL(k_advance, dim3((p.count + 63) / 64), dim3(64), 0, s, p.eps, p.count);
    return hipGetLastError();
}

hipError_t launch_step_timed(const StepPlan& p, hipStream_t s, hipEvent_t* ev) {
    hipError_t e;
    if ((e = hipEventRecord(ev[0], s)) != hipSuccess) return e;
    hipLaunchKernelGGL(k_flags, dim3(p.blocks_flags), dim3(kFlagTile), 0, s, p.eps, p.off_flags, p.count);
    if ((e = hipEventRecord(ev[1], s)) != hipSuccess) return e;
    hipLaunchKernelGGL(k_compact, dim3(p.blocks_compact), dim3(256), 0, s, p.eps, p.off_compact, p.count);
    if ((e = hipEventRecord(ev[2], s)) != hipSuccess) return e;
    hipLaunchKernelGGL(k_solve1, dim3(p.blocks_solve), dim3(256), 0, s, p.eps, p.off_solve, p.count);
    if ((e = hipEventRecord(ev[3], s)) != hipSuccess) return e;
    hipLaunchKernelGGL(k_reduced, dim3(p.blocks_reduced), dim3(256), 0, s, p.eps, p.off_reduced, p.count);
    if ((e = hipEventRecord(ev[4], s)) != hipSuccess) return e;
    hipLaunchKernelGGL(k_solve3, dim3(p.blocks_solve), dim3(256), 0, s, p.eps, p.off_solve, p.count);
    if ((e = hipEventRecord(ev[5], s)) != hipSuccess) return e;
    hipLaunchKernelGGL(k_eval, dim3(p.blocks_eval), dim3(64), 0, s, p.eps, p.off_eval, p.count);
    if ((e = hipEventRecord(ev[6], s)) != hipSuccess) return e;
    hipLaunchKernelGGL(k_advance, dim3((p.count + 63) / 64), dim3(64), 0, s, p.eps, p.count);
    if ((e = hipEventRecord(ev[7], s)) != hipSuccess) return e;
    return hipGetLastError();
}

hipError_t launch_check_coords(const int32_t* x, int T, int* flag, hipStream_t s) {
    hipLaunchKernelGGL(k_check_coords, dim3((T + 255) / 256), dim3(256), 0, s, x, T, flag);
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
