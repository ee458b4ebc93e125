// Internal interface between the C-ABI layer (capi.hip) and the gfx950 kernels
// (kernels.hip).  Not installed; callers use include/memd_amd.h.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stddef.h>

namespace memd {

// Rows per block of the blocked tridiagonal solver, separator included.
// Part of the arithmetic contract: the FP64 oracle uses the same value.
constexpr int kSolverBlock = 32;
// Samples per projection/extrema tile (4 wavefronts of 64).
constexpr int kFlagTile = 256;
// The fused evaluate/mean/update kernel builds the spline records of each tile
// in LDS (no record array in HBM); tiles of 64 * TW samples x 16 / TW channels,
// TW = EpochDesc::etw in {1, 2, 4}.
// Intervals meeting a tile of 64 * TW samples: at most 32 * TW knots of one kind
// lie in its first 64 * TW - 1 samples (strict extrema are never adjacent; the
// forced knot at t = 0 next to an extremum at t = 1 still leaves at most 32 in
// samples 0..62), so at most 32 * TW + 1 intervals.
__host__ __device__ inline int eval_max_int(int tw) { return 32 * tw + 1; }
// LDS budget of one k_eval_fused workgroup.  kEvalCap records hold one
// direction's worst case at every tile shape: 2 envelopes x channels x odd row
// stride, at most 2 x 16 x 33 (TW = 1), 2 x 8 x 65, 2 x 4 x 129.
constexpr int kEvalCap = 1056;
constexpr int kEvalXkCap = 384;    // knot positions (worst direction: 2 x 131)
constexpr int kEvalBatch = 16;     // directions per batch (2 segments each: 32 formation lanes)
constexpr int kEvalChunk = 32;     // directions whose formation inputs are staged together
// Flag words of the tile held for a batch's segments and for a staged chunk: a
// tile of TW words takes batches of at most min(2 kEvalBatch, kEvalSegWords / TW)
// segments and chunks of min(kEvalChunk, kEvalChunkWords / (2 TW)) directions
// (TW = 1, 2: 16 directions per batch, 32 per chunk; TW = 4: 8 and 16)
constexpr int kEvalSegWords = 64;
constexpr int kEvalChunkWords = 128;
static_assert(kEvalChunk >= kEvalBatch && kEvalChunkWords >= kEvalSegWords, "a batch fits in one staged chunk");
// Projection/extrema: the LDS-staged kernel holds n x 258 doubles up to
// kFlagLdsMaxCh channels and runs all directions kFlagGroup at a time; wider
// epochs pass kFlagChunk channels at a time through the same tile, kFlagGroup
// directions per workgroup.
constexpr int kFlagLdsMaxCh = 16;
constexpr int kFlagChunk = 16;
constexpr int kFlagGroup = 32;
// Directions rounded up to a whole k_flags group: the row length of the
// channel-major copy of the directions (EpochDesc::vt).
__host__ __device__ constexpr int dir_pad(int K) { return (K + kFlagGroup - 1) / kFlagGroup * kFlagGroup; }
// Knot-value / right-hand-side kernel: 256 knots per workgroup, 4 channels per pass.
// Coefficient kernel: items (channel, interval) per thread.
// Solver blocks per k_block work item (EpochDesc::bpw).  Default: at most 4 (<= 128
// rows) and n * blocks <= 128, so the LDS tile (rows p0-1 .. qlast+2 of every
// channel, odd row stride) stays near 35 KB for n <= 128 and four workgroups fit
// a CU; tuning may choose up to kBlockBpwMax.
__host__ __device__ inline int blocks_per_item(int n) {
    int b = 128 / n;
    return b < 1 ? 1 : (b > 4 ? 4 : b);
}
constexpr int kBlockBpwMax = 8;
constexpr int kBlockThreads = 256;   // k_block workgroup size
// k_block's dynamic LDS: per-row factors for rmax rows, knot positions and the
// tile of the widest epoch.  rmax = 32 * max bpw over the batch.
struct BlockLds {
    int rmax;
    size_t tile_off;   // bytes: start of the tile
    size_t bytes;
};
BlockLds block_lds(const int* n, const int* bpw, int count);
constexpr int kBlockLdsMax = 160 * 1024;   // gfx950 LDS per workgroup
// Largest bpw <= want (>= 1) whose k_block LDS for n channels fits kBlockLdsMax.
int block_bpw_fit(int n, int want);
constexpr int kCoefItems = 16;     // (channel, knot) pairs per k_coef lane and work item
// Persistent grid size of the list-driven solver kernels (workgroups).
constexpr long kListGrid = 2048;
// Factor lists of at most this many items (8 epochs at K = 64) use k_factor's
// 256-thread form.
constexpr long kFactorWideItems = 1024;
// Per knot row u of one (direction, envelope), shared by its channels: three
// planes of Ecap pairs.  kFacAB: al, be, the phase-3 coefficients of the blocked
// solve (the block's spike entries; 0 where a block has no left or right spike
// and at the rows outside blocks: the end rows and the separators); kFacIH:
// ih = 1/h_u, i3h = 1/(3 h_u) of interval u; kFacRC: rd, cp, the block factors.
enum : int { kFacAB = 0, kFacIH = 1, kFacRC = 2, kFacPlanes = 3 };
// Largest channel count: one solver workgroup of 256 lanes holds >= 1 block.
constexpr int kMaxChannels = 256;

enum : int { kModeFix = 0, kModeStop = 1 };
// Stop-mode decisions: subtract the mean (and finish the IMF after S sifts),
// finish the IMF without subtracting, or end the decomposition.
enum : int { kActContinue = 0, kActFinishImf = 1, kActEnd = 2, kActNone = 3 };
// Stop mode keeps per-sample channel sums in LDS: 64 samples x n channels.
constexpr int kStopMaxChannels = 128;   // the record path (mirrored ends)
enum : int { kWindowRefWarp32 = 0, kWindowFull = 1 };

// Per-epoch device state, advanced on the device after every sift step.
struct EpochState {
    int imf;        // IMF being sifted (its slot in imfs holds the current signal)
    int nsift;      // subtractions done for this IMF
    int done;       // 1 once the residual has been written
    int pass;       // envelope passes performed (trace index)
    int n_imf_out;  // output slots used (IMFs + residual), valid when done
    int action;     // stop mode: decision of the last step (kActContinue, ...)
    int act_imf;    // stop mode: IMF index the decision applies to
    int pad;
};

// Everything a kernel needs to know about one epoch.  Lives in device memory.
struct EpochDesc {
    int n, T, K, W;        // W = ceil(T/64) bitmask words per projection row
    int np;                // row stride of curt, yend, ysol, sep: n rounded up to a multiple of 4
    int Ecap;              // knot capacity per (direction, envelope) = T/2 + 4 (mirrored ends add
                           //   up to 2 knots; their lists are compacted one slot in)
    int nbmax;             // solver blocks per (direction, envelope) at Ecap knots
    int N_imf, S, mode, window;
    double sd, sd2, tol;
    const double* y;       // [n][T]
    const double* V;       // [K][n]
    double* vt;            // [n][dir_pad(K)] the directions channel-major, zero past K (k_init):
                           //   a k_flags group's coefficients of one channel are adjacent words
    double* imfs;          // [N_imf][n][T]
    double* running;       // [n][T]
    double* curt;          // [T][np] current signal, sample-major (written by k_flags every step),
                           //   followed by yend
    unsigned long long* mask; // [K][2][W] extrema flags (bit t%64 of word t/64)
    int* pre;              // [K][2][W] knots before each word
    int* E;                // [K][2] knot counts of the lists the solver uses
    int* Eraw;             // [K][2] knot counts of the clamp lists (interior extrema + 2): the
                           //   trace and stop mode's nem; equal to E unless ends are mirrored
    int* X;                // [K][2][Ecap] knot positions (samples; mirrored ends may lie left of
                           //   0 or right of T-1)
    int4* mir;             // [K][2] mirrored ends: (knots added left, knots added right, left mirror
                           //   point, right mirror point); an added knot at position p carries
                           //   the signal at sample 2 * point - p
    double* fac;           // [K][2][kFacPlanes][Ecap][2] per knot row, shared by channels (kFacAB ..)
    double* red;           // [K][2][3][nbmax] reduced system per separator: lo, rd, cp
    // Knot-major and separator-major arrays: rows of np = n rounded up to a
    // multiple of 4 words (32-byte aligned 4-channel groups; padding is never read
    // into results)
    double* yend;          // [K][2][2][np] clamped end-knot values Y_0, Y_{E-1} (E >= 6): rows T ..
                           //   T + 4K - 1 of curt; the other knot values are the signal at the knot
    double* ysol;          // [K][2][Ecap][np] per knot row the value y^ of M_u = fma(-Z[i1], be_u,
                           //   fma(-Z[i0], al_u, y^_u)): the block solution at block rows (i0, i1 =
                           //   b, b + 1 for block b), the separator value at separator rows and 0 (or the
                           //   not-a-knot end value) at the end rows (i0 = i1 = 0 at both)
    double* coef;          // [K][2][n][Ecap][4] spline record (a,b,c,d) per interval
    double* sep;           // [K][2][nbmax + 1][np] Z: Z[0] = Z[J+1] = 0, Z[j] = separator value j
    double* sepin;         // [K][2][nbmax][np] right-hand side of each separator's row (the y^ rows
                           //   beside it are read from ysol)
    double* mean;          // [n][T] stop mode: envelope mean
    double* ampb;          // [K][ceil(n/16)][T] stop mode, fused evaluation: per valid direction and
                           //   block of 16 channels the sum of (e+ - e-)^2 (groups of 4 in order)
    double* tstat;         // [T][2] stop mode: |mean(t)|^2, amplitude(t)
    int* stopred;          // [4] stop mode: reduction scratch
    int* d_trace;          // optional [max_passes][K][2]
    int max_passes;
    int* sift_counts;      // device [N_imf]
    EpochState* st;
    unsigned long long* stats; // optional [kStatSlots][4]: knots, knot*channels, eval points, samples
    int* ndone;                // shared by the batch: epochs finished (stop mode)
    int coef_records;          // 1: k_coef writes the record array (stop mode)
    int xcd;                   // XCD group: work items go to sub-list xcd (when the plan has groups);
                               //   -1: spread, direction k's items go to sub-list k mod 8
    int bpw;                   // solver blocks per k_block work item
    // k_block work items (ke, group, E, 0) of this epoch, appended by k_compact and
    // drained by the epoch's bblocks workgroups (an XCD works through one epoch's
    // items at a time, so the epoch's signal stays in its L2)
    int4* bitem;
    int* bcount;
    int bcap;
    int bblocks;
    int etw;                   // k_eval_fused tile: 64 * etw samples x 16 / etw channels (etw in {1, 2, 4})
    int spline;                // 0 natural ends, 1 not-a-knot
    int boundary;              // 0 the reference's clamped ends, 1 mirrored ends (record path only)
};

// Fixed-size launches: block b of a launch belongs to epoch tab[b].x as its
// local block tab[b].y.
enum : int { kMapInit = 0, kMapFlags, kMapCompact, kMapEval, kMapBlock, kMapAmp, kNumMaps };

// List-driven launches: k_compact appends the exact work items (epoch, ke, sub)
// of every valid direction; persistent grids drain the lists.
enum : int { kListFactor = 0, kListReduced, kListCoef, kNumLists };

// Workgroups are dispatched to the 8 XCDs round-robin (workgroup id mod 8).
// With >= kXcds epochs in a batch, every epoch is pinned to XCD (epoch mod 8):
// its table blocks sit at ids = XCD (mod 8) and its work items go to that
// XCD's sub-list, so its signal and knot data stay in one XCD's L2.
constexpr int kXcds = 8;
// Work-list counters sit kCountStride ints (one 128-byte line) apart, so that the
// appends of different lists and XCD groups do not contend for one line.
constexpr int kCountStride = 32;
__host__ __device__ inline int count_slot(int list, int g) { return (list * kXcds + g) * kCountStride; }
constexpr size_t kCountBytes = sizeof(int) * kCountStride * kNumLists * kXcds;
constexpr size_t kBcountStride = 128;   // bytes between the per-epoch k_block item counters

struct WorkLists {
    int4* item[kNumLists];
    int* count;                     // [kNumLists][kXcds] slots of kCountStride ints, cleared every step
    int off[kNumLists][kXcds];      // sub-list start in item[list]
    int cap[kNumLists][kXcds];
    int groups;                     // kXcds (XCD affinity) or 1
};

struct StepPlan {
    const EpochDesc* eps;   // device array
    int count;
    const int2* tab[kNumMaps];   // device block tables
    int blocks[kNumMaps];
    WorkLists wl;
    int grid[kNumLists];    // persistent grid size per list
    int mode;               // kModeFix or kModeStop (all epochs of one call)
    int nmax;               // largest channel count in the batch
    int nlds;               // channels of k_flags' LDS tile: the largest n <= kFlagLdsMaxCh, at least kFlagChunk if an epoch is wider, 0 if none
    int coef_records;       // launch k_coef (stop mode: the evaluation reads the record array)
    int factor_wide;        // k_factor with 256 threads per item (few items: small batches)
    int block_rmax;         // k_block: rows of the longest work item
    size_t block_tile_off;  // k_block: LDS byte offset of the tile
    size_t block_lds;       // k_block: dynamic LDS bytes
};

// Host-side block counts for one epoch (must match the kernels' decoding).
int blocks_for(int map, const EpochDesc& e);
// Worst-case work items of one epoch in a list.
long list_capacity(int list, const EpochDesc& e);
// Kernel slots of one sift step (profiling order): flags, compact, factor,
// block, reduced, coef, eval | eval_stop, advance | stop_decide, stop_apply
// (stop mode only); slot kNumStepKernels is the one-off init.
constexpr int kNumStepKernels = 9;
// Profiling counters are summed over this many slots (one per workgroup index
// modulo kStatSlots), so the profiled launches do not serialise on one address.
constexpr int kStatSlots = 64;
int knot_capacity(int T);
inline int channel_stride(int n) { return (n + 3) & ~3; }
int solver_block_capacity(int Ecap);

// Per-device kernel attributes (large dynamic LDS); call once per context.
hipError_t configure_kernels();
hipError_t launch_init(const StepPlan& p, hipStream_t s);
// k_init over an explicit block table (epochs joining a running set).
hipError_t launch_init_table(const StepPlan& p, const int2* tab, int blocks, hipStream_t s);
hipError_t launch_step(const StepPlan& p, hipStream_t s);
// Same launches with an event recorded after each kernel (ev[0] before the first).
hipError_t launch_step_timed(const StepPlan& p, hipStream_t s, hipEvent_t* ev /*[kNumStepKernels+1]*/);
hipError_t launch_check_coords(const int32_t* x, int T, int* flag, hipStream_t s);
// p[0 .. count) = 0 by a kernel (no memset node in a caller's captured graph)
hipError_t launch_zero_ints(int* p, int count, hipStream_t s);
// Server probe: every epoch's state into states[e] and its sift counts into
// sift[sift_off[e] ..] (one copy to the host carries both).
hipError_t launch_gather_probe(const StepPlan& p, EpochState* states, int* sift, const int* sift_off, hipStream_t s);
// Withdraw one epoch: its state becomes done with no pending action, so no later
// step touches its buffers.
hipError_t launch_mark_done(EpochState* st, hipStream_t s);
hipError_t launch_f32_to_f64(const float* in, double* out, size_t count, hipStream_t s);
hipError_t launch_f64_to_f32(const double* in, float* out, size_t count, hipStream_t s);

}  // namespace memd
