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
// Channels per lane in the fused evaluate/mean/update kernel.
constexpr int kEvalCPL = 4;
// Largest channel count: one solver workgroup of 256 lanes holds >= 1 block.
constexpr int kMaxChannels = 256;

enum : int { kModeFix = 0, kModeStop = 1 };
enum : int { kWindowRefWarp32 = 0, kWindowFull = 1 };

// Per-epoch device state, advanced on the device after every sift step.
struct EpochState {
    int imf;        // IMF being sifted (its slot in imfs holds the current signal)
    int nsift;      // subtractions done for this IMF
    int done;       // 1 once the residual has been written
    int pass;       // envelope passes performed (trace index)
    int n_imf_out;  // output slots used (IMFs + residual), valid when done
    int action;     // stop mode: decision of the last step (see kernels.hip)
    int pad[2];
};

// Everything a kernel needs to know about one epoch.  Lives in device memory.
struct EpochDesc {
    int n, T, K, W;        // W = ceil(T/64) bitmask words per projection row
    int Ecap;              // knot capacity per (direction, envelope) = T/2 + 2
    int nbmax;             // solver blocks per (direction, envelope) at Ecap knots
    int N_imf, S, mode, window;
    double sd, sd2, tol;
    const double* y;       // [n][T]
    const double* V;       // [K][n]
    double* imfs;          // [N_imf][n][T]
    double* running;       // [n][T]
    unsigned long long* mask; // [K][2][W] extrema flags (bit t%64 of word t/64)
    int* pre;              // [K][2][W] knots before each word
    int* E;                // [K][2] knot counts
    int* X;                // [K][2][Ecap] knot sample positions
    double* coef;          // [K][2][n][Ecap][4] spline (a,b,c,d) per interval
    double* bend;          // [K][2][nbmax][n][2] block solution ends (first, last)
    double* bspk;          // [K][2][nbmax][4] spikes (alpha_first, alpha_last, beta_first, beta_last)
    double* sep;           // [K][2][nbmax][n][2] reduced system (z, cp) then x in slot 0
    double* mean;          // [n][T] stop mode: envelope mean
    double* tstat;         // [T][2] stop mode: |mean(t)|^2, amplitude(t)
    int* red;              // [4] stop mode: reduction scratch
    int* d_trace;          // optional [max_passes][K][2]
    int max_passes;
    int* sift_counts;      // device [N_imf]
    EpochState* st;
    unsigned long long* stats; // optional [4]: knots, knot*channels, eval points, samples
};

// Work map for one kernel: block b of the launch belongs to the epoch e with
// off[e] <= b < off[e+1].
struct WorkMap {
    const int* off;
    int count;
};

struct StepPlan {
    const EpochDesc* eps;   // device array
    int count;
    // block prefix arrays (device), one per kernel
    const int* off_init;
    const int* off_flags;
    const int* off_compact;
    const int* off_solve;
    const int* off_reduced;
    const int* off_eval;
    int blocks_init, blocks_flags, blocks_compact, blocks_solve, blocks_reduced, blocks_eval;
    int any_stop;
};

// Host-side block counts for one epoch (must match the kernels' decoding).
int blocks_init(const EpochDesc& e);
int blocks_flags(const EpochDesc& e);
int blocks_compact(const EpochDesc& e);
int blocks_solve(const EpochDesc& e);
int blocks_reduced(const EpochDesc& e);
int blocks_eval(const EpochDesc& e);
int solver_blocks_per_wg(int n);
int knot_capacity(int T);
int solver_block_capacity(int Ecap);

hipError_t launch_init(const StepPlan& p, hipStream_t s);
hipError_t launch_step(const StepPlan& p, hipStream_t s);
// Same launches with an event recorded after each kernel (ev[0] before the first).
hipError_t launch_step_timed(const StepPlan& p, hipStream_t s, hipEvent_t* ev /*[8]*/);
hipError_t launch_check_coords(const int32_t* x, int T, int* flag, hipStream_t s);
hipError_t launch_f32_to_f64(const float* in, double* out, size_t count, hipStream_t s);
hipError_t launch_f64_to_f32(const double* in, float* out, size_t count, hipStream_t s);

}  // namespace memd
