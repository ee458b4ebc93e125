/*
 * memd_amd.h -- C ABI of the MI355X (gfx950) MEMD engine.
 *
 * Plain C: pointers, sizes and status codes only.  Every entry point that
 * replaces a reference interface cites it (file:line into EEGLab-Pannon/MEMD-GPU).
 *
 * Arithmetic is FP64 end to end (the *_f32 entry points convert at the
 * boundary).  Results are bit-identical to the FP64 host oracle that restates
 * the reference's algorithm; see the engine's design notes for the contract.
 *
 * Threading: a memd_ctx is not thread-safe; use one per host thread.  All
 * device pointers must belong to the context's device.
 */
#ifndef MEMD_AMD_H
#define MEMD_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum memd_status {
    MEMD_OK = 0,
    MEMD_ERR_ARG = 1,        /* invalid shape / argument */
    MEMD_ERR_HIP = 2,        /* HIP runtime failure (message in memd_last_error) */
    MEMD_ERR_NOMEM = 3,      /* device allocation failed */
    MEMD_ERR_COORDS = 4,     /* sample coordinates are not 0..T-1 */
    MEMD_ERR_UNSUPPORTED = 5,/* shape outside the engine's limits (e.g. n > 256) */
    MEMD_ERR_ADMISSION = 6,  /* request refused by the HBM admission budget */
    MEMD_ERR_NOT_FOUND = 7,  /* unknown request id */
    MEMD_ERR_IO = 8          /* file could not be opened, read or written */
} memd_status;

/* Sifting policy. */
enum {
    MEMD_MODE_FIX = 0,  /* exactly S sifts per IMF (memReducedEpochs_dev.cu:1176-1179) */
    MEMD_MODE_STOP = 1  /* memd.m stop criterion [sd, sd2, tol], at most S sifts per IMF */
};
/* Spline end conditions. */
enum {
    MEMD_SPLINE_NATURAL = 0,   /* the reference's (memReducedEpochs_dev.cu:704-736) */
    MEMD_SPLINE_NOT_A_KNOT = 1 /* third-derivative continuity at the second and second-to-last
                                  knots: the reference's unused tridiagonal_setup_nak (:738-773)
                                  restated with the end unknowns eliminated, so the system is
                                  diagonally dominant (the reference's form has a zero pivot
                                  whenever the first two knot gaps are equal) */
};
/* End treatment of the knot lists. */
enum {
    MEMD_BOUNDARY_CLAMP = 0,   /* the reference's: samples 0 and T-1 are knots, raised / lowered to
                                  the linear extrapolation of the next two knots when a list has
                                  >= 6 knots (memReducedEpochs_dev.cu:99-104, 447-486, 545-584) */
    MEMD_BOUNDARY_MIRROR = 1   /* memd.m's boundary_conditions (nbsym = 2), not in the reference:
                                  up to two extrema per end and list mirrored about the end
                                  extremum or the end sample, no clamp.  Runs the record-path
                                  kernels in both sifting modes: n <= 128 channels */
};
/* Extrema window. */
enum {
    MEMD_WINDOW_REFWARP32 = 0, /* the reference's coverage: interior t <= 240*ceil(T/242)
                                  (memReducedEpochs_dev.cu:77-141 launched at :1191-1193) */
    MEMD_WINDOW_FULL = 1       /* every interior sample */
};

typedef struct memd_opts {
    int32_t mode;            /* MEMD_MODE_* (default FIX) */
    int32_t window;          /* MEMD_WINDOW_* (default REFWARP32) */
    double sd, sd2, tol;     /* stop thresholds (default 0.075, 0.75, 0.075) */
    int32_t use_graph;       /* 1: replay one captured sift step as a hipGraph (default 1) */
    /* optional diagnostics (NULL = off) */
    int32_t* d_knot_counts;  /* device [max_passes][K][2]: E+ / E- per envelope pass */
    int32_t max_passes;
    int32_t* h_sift_counts;  /* host [N_imf]: sifts per IMF */
    int32_t* h_n_imf_out;    /* host: output slots used (IMFs + residual) */
    int32_t spline;          /* MEMD_SPLINE_* (default NATURAL) */
    int32_t boundary;        /* MEMD_BOUNDARY_* (default CLAMP) */
} memd_opts;

typedef struct memd_ctx memd_ctx;

/* Context: owns the stream binding, the cached device workspace and the
 * direction cache.  device < 0 selects the current device. */
memd_status memd_ctx_create(int device, memd_ctx** out);
memd_status memd_ctx_destroy(memd_ctx* ctx);
/* Run on the given hipStream_t (passed as void*); NULL = legacy default stream. */
memd_status memd_ctx_set_stream(memd_ctx* ctx, void* hip_stream);
/* Direction cache (no reference counterpart: the reference regenerates the set on
 * every call, memReducedEpochs_dev.cu:1029-1072).  Sets are kept per (n, K,
 * primes) and the least recently used are freed once they would exceed
 * max_bytes (default 256 MiB; a single larger set is still kept).  Setting a
 * smaller cap evicts at once. */
memd_status memd_ctx_set_direction_cache(memd_ctx* ctx, size_t max_bytes);
size_t memd_ctx_direction_cache_bytes(memd_ctx* ctx);
/* Testing hook (no reference counterpart): the next `count` device allocations
 * the library makes for this context fail as out of memory, so callers can
 * exercise their MEMD_ERR_NOMEM handling.  0 clears it. */
memd_status memd_debug_fail_allocations(memd_ctx* ctx, int32_t count);
/* Message of the last failure on this context (never NULL).  With ctx == NULL:
 * the last failure of a context-free call (file I/O, memd_ctx_create) on this
 * host thread. */
const char* memd_last_error(memd_ctx* ctx);
void memd_opts_default(memd_opts* o);
/* Library version string. */
const char* memd_version(void);

/*
 * Direction vectors, FP64, written to device d_V [K][n].
 * Replaces GenerateDirectionVector<T>(SignalDim, NumDirVector, d_DirectionVector)
 * (memReducedEpochs_dev.cu:1029-1072).  h_primes == NULL uses the reference's
 * table [-K, 2, 3, 5, ...] (:1033-1046); a host array of n primes selects the
 * MEX semantics (memdMatlabCuda.cu:1131-1137: h[k][0] = (k+1.5)/(-P[0])).
 */
memd_status memd_directions_f64(memd_ctx* ctx, size_t n, size_t K, const int32_t* h_primes,
                                double* d_V);
/* The same FP64 set rounded to float, for the GenerateDirectionVector<float>
 * instantiation the reference's main() uses (memReducedEpochs_dev.cu:1513-1515). */
memd_status memd_directions_f32(memd_ctx* ctx, size_t n, size_t K, const int32_t* h_primes,
                                float* d_V);

/*
 * One epoch, all pointers on the device.
 * Replaces template memd<coord_t,real_t>(SignalDim, SignalLength, NumDirVector,
 * num_IMFs, max_iter, d_x, d_y, d_DirectionVectors, d_IMFs)
 * (memReducedEpochs_dev.cu:1074-1075): d_x [T] sample coordinates (NULL, or
 * exactly 0..T-1 as every reference caller passes, :1504-1506), d_y [n][T],
 * d_V [K][n], d_IMFs [N_imf][n][T] with the residual in the last used slot.
 * elapsed_s (optional) receives the device time of the sifting loop, like the
 * reference's return value (:1336-1337,1367).  Blocking.
 */
memd_status memd_run_f64(memd_ctx* ctx, size_t n, size_t T, size_t K, size_t N_imf, size_t S,
                         const int32_t* d_x, const double* d_y, const double* d_V, double* d_IMFs,
                         const memd_opts* opts, double* elapsed_s);

/*
 * memd_run_f64 without any device allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate (code running under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc).  The
 * workspace is a static per-device array of MEMD_NOALLOC_SCRATCH_BYTES that the
 * library holds on every device it runs on until the process ends; the step is
 * launched directly (no graph: instantiating one allocates) and no hipMalloc,
 * hipFree or other allocating call is made.  Shapes whose workspace does not fit
 * (memd_workspace_bytes_opts plus the plan region, e.g. up to 32 ch x 4096, K =
 * 64 in fix mode) are refused with MEMD_ERR_UNSUPPORTED.  Results are
 * bit-identical to memd_run_f64.  Every call on one device shares that device's
 * array: calls that overlap on one device (any context, any stream) race and
 * must be serialised by the caller.
 */
#define MEMD_NOALLOC_SCRATCH_BYTES (128u * 1024u * 1024u)
memd_status memd_run_noalloc_f64(memd_ctx* ctx, size_t n, size_t T, size_t K, size_t N_imf, size_t S,
                                 const int32_t* d_x, const double* d_y, const double* d_V, double* d_IMFs,
                                 const memd_opts* opts, double* elapsed_s);

/* Same call with the reference's memd<int,float> argument types
 * (memReducedEpochs_dev.cu:1522); converts to FP64 on the device. */
memd_status memd_run_f32(memd_ctx* ctx, size_t n, size_t T, size_t K, size_t N_imf, size_t S,
                         const int32_t* d_x, const float* d_y, const float* d_V, float* d_IMFs,
                         const memd_opts* opts, double* elapsed_s);

/*
 * MEX argument layout (memdMatlabCuda.cu:996-1039, assistMatlabCuda.m:3-17):
 * prhs[0..3] = n, K, N_imf, S (host int32); prhs[4] = data (device, numel
 * values, memory [n][T], T = numel/n); prhs[5] = sample index (device int32
 * [T]); prhs[6] = primes (device int32 [n], P[0] used as -divisor);
 * plhs[0] = d_out (device [N_imf][n][T]).  Directions are generated from the
 * supplied primes and cached per (n, K, primes).
 */
memd_status memd_mex_compat_f32(memd_ctx* ctx, int32_t n, int32_t K, int32_t N_imf, int32_t S,
                                size_t numel, const float* d_data, const int32_t* d_index,
                                const int32_t* d_primes, float* d_out);
memd_status memd_mex_compat_f64(memd_ctx* ctx, int32_t n, int32_t K, int32_t N_imf, int32_t S,
                                size_t numel, const double* d_data, const int32_t* d_index,
                                const int32_t* d_primes, double* d_out);

/*
 * Batched: many independent epochs (own n, T, K, N_imf, S) advanced together,
 * one launch per kernel per sift step over all (epoch, direction) work items.
 * No reference counterpart (the reference runs one epoch per call,
 * memReducedEpochs_dev.cu:1074).  Per-epoch arithmetic is unchanged, so each
 * epoch's result is bit-identical to memd_run_f64 on it alone.
 */
typedef struct memd_epoch {
    size_t n, T, K, N_imf, S;
    const double* d_y;     /* [n][T] */
    const double* d_V;     /* [K][n] */
    double* d_IMFs;        /* [N_imf][n][T] */
    int32_t* h_sift_counts;/* optional host [N_imf] */
    int32_t* h_n_imf_out;  /* optional host */
} memd_epoch;

memd_status memd_run_batch_f64(memd_ctx* ctx, const memd_epoch* epochs, size_t count,
                               const memd_opts* opts, double* elapsed_s);

/*
 * Diagnostics (no reference counterpart): runs the batch once with direct
 * launches and HIP events between kernels, and returns the average device time
 * per launch of every kernel plus the algorithmic work counters the roofline
 * figures are computed from.  Results are the same as memd_run_batch_f64's.
 * Kernel slots: 0 flags, 1 compact, 2 factor, 3 block (knot values + block
 * sweeps), 4 reduced, 5 coef, 6 eval (stop mode: eval_stop), 7 advance (stop
 * mode: stop_decide), 8 stop_apply (stop mode only), 9 init.
 */
#define MEMD_NUM_KERNELS 10
typedef struct memd_kernel_profile {
    double ms_per_launch[MEMD_NUM_KERNELS];
    int64_t launches[MEMD_NUM_KERNELS];
    int64_t steps;          /* sift steps launched */
    int64_t knots;          /* sum over epochs, passes, directions, envelopes of E */
    int64_t knot_channels;  /* same, times the epoch's channel count */
    int64_t eval_points;    /* sum over epochs and passes of nvalid * n * T */
    int64_t samples;        /* sum over epochs and passes of n * T */
} memd_kernel_profile;

memd_status memd_profile_batch_f64(memd_ctx* ctx, const memd_epoch* epochs, size_t count,
                                   const memd_opts* opts, memd_kernel_profile* out);

/*
 * Tile shapes of the engine's kernels for epochs of n channels.  No counterpart
 * in the reference (its kernels have fixed launch shapes); the choice never
 * changes results, only speed.
 *   block_groups     solver blocks (32 knot rows) per work item of the block
 *                    solve, 1..8, limited by 160 KB of LDS for n channels;
 *   eval_tile_words  evaluation tile of 64 * w samples x 16 / w channels,
 *                    w in {1, 2, 4}.
 * memd_autotune times the candidates on a synthetic batch of (n, T, K) epochs
 * on this context's device and installs the fastest for n; memd_ctx_set_tuning
 * installs a choice (e.g. one saved from an earlier autotune);
 * memd_ctx_get_tuning reports the choice used for n (the built-in policy when
 * none is installed).  block_ms / eval_ms: the chosen shapes' kernel times per
 * sift step in memd_autotune (0 otherwise).
 */
typedef struct memd_tuning {
    int32_t n;
    int32_t block_groups;
    int32_t eval_tile_words;
    int32_t reserved;
    double block_ms;
    double eval_ms;
} memd_tuning;

memd_status memd_autotune(memd_ctx* ctx, size_t n, size_t T, size_t K, memd_tuning* out);
memd_status memd_ctx_set_tuning(memd_ctx* ctx, const memd_tuning* t);
memd_status memd_ctx_get_tuning(memd_ctx* ctx, int32_t n, memd_tuning* out);

/* Device bytes the engine needs for one fix-mode epoch of this shape with at
 * most 64 output slots (workspace only; no reference counterpart: the reference
 * allocates inside every call, memReducedEpochs_dev.cu:1077-1165). */
size_t memd_workspace_bytes(size_t n, size_t T, size_t K);
/* The same for a given number of output slots and options (opts == NULL: the
 * defaults).  Stop mode adds its per-interval record array, so size a stop-mode
 * server pool with this call. */
size_t memd_workspace_bytes_opts(size_t n, size_t T, size_t K, size_t N_imf, const memd_opts* opts);

/*
 * Raw sample files: float32, channel-major [n][T_file], no header -- the
 * reference demo's input (EEGsample_32CH_30504.bin) and output (mode_matrix.bin,
 * [N_imf][n][T]).  Replace getBinSize / readBin / writeBin
 * (memReducedEpochs_dev.cu:896-932) as main() (:1484-1501, :1526-1527) and
 * runBenchmark() (:1405-1426) use them.  Host memory only, no device needed;
 * failures return MEMD_ERR_IO / MEMD_ERR_ARG (the reference exits) with the
 * message in memd_last_error(NULL).
 */
/* T_file = bytes / sizeof(float) / n, floored like the reference (:1486). */
memd_status memd_bin_length(const char* path, size_t n, size_t* T_file);
/* h_y [n][T] <- the first T samples of each channel of an [n][T_file] file
 * (T == T_file: the whole file, as main(); T < T_file: runBenchmark's slicing). */
memd_status memd_read_bin_f32(const char* path, size_t n, size_t T_file, size_t T, float* h_y);
/* Write count floats (writeBin). */
memd_status memd_write_bin_f32(const char* path, const float* h, size_t count);

/*
 * Served continuous batching.  No reference counterpart: the reference runs one
 * epoch per blocking call (memReducedEpochs_dev.cu:1074) and allocates its
 * workspace inside it (:1077-1165).  A server owns a fixed HBM pool.  Between
 * sift steps it admits queued epochs whose workspace fits the pool (first fit,
 * FIFO order with bounded backfill), advances every active epoch by one sift
 * step per launch sequence, and retires each epoch as soon as it finishes,
 * returning its workspace to the pool.  Per-epoch arithmetic is unchanged, so
 * every result is bit-identical to memd_run_f64 on that epoch alone.  The IMF
 * stack is written into the request's d_IMFs on the device as sifting
 * proceeds; read it once the request is MEMD_REQ_DONE.
 */
typedef struct memd_server memd_server;
enum {
    MEMD_REQ_QUEUED = 0,
    MEMD_REQ_RUNNING = 1,
    MEMD_REQ_DONE = 2,
    MEMD_REQ_CANCELLED = 3,   /* withdrawn by memd_server_cancel */
    MEMD_REQ_TIMED_OUT = 4    /* withdrawn at its deadline (memd_server_set_deadline) */
};
/* Server counters (memd_server_get_info). */
typedef struct memd_server_info {
    int64_t steps;      /* sift steps queued (graph replays) */
    int64_t probes;     /* state probes taken (one per memd_server_step that queued steps) */
    int64_t rebuilds;   /* plans built, each with its captured step graph */
    int64_t retired;    /* requests finished */
    int64_t withdrawn;  /* requests cancelled or timed out */
    int64_t deadlines;  /* deadlines set */
    double rebuild_s;   /* host seconds spent building plans and capturing their graphs */
} memd_server_info;

/* hbm_pool_bytes: workspace pool (sum of memd_workspace_bytes_opts of the
 * epochs active at once, with the server's options); max_active: cap on epochs advanced together (0 = only the
 * pool limits it).  The options' mode, window and thresholds apply to every
 * request; the diagnostics fields must be unset.  Uses the context's stream. */
memd_status memd_server_create(memd_ctx* ctx, size_t hbm_pool_bytes, int32_t max_active, const memd_opts* opts,
                               memd_server** out);
memd_status memd_server_destroy(memd_server* srv);
/* Queue one epoch.  Its device buffers and optional host outputs
 * (h_sift_counts, h_n_imf_out, written at completion) must stay valid until the
 * request is done.  MEMD_ERR_ADMISSION: its workspace can never fit the pool. */
memd_status memd_server_submit(memd_server* srv, const memd_epoch* ep, int64_t* id);
/* Withdraw expired requests, admit what fits (rebuilding the plan and its step
 * graph when epochs join), queue `steps` replays of the step graph over the
 * active set with a state probe behind them, then wait for the PREVIOUS call's
 * probe and retire the epochs it shows finished.  The call returns while its own
 * steps run, so a loop of calls keeps the device busy; an epoch is reported done
 * by the call after the one whose steps finished it (or by a call with
 * steps = 0, which waits for every probe).  Optional outputs: active and queued
 * counts afterwards. */
memd_status memd_server_step(memd_server* srv, int32_t steps, int32_t* n_active, int32_t* n_queued);
/* Withdraw a request.  Queued: removed at once.  Running: the device is told to
 * skip it behind the queued work and the call waits for that, so its buffers
 * are no longer touched when it returns; its workspace returns to the pool and
 * its IMF buffer holds a partial result.  It then reads MEMD_REQ_CANCELLED and
 * is listed by memd_server_completed.  MEMD_ERR_ARG if it has already ended. */
memd_status memd_server_cancel(memd_server* srv, int64_t id);
/* Deadline `seconds` from now (host steady clock): a request that has not ended
 * by then is withdrawn as by memd_server_cancel at the next memd_server_step and
 * reads MEMD_REQ_TIMED_OUT.  seconds <= 0 clears the deadline. */
memd_status memd_server_set_deadline(memd_server* srv, int64_t id, double seconds);
memd_status memd_server_get_info(memd_server* srv, memd_server_info* out);
/* Pop up to cap ids of requests that ended (done, cancelled or timed out) since
 * the last call, in order. */
memd_status memd_server_completed(memd_server* srv, int64_t* ids, int32_t cap, int32_t* n);
/* MEMD_REQ_* state of a request; MEMD_ERR_NOT_FOUND for unknown or released ids. */
memd_status memd_server_poll(memd_server* srv, int64_t id, int32_t* state);
/* Forget a request that has ended. */
memd_status memd_server_release(memd_server* srv, int64_t id);
/* Pool bytes currently held by active epochs. */
size_t memd_server_pool_used(memd_server* srv);

#ifdef __cplusplus
}
#endif
#endif /* MEMD_AMD_H */
