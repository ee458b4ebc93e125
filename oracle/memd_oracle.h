/*
 * FP64 host oracle for MEMD (test infrastructure only).
 *
 * This library is the checker for the HIP engine: only the test-suite, the
 * smoke check and the CPU-baseline leg of the benchmark load it.  It is never
 * part of the product path.
 *
 * It restates, step by step, the epoch-level MEMD of
 * memReducedEpochs_dev.cu:1074-1368 (memd<coord_t,real_t>) in FP64, with the
 * arithmetic contract documented in memd_oracle.cpp (explicit fma where noted,
 * everything else plain IEEE +,-,*,/ with contraction disabled).
 */
#ifndef MEMD_ORACLE_H
#define MEMD_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Sifting policy. FIX = the reference's fixed S sifts per IMF
 * (memReducedEpochs_dev.cu:1176-1179); STOP = memd.m-style stop criterion. */
enum { ORACLE_MODE_FIX = 0, ORACLE_MODE_STOP = 1 };
/* Extrema window. REFWARP32 = the reference's 30-of-32-lane coverage
 * (memReducedEpochs_dev.cu:77-141 with the grid of :1191-1193); FULL = every
 * interior sample. */
enum { ORACLE_WINDOW_REFWARP32 = 0, ORACLE_WINDOW_FULL = 1 };

/* Spline end conditions. NATURAL = the reference's (memReducedEpochs_dev.cu:704-736);
 * NOT_A_KNOT = third-derivative continuity at the second and second-to-last knots
 * (the reference's dead tridiagonal_setup_nak, :738-773, in a pivot-free form). */
enum { ORACLE_SPLINE_NATURAL = 0, ORACLE_SPLINE_NOT_A_KNOT = 1 };

/* End treatment of the knot lists. CLAMP = the reference's: samples 0 and T-1 are
 * knots, raised / lowered to the linear extrapolation of the next two knots when
 * E >= 6 (memReducedEpochs_dev.cu:99-104, 447-486, 545-584). MIRROR = memd.m's
 * boundary_conditions with nbsym = 2 (not in the reference): the interior extrema
 * are extended by up to two mirror images per end, reflected about the end
 * extremum or the end sample, with no clamp. */
enum { ORACLE_BOUNDARY_CLAMP = 0, ORACLE_BOUNDARY_MIRROR = 1 };

typedef struct oracle_opts {
    int mode;        /* ORACLE_MODE_* */
    int window;      /* ORACLE_WINDOW_* */
    double sd, sd2, tol; /* stop thresholds (stop mode only) */
    int threads;     /* OpenMP threads; <=1 runs the serial code path */
    int spline;      /* ORACLE_SPLINE_* */
    int boundary;    /* ORACLE_BOUNDARY_* */
} oracle_opts;

typedef struct oracle_trace {
    /* optional outputs, any may be NULL */
    int32_t* knot_counts;    /* [max_sifts][K][2]: E+ (max) and E- (min) per direction per sift pass */
    int32_t max_sifts;       /* capacity of knot_counts in sift passes */
    int32_t* sift_counts;    /* [N_imf]: sifts (subtractions) per IMF */
    int32_t* n_imf_out;      /* number of output slots used (IMFs + residual) */
    int32_t* n_passes;       /* number of envelope passes performed */
} oracle_trace;

/* ---- directions (memReducedEpochs_dev.cu:934-1072; memdMatlabCuda.cu:915-978) ---- */
/* primes == NULL selects the C++ path's prime table [-K, 2, 3, 5, ...]
 * (memReducedEpochs_dev.cu:1033-1046); otherwise primes[0..n-1] is used as
 * given (MEX path, memdMatlabCuda.cu:1131-1137). V is [K][n]. */
int oracle_directions(int n, int K, const int32_t* primes, double* V);
double oracle_radical_inverse(int base, unsigned i);
int oracle_prime_table(int n, int K, int32_t* P);

/* ---- per-stage restatements (exposed for stage-level tests) ---- */
void oracle_project(int n, int T, int K, const double* cur, const double* V, double* P);
/* flags of one projection row: env 0 = maxima, 1 = minima. Returns E. X gets the
 * knot sample positions (size >= T). */
int oracle_extrema(int T, const double* Pk, int env, int window, int32_t* X);
/* end-knot clamp of memReducedEpochs_dev.cu:447-486 / :545-584; Y has E values */
void oracle_clamp(int E, const int32_t* X, double* Y, int env);
/* mirrored knot lists of one direction (ORACLE_BOUNDARY_MIRROR). Xmax / Xmin are
 * the clamp-mode lists of oracle_extrema (Emax, Emin >= 4: both ends included),
 * Pk the projection row. Writes per envelope e the positions P[e] and source
 * samples S[e] (capacity Emax + 2 / Emin + 2; a knot at position p carries the
 * signal at sample S) and returns the counts in E[2]; sym[2] gets the left and
 * right mirror points (samples). Returns 0, or -1 when a list has fewer than 4 knots. */
int oracle_mirror_knots(int T, const double* Pk, const int32_t* Xmax, int Emax, const int32_t* Xmin, int Emin,
                        int32_t* Pmax, int32_t* Smax, int32_t* Pmin, int32_t* Smin, int32_t* E, int32_t* sym);
/* natural-spline second-derivative-type solve for nrhs channels:
 * Y is [nrhs][E] (already clamped), M is [nrhs][E] output. */
void oracle_spline_solve(int E, const int32_t* X, int nrhs, const double* Y, double* M);
/* the same with a choice of end conditions (ORACLE_SPLINE_*); not-a-knot needs E >= 4 */
void oracle_spline_solve_bc(int E, const int32_t* X, int nrhs, const double* Y, double* M, int spline);
/* coefficients [E-1][4] = (a,b,c,d) per interval of one channel */
void oracle_spline_coefs(int E, const int32_t* X, const double* Y, const double* M, double* coef);
/* envelope of one channel at t = 0..T-1 */
void oracle_envelope(int T, int E, const int32_t* X, const double* coef, double* env);
/* plain tridiagonal solve with the oracle's blocked algorithm (rows 1..N of
 * arrays indexed from 0; lo/di/up/r length N+2, x length N+2, x[0]=x[N+1]=0) */
void oracle_tridiag_blocked(int N, const double* lo, const double* di, const double* up,
                            const double* r, double* x);

/* ---- one sift pass: envelopes of cur, mean over valid directions.
 * mean [n][T] receives sum/nvalid (NaN when no direction is valid);
 * amp [T] (may be NULL) receives memd.m's amplitude; counts [K][2] knot counts. */
void oracle_envelope_mean(int n, int T, int K, const double* cur, const double* V,
                          const oracle_opts* o, double* mean, double* amp, int32_t* counts);

/* ---- stop mode's decision for one pass (memd.m stop(); not in the reference):
 * mean [n][T] (sum over valid directions / nvalid), amp [T] (memd.m amplitude),
 * counts [K][2] knot counts (E+, E-, endpoints included).  Returns 1 to stop
 * sifting, 0 to subtract the mean and sift again.  oracle_memd calls it on every
 * pass with at least one valid direction. */
int oracle_stop_decision(int n, int T, int K, const double* mean, const double* amp, const int32_t* counts,
                         double sd, double sd2, double tol);

/* ---- whole epoch (memReducedEpochs_dev.cu:1074-1368) ----
 * y [n][T], V [K][n], imfs [N_imf][n][T]. Returns 0 on success. */
int oracle_memd(int n, int T, int K, int N_imf, int S, const double* y, const double* V,
                double* imfs, const oracle_opts* o, oracle_trace* tr);

/* ---- many epochs at once: OpenMP over epochs (threads), each epoch on one thread
 * through the serial path, so every result equals oracle_memd's for that epoch
 * alone (the CPU counterpart of a batched call).  o->threads is ignored. */
int oracle_memd_batch(int count, const int* n, const int* T, const int* K, const int* N_imf, const int* S,
                      const double* const* y, const double* const* V, double* const* imfs, const oracle_opts* o,
                      int threads);

/* the amplitude's channel reduction sum_c d[c]^2 in the engine's order: groups of
 * 4 channels, 4 groups to a block of 16, blocks in order */
double oracle_amp_block_sum(const double* d, int n);

/* blocked-solver block length (rows per block incl. the separator) */
int oracle_solver_block(void);
/* last interior sample examined by the extrema test */
int oracle_window_limit(int T, int window);

#ifdef __cplusplus
}
#endif
#endif
