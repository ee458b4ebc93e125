/*
 * FP64 host oracle for MEMD -- TEST INFRASTRUCTURE, NOT THE PRODUCT.
 *
 * Only tests/, the smoke check and the benchmark's CPU-baseline leg load this
 * library, and only as the checker.  The HIP engine never calls it.
 *
 * What it restates (file:line into the reference, EEGLab-Pannon/MEMD-GPU):
 *   directions          memReducedEpochs_dev.cu:934-1072 (MEX twin memdMatlabCuda.cu:915-978)
 *   projection          memReducedEpochs_dev.cu:1182-1184 (cublasSgemm, P = V * cur)
 *   extrema + window    memReducedEpochs_dev.cu:77-141, launch geometry :1191-1193
 *   exclusive scan      memReducedEpochs_dev.cu:143-382 (knot slot = #flags before t)
 *   compaction          memReducedEpochs_dev.cu:384-445 (E = scan[T-1] + 1)
 *   end-knot clamp      memReducedEpochs_dev.cu:447-486 (max), :545-584 (min)
 *   spline system       memReducedEpochs_dev.cu:704-736 (natural end conditions)
 *   tridiagonal solve   memReducedEpochs_dev.cu:1227-1241 (cusparseSgtsv2_nopivot, one
 *                       system per direction, n right-hand sides)
 *   coefficients        memReducedEpochs_dev.cu:775-793
 *   evaluation          memReducedEpochs_dev.cu:795-830
 *   mean + update       memReducedEpochs_dev.cu:832-864
 *   IMF loop            memReducedEpochs_dev.cu:1174-1179, 1328-1338
 *
 * Arithmetic contract (shared bit-for-bit with the HIP kernels):
 *   - FP64 throughout; compiled with -ffp-contract=off, so a*b+c is two roundings
 *     unless written as fma() below.
 *   - projection: acc = fma(V[k][c], cur[c][t], acc), c ascending, acc starts at 0.
 *   - envelope evaluation: fma(fma(fma(d,tau,c),tau,b),tau,a) (Horner with fma).
 *   - tridiagonal: the blocked Thomas algorithm of struct Blocked below (block length
 *     SOLVER_BLOCK), replacing the closed cuSPARSE gtsv2 solver whose internal
 *     order is not published.  Exact in exact arithmetic; pinned against dense
 *     solves and scipy's natural CubicSpline by the golden tests.
 *   - slopes and coefficients (memReducedEpochs_dev.cu:729-731, 775-793) use the
 *     per-interval reciprocals ih = 1/h and i3h = 1/(3h), shared by all channels:
 *     sl = (Y_{i+1}-Y_i)*ih, b = sl - (h*(2M_i+M_{i+1}))*(1/3), d = (M_{i+1}-M_i)*i3h
 *     (the reference divides; the results differ by at most an ulp per operation).
 *   - mean over directions: sum = sum + (e_max + e_min) * 0.5 in ascending k over
 *     valid directions, then cur - sum / nvalid (0/0 -> NaN as the reference).
 *   - not-a-knot ends (option; the reference's tridiagonal_setup_nak, :738-773, is
 *     dead code): its row-0 form has diagonal h0^2 - h1^2, zero whenever the first
 *     two gaps are equal, so a pivot-free solve of it divides by zero.  Here M_0
 *     and M_{E-1} are eliminated instead (de Boor's form): row 1 becomes
 *     [(h0+h1)(h0+2h1), h1^2-h0^2 | h1 r_1], row N = E-2 becomes
 *     [a^2-b^2, (a+b)(2a+b) | a r_N] with a = h_{N-1}, b = h_N (strictly
 *     diagonally dominant), and after the solve M_0 = ((h0+h1) M_1 - h0 M_2) * ih_1,
 *     M_{E-1} = ((a+b) M_{N} - b M_{N-1}) * ih_{N-1}.  The integer-valued row
 *     coefficients are exact in FP64.
 *
 * Stop mode (not in the reference, which only has fix mode) follows the public
 * memd.m criterion (oracle_stop_decision): sx(t) = |mean(t)|_2 / amp(t), stop when
 * !((frac(sx > sd) > tol || any(sx > sd2)) && any(nem > 2)); an IMF is not started
 * when every direction has fewer than 3 interior extrema.  The reference cannot
 * pin it (it has no stop mode); the tests pin it against an independent numpy/scipy
 * restatement of memd.m's loop (tests/golden/make_golden.py), known-answer cases
 * at every threshold and the limit cases sd = sd2 = tol = 0 (= fix mode) and
 * sd = sd2 = inf (no sift).  The amplitude's sum over channels of (e+ - e-)^2
 * runs in a fixed order of groups of 4 and blocks of 16 channels (amp_block_sum;
 * memd.m's sum() order is MATLAB's own), pinned by the block-sum known answers
 * and the numpy amplitude to 1e-11.
 *
 * Mirrored ends (option, not in the reference) follow memd.m's
 * boundary_conditions with nbsym = 2 (mirror_ends below): on the reference's
 * own interior extrema, with the reference's direction validity (both lists with
 * at least two interior extrema), no clamp, and the spline through the mirrored
 * knots (positions may lie left of 0 or right of T-1; a knot carries the signal
 * at its source sample).  Knot counts reported per pass stay the clamp lists'
 * (interior extrema + 2), so stop mode's nem is unchanged.  Pinned by a 1-based
 * MATLAB-style transcription of boundary_conditions knot by knot over every
 * reachable branch, scipy CubicSplines through the mirrored knots (one pass,
 * the fixed-count loop, and memd.m's own stop / mirror / not-a-knot setting),
 * the reflection invariants, and flat envelopes for level extrema.
 */
#include "memd_oracle.h"

#include <cmath>
#include <cstdlib>
#include <cstring>
#include <vector>
#include <algorithm>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

constexpr int SOLVER_BLOCK = 32;   // rows per block, separator included

inline double fma_(double a, double b, double c) { return std::fma(a, b, c); }

/* ---------------------------------------------------------------- directions */

bool is_composite(int v) {   // memReducedEpochs_dev.cu:1014-1027 (returns 1 for composite)
    for (int i = 2; i <= v / 2; ++i)
        if (v % i == 0) return true;
    return false;
}

/* ------------------------------------------------------------ tridiagonal core */

/* Factor rows p..q of (lo,di,up) as one Thomas block: den_p = di_p,
 * den_u = fma(-lo_u, cp_{u-1}, di_u); rd_u = 1/den_u; cp_u = up_u * rd_u. */
void tridiag_factor(int p, int q, const double* lo, const double* di, const double* up,
                    double* rd, double* cp) {
    for (int u = p; u <= q; ++u) {
        double den = (u == p) ? di[u] : fma_(-lo[u], cp[u - 1], di[u]);
        rd[u] = 1.0 / den;
        cp[u] = up[u] * rd[u];
    }
}

/* Solve rows p..q with the factors above: z_p = g_p*rd_p,
 * z_u = fma(-lo_u, z_{u-1}, g_u) * rd_u, x_q = z_q, x_u = fma(-cp_u, x_{u+1}, z_u). */
void tridiag_apply(int p, int q, const double* lo, const double* rd, const double* cp,
                   const double* g, double* z, double* x) {
    z[p] = g[p] * rd[p];
    for (int u = p + 1; u <= q; ++u) z[u] = fma_(-lo[u], z[u - 1], g[u]) * rd[u];
    x[q] = z[q];
    for (int u = q - 1; u >= p; --u) x[u] = fma_(-cp[u], x[u + 1], z[u]);
}

/* Blocked Thomas (shared by all right-hand sides of one direction).
 * Rows 1..N; x[0] = x[N+1] = 0 are the natural end conditions.
 * Separators sit at rows jL (j = 1..J, J = (N-1)/L); block b holds rows
 * p_b = bL+1 .. q_b = (b<J ? (b+1)L-1 : N).
 *   factor : per block, forward factors (rd, cp) and the two spike vectors
 *            alpha = A_b^{-1}(lo_p e_first) (b >= 1) and
 *            beta  = A_b^{-1}(up_q e_last)  (b < J), each by tridiag_apply;
 *   phase 1: per block and right-hand side, y = A_b^{-1} r by tridiag_apply;
 *   phase 2: the J-row reduced tridiagonal for the separator values X_j,
 *            built from the block-end entries of y, alpha, beta and solved by
 *            tridiag_apply;
 *   phase 3: x_u = fma(-X_{b+1}, beta_u, fma(-X_b, alpha_u, y_u)), each term
 *            present only where its spike exists.
 * With N <= L there is one block and x = y (plain Thomas). */
struct Blocked {
    int N = 0, L = SOLVER_BLOCK, J = 0;
    const double *lo = nullptr, *di = nullptr, *up = nullptr;
    std::vector<double> rd, cp;                  // per row, block-local forward factors
    std::vector<double> al, be;                  // spike vectors, per row of their block
    std::vector<double> Rlo, Rdi, Rup, Rrd, Rcp; // reduced system rows 1..J
    std::vector<double> z, g;                    // factor() scratch

    int bp_(int b) const { return b * L + 1; }
    int bq_(int b) const { return (b < J) ? (b + 1) * L - 1 : N; }

    void factor(int N_, const double* lo_, const double* di_, const double* up_) {
        N = N_; lo = lo_; di = di_; up = up_;
        J = (N >= 1) ? (N - 1) / L : 0;
        rd.assign(N + 2, 0.0); cp.assign(N + 2, 0.0);
        al.assign(N + 2, 0.0); be.assign(N + 2, 0.0);
        z.assign(N + 2, 0.0); g.assign(N + 2, 0.0);
        if (N <= 0) return;
        for (int b = 0; b <= J; ++b) tridiag_factor(bp_(b), bq_(b), lo, di, up, rd.data(), cp.data());
        if (J == 0) return;
        for (int b = 0; b <= J; ++b) {
            int p = bp_(b), q = bq_(b);
            if (b >= 1) {   // left spike
                for (int u = p; u <= q; ++u) g[u] = 0.0;
                g[p] = lo[p];
                tridiag_apply(p, q, lo, rd.data(), cp.data(), g.data(), z.data(), al.data());
            }
            if (b < J) {    // right spike
                for (int u = p; u <= q; ++u) g[u] = 0.0;
                g[q] = up[q];
                tridiag_apply(p, q, lo, rd.data(), cp.data(), g.data(), z.data(), be.data());
            }
        }
        // reduced system (rows 1..J); coefficients of X_{j-1}, X_j, X_{j+1}
        Rlo.assign(J + 2, 0.0); Rdi.assign(J + 2, 0.0); Rup.assign(J + 2, 0.0);
        Rrd.assign(J + 2, 0.0); Rcp.assign(J + 2, 0.0);
        for (int j = 1; j <= J; ++j) {
            int s = j * L;
            int qL = s - 1, pR = s + 1;   // last row of block j-1, first row of block j
            Rlo[j] = (j >= 2) ? -(lo[s] * al[qL]) : 0.0;
            Rdi[j] = fma_(-up[s], al[pR], fma_(-lo[s], be[qL], di[s]));
            Rup[j] = (j <= J - 1) ? -(up[s] * be[pR]) : 0.0;
        }
        tridiag_factor(1, J, Rlo.data(), Rdi.data(), Rup.data(), Rrd.data(), Rcp.data());
    }

    /* r: rows 1..N (index 0 and N+1 ignored); out: x[0..N+1] with x[0]=x[N+1]=0.
     * Reads only the factorisation: concurrent solves of one system (one per
     * channel) each bring their own scratch. */
    void solve(const double* r, double* out) const {
        out[0] = 0.0; out[N + 1] = 0.0;
        if (N <= 0) return;
        std::vector<double> z(N + 2, 0.0);
        if (J == 0) {
            tridiag_apply(1, N, lo, rd.data(), cp.data(), r, z.data(), out);
            return;
        }
        std::vector<double> y(N + 2, 0.0);
        std::vector<double> D(J + 2, 0.0), Zr(J + 2), Xs(J + 2, 0.0);
        for (int b = 0; b <= J; ++b)     // phase 1
            tridiag_apply(bp_(b), bq_(b), lo, rd.data(), cp.data(), r, z.data(), y.data());
        for (int j = 1; j <= J; ++j) {   // phase 2
            int s = j * L;
            D[j] = fma_(-up[s], y[s + 1], fma_(-lo[s], y[s - 1], r[s]));
        }
        tridiag_apply(1, J, Rlo.data(), Rrd.data(), Rcp.data(), D.data(), Zr.data(), Xs.data());
        for (int b = 0; b <= J; ++b) {   // phase 3
            int p = bp_(b), q = bq_(b);
            for (int u = p; u <= q; ++u) {
                double x = y[u];
                if (b >= 1) x = fma_(-Xs[b], al[u], x);
                if (b < J) x = fma_(-Xs[b + 1], be[u], x);
                out[u] = x;
            }
        }
        for (int j = 1; j <= J; ++j) out[j * L] = Xs[j];
    }
};

/* ----------------------------------------------------------- spline per env */

/* Shared (per direction, per envelope) spline system: h, lo, di, up from the
 * knot positions, natural rows 0 and m-1 (memReducedEpochs_dev.cu:704-736),
 * and the per-interval reciprocals ih = 1/h, i3h = 1/(3h) that every channel
 * shares (slopes and the d coefficient multiply by them). */
constexpr double kThird = 1.0 / 3.0;

struct KnotSystem {
    int E = 0;
    bool nak = false;
    std::vector<double> h, ih, i3h, lo, di, up;
    Blocked solver;
    void build(int E_, const int32_t* X, int spline = ORACLE_SPLINE_NATURAL) {
        E = E_;
        nak = (spline == ORACLE_SPLINE_NOT_A_KNOT) && E >= 4;
        h.assign(std::max(E - 1, 1), 0.0);
        ih.assign(std::max(E - 1, 1), 0.0);
        i3h.assign(std::max(E - 1, 1), 0.0);
        for (int i = 0; i + 1 < E; ++i) {
            h[i] = (double)(X[i + 1] - X[i]);
            ih[i] = 1.0 / h[i];
            i3h[i] = 1.0 / (3.0 * h[i]);
        }
        int N = E - 2;
        lo.assign(E + 1, 0.0); di.assign(E + 1, 0.0); up.assign(E + 1, 0.0);
        for (int u = 1; u <= N; ++u) {
            lo[u] = h[u - 1];
            di[u] = (double)(2 * ((X[u] - X[u - 1]) + (X[u + 1] - X[u])));
            up[u] = h[u];
        }
        if (nak) {
            long long h0 = X[1] - X[0], h1 = X[2] - X[1];
            lo[1] = 0.0;
            di[1] = (double)((h0 + h1) * (h0 + 2 * h1));
            up[1] = (double)(h1 * h1 - h0 * h0);
            long long a = X[N] - X[N - 1], b = X[N + 1] - X[N];
            lo[N] = (double)(a * a - b * b);
            di[N] = (double)((a + b) * (2 * a + b));
            up[N] = 0.0;
        }
        solver.factor(N, lo.data(), di.data(), up.data());
    }
    /* one channel: Y (clamped, E values) -> coefficients [E-1][4]; sl, r, M are the
     * caller's scratch (one set per thread) */
    void channel(const double* Y, double* coef, std::vector<double>& sl, std::vector<double>& r,
                 std::vector<double>& M) const {
        int N = E - 2;
        sl.resize(std::max(E - 1, 1));
        r.assign(E + 1, 0.0);
        M.assign(E + 1, 0.0);
        for (int i = 0; i + 1 < E; ++i) sl[i] = (Y[i + 1] - Y[i]) * ih[i];
        for (int u = 1; u <= N; ++u) r[u] = 3.0 * (sl[u] - sl[u - 1]);
        if (nak) {
            r[1] = r[1] * h[1];
            r[N] = r[N] * h[N - 1];
        }
        solver.solve(r.data(), M.data());   // M[0] = M[E-1] = 0
        if (nak) {
            M[0] = ((h[0] + h[1]) * M[1] - h[0] * M[2]) * ih[1];
            M[E - 1] = ((h[N - 1] + h[N]) * M[N] - h[N] * M[N - 1]) * ih[N - 1];
        }
        for (int j = 0; j + 1 < E; ++j) {
            coef[4 * j + 0] = Y[j];
            coef[4 * j + 1] = sl[j] - (h[j] * (2.0 * M[j] + M[j + 1])) * kThird;
            coef[4 * j + 2] = M[j];
            coef[4 * j + 3] = (M[j + 1] - M[j]) * i3h[j];
        }
    }
};

inline double horner(const double* cf, double tau) {
    return fma_(fma_(fma_(cf[3], tau, cf[2]), tau, cf[1]), tau, cf[0]);
}

void clamp_ends(int E, const int32_t* X, double* Y, int env) {
    // memReducedEpochs_dev.cu:451-485 (max) / :549-583 (min): only when scan[T-1] > 4, i.e. E >= 6
    if (E < 6) return;
    {
        double s = (Y[2] - Y[1]) / (double)(X[2] - X[1]);
        double tv = Y[1] - s * (double)(X[1] - X[0]);
        if (env == 0 ? (tv > Y[0]) : (tv < Y[0])) Y[0] = tv;
    }
    {
        int m = E - 1;
        double s = (Y[m - 1] - Y[m - 2]) / (double)(X[m - 1] - X[m - 2]);
        double tv = Y[m - 1] + s * (double)(X[m] - X[m - 1]);
        if (env == 0 ? (tv > Y[m]) : (tv < Y[m])) Y[m] = tv;
    }
}

int window_limit(int T, int window) {
    if (window == ORACLE_WINDOW_FULL) return T - 2;
    // grid of ceil(T/242) blocks x 8 warps; warp w tests samples 30w+1 .. 30w+30
    long G = ((long)T + 241) / 242;
    long lim = 240 * G;
    return (int)std::min<long>(T - 2, lim);
}

int extrema(int T, const double* P, int env, int window, int32_t* X) {
    int lim = window_limit(T, window);
    int E = 0;
    X[E++] = 0;
    for (int t = 1; t <= lim; ++t) {
        bool f = (env == 0) ? (P[t] > P[t - 1] && P[t] > P[t + 1])
                            : (P[t] < P[t - 1] && P[t] < P[t + 1]);
        if (f) X[E++] = t;
    }
    if (T - 1 > 0) X[E++] = T - 1;
    return E;
}

/* memd.m boundary_conditions (Rehman & Mandic's MEMD toolbox, after Rilling,
 * Flandrin & Goncalves' EMD boundary handling), nbsym = 2, in 0-based samples
 * (t(i) = i, so the first sample is 0 and the last T-1).  imax[0..a-1] and
 * imin[0..b-1] are the interior extrema (a, b >= 2), x the projection.  Per end:
 * up to two extrema of each kind, nearest the end, are reflected about the end
 * extremum (when the end sample lies beyond the first extremum of the other
 * kind) or about the end sample, which then joins one list as a knot; when a
 * reflection about an extremum does not reach the end sample the reflection
 * point becomes the end sample.  The output lists hold the sources (samples)
 * of the added knots in position order (position = 2 * sym - source). */
struct MirrorEnds {
    int nl[2], nr[2];        // knots added at the left / right end, per envelope (0 max, 1 min)
    int ls[2][3], rs[2][3];  // their sources, in position order
    int lsym, rsym;
};

int take_reversed(const int32_t* v, int lo, int hi, int* out) {   // v[max(lo,0) .. hi-1], last first
    int m = 0;
    for (int i = hi - 1; i >= std::max(lo, 0); --i) out[m++] = v[i];
    return m;
}

void mirror_ends(int T, const double* x, const int32_t* imax, int a, const int32_t* imin, int b, MirrorEnds& r) {
    const int last = T - 1;
    int* lmax = r.ls[0]; int* lmin = r.ls[1];
    int* rmax = r.rs[0]; int* rmin = r.rs[1];
    int& nlmax = r.nl[0]; int& nlmin = r.nl[1];
    int& nrmax = r.nr[0]; int& nrmin = r.nr[1];
    if (imax[0] < imin[0]) {
        if (x[0] > x[imin[0]]) {
            nlmax = take_reversed(imax, 1, std::min(a, 3), lmax);
            nlmin = take_reversed(imin, 0, std::min(b, 2), lmin);
            r.lsym = imax[0];
        } else {
            nlmax = take_reversed(imax, 0, std::min(a, 2), lmax);
            nlmin = take_reversed(imin, 0, std::min(b, 1), lmin);
            lmin[nlmin++] = 0;
            r.lsym = 0;
        }
    } else {
        if (x[0] < x[imax[0]]) {
            nlmax = take_reversed(imax, 0, std::min(a, 2), lmax);
            nlmin = take_reversed(imin, 1, std::min(b, 3), lmin);
            r.lsym = imin[0];
        } else {
            nlmax = take_reversed(imax, 0, std::min(a, 1), lmax);
            lmax[nlmax++] = 0;
            nlmin = take_reversed(imin, 0, std::min(b, 2), lmin);
            r.lsym = 0;
        }
    }
    // the first mirrored knot of either list must not lie right of sample 0
    if (2 * r.lsym - lmin[0] > 0 || 2 * r.lsym - lmax[0] > 0) {
        if (r.lsym == imax[0]) nlmax = take_reversed(imax, 0, std::min(a, 2), lmax);
        else nlmin = take_reversed(imin, 0, std::min(b, 2), lmin);
        r.lsym = 0;
    }
    if (imax[a - 1] < imin[b - 1]) {
        if (x[last] < x[imax[a - 1]]) {
            nrmax = take_reversed(imax, a - 2, a, rmax);
            nrmin = take_reversed(imin, b - 3, b - 1, rmin);
            r.rsym = imin[b - 1];
        } else {
            rmax[0] = last;
            nrmax = 1 + take_reversed(imax, a - 1, a, rmax + 1);
            nrmin = take_reversed(imin, b - 2, b, rmin);
            r.rsym = last;
        }
    } else {
        if (x[last] > x[imin[b - 1]]) {
            nrmax = take_reversed(imax, a - 3, a - 1, rmax);
            nrmin = take_reversed(imin, b - 2, b, rmin);
            r.rsym = imax[a - 1];
        } else {
            nrmax = take_reversed(imax, a - 2, a, rmax);
            rmin[0] = last;
            nrmin = 1 + take_reversed(imin, b - 1, b, rmin + 1);
            r.rsym = last;
        }
    }
    // the last mirrored knot of either list must not lie left of sample T-1
    if (2 * r.rsym - rmin[nrmin - 1] < last || 2 * r.rsym - rmax[nrmax - 1] < last) {
        if (r.rsym == imax[a - 1]) nrmax = take_reversed(imax, a - 2, a, rmax);
        else nrmin = take_reversed(imin, b - 2, b, rmin);
        r.rsym = last;
    }
}

/* Mirrored knot lists of one valid direction from the clamp-mode lists X[e]
 * (E[e] >= 4, ends included): positions P[e] and sources S[e]; returns E'. */
void mirror_lists(int T, const double* Pk, const int32_t* const X[2], const int E[2], int32_t* const P[2],
                  int32_t* const S[2], int Em[2], int sym[2]) {
    MirrorEnds m;
    mirror_ends(T, Pk, X[0] + 1, E[0] - 2, X[1] + 1, E[1] - 2, m);
    sym[0] = m.lsym;
    sym[1] = m.rsym;
    for (int e = 0; e < 2; ++e) {
        int k = 0;
        for (int i = 0; i < m.nl[e]; ++i, ++k) { S[e][k] = m.ls[e][i]; P[e][k] = 2 * m.lsym - m.ls[e][i]; }
        for (int i = 1; i + 1 < E[e]; ++i, ++k) { S[e][k] = X[e][i]; P[e][k] = X[e][i]; }
        for (int i = 0; i < m.nr[e]; ++i, ++k) { S[e][k] = m.rs[e][i]; P[e][k] = 2 * m.rsym - m.rs[e][i]; }
        Em[e] = k;
    }
}

/* The channel reduction of memd.m's amplitude, sum_c d_c^2 (d = e+ - e-), in the
 * order the fused GPU evaluation forms it: G_g = sequential sum over channels
 * 4g .. 4g+3, B_b = ((G_4b + G_4b+1) + G_4b+2) + G_4b+3 over the groups of block
 * b (channels 16b .. 16b+15), then the blocks in order.  For n <= 4 this is the
 * plain sequential sum.  d[c * stride] is channel c. */
double amp_block_sum(const double* d, size_t stride, int n) {
    double s = 0.0;
    for (int b = 0; 16 * b < n; ++b) {
        double B = 0.0;
        for (int g = 0; g < 4 && 16 * b + 4 * g < n; ++g) {
            double G = 0.0;
            const int c0 = 16 * b + 4 * g, c1 = std::min(n, c0 + 4);
            for (int c = c0; c < c1; ++c) {
                const double x = d[(size_t)c * stride];
                G = G + x * x;
            }
            B = (g == 0) ? G : B + G;
        }
        s = (b == 0) ? B : s + B;
    }
    return s;
}

struct Workspace {
    std::vector<double> P;
    std::vector<int32_t> X[2];
    int E[2] = {0, 0};
};

/* Envelopes of cur for all directions; accumulate mean into sum (ascending k),
 * optional memd.m amplitude.  nthreads>1 parallelises over channels / samples. */
void envelope_pass(int n, int T, int K, const double* cur, const double* V, int window, int spline, int boundary,
                   bool want_amp, int nthreads, double* sum, double* ampsum, double* nvalid_out,
                   int32_t* counts) {
    std::vector<double> Pk(T);
    std::vector<int32_t> X[2], MP[2], MS[2];
    X[0].resize(T); X[1].resize(T);
    const bool mirror = boundary == ORACLE_BOUNDARY_MIRROR;
    if (mirror) for (int e = 0; e < 2; ++e) { MP[e].resize(T + 4); MS[e].resize(T + 4); }
    KnotSystem sys[2];
    std::vector<double> diff;
    if (want_amp) diff.assign((size_t)n * T, 0.0);
    for (size_t i = 0; i < (size_t)n * T; ++i) sum[i] = 0.0;
    if (want_amp) for (int t = 0; t < T; ++t) ampsum[t] = 0.0;
    double nvalid = 0.0;
    (void)nthreads;
    for (int k = 0; k < K; ++k) {
        const double* Vk = V + (size_t)k * n;
#pragma omp parallel for schedule(static) num_threads(nthreads) if (nthreads > 1)
        for (int t = 0; t < T; ++t) {
            double acc = 0.0;
            for (int c = 0; c < n; ++c) acc = fma_(Vk[c], cur[(size_t)c * T + t], acc);
            Pk[t] = acc;
        }
        int E[2];
        for (int e = 0; e < 2; ++e) E[e] = extrema(T, Pk.data(), e, window, X[e].data());
        if (counts) { counts[2 * k] = E[0]; counts[2 * k + 1] = E[1]; }
        bool valid = (E[0] > 3) && (E[1] > 3);   // memReducedEpochs_dev.cu:840,857
        if (!valid) continue;
        nvalid = nvalid + 1.0;
        // knot positions (Xp) and the samples whose values they carry (Xs)
        const int32_t* Xp[2] = {X[0].data(), X[1].data()};
        const int32_t* Xs[2] = {X[0].data(), X[1].data()};
        int Ek[2] = {E[0], E[1]};
        if (mirror) {
            const int32_t* Xr[2] = {X[0].data(), X[1].data()};
            int32_t* P2[2] = {MP[0].data(), MP[1].data()};
            int32_t* S2[2] = {MS[0].data(), MS[1].data()};
            int sym[2];
            mirror_lists(T, Pk.data(), Xr, E, P2, S2, Ek, sym);
            for (int e = 0; e < 2; ++e) { Xp[e] = P2[e]; Xs[e] = S2[e]; }
        }
        for (int e = 0; e < 2; ++e) sys[e].build(Ek[e], Xp[e], spline);
#pragma omp parallel num_threads(nthreads) if (nthreads > 1)
        {
            std::vector<double> Y[2], coef[2], sl, r, M, env[2];
            for (int e = 0; e < 2; ++e) {
                Y[e].resize(Ek[e]); coef[e].resize(4 * (size_t)std::max(Ek[e] - 1, 1)); env[e].resize(T);
            }
#pragma omp for schedule(static)
            for (int c = 0; c < n; ++c) {
                const double* cc = cur + (size_t)c * T;
                for (int e = 0; e < 2; ++e) {
                    for (int i = 0; i < Ek[e]; ++i) Y[e][i] = cc[Xs[e][i]];
                    if (!mirror) clamp_ends(Ek[e], Xp[e], Y[e].data(), e);
                    sys[e].channel(Y[e].data(), coef[e].data(), sl, r, M);
                    // evaluation: interval j with X_j < t <= X_{j+1} (j = 0 for t <= X_1):
                    // the first knot is at or left of sample 0 in both end treatments
                    int j = 0;
                    const int32_t* Xe = Xp[e];
                    for (int t = 0; t < T; ++t) {
                        while (j + 1 < Ek[e] - 1 && Xe[j + 1] < t) ++j;
                        double tau = (double)(t - Xe[j]);
                        env[e][t] = horner(&coef[e][4 * (size_t)j], tau);
                    }
                }
                double* sc = sum + (size_t)c * T;
                for (int t = 0; t < T; ++t) sc[t] = sc[t] + (env[0][t] + env[1][t]) * 0.5;
                if (want_amp) {
                    double* dc = diff.data() + (size_t)c * T;
                    for (int t = 0; t < T; ++t) dc[t] = env[0][t] - env[1][t];
                }
            }
        }
        if (want_amp) {
#pragma omp parallel for schedule(static) num_threads(nthreads) if (nthreads > 1)
            for (int t = 0; t < T; ++t) {
                // sum_c (e+ - e-)^2 in a fixed order: sequentially within groups of 4
                // channels, the groups added 4 at a time into blocks of 16 channels, the
                // blocks added in channel order (amp_block_sum)
                ampsum[t] = ampsum[t] + std::sqrt(amp_block_sum(diff.data() + t, (size_t)T, n)) * 0.5;
            }
        }
    }
    *nvalid_out = nvalid;
}

}  // namespace

/* ============================================================= C interface */

extern "C" {

int oracle_solver_block(void) { return SOLVER_BLOCK; }
double oracle_amp_block_sum(const double* d, int n) { return amp_block_sum(d, 1, n); }
int oracle_window_limit(int T, int window) { return window_limit(T, window); }

double oracle_radical_inverse(int base, unsigned i) {
    // memReducedEpochs_dev.cu:934-946, in FP64 instead of float
    double digit, radical, inv;
    digit = radical = 1.0 / (double)base;
    inv = 0.0;
    while (i) {
        inv = inv + digit * (double)(i % (unsigned)base);
        digit = digit * radical;
        i /= (unsigned)base;
    }
    return inv;
}

int oracle_prime_table(int n, int K, int32_t* P) {
    // memReducedEpochs_dev.cu:1033-1046
    if (n < 1 || K < 1) return -1;
    int cnt = 1, cand = 2;
    while (cnt < n) {
        if (!is_composite(cand)) P[cnt++] = cand;
        ++cand;
    }
    P[0] = -K;
    return 0;
}

int oracle_directions(int n, int K, const int32_t* primes, double* V) {
    if (n < 1 || K < 1) return -1;
    std::vector<int32_t> P(n);
    if (primes) {
        for (int j = 0; j < n; ++j) P[j] = primes[j];
    } else {
        oracle_prime_table(n, K, P.data());
    }
    if (P[0] == 0) return -1;
    for (int j = 1; j < n; ++j) if (P[j] < 2) return -1;
    std::vector<double> b(n), tht(n > 1 ? n - 1 : 1);
    for (int k = 0; k < K; ++k) {
        // Hammersley point (memReducedEpochs_dev.cu:992-1005)
        b[0] = ((double)k + 1.5) / (double)(-P[0]);
        for (int j = 1; j < n; ++j) b[j] = oracle_radical_inverse(P[j], (unsigned)(k + 1));
        // Sequence2Vector (memReducedEpochs_dev.cu:948-990)
        for (int i = 0; i < n; ++i) b[i] = 2.0 * b[i] - 1.0;
        for (int i = 0; i < n - 1; ++i) {
            double tem = 0.0;
            for (int j = n - 1; j > i; --j) tem = tem + b[j] * b[j];
            tht[i] = std::atan2(std::sqrt(tem), b[i]);
        }
        double* v = V + (size_t)k * n;
        for (int i = 0; i < n; ++i) {
            if (i == 0) { v[i] = 1.0; continue; }
            double tem = 1.0;
            for (int j = 0; j < i; ++j) tem = tem * std::sin(tht[j]);
            v[i] = tem;
        }
        for (int i = 0; i < n - 1; ++i) v[i] = v[i] * std::cos(tht[i]);
    }
    return 0;
}

void oracle_project(int n, int T, int K, const double* cur, const double* V, double* P) {
    for (int k = 0; k < K; ++k)
        for (int t = 0; t < T; ++t) {
            double acc = 0.0;
            for (int c = 0; c < n; ++c) acc = fma_(V[(size_t)k * n + c], cur[(size_t)c * T + t], acc);
            P[(size_t)k * T + t] = acc;
        }
}

int oracle_extrema(int T, const double* Pk, int env, int window, int32_t* X) {
    return extrema(T, Pk, env, window, X);
}

void oracle_clamp(int E, const int32_t* X, double* Y, int env) { clamp_ends(E, X, Y, env); }

int oracle_mirror_knots(int T, const double* Pk, const int32_t* Xmax, int Emax, const int32_t* Xmin, int Emin,
                        int32_t* Pmax, int32_t* Smax, int32_t* Pmin, int32_t* Smin, int32_t* E, int32_t* sym) {
    if (Emax < 4 || Emin < 4) return -1;
    const int32_t* X[2] = {Xmax, Xmin};
    const int Er[2] = {Emax, Emin};
    int32_t* P[2] = {Pmax, Pmin};
    int32_t* S[2] = {Smax, Smin};
    int Em[2], sy[2];
    mirror_lists(T, Pk, X, Er, P, S, Em, sy);
    E[0] = Em[0]; E[1] = Em[1];
    sym[0] = sy[0]; sym[1] = sy[1];
    return 0;
}

void oracle_spline_solve(int E, const int32_t* X, int nrhs, const double* Y, double* M) {
    oracle_spline_solve_bc(E, X, nrhs, Y, M, ORACLE_SPLINE_NATURAL);
}

void oracle_spline_solve_bc(int E, const int32_t* X, int nrhs, const double* Y, double* M, int spline) {
    KnotSystem sys;
    sys.build(E, X, spline);
    std::vector<double> sl, r, Mv, coef(4 * (size_t)std::max(E - 1, 1));
    for (int c = 0; c < nrhs; ++c) {
        sys.channel(Y + (size_t)c * E, coef.data(), sl, r, Mv);
        for (int i = 0; i < E; ++i) M[(size_t)c * E + i] = Mv[i];
    }
}

void oracle_spline_coefs(int E, const int32_t* X, const double* Y, const double* M, double* coef) {
    for (int j = 0; j + 1 < E; ++j) {
        double hj = (double)(X[j + 1] - X[j]);
        double ih = 1.0 / hj, i3h = 1.0 / (3.0 * hj);
        double sl = (Y[j + 1] - Y[j]) * ih;
        coef[4 * j + 0] = Y[j];
        coef[4 * j + 1] = sl - (hj * (2.0 * M[j] + M[j + 1])) * kThird;
        coef[4 * j + 2] = M[j];
        coef[4 * j + 3] = (M[j + 1] - M[j]) * i3h;
    }
}

void oracle_envelope(int T, int E, const int32_t* X, const double* coef, double* env) {
    int j = 0;
    for (int t = 0; t < T; ++t) {
        while (j + 1 < E - 1 && X[j + 1] < t) ++j;
        env[t] = horner(&coef[4 * (size_t)j], (double)(t - X[j]));
    }
}

void oracle_tridiag_blocked(int N, const double* lo, const double* di, const double* up,
                            const double* r, double* x) {
    Blocked b;
    b.factor(N, lo, di, up);
    b.solve(r, x);
}

void oracle_envelope_mean(int n, int T, int K, const double* cur, const double* V,
                          const oracle_opts* o, double* mean, double* amp, int32_t* counts) {
    int window = o ? o->window : ORACLE_WINDOW_REFWARP32;
    int spline = o ? o->spline : ORACLE_SPLINE_NATURAL;
    int boundary = o ? o->boundary : ORACLE_BOUNDARY_CLAMP;
    int nth = (o && o->threads > 1) ? o->threads : 1;
    std::vector<double> ampsum(T);
    double nv = 0.0;
    envelope_pass(n, T, K, cur, V, window, spline, boundary, amp != nullptr, nth, mean, ampsum.data(), &nv, counts);
    for (size_t i = 0; i < (size_t)n * T; ++i) mean[i] = mean[i] / nv;
    if (amp) for (int t = 0; t < T; ++t) amp[t] = ampsum[t] / nv;
}

int oracle_stop_decision(int n, int T, int K, const double* mean, const double* amp, const int32_t* counts,
                         double sd, double sd2, double tol) {
    // memd.m stop(): sx(t) = |mean(t)|_2, divided by amp(t) only when every amp(t)
    // is nonzero; keep sifting while (frac(sx > sd) > tol || any(sx > sd2)) holds
    // and some direction has more than 2 interior extrema
    bool amp_nonzero = true;
    for (int t = 0; t < T; ++t)
        if (!(amp[t] != 0.0)) amp_nonzero = false;
    long cnt1 = 0;
    bool any2 = false;
    for (int t = 0; t < T; ++t) {
        double s = 0.0;
        for (int c = 0; c < n; ++c) {
            double m = mean[(size_t)c * T + t];
            s = s + m * m;
        }
        double sx = std::sqrt(s);
        if (amp_nonzero) sx = sx / amp[t];
        if (sx > sd) ++cnt1;
        if (sx > sd2) any2 = true;
    }
    bool anynem = false;
    for (int k = 0; k < K; ++k)
        if ((counts[2 * k] - 2) + (counts[2 * k + 1] - 2) > 2) anynem = true;
    return !((((double)cnt1 / (double)T) > tol || any2) && anynem) ? 1 : 0;
}

int oracle_memd(int n, int T, int K, int N_imf, int S, const double* y, const double* V,
                double* imfs, const oracle_opts* o, oracle_trace* tr) {
    if (n < 1 || T < 4 || K < 1 || N_imf < 1 || S < 1) return -1;
    int mode = o ? o->mode : ORACLE_MODE_FIX;
    int window = o ? o->window : ORACLE_WINDOW_REFWARP32;
    int spline = o ? o->spline : ORACLE_SPLINE_NATURAL;
    int boundary = o ? o->boundary : ORACLE_BOUNDARY_CLAMP;
    int nth = (o && o->threads > 1) ? o->threads : 1;
    double sd = o ? o->sd : 0.075, sd2 = o ? o->sd2 : 0.75, tol = o ? o->tol : 0.075;
    size_t nT = (size_t)n * T;
    std::vector<double> running(y, y + nT), sum(nT), ampsum(T), meanv(mode == ORACLE_MODE_STOP ? nT : 0),
        ampv(mode == ORACLE_MODE_STOP ? T : 0);
    std::vector<int32_t> counts((size_t)2 * K);
    int pass = 0;
    auto record = [&]() {
        if (tr && tr->knot_counts && pass < tr->max_sifts)
            std::memcpy(tr->knot_counts + (size_t)pass * 2 * K, counts.data(), sizeof(int32_t) * 2 * K);
        ++pass;
    };
    if (tr && tr->sift_counts) for (int i = 0; i < N_imf; ++i) tr->sift_counts[i] = 0;
    int imf = 0;
    if (mode == ORACLE_MODE_FIX) {
        // memReducedEpochs_dev.cu:1176-1338
        for (; imf < N_imf - 1; ++imf) {
            double* cur = imfs + (size_t)imf * nT;
            std::memcpy(cur, running.data(), sizeof(double) * nT);
            for (int s = 0; s < S; ++s) {
                double nv = 0.0;
                envelope_pass(n, T, K, cur, V, window, spline, boundary, false, nth, sum.data(), nullptr, &nv, counts.data());
                record();
                for (size_t i = 0; i < nT; ++i) cur[i] = cur[i] - sum[i] / nv;
            }
            if (tr && tr->sift_counts) tr->sift_counts[imf] = S;
            for (size_t i = 0; i < nT; ++i) running[i] = running[i] - cur[i];
        }
    } else {
        bool finished = false;
        for (; imf < N_imf - 1 && !finished; ++imf) {
            double* cur = imfs + (size_t)imf * nT;
            std::memcpy(cur, running.data(), sizeof(double) * nT);
            int nsift = 0;
            for (;;) {
                double nv = 0.0;
                envelope_pass(n, T, K, cur, V, window, spline, boundary, true, nth, sum.data(), ampsum.data(), &nv,
                              counts.data());
                record();
                if (nsift == 0) {   // memd.m stop_emd: every direction < 3 interior extrema
                    bool all_small = true;
                    for (int k = 0; k < K; ++k)
                        if ((counts[2 * k] - 2) + (counts[2 * k + 1] - 2) >= 3) all_small = false;
                    if (all_small) { finished = true; break; }
                }
                bool stop;
                if (nv == 0.0) {
                    stop = true;   // memd.m: no valid direction -> zero mean, stop
                } else {
                    for (size_t i = 0; i < nT; ++i) meanv[i] = sum[i] / nv;
                    for (int t = 0; t < T; ++t) ampv[t] = ampsum[t] / nv;
                    stop = oracle_stop_decision(n, T, K, meanv.data(), ampv.data(), counts.data(), sd, sd2, tol) != 0;
                }
                if (stop) break;
                for (size_t i = 0; i < nT; ++i) cur[i] = cur[i] - sum[i] / nv;
                ++nsift;
                if (nsift == S) break;
            }
            if (finished) break;
            if (tr && tr->sift_counts) tr->sift_counts[imf] = nsift;
            for (size_t i = 0; i < nT; ++i) running[i] = running[i] - cur[i];
        }
    }
    std::memcpy(imfs + (size_t)imf * nT, running.data(), sizeof(double) * nT);
    if (tr && tr->n_imf_out) *tr->n_imf_out = imf + 1;
    if (tr && tr->n_passes) *tr->n_passes = pass;
    return 0;
}

int oracle_memd_batch(int count, const int* n, const int* T, const int* K, const int* N_imf, const int* S,
                      const double* const* y, const double* const* V, double* const* imfs, const oracle_opts* o,
                      int threads) {
    // epochs in parallel, each on one thread through the serial path: every epoch's
    // result is the one oracle_memd gives it alone
    oracle_opts serial;
    if (o) {
        serial = *o;
    } else {
        serial.mode = ORACLE_MODE_FIX;
        serial.window = ORACLE_WINDOW_REFWARP32;
        serial.sd = 0.075; serial.sd2 = 0.75; serial.tol = 0.075;
        serial.spline = ORACLE_SPLINE_NATURAL;
        serial.boundary = ORACLE_BOUNDARY_CLAMP;
    }
    serial.threads = 1;
    int bad = 0;
    (void)threads;
#pragma omp parallel for schedule(dynamic, 1) num_threads(threads > 1 ? threads : 1) reduction(| : bad)
    for (int e = 0; e < count; ++e)
        bad |= oracle_memd(n[e], T[e], K[e], N_imf[e], S[e], y[e], V[e], imfs[e], &serial, nullptr) != 0 ? 1 : 0;
    return bad ? -1 : 0;
}

}  // extern "C"
