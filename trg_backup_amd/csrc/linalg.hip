// Float64 dense linear algebra for the isometry / decomposition steps.
//
//  * syevj: cyclic Jacobi eigensolver, one workgroup per matrix, parallel
//    (round-robin) ordering so the n/2 rotations of a round are applied
//    together; the matrix lives in LDS when it fits.  Jacobi keeps high
//    relative accuracy, which matters for the small end of a spectrum that
//    spans many decades.
//  * chol_inv: shifted Cholesky of an l x l Gram matrix and the inverse of
//    its upper-triangular factor: the building block of (shifted)
//    CholeskyQR, used instead of Householder QR because it is two GEMMs
//    plus an l^3 kernel.
//  * eigh_topk: top-k eigenpairs of an n x n PSD matrix by block subspace
//    iteration (block l > k), shift-regularised CholeskyQR3 orthonormalisation
//    and a Rayleigh-Ritz step -- only the k kept eigenvectors of the
//    environment are ever needed by HOTRG.
#include <math.h>

#include <mutex>
#include <utility>

#include "common.h"

namespace trg {

static int g_lds_limit[64];
static std::once_flag g_lds_once[64];

static int lds_limit() {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return 65536;
    std::call_once(g_lds_once[dev], [dev]() {
        int v = 65536;
        if (hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess) v = 65536;
        g_lds_limit[dev] = v;
    });
    return g_lds_limit[dev];
}

// Raise a kernel's dynamic-LDS limit to the device maximum once per process
// and device, so that later launches (including ones recorded into a HIP
// graph) issue no attribute calls.  Returns non-zero on failure.
static int ensure_dyn_lds(const void *fn, int slot) {
    static std::once_flag once[8][64];
    static int status[8][64];
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64 || slot < 0 || slot >= 8) return 1;
    std::call_once(once[slot][dev], [&]() {
        const int lim = lds_limit() - 16384;
        status[slot][dev] = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, lim) == hipSuccess ? 0 : 1;
    });
    return status[slot][dev];
}

// ---------------------------------------------------------------------------
// Jacobi eigensolver.  Work matrix W (n x n, row-major) either in LDS or in
// the caller's global buffer.  V accumulates the rotations.
// ---------------------------------------------------------------------------
constexpr int JT = 256;   // threads per workgroup
constexpr int JMAX = 256; // largest n

// Round-robin (circle method) position of slot j in round r of N players:
// slot 0 stays, the others rotate.  0 <= j - 1 + r < 2 (N - 1), so one
// conditional subtraction replaces the modulo.
__device__ static inline int rr_pos(int j, int r, int N) {
    if (j == 0) return 0;
    int x = j - 1 + r;
    if (x >= N - 1) x -= N - 1;
    return x + 1;
}

// LDS (template): the work matrix lives in LDS; the pointer then carries the
// LDS address space, so the compiler emits LDS instructions and can reorder
// them freely around the global accesses to V.
template <bool LDS>
__global__ void __launch_bounds__(JT) syevj_kernel(int n, double *Ag, long long sA, double *wout, double *Vout,
                                                   int max_sweeps, int *info) {
    extern __shared__ double lds[];
    __shared__ double cs[JMAX / 2], sn[JMAX / 2];
    __shared__ int pp[JMAX / 2], qq[JMAX / 2];
    __shared__ int rotated;
    __shared__ double red[JT];
    __shared__ double wv[JMAX];
    __shared__ int rank[JMAX];

    const int b = blockIdx.x;
    const int tid = threadIdx.x;
    double *Ab = Ag + (long long)b * sA;
    double *V = Vout + (long long)b * n * n;
    double *W;
    if constexpr (LDS) W = lds;
    else W = Ab;

    // symmetrised copy W = (A + A^T)/2, V = I, Frobenius norm for the floor
    double fro = 0.0;
    for (int e = tid; e < n * n; e += JT) {
        const int i = e / n, j = e % n;
        const double v = 0.5 * (Ab[i * n + j] + Ab[j * n + i]);
        fro += v * v;
        V[e] = (i == j) ? 1.0 : 0.0;
        if constexpr (LDS) W[e] = v;
    }
    red[tid] = fro;
    __syncthreads();
    if constexpr (!LDS) {
        // in place: symmetrise the upper/lower halves (reads above used the old values)
        for (int e = tid; e < n * n; e += JT) {
            const int i = e / n, j = e % n;
            if (i < j) {
                const double v = 0.5 * (W[i * n + j] + W[j * n + i]);
                W[i * n + j] = v;
                W[j * n + i] = v;
            }
        }
    }
    for (int s = JT / 2; s > 0; s >>= 1) {
        if (tid < s) red[tid] += red[tid + s];
        __syncthreads();
    }
    const double floor_abs = 1e-300 + 1e-17 * sqrt(red[0]);
    __syncthreads();

    const int N = n + (n & 1);
    const int npair = N / 2;
    // Rotation loops over (pair i, column j): each thread keeps one column j
    // and walks the pairs with a fixed stride, so the loops carry no integer
    // division (a runtime-divisor div/mod per element dominated them).
    const int groups = JT / n;                 // n <= JMAX == JT
    const int jcol = tid % n, grp = tid / n;
    const int cgroups = JT / npair;            // npair <= JT / 2
    const int cpair = tid % npair, cgrp = tid / npair;
    int sweep = 0;
    for (; sweep < max_sweeps; ++sweep) {
        if (tid == 0) rotated = 0;
        __syncthreads();
        for (int r = 0; r < N - 1; ++r) {
            if (tid < npair) {
                int p = rr_pos(tid, r, N), q = rr_pos(N - 1 - tid, r, N);
                if (p > q) { const int t = p; p = q; q = t; }
                double c = 1.0, s = 0.0;
                if (q < n) {
                    const double app = W[p * n + p], aqq = W[q * n + q], apq = W[p * n + q];
                    const double thr = fmax(2.2e-16 * sqrt(fabs(app * aqq)), floor_abs);
                    if (fabs(apq) > thr) {
                        const double tau = (aqq - app) / (2.0 * apq);
                        const double t = (tau >= 0 ? 1.0 : -1.0) / (fabs(tau) + sqrt(1.0 + tau * tau));
                        c = 1.0 / sqrt(1.0 + t * t);
                        s = t * c;
                        rotated = 1;
                    }
                }
                cs[tid] = c; sn[tid] = s; pp[tid] = p; qq[tid] = q;
            }
            __syncthreads();
            // rows: a_pj, a_qj <- c a_pj - s a_qj, s a_pj + c a_qj
            if (grp < groups) {
                const int j = jcol;
                for (int i = grp; i < npair; i += groups) {
                    const double s = sn[i];
                    if (s != 0.0) {
                        const int p = pp[i], q = qq[i];
                        const double c = cs[i];
                        const double ap = W[p * n + j], aq = W[q * n + j];
                        W[p * n + j] = c * ap - s * aq;
                        W[q * n + j] = s * ap + c * aq;
                    }
                }
            }
            __syncthreads();
            // columns of W and of V: the pair index runs fastest across the
            // threads, so a wave touches many columns of one row (distinct banks)
            if (cgrp < cgroups) {
                const int i = cpair;
                const double s = sn[i];
                if (s != 0.0) {
                    const int p = pp[i], q = qq[i];
                    const double c = cs[i];
                    for (int j = cgrp; j < n; j += cgroups) {
                        const double ap = W[j * n + p], aq = W[j * n + q];
                        W[j * n + p] = c * ap - s * aq;
                        W[j * n + q] = s * ap + c * aq;
                        const double vp = V[j * n + p], vq = V[j * n + q];
                        V[j * n + p] = c * vp - s * vq;
                        V[j * n + q] = s * vp + c * vq;
                    }
                }
            }
            __syncthreads();
        }
        if (!rotated) break;
        __syncthreads();
    }
    // eigenvalues, stable ascending rank, permuted eigenvector columns
    for (int i = tid; i < n; i += JT) wv[i] = W[i * n + i];
    __syncthreads();
    for (int i = tid; i < n; i += JT) {
        int rk = 0;
        const double wi = wv[i];
        for (int j = 0; j < n; ++j) {
            const double wj = wv[j];
            rk += (wj < wi) || (wj == wi && j < i);
        }
        rank[i] = rk;
        wout[(long long)b * n + rk] = wi;
    }
    __syncthreads();
    for (int e = tid; e < n * n; e += JT) W[e] = V[e];
    __syncthreads();
    for (int e = tid; e < n * n; e += JT) {
        const int j = e / n, i = e % n;
        V[j * n + rank[i]] = W[j * n + i];
    }
    if (tid == 0 && info) info[b] = (sweep >= max_sweeps) ? 1 : 0;
}

// ---------------------------------------------------------------------------
// Shifted Cholesky  R^T R = G + shift_rel * tr(G) * I  and  Rinv = R^{-1}.
// ---------------------------------------------------------------------------
// psd_tol > 0 selects the positive-semidefinite mode: a pivot below
// psd_tol * scale[b] zeroes its row of R (and the matching column of the
// generalised inverse) instead of flagging an error.  Rout (optional)
// receives R itself.
constexpr int CHOL_CT = 4;   // column slots of 64 per lane in the Cholesky update (l <= 256)

template <bool LDS>
__global__ void __launch_bounds__(256) chol_inv_kernel(int l, double *Gg, long long sG, double shift_rel,
                                                       double *Rinv, int *info, double psd_tol,
                                                       const double *scale, double *Rout) {
    extern __shared__ double lds[];
    __shared__ double red[256];
    __shared__ int bad;
    const int b = blockIdx.x, tid = threadIdx.x;
    double *Gb = Gg + (long long)b * sG;
    double *G;
    if constexpr (LDS) G = lds;
    else G = Gb;
    double *X = Rinv + (long long)b * l * l;
    if (tid == 0) bad = 0;
    double tr = 0.0;
    for (int e = tid; e < l * l; e += 256) {
        const int i = e / l, j = e % l;
        const double v = 0.5 * (Gb[i * l + j] + Gb[j * l + i]);
        if constexpr (LDS) G[e] = v;
        if (i == j) tr += v;
    }
    red[tid] = tr;
    __syncthreads();
    if constexpr (!LDS) {
        for (int e = tid; e < l * l; e += 256) {
            const int i = e / l, j = e % l;
            if (i < j) {
                const double v = 0.5 * (G[i * l + j] + G[j * l + i]);
                G[i * l + j] = v;
                G[j * l + i] = v;
            }
        }
    }
    for (int s = 128; s > 0; s >>= 1) {
        if (tid < s) red[tid] += red[tid + s];
        __syncthreads();
    }
    const double shift = shift_rel * red[0];
    for (int i = tid; i < l; i += 256) G[i * l + i] += shift;
    __syncthreads();
    const double tiny = 1e-300 + 1e-30 * fabs(red[0]);
    const double zero_below = psd_tol > 0 ? psd_tol * (scale ? scale[b] : fabs(red[0])) : -1.0;
    const int wv = tid >> 6, ln = tid & 63;
    // pivot of a diagonal value d; every lane that evaluates it gets the same r
    auto pivot = [&](double d, bool record) -> double {
        if (zero_below >= 0) {
            if (record && d != d) bad = 1;
            return (d > zero_below) ? sqrt(d) : 0.0;
        }
        if (!(d > tiny)) {   // also catches NaN
            if (record) bad = 1;
            d = tiny > 0 ? tiny : 1e-300;
        }
        return sqrt(d);
    };
    // Right-looking Cholesky, one barrier per column.  Row j of R is final
    // (divided by its pivot) when column j starts; the update of column j
    // leaves row j + 1 final too: wave 0 owns that row, takes its pivot from
    // the diagonal lane and divides the row before storing it.
    if (wv == 0) {
        const double r = pivot(G[0], ln == 0);
#pragma unroll
        for (int t = 0; t < CHOL_CT; ++t) {
            const int c = ln + 64 * t;
            if (c > 0 && c < l) G[c] = r > 0.0 ? G[c] / r : 0.0;
        }
        if (ln == 0) G[0] = r;
    }
    __syncthreads();
    for (int j = 0; j + 1 < l; ++j) {
        // trailing update G[a][c] -= R[j][a] R[j][c] (j < a <= c): each wave
        // takes every fourth row, each lane the columns ln, ln + 64, ... of it
        // (one LDS row segment per access, no integer division).  The row-j
        // factors of the lane's columns are read once; rows go in groups of
        // four whose loads are all issued before their updates.
        double rc[CHOL_CT];
#pragma unroll
        for (int t = 0; t < CHOL_CT; ++t) rc[t] = (ln + 64 * t < l) ? G[j * l + ln + 64 * t] : 0.0;
        for (int a0 = j + 1 + wv; a0 < l; a0 += 16) {
            double ra[4], g[4][CHOL_CT];
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const int a = a0 + 4 * u;
                ra[u] = a < l ? G[j * l + a] : 0.0;
#pragma unroll
                for (int t = 0; t < CHOL_CT; ++t) {
                    const int c = ln + 64 * t;
                    g[u][t] = (a < l && c >= a && c < l) ? G[a * l + c] : 0.0;
                }
            }
#pragma unroll
            for (int u = 0; u < 4; ++u)
#pragma unroll
                for (int t = 0; t < CHOL_CT; ++t) g[u][t] = g[u][t] - ra[u] * rc[t];
            if (a0 == j + 1) {
                // wave 0, row j + 1: its pivot, then the row divided by it
                const int dl = (j + 1) & 63, dt = (j + 1) >> 6;
                double mine = 0.0;
#pragma unroll
                for (int t = 0; t < CHOL_CT; ++t)
                    if (t == dt) mine = g[0][t];
                const double r = pivot(__shfl(mine, dl), ln == dl);
#pragma unroll
                for (int t = 0; t < CHOL_CT; ++t) {
                    const int c = ln + 64 * t;
                    if (c == j + 1) g[0][t] = r;
                    else if (c > j + 1) g[0][t] = r > 0.0 ? g[0][t] / r : 0.0;
                }
            }
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const int a = a0 + 4 * u;
#pragma unroll
                for (int t = 0; t < CHOL_CT; ++t) {
                    const int c = ln + 64 * t;
                    if (a < l && c >= a && c < l) G[a * l + c] = g[u][t];
                }
            }
        }
        __syncthreads();
    }
    if (Rout) {
        for (int e = tid; e < l * l; e += 256) {
            const int i = e / l, c = e % l;
            Rout[(long long)b * l * l + e] = (i <= c) ? G[e] : 0.0;
        }
    }
    // Inverse of the upper factor, column by column (X R = I): with the
    // leading j x j block already inverted (X), column j of X is
    // -X[0:j,0:j] R[0:j,j] / R[j][j] and X[j][j] = 1 / R[j][j].  X is kept
    // transposed in the lower triangle (X[i][c] at G[c][i], the diagonal shared
    // with R's), so thread i reads X[i][k] at G[k][i]: consecutive threads hit
    // consecutive words instead of one LDS bank.
    for (int j = 0; j < l; ++j) {
        const double gjj = G[j * l + j];
        const double djj = gjj != 0.0 ? 1.0 / gjj : 0.0;
        for (int i0 = 0; i0 < j; i0 += 256) {
            // every lane walks k = 0 .. j-1 (a uniform trip count, so the loads
            // of eight steps are issued together) and adds only its terms k >= i,
            // in ascending k: the same sum as a loop from k = i
            const int i = i0 + tid;
            const int ii = i < j ? i : j - 1;
            double v = 0.0;
            int k = 0;
            for (; k + 8 <= j; k += 8) {
                double x[8], r[8];
#pragma unroll
                for (int u = 0; u < 8; ++u) {
                    x[u] = G[(k + u) * l + ii];
                    r[u] = G[(k + u) * l + j];
                }
#pragma unroll
                for (int u = 0; u < 8; ++u)
                    if (k + u >= i) v += x[u] * r[u];
            }
            for (; k < j; ++k)
                if (k >= i) v += G[k * l + ii] * G[k * l + j];
            if (i < j) G[j * l + i] = -v * djj;
        }
        __syncthreads();
        if (tid == 0) G[j * l + j] = djj;
        __syncthreads();
    }
    // X[i][c] = G[c][i] for i <= c: consecutive threads read consecutive LDS words
    for (int e = tid; e < l * l; e += 256) {
        const int c = e / l, i = e % l;
        X[i * l + c] = (i <= c) ? G[e] : 0.0;
    }
    if (tid == 0 && info) info[b] = bad;
}

// ---------------------------------------------------------------------------
// Register-blocked variant of chol_inv for l <= 16 S (S <= 8): the 256 threads
// form a 16 x 16 grid and thread (ty, tx) keeps the S x S block of rows
// [S ty, S ty + S) and columns [S tx, S tx + S) in registers.  Column j of the
// Cholesky costs two barriers: the owner of the diagonal publishes the pivot,
// the owners of row j publish the divided row through LDS, and every thread
// updates its block (G[a][c] -= R[j][a] R[j][c], kept symmetric).  The inverse
// solves X R = I column by column with the same ownership: the owners of
// column c publish X[.][c], every thread subtracts X[i][c] R[c][q] from its
// columns q > c.  Same pivot rule and the same terms in the same order as
// chol_inv_kernel, so R and R^{-1} are the same bits.
// ---------------------------------------------------------------------------
template <int S>
__global__ void __launch_bounds__(256) chol_inv_reg_kernel(int l, double *Gg, long long sG, double shift_rel,
                                                           double *Rinv, int *info, double psd_tol,
                                                           const double *scale, double *Rout) {
    extern __shared__ double lds[];              // R (l x l) for the inverse
    __shared__ double rowbuf[16 * S];
    __shared__ double red[4];
    __shared__ double pivs;
    __shared__ int bad;
    const int b = blockIdx.x, tid = threadIdx.x;
    const int ty = tid >> 4, tx = tid & 15, r0 = S * ty, c0 = S * tx;
    const double *Gb = Gg + (long long)b * sG;
    double *X = Rinv + (long long)b * l * l;
    if (tid == 0) bad = 0;
    double a[S][S];
    double tr = 0.0;
#pragma unroll
    for (int i = 0; i < S; ++i)
#pragma unroll
        for (int j = 0; j < S; ++j) {
            const int ri = r0 + i, cj = c0 + j;
            const double v = (ri < l && cj < l) ? 0.5 * (Gb[ri * l + cj] + Gb[cj * l + ri]) : 0.0;
            a[i][j] = v;
            if (ri == cj) tr += v;
        }
    for (int off = 32; off > 0; off >>= 1) tr += __shfl_xor(tr, off);
    if ((tid & 63) == 0) red[tid >> 6] = tr;
    __syncthreads();
    const double trace = red[0] + red[1] + red[2] + red[3];
    const double shift = shift_rel * trace;
    if (tx == ty) {
#pragma unroll
        for (int i = 0; i < S; ++i)
            if (r0 + i < l) a[i][i] += shift;
    }
    const double tiny = 1e-300 + 1e-30 * fabs(trace);
    const double zero_below = psd_tol > 0 ? psd_tol * (scale ? scale[b] : fabs(trace)) : -1.0;
    for (int j = 0; j < l; ++j) {
        const int jb = j / S, jo = j - jb * S;
        if (ty == jb && tx == jb) {
            double d = 0.0;
#pragma unroll
            for (int i = 0; i < S; ++i)
                if (i == jo) d = a[i][i];
            double piv;
            if (zero_below >= 0) {
                piv = (d > zero_below) ? sqrt(d) : 0.0;
                if (d != d) bad = 1;
            } else {
                if (!(d > tiny)) {   // also catches NaN
                    bad = 1;
                    d = tiny > 0 ? tiny : 1e-300;
                }
                piv = sqrt(d);
            }
            pivs = piv;
        }
        __syncthreads();
        const double r = pivs;
        if (ty == jb) {
            // row j of R: the pivot on the diagonal, the rest divided by it
#pragma unroll
            for (int i = 0; i < S; ++i)
                if (i == jo) {
#pragma unroll
                    for (int q = 0; q < S; ++q) {
                        const int c = c0 + q;
                        double v = a[i][q];
                        if (c == j) v = r;
                        else if (c > j) v = r > 0.0 ? v / r : 0.0;
                        a[i][q] = v;
                        if (c < l) rowbuf[c] = c >= j ? v : 0.0;
                    }
                }
        }
        __syncthreads();
        if (r0 + S - 1 > j && c0 + S - 1 > j) {
            double ra[S], rc[S];
#pragma unroll
            for (int i = 0; i < S; ++i) {
                ra[i] = (r0 + i < l) ? rowbuf[r0 + i] : 0.0;
                rc[i] = (c0 + i < l) ? rowbuf[c0 + i] : 0.0;
            }
#pragma unroll
            for (int i = 0; i < S; ++i)
#pragma unroll
                for (int q = 0; q < S; ++q)
                    if (r0 + i > j && c0 + q > j) a[i][q] = a[i][q] - ra[i] * rc[q];
        }
    }
    // R (upper) to LDS, and to Rout when asked
#pragma unroll
    for (int i = 0; i < S; ++i)
#pragma unroll
        for (int q = 0; q < S; ++q) {
            const int ri = r0 + i, cj = c0 + q;
            if (ri < l && cj < l) {
                const double v = ri <= cj ? a[i][q] : 0.0;
                lds[ri * l + cj] = v;
                if (Rout) Rout[(long long)b * l * l + ri * l + cj] = v;
            }
        }
    // X R = I column by column, left to right, on B = I: column c of X is
    // B[.][c] / R[c][c] once columns 0 .. c-1 have been subtracted
    // (B[.][q] -= X[.][c] R[c][q] for q > c), i.e. X[i][c] =
    // -(sum_{k<c} X[i][k] R[k][c]) / R[c][c] with the terms in ascending k.
#pragma unroll
    for (int i = 0; i < S; ++i)
#pragma unroll
        for (int q = 0; q < S; ++q) a[i][q] = (r0 + i == c0 + q) ? 1.0 : 0.0;
    __syncthreads();
    for (int c = 0; c < l; ++c) {
        const int cb = c / S, co = c - cb * S;
        const double rcc = lds[c * l + c];
        const double dinv = rcc != 0.0 ? 1.0 / rcc : 0.0;
        if (tx == cb) {
#pragma unroll
            for (int q = 0; q < S; ++q)
                if (q == co) {
#pragma unroll
                    for (int i = 0; i < S; ++i) {
                        const int ri = r0 + i;
                        const double v = ri <= c ? a[i][q] * dinv : 0.0;
                        a[i][q] = v;
                        if (ri < l) rowbuf[ri] = v;
                    }
                }
        }
        __syncthreads();
        if (c0 + S - 1 > c) {
            double xr[S], rq[S];
#pragma unroll
            for (int i = 0; i < S; ++i) xr[i] = (r0 + i < l) ? rowbuf[r0 + i] : 0.0;
#pragma unroll
            for (int q = 0; q < S; ++q) rq[q] = (c0 + q > c && c0 + q < l) ? lds[c * l + c0 + q] : 0.0;
#pragma unroll
            for (int i = 0; i < S; ++i)
#pragma unroll
                for (int q = 0; q < S; ++q)
                    if (c0 + q > c) a[i][q] = a[i][q] - xr[i] * rq[q];
        }
        __syncthreads();
    }
#pragma unroll
    for (int i = 0; i < S; ++i)
#pragma unroll
        for (int q = 0; q < S; ++q) {
            const int ri = r0 + i, cj = c0 + q;
            if (ri < l && cj < l) X[ri * l + cj] = ri <= cj ? a[i][q] : 0.0;
        }
    if (tid == 0 && info) info[b] = bad;
}

// Launch of the shifted / PSD Cholesky with inverse: the register-blocked
// kernel for l <= 128, the LDS kernel (or, beyond the LDS, the in-place one)
// otherwise.
static int launch_chol_inv(int batch, int l, double *G, long long sG, double shift_rel, double *Rinv, int *info,
                           double psd_tol, const double *scale, double *R, hipStream_t s) {
    if (l <= 128) {
        const size_t bytes = (size_t)l * l * sizeof(double);
        const void *fn = l <= 32 ? (const void *)chol_inv_reg_kernel<2>
                                 : (l <= 64 ? (const void *)chol_inv_reg_kernel<4> : (const void *)chol_inv_reg_kernel<8>);
        if (bytes > 65536 && ensure_dyn_lds(fn, 3)) {
            g_last_hip_error = (int)hipGetLastError();
            return TRG_HIP_ERROR;
        }
        if (l <= 32)
            hipLaunchKernelGGL(chol_inv_reg_kernel<2>, batch, 256, bytes, s, l, G, sG, shift_rel, Rinv, info, psd_tol,
                               scale, R);
        else if (l <= 64)
            hipLaunchKernelGGL(chol_inv_reg_kernel<4>, batch, 256, bytes, s, l, G, sG, shift_rel, Rinv, info, psd_tol,
                               scale, R);
        else
            hipLaunchKernelGGL(chol_inv_reg_kernel<8>, batch, 256, bytes, s, l, G, sG, shift_rel, Rinv, info, psd_tol,
                               scale, R);
        return check_launch();
    }
    const size_t bytes = (size_t)l * l * sizeof(double);
    const int use_lds = (bytes + 16384 <= (size_t)lds_limit()) ? 1 : 0;
    if (use_lds && bytes > 65536) {
        if (ensure_dyn_lds((const void *)chol_inv_kernel<true>, 1)) {
            g_last_hip_error = (int)hipGetLastError();
            return TRG_HIP_ERROR;
        }
    }
    if (use_lds)
        hipLaunchKernelGGL(chol_inv_kernel<true>, batch, 256, bytes, s, l, G, sG, shift_rel, Rinv, info, psd_tol,
                           scale, R);
    else
        hipLaunchKernelGGL(chol_inv_kernel<false>, batch, 256, 0, s, l, G, sG, shift_rel, Rinv, info, psd_tol, scale,
                           R);
    return check_launch();
}

// Y[b] += shift_rel * t[b] * Q[b]   (t = trace of the operator)
__global__ void axpy_scaled_kernel(long long n, double shift_rel, const double *t, const double *Q, long long sQ,
                                   double *Y, long long sY) {
    const int b = blockIdx.y;
    const double a = shift_rel * t[b];
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
        Y[(long long)b * sY + i] += a * Q[(long long)b * sQ + i];
}

// eps[b] = t[b] - sum w[b][0..k)  (fixed order, one thread per instance)
__global__ void discarded_kernel(int batch, int k, const double *t, const double *w, long long sw, double *eps) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= batch) return;
    double s = 0.0;
    for (int i = 0; i < k; ++i) s += w[(long long)b * sw + i];
    eps[b] = t[b] - s;
}

// w[b][i] = theta[b][l-k+i]
__global__ void tail_copy_kernel(int batch, int l, int k, const double *theta, double *w) {
    const int b = blockIdx.y;
    for (int i = threadIdx.x; i < k; i += blockDim.x) w[(long long)b * k + i] = theta[(long long)b * l + (l - k) + i];
}

__global__ void or_info_kernel(int batch, const int *src, int *dst) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b < batch && src[b]) dst[b] = 1;
}

}  // namespace trg

using namespace trg;

extern "C" int trg_syevj_f64(int batch, int n, double *A, long long sA, double *w, double *V, int max_sweeps,
                             int *info, void *stream) {
    clear_hip_error();
    if (batch < 0 || n < 0 || max_sweeps < 1) return TRG_BAD_ARG;
    if (batch == 0 || n == 0) return TRG_OK;
    if (n > JMAX) return TRG_SHAPE;
    if (!A || !w || !V) return TRG_BAD_ARG;
    if (batch > 65535) return TRG_SHAPE;
    const size_t bytes = (size_t)n * n * sizeof(double);
    const int use_lds = (bytes + 16384 <= (size_t)lds_limit()) ? 1 : 0;
    if (use_lds && bytes > 65536) {
        if (ensure_dyn_lds((const void *)syevj_kernel<true>, 0)) {
            g_last_hip_error = (int)hipGetLastError();
            return TRG_HIP_ERROR;
        }
    }
    if (use_lds)
        hipLaunchKernelGGL(syevj_kernel<true>, batch, JT, bytes, as_stream(stream), n, A, sA, w, V, max_sweeps, info);
    else
        hipLaunchKernelGGL(syevj_kernel<false>, batch, JT, 0, as_stream(stream), n, A, sA, w, V, max_sweeps, info);
    return check_launch();
}

extern "C" int trg_chol_inv_f64(int batch, int l, double *G, long long sG, double shift_rel, double *Rinv, int *info,
                                void *stream) {
    clear_hip_error();
    if (batch < 0 || l < 0) return TRG_BAD_ARG;
    if (batch == 0 || l == 0) return TRG_OK;
    if (l > 256) return TRG_SHAPE;
    if (!G || !Rinv) return TRG_BAD_ARG;
    if (batch > 65535) return TRG_SHAPE;
    return launch_chol_inv(batch, l, G, sG, shift_rel, Rinv, info, 0.0, nullptr, nullptr, as_stream(stream));
}

// Workspace (structure of arrays over the batch, doubles): Q, Y (batch*n*l
// each), G, R, H, W (batch*l*l each), theta (batch*l), trace (batch); then
// int flags (batch).
extern "C" long long trg_eigh_topk_workspace(int batch, int n, int l) {
    if (batch <= 0 || n <= 0 || l <= 0) return 0;
    const long long dbl = (long long)batch * (2LL * n * l + 4LL * l * l + l + 1);
    return dbl * (long long)sizeof(double) + 4LL * batch + 256;
}

constexpr int GRAM_SLABS_MAX = 64;

// out = sum of `count` consecutive slabs of `len` doubles (fixed order).
__global__ void sum_slabs_kernel(int count, long long len, const double *__restrict__ P, double *__restrict__ out) {
    for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < len; e += (long long)gridDim.x * blockDim.x) {
        double acc = 0.0;
        for (int j = 0; j < count; ++j) acc += P[j * len + e];
        out[e] = acc;
    }
}

// Slab count of the split Gram of an n x l block: a function of the shape only,
// so an instance's summation order never depends on the batch it runs in.
static int gram_slabs(int n, int l) {
    const long long tiles = cdiv(l, 64) * cdiv(l, 64);
    if (tiles >= 16) return 1;
    return (int)std::max<long long>(1, std::min<long long>(GRAM_SLABS_MAX, n / 256));
}

// G[b] = X[b]^T X[b] for a tall-skinny n x l block.  The l x l product has
// only a handful of output tiles, so with `part` (GRAM_SLABS_MAX * l * l
// doubles per instance) the n rows are split into slabs whose partial Grams
// run as one batched GEMM and are summed in slab order.
static int gram_f64(int batch, int n, int l, const double *X, long long sX, double *G, double *part,
                    hipStream_t s) {
    const long long sG = (long long)l * l;
    const int S = part ? gram_slabs(n, l) : 1;
    if (S < 2) return trg_gemm(TRG_F64, TRG_F64, batch, l, l, n, 1.0, X, sX, 1, l, X, sX, l, 1, 0.0, G, sG, l, 1, s);
    const int kc = n / S;
    const int kl = n - (S - 1) * kc;
    for (int b = 0; b < batch; ++b) {
        const double *Xb = X + (long long)b * sX;
        double *Pb = part + (long long)b * S * sG;
        TRG_TRY(trg_gemm(TRG_F64, TRG_F64, S - 1, l, l, kc, 1.0, Xb, (long long)kc * l, 1, l, Xb, (long long)kc * l,
                         l, 1, 0.0, Pb, sG, l, 1, s));
        const double *Xt = Xb + (long long)(S - 1) * kc * l;
        TRG_TRY(trg_gemm(TRG_F64, TRG_F64, 1, l, l, kl, 1.0, Xt, 0, 1, l, Xt, 0, l, 1, 0.0, Pb + (S - 1) * sG, 0, l, 1,
                         s));
        hipLaunchKernelGGL(sum_slabs_kernel, (unsigned)std::min<long long>(cdiv(sG, 256), 1024), 256, 0, s, S, sG, Pb,
                           G + (long long)b * sG);
        TRG_TRY(check_launch());
    }
    return TRG_OK;
}

// Orthonormalise X (batch of n x l) with shifted CholeskyQR3 (one shifted
// pass, two plain passes; Fukaya et al., SIAM J. Sci. Comput. 42 (2020)), or
// with the shifted pass alone when passes == 1 (a well-conditioned basis of
// the same span, enough inside a power iteration).  The result lands in Y
// (X is clobbered).
static int cholqr3(int batch, int n, int l, double *X, double *Y, double *G, double *R, int *flags, int *info,
                   hipStream_t s, int passes = 3, double *part = nullptr) {
    const double shift0 = 11.0 * ((double)n * l + (double)l * (l + 1)) * 1.1102230246251565e-16;
    const long long sX = (long long)n * l, sG = (long long)l * l;
    double *src = X, *dst = Y;
    for (int pass = 0; pass < passes; ++pass) {
        // G = src^T src
        TRG_TRY(gram_f64(batch, n, l, src, sX, G, part, s));
        TRG_TRY(trg_chol_inv_f64(batch, l, G, sG, pass == 0 ? shift0 : 0.0, R, flags, s));
        if (info) {
            hipLaunchKernelGGL(or_info_kernel, (unsigned)cdiv(batch, 256), 256, 0, s, batch, flags, info);
            TRG_TRY(check_launch());
        }
        // dst = src R^{-1}
        TRG_TRY(trg_gemm(TRG_F64, TRG_F64, batch, n, l, l, 1.0, src, sX, l, 1, R, sG, l, 1, 0.0, dst, sX, l, 1, s));
        double *t = src; src = dst; dst = t;
    }
    return TRG_OK;
}

extern "C" int trg_eigh_topk_f64(int batch, int n, int k, int l, int q, const double *A, long long sA, double *w,
                                 double *V, double *eps, unsigned long long seed, void *ws, long long ws_bytes,
                                 int *info, void *stream) {
    clear_hip_error();
    if (batch < 0 || n < 0 || k < 0 || l < 0 || q < 0) return TRG_BAD_ARG;
    if (batch == 0 || n == 0 || k == 0) return TRG_OK;
    if (k > l || l > n || l > 256) return TRG_SHAPE;
    if (!A || !w || !V || !eps || !ws) return TRG_BAD_ARG;
    if (ws_bytes < trg_eigh_topk_workspace(batch, n, l)) return TRG_OOM;
    if (batch > 65535) return TRG_SHAPE;
    hipStream_t s = as_stream(stream);
    const long long sX = (long long)n * l, sG = (long long)l * l;
    double *Q = static_cast<double *>(ws);
    double *Y = Q + batch * sX;
    double *G = Y + batch * sX;
    double *R = G + batch * sG;
    double *H = R + batch * sG;
    double *W = H + batch * sG;
    double *theta = W + batch * sG;
    double *tr = theta + (long long)batch * l;
    int *flags = reinterpret_cast<int *>(tr + batch);
    if (info) TRG_TRY(trg_fill(3, batch, 0.0, info, s));

    long long tdims[1] = {n}, tst[1] = {(long long)n + 1};
    TRG_TRY(trg_trace(TRG_F64, batch, 1, tdims, tst, A, sA, tr, s));
    // Gaussian start block, the same draw for every instance (batch-invariant)
    TRG_TRY(trg_gaussian(TRG_F64, batch, sX, seed, 0u, 0u, 0ull, Q, sX, s));
    TRG_TRY(cholqr3(batch, n, l, Q, Y, G, R, flags, info, s));
    std::swap(Q, Y);
    // (A + sigma I) keeps every iterate of full column rank; sigma is far
    // below the kept eigenvalues and the Ritz step uses A itself.
    const double shift_rel = 1e-13;
    for (int it = 0; it < q; ++it) {
        TRG_TRY(trg_gemm(TRG_F64, TRG_F64, batch, n, l, n, 1.0, A, sA, n, 1, Q, sX, l, 1, 0.0, Y, sX, l, 1, s));
        const dim3 grid((unsigned)std::min<long long>(cdiv(sX, 256), 4096), batch);
        hipLaunchKernelGGL(axpy_scaled_kernel, grid, 256, 0, s, sX, shift_rel, tr, Q, sX, Y, sX);
        TRG_TRY(check_launch());
        // every iterate gets all three passes: after a single shifted pass the
        // next product with a wide-spectrum A is too ill-conditioned for the
        // unshifted passes (the Cholesky breaks down)
        TRG_TRY(cholqr3(batch, n, l, Y, Q, G, R, flags, info, s));   // result lands in Q
    }
    // Rayleigh-Ritz: H = Q^T A Q, H = W diag(theta) W^T
    TRG_TRY(trg_gemm(TRG_F64, TRG_F64, batch, n, l, n, 1.0, A, sA, n, 1, Q, sX, l, 1, 0.0, Y, sX, l, 1, s));
    TRG_TRY(trg_gemm(TRG_F64, TRG_F64, batch, l, l, n, 1.0, Q, sX, 1, l, Y, sX, l, 1, 0.0, H, sG, l, 1, s));
    TRG_TRY(trg_syevj_f64(batch, l, H, sG, theta, W, 40, flags, s));
    if (info) {
        hipLaunchKernelGGL(or_info_kernel, (unsigned)cdiv(batch, 256), 256, 0, s, batch, flags, info);
        TRG_TRY(check_launch());
    }
    // V = Q W[:, l-k:]  (eigenvectors of the k largest, ascending like eigh)
    TRG_TRY(trg_gemm(TRG_F64, TRG_F64, batch, n, k, l, 1.0, Q, sX, l, 1, W + (l - k), sG, l, 1, 0.0, V,
                     (long long)n * k, k, 1, s));
    hipLaunchKernelGGL(tail_copy_kernel, dim3(1, batch), 256, 0, s, batch, l, k, theta, w);
    TRG_TRY(check_launch());
    hipLaunchKernelGGL(discarded_kernel, (unsigned)cdiv(batch, 256), 256, 0, s, batch, k, tr, w, (long long)k, eps);
    return check_launch();
}

// ---------------------------------------------------------------------------
// Allocation-free top-k eigensolver.  The workspace of trg_eigh_topk_f64
// lives in a file-scope device array (each GPU holds its own copy, at its
// own address, for the life of the process).  The array's name is a host
// stand-in, not a device address: host code only ever passes the pointer
// hipGetSymbolAddress returned for the current device.  Every workspace
// word is written before it is read (Gaussian start block, Grams, Cholesky
// factors, per-instance flags), so nothing needs clearing between calls.
// Calls that overlap on one GPU share the array and race.
// ---------------------------------------------------------------------------
constexpr long long EIGH_SCRATCH_BYTES = 64LL << 20;
static __device__ __align__(256) unsigned char scratch[EIGH_SCRATCH_BYTES];

static void *g_scratch_ptr[64];
static int g_scratch_status[64];
static std::once_flag g_scratch_once[64];

// Device address of `scratch` on the current device (nullptr on failure).
static void *scratch_ptr() {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return nullptr;
    std::call_once(g_scratch_once[dev], [dev]() {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        g_scratch_status[dev] = (int)e;
        g_scratch_ptr[dev] = (e == hipSuccess) ? p : nullptr;
    });
    if (!g_scratch_ptr[dev]) g_last_hip_error = g_scratch_status[dev];
    return g_scratch_ptr[dev];
}

extern "C" long long trg_eigh_topk_noalloc_max_bytes(void) { return EIGH_SCRATCH_BYTES; }

extern "C" int trg_eigh_topk_f64_noalloc(int batch, int n, int k, int l, int q, const double *A, long long sA,
                                         double *w, double *V, double *eps, unsigned long long seed, int *info,
                                         void *stream) {
    clear_hip_error();
    if (batch < 0 || n < 0 || k < 0 || l < 0 || q < 0) return TRG_BAD_ARG;
    if (batch == 0 || n == 0 || k == 0) return TRG_OK;
    if (k > l || l > n || l > 256) return TRG_SHAPE;
    const long long need = trg_eigh_topk_workspace(batch, n, l);
    if (need > EIGH_SCRATCH_BYTES) return TRG_OOM;
    void *ws = scratch_ptr();
    if (!ws) return TRG_HIP_ERROR;
    return trg_eigh_topk_f64(batch, n, k, l, q, A, sA, w, V, eps, seed, ws, EIGH_SCRATCH_BYTES, info, stream);
}

// ---------------------------------------------------------------------------
// One-sided (Hestenes) Jacobi SVD of an m x n matrix, n <= m <= 256 per
// instance, one workgroup each.  Columns are orthogonalised pairwise in the
// round-robin order; singular values are the final column norms (high
// relative accuracy), U = W diag(1/s), V accumulates the rotations.
// Output: s descending (stride n), U (m x n row-major, stride m*n), V
// (n x n, stride n*n, columns are right singular vectors).
// ---------------------------------------------------------------------------
// Layout: the work columns live in LDS row-major with an odd row pitch
// (n + 1 doubles), so the lanes of a wave walking one column down its rows hit
// distinct banks; V is kept transposed in the output buffer during the sweeps
// (rows of Vt are the columns of V), so the lanes' accesses are contiguous.
// Each pair's dot products are split over eight lanes (three shuffle steps to
// reduce); the rotation pass issues the loads of eight pairs before their
// stores.
constexpr int GJ_ROT = 8;      // pairs per batch of rotation loads

__device__ inline int gesvj_pitch(int n) { return n | 1; }

__global__ void __launch_bounds__(256) gesvj_kernel(int m, int n, const double *Ag, long long sA, double *sout,
                                                    double *Uout, double *Vout, int max_sweeps, int *info) {
    extern __shared__ double lds[];
    __shared__ double cs[JMAX / 2], sn[JMAX / 2];
    __shared__ int pp[JMAX / 2], qq[JMAX / 2];
    __shared__ int rotated;
    __shared__ double nrm[JMAX];
    __shared__ int rank[JMAX];
    const int b = blockIdx.x, tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const double *A = Ag + (long long)b * sA;
    double *U = Uout + (long long)b * m * n;
    double *Vt = Vout + (long long)b * n * n;      // V^T during the sweeps
    double *W = lds;                               // work columns, row pitch P
    const int P = gesvj_pitch(n);
    __shared__ double red[4];
    double fro = 0.0;
    for (int e = tid; e < m * n; e += 256) {
        const int k = e / n, j = e - k * n;
        const double x = A[e];
        W[k * P + j] = x;
        fro += x * x;
    }
    for (int e = tid; e < n * n; e += 256) Vt[e] = (e / n == e % n) ? 1.0 : 0.0;
    for (int off = 32; off > 0; off >>= 1) fro += __shfl_xor(fro, off);
    if (lane == 0) red[wave] = fro;
    __syncthreads();
    // Pairs whose inner product sits below (eps ||A||_F)^2 are noise-level
    // columns of a rank-deficient matrix: rotating them changes nothing
    // above rounding, and a purely relative test would keep rotating them
    // for ever (no convergence); they are left alone.
    const double floor_g = 4.93e-32 * (red[0] + red[1] + red[2] + red[3]);
    // relative rotation threshold sqrt(m) eps: below it the cosine of the two
    // columns is at the rounding level of their length-m dot product, so a
    // rotation cannot make them more orthogonal (a threshold of eps alone
    // leaves convergence to rounding luck)
    const double ctol = 2.2e-16 * sqrt((double)m);
    const int N = n + (n & 1);
    const int npair = N / 2;
    const int mn = m + n;
    const int ugroups = mn >= 256 ? 1 : 256 / mn, ustride = mn >= 256 ? 256 : mn;
    const int uk0 = mn >= 256 ? tid : tid % mn, ugrp = mn >= 256 ? 0 : tid / mn;
    // rotation of columns p, q of one row (of W in LDS or of V^T's transposed row)
    auto rotate = [&](double *row, int step) {
        for (int i0 = ugrp; i0 < npair; i0 += GJ_ROT * ugroups) {
            double x[GJ_ROT], y[GJ_ROT];
#pragma unroll
            for (int u = 0; u < GJ_ROT; ++u) {
                const int i = i0 + u * ugroups;
                x[u] = y[u] = 0.0;
                if (i < npair && sn[i] != 0.0) {
                    x[u] = row[pp[i] * step];
                    y[u] = row[qq[i] * step];
                }
            }
#pragma unroll
            for (int u = 0; u < GJ_ROT; ++u) {
                const int i = i0 + u * ugroups;
                if (i < npair) {
                    const double s = sn[i];
                    if (s != 0.0) {
                        const double c = cs[i];
                        row[pp[i] * step] = c * x[u] - s * y[u];
                        row[qq[i] * step] = s * x[u] + c * y[u];
                    }
                }
            }
        }
    };
    int sweep = 0;
    for (; sweep < max_sweeps; ++sweep) {
        if (tid == 0) rotated = 0;
        __syncthreads();
        for (int r = 0; r < N - 1; ++r) {
            // wave w handles pairs w, w + 4, ...: eight at a time, each by a
            // group of eight lanes that splits the m rows and reduces with
            // three shuffle steps; the group's first lane forms the rotation
            {
                const int gq = lane >> 3, gt = lane & 7;
                for (int c0 = wave; c0 < npair; c0 += 32) {
                    const int i = c0 + 4 * gq;
                    double a = 0.0, bq = 0.0, g = 0.0;
                    int p = 0, q = 0;
                    if (i < npair) {
                        p = rr_pos(i, r, N);
                        q = rr_pos(N - 1 - i, r, N);
                        if (p > q) { const int t = p; p = q; q = t; }
                        if (q < n) {
                            for (int k = gt; k < m; k += 8) {
                                const double x = W[k * P + p], y = W[k * P + q];
                                a += x * x; bq += y * y; g += x * y;
                            }
                        }
                    }
                    for (int off = 4; off > 0; off >>= 1) {
                        a += __shfl_xor(a, off);
                        bq += __shfl_xor(bq, off);
                        g += __shfl_xor(g, off);
                    }
                    if (gt == 0 && i < npair) {
                        double c = 1.0, s = 0.0;
                        if (q < n && fabs(g) > fmax(ctol * sqrt(a * bq), floor_g) && g != 0.0) {
                            const double zeta = (bq - a) / (2.0 * g);
                            const double t = (zeta >= 0 ? 1.0 : -1.0) / (fabs(zeta) + sqrt(1.0 + zeta * zeta));
                            c = 1.0 / sqrt(1.0 + t * t);
                            s = c * t;
                            rotated = 1;
                        }
                        cs[i] = c; sn[i] = s; pp[i] = p; qq[i] = q;
                    }
                }
            }
            __syncthreads();
            // (row k of W or of V^T's columns, pairs with a fixed stride)
            if (ugrp < ugroups) {
                for (int k = uk0; k < mn; k += ustride) {
                    if (k < m) rotate(W + k * P, 1);
                    else rotate(Vt + (k - m), n);
                }
            }
            __syncthreads();
        }
        if (!rotated) break;
        __syncthreads();
    }
    // column norms, descending rank
    for (int j = wave; j < n; j += 4) {
        double a = 0.0;
        for (int k = lane; k < m; k += 64) a += W[k * P + j] * W[k * P + j];
        for (int off = 32; off > 0; off >>= 1) a += __shfl_xor(a, off);
        if (lane == 0) nrm[j] = sqrt(a);
    }
    __syncthreads();
    for (int i = tid; i < n; i += 256) {
        int rk = 0;
        const double si = nrm[i];
        for (int j = 0; j < n; ++j) rk += (nrm[j] > si) || (nrm[j] == si && j < i);
        rank[i] = rk;
        sout[(long long)b * n + rk] = si;
    }
    __syncthreads();
    // U[:, rank[j]] = W[:, j] / s_j (zero column for s_j = 0)
    for (int e = tid; e < m * n; e += 256) {
        const int k = e / n, j = e - k * n;
        U[k * n + rank[j]] = nrm[j] > 0.0 ? W[k * P + j] / nrm[j] : 0.0;
    }
    __syncthreads();
    // V[k][rank[j]] = Vt[j][k], staged through LDS (W is no longer needed)
    for (int e = tid; e < n * n; e += 256) W[e] = Vt[e];
    __syncthreads();
    for (int e = tid; e < n * n; e += 256) {
        const int k = e / n, j = e - k * n;
        Vt[k * n + rank[j]] = W[j * n + k];
    }
    if (tid == 0 && info) info[b] = (sweep >= max_sweeps) ? 1 : 0;
}

// Rank selection of utility/truncated_svd.py:17-78 on the device for one
// descending vector s (length n):  writes chi and the tail-norm error.
constexpr int TRUNC_MAX = 4096;   // longest spectrum the rank rule accepts

__global__ void truncate_kernel(int n, const double *s, int k, double deg_eps, double trunc_eps, int *chi_out,
                                double *err_out) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    if (deg_eps == 0.0 && trunc_eps == 0.0) { *chi_out = k; *err_out = 0.0; return; }
    if (k > n) { *chi_out = n; *err_out = 0.0; return; }
    const double s0 = s[0];
    auto delta = [&](int i) -> int {   // diff_i > deg_eps ?  (diff_0 = 0, diff_n = 1)
        if (i == 0) return 0 > deg_eps ? 1 : 0;
        if (i >= n) return 1.0 > deg_eps ? 1 : 0;
        const double a = s[i] / s0, p = s[i - 1] / s0;
        const double d = fabs(a - p) / fabs(a);
        return d > deg_eps ? 1 : 0;    // NaN compares false
    };
    int chi1 = k;
    for (chi1 = k; chi1 > 0; --chi1)
        if (delta(chi1) == 1) break;
    if (chi1 == 0) chi1 = 1;
    // suffix sums suf[i] = sum_{j>=i} (s_j/s_0)^2 (no cancellation for tiny tails)
    __shared__ double suf[TRUNC_MAX + 1];
    suf[n] = 0.0;
    for (int i = n - 1; i >= 0; --i) suf[i] = suf[i + 1] + (s[i] / s0) * (s[i] / s0);
    double tot = 0.0;
    for (int i = 0; i < n; ++i) tot += (s[i] / s0) * (s[i] / s0);
    // errors[i] = sqrt(suf[i] / tot), decreasing in i
    int chi2 = 0;
    for (int i = 0; i < chi1; ++i)
        if (sqrt(suf[i] / tot) > trunc_eps) ++chi2;
    int chi = chi1;
    if (chi2 < chi1) {
        for (chi = chi2; chi <= chi1; ++chi)
            if (delta(chi) == 1) break;
        if (chi > chi1) chi = chi1;
    }
    const double err = chi < n ? sqrt(suf[chi] / tot) : 0.0;
    *chi_out = chi;
    *err_out = err;
}

extern "C" int trg_gesvj_f64(int batch, int m, int n, const double *A, long long sA, double *s, double *U, double *V,
                             int max_sweeps, int *info, void *stream) {
    clear_hip_error();
    if (batch < 0 || m < 0 || n < 0 || max_sweeps < 1) return TRG_BAD_ARG;
    if (batch == 0 || m == 0 || n == 0) return TRG_OK;
    if (n > m || m > JMAX) return TRG_SHAPE;
    if (!A || !s || !U || !V) return TRG_BAD_ARG;
    if (batch > 65535) return TRG_SHAPE;
    const size_t bytes = (size_t)m * (n | 1) * sizeof(double);
    if (bytes + 16384 > (size_t)lds_limit()) return TRG_SHAPE;
    if (bytes > 65536) {
        if (ensure_dyn_lds((const void *)gesvj_kernel, 2)) {
            g_last_hip_error = (int)hipGetLastError();
            return TRG_HIP_ERROR;
        }
    }
    hipLaunchKernelGGL(gesvj_kernel, batch, 256, bytes, as_stream(stream), m, n, A, sA, s, U, V, max_sweeps, info);
    return check_launch();
}

extern "C" int trg_truncate_rank(int n, const double *s, int k, double degeneracy_eps, double truncate_eps, int *chi,
                                 double *err, void *stream) {
    clear_hip_error();
    if (n < 0 || k < 0) return TRG_BAD_ARG;
    if (!s || !chi || !err) return TRG_BAD_ARG;
    if (n == 0 || n > TRUNC_MAX) return TRG_SHAPE;
    hipLaunchKernelGGL(truncate_kernel, 1, 64, 0, as_stream(stream), n, s, k, degeneracy_eps, truncate_eps, chi, err);
    return check_launch();
}

extern "C" int trg_potrf_psd_f64(int batch, int l, double *G, long long sG, double psd_tol, const double *scale,
                                 double *R, double *Rinv, void *stream) {
    clear_hip_error();
    if (batch < 0 || l < 0 || psd_tol <= 0) return TRG_BAD_ARG;
    if (batch == 0 || l == 0) return TRG_OK;
    if (l > 256) return TRG_SHAPE;
    if (!G || !Rinv) return TRG_BAD_ARG;
    if (batch > 65535) return TRG_SHAPE;
    return launch_chol_inv(batch, l, G, sG, 0.0, Rinv, nullptr, psd_tol, scale, R, as_stream(stream));
}

extern "C" long long trg_orthonormalize_workspace(int batch, int m, int l) {
    if (batch <= 0 || m <= 0 || l <= 0) return 0;
    const long long slabs = gram_slabs(m, l) > 1 ? (long long)gram_slabs(m, l) * l * l : 0;
    return (long long)batch * ((long long)m * l + 2LL * l * l + slabs) * (long long)sizeof(double) + 4LL * batch +
           256;
}

extern "C" int trg_orthonormalize_f64(int batch, int m, int l, double *X, long long sX, double *Q, void *ws,
                                      long long ws_bytes, int passes, int *info, void *stream) {
    clear_hip_error();
    // X (batch x m x l, stride sX) is clobbered; the orthonormal basis of its
    // column span lands in Q (stride m*l).  Shifted CholeskyQR3 (passes = 3)
    // or its shifted first pass alone (passes = 1).
    if (batch < 0 || m < 0 || l < 0 || (passes != 1 && passes != 3)) return TRG_BAD_ARG;
    if (batch == 0 || m == 0 || l == 0) return TRG_OK;
    if (l > 256 || l > m) return TRG_SHAPE;
    if (!X || !Q || !ws) return TRG_BAD_ARG;
    if (ws_bytes < trg_orthonormalize_workspace(batch, m, l)) return TRG_OOM;
    hipStream_t s = as_stream(stream);
    const long long mx = (long long)m * l, sG = (long long)l * l;
    double *Xc = static_cast<double *>(ws);
    double *G = Xc + batch * mx;
    double *R = G + batch * sG;
    double *part = R + batch * sG;
    const long long slabs = gram_slabs(m, l) > 1 ? (long long)gram_slabs(m, l) * sG : 0;
    int *flags = reinterpret_cast<int *>(part + batch * slabs);
    long long dims[2] = {batch, mx}, si[2] = {sX, 1}, so[2] = {mx, 1};
    TRG_TRY(trg_copy_strided(TRG_F64, TRG_F64, 2, dims, X, si, Xc, so, s));
    if (info) TRG_TRY(trg_fill(TRG_I32, batch, 0.0, info, s));
    // odd pass counts: Xc -> Q (-> Xc -> Q)
    TRG_TRY(cholqr3(batch, m, l, Xc, Q, G, R, flags, info, s, passes, slabs ? part : nullptr));
    return TRG_OK;
}

// Y[b] += alpha_rel * scal[b] * X[b]
extern "C" int trg_axpy_dev_f64(int batch, long long n, double alpha_rel, const double *scal, const double *X,
                                long long sX, double *Y, long long sY, void *stream) {
    clear_hip_error();
    if (batch < 0 || n < 0) return TRG_BAD_ARG;
    if (batch == 0 || n == 0) return TRG_OK;
    if (!scal || !X || !Y) return TRG_BAD_ARG;
    if (batch > 65535) return TRG_SHAPE;
    const dim3 grid((unsigned)std::min<long long>(cdiv(n, 256), 4096), batch);
    hipLaunchKernelGGL(axpy_scaled_kernel, grid, 256, 0, as_stream(stream), n, alpha_rel, scal, X, sX, Y, sY);
    return check_launch();
}
