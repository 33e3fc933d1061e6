// Float64 GEMM on the int8 matrix cores (v_mfma_i32_32x32x32_i8, 5 POP/s
// dense on MI355X against 78.6 TF/s for the fp64 MFMA).  The method is the
// error-free splitting of Ozaki, Ogita, Oishi and Rump (Numer. Algorithms 59,
// 2012) in its integer form (Ootomo, Ozaki and Yokota, "DGEMM on integer
// matrix multiplication unit", IJHPCA 2024): fixed-point slices per row,
// exact integer products, the slice pairs grouped by weight.
//
// Each row of A (and each column of B) is written as 2^e times a value in
// (-1, 1), and that value as S signed 7-bit slices: r_0 = x 2^-e,
// q_i = trunc(128 r_{i-1}), r_i = 128 r_{i-1} - q_i (every step exact in
// fp64), so x 2^-e = sum_i q_i 2^-7i + r_S 2^-7S.  The product is
//   C[m][n] = 2^(e_m + f_n) sum_{d=2}^{S+1} 2^-7d P_d[m][n],
//   P_d = sum_{i+j=d} sum_k qA_i[m][k] qB_j[n][k],
// the pairs with i + j > S + 1 (below 2^-7(S+1) of the row and column
// maxima) dropped.  Each P_d is an integer sum of int8 products, exact in
// the MFMA's int32 accumulator while S 127^2 K < 2^31 (the host splits K
// beyond that); the diagonals are combined in fp64 in the epilogue.  With
// S = 6 the operands carry 42 bits and the product is within about 2^-42 of
// the row-max x column-max scale, 21 int8 MFMAs per product.  The result is a
// fixed function of the inputs: no configuration, batch or K order enters it
// beyond the K chunks above.
#include <cstdint>
#include <cmath>

#include "gemm_kernels.h"

namespace trg {

typedef int i32x4 __attribute__((ext_vector_type(4)));
typedef int i32x16 __attribute__((ext_vector_type(16)));

// Row exponents of a non-finite row (its products are NaN).
constexpr int EXP_NONFINITE = 1 << 20;
constexpr int SLICE_KB = 64;   // row pitch of the slices: K rounded up to this

// ---------------------------------------------------------------------------
// Slicing: X (batch, R, K), element (r, k) at X + b sXb + r rsX + k csX.
// Out[b][i][r][k] (int8, row pitch Kp, zero for k >= K), E[b][r].  One
// workgroup per 64 rows; the 64 x 64 tiles go through LDS so that global
// reads run along whichever of r / k is contiguous.
// ---------------------------------------------------------------------------
template <int S>
__global__ void __launch_bounds__(256)
slice_rows_kernel(int R, int K, int Kp, const double *__restrict__ X, long long sXb, long long rsX, long long csX,
                  int8_t *__restrict__ Out, int *__restrict__ E) {
    __shared__ double tile[64][65];
    const int b = blockIdx.y;
    const int r0 = blockIdx.x * 64;
    X += (long long)b * sXb;
    Out += (long long)b * S * R * Kp;
    E += (long long)b * R;
    const int tid = threadIdx.x;
    const bool kcont = (csX == 1);

    auto load_tile = [&](int k0) {
#pragma unroll 4
        for (int i = 0; i < 16; ++i) {
            const int e = tid + 256 * i;
            const int rr = kcont ? e / 64 : e % 64;
            const int kk = kcont ? e % 64 : e / 64;
            const int r = r0 + rr, k = k0 + kk;
            tile[rr][kk] = (r < R && k < K) ? X[(long long)r * rsX + (long long)k * csX] : 0.0;
        }
    };

    // thread t owns row t / 4 and the 16 tile columns from 16 (t % 4)
    const int rr = tid >> 2, c0 = (tid & 3) * 16;
    double m = 0.0;
    bool finite = true;
    for (int k0 = 0; k0 < K; k0 += 64) {
        load_tile(k0);
        __syncthreads();
#pragma unroll
        for (int c = 0; c < 16; ++c) {
            const double v = tile[rr][c0 + c];
            finite = finite && __builtin_isfinite(v);
            m = fmax(m, fabs(v));
        }
        __syncthreads();
    }
    // the four threads of a row are consecutive lanes of one wave
    m = fmax(m, __shfl_xor(m, 1));
    m = fmax(m, __shfl_xor(m, 2));
    int fin = finite ? 1 : 0;
    fin = fin & __shfl_xor(fin, 1);
    fin = fin & __shfl_xor(fin, 2);
    int e = 0;
    if (!fin) e = EXP_NONFINITE;
    else if (m > 0.0) (void)frexp(m, &e);      // m = f 2^e, f in [0.5, 1): every |x| < 2^e
    if ((tid & 3) == 0 && r0 + rr < R) E[r0 + rr] = e;
    const int esh = fin ? -e : 0;

    for (int k0 = 0; k0 < Kp; k0 += 64) {
        load_tile(k0);
        __syncthreads();
        int8_t q[S][16];
#pragma unroll
        for (int c = 0; c < 16; ++c) {
            double r = ldexp(tile[rr][c0 + c], esh);
#pragma unroll
            for (int i = 0; i < S; ++i) {
                const double t = r * 128.0;
                const double qi = trunc(t);
                r = t - qi;
                q[i][c] = (int8_t)(int)qi;
            }
        }
        if (r0 + rr < R) {
#pragma unroll
            for (int i = 0; i < S; ++i) {
                i32x4 v;
#pragma unroll
                for (int w = 0; w < 4; ++w)
                    v[w] = (int)(uint8_t)q[i][4 * w] | ((int)(uint8_t)q[i][4 * w + 1] << 8) |
                           ((int)(uint8_t)q[i][4 * w + 2] << 16) | ((int)(uint8_t)q[i][4 * w + 3] << 24);
                *reinterpret_cast<i32x4 *>(&Out[(long long)i * R * Kp + (long long)(r0 + rr) * Kp + k0 + c0]) = v;
            }
        }
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// GEMM on the slices.  A slices: (batch, S, M, Kp), B slices (batch, S, N, Kp)
// (B's columns as rows), both K-contiguous; row exponents eA (batch, M), eB
// (batch, N).  g carries M, N, K = Kp, the batch strides of the slices
// (sAb, sBb; 0 = shared), C and its (two-level) layout, alpha, beta.
// ---------------------------------------------------------------------------
struct I8Args {
    GemmArgs g;
    const int *eA;
    long long seA;   // batch stride of eA (0 = shared)
    const int *eB;
    long long seB;
};

template <int S, int BM, int BN, int WM, int WN>
__global__ void __launch_bounds__((BM / WM) * (BN / WN) * 64)
gemm_i8s_kernel(I8Args a) {
    const GemmArgs &g = a.g;
    constexpr int WAVES_M = BM / WM;
    constexpr int NT = (BM / WM) * (BN / WN) * 64;
    constexpr int TM = WM / 32, TN = WN / 32;
    constexpr int BK = 32;             // one MFMA k step per stage
    constexpr int KP = BK + 16;        // LDS row pitch in bytes
    constexpr int AT = S * BM * 2, BT = S * BN * 2;             // 16-byte vectors of one stage
    constexpr int AV = (AT + NT - 1) / NT, BV = (BT + NT - 1) / NT;   // per thread
    constexpr int ASZ = S * BM * KP, BSZ = S * BN * KP;        // bytes of one stage
    __shared__ __attribute__((aligned(16))) int8_t As[2 * ASZ];
    __shared__ __attribute__((aligned(16))) int8_t Bs[2 * BSZ];

    int tm, tn;
    tile_coords(g, tm, tn);
    const int b = blockIdx.y;
    const int m0 = tm * BM, n0 = tn * BN;
    const int Kp = g.K;
    const int8_t *__restrict__ A = static_cast<const int8_t *>(g.A) + (long long)b * g.sAb;
    const int8_t *__restrict__ B = static_cast<const int8_t *>(g.B) + (long long)b * g.sBb;
    const long long slA = (long long)g.M * Kp, slB = (long long)g.N * Kp;

    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = wave % WAVES_M, wn = wave / WAVES_M;

    i32x16 acc[S][TM][TN];
#pragma unroll
    for (int d = 0; d < S; ++d)
#pragma unroll
        for (int i = 0; i < TM; ++i)
#pragma unroll
            for (int j = 0; j < TN; ++j)
#pragma unroll
                for (int r = 0; r < 16; ++r) acc[d][i][j][r] = 0;

    i32x4 ra[AV], rb[BV];
    // vector e of a stage: slice e / (2 BM), row (e / 2) % BM, half e % 2
    auto load = [&](int k0) {
#pragma unroll
        for (int i = 0; i < AV; ++i) {
            const int e = tid + i * NT;
            const int s = e / (2 * BM), row = (e >> 1) % BM, h = e & 1;
            const int m = m0 + row;
            ra[i] = (AT % NT == 0 || e < AT) && m < g.M
                        ? *reinterpret_cast<const i32x4 *>(A + s * slA + (long long)m * Kp + k0 + 16 * h)
                        : i32x4{0, 0, 0, 0};
        }
#pragma unroll
        for (int i = 0; i < BV; ++i) {
            const int e = tid + i * NT;
            const int s = e / (2 * BN), col = (e >> 1) % BN, h = e & 1;
            const int n = n0 + col;
            rb[i] = (BT % NT == 0 || e < BT) && n < g.N
                        ? *reinterpret_cast<const i32x4 *>(B + s * slB + (long long)n * Kp + k0 + 16 * h)
                        : i32x4{0, 0, 0, 0};
        }
    };
    auto store = [&](int stage) {
#pragma unroll
        for (int i = 0; i < AV; ++i) {
            const int e = tid + i * NT;
            if (AT % NT != 0 && e >= AT) break;
            const int s = e / (2 * BM), row = (e >> 1) % BM, h = e & 1;
            *reinterpret_cast<i32x4 *>(&As[stage * ASZ + (s * BM + row) * KP + 16 * h]) = ra[i];
        }
#pragma unroll
        for (int i = 0; i < BV; ++i) {
            const int e = tid + i * NT;
            if (BT % NT != 0 && e >= BT) break;
            const int s = e / (2 * BN), col = (e >> 1) % BN, h = e & 1;
            *reinterpret_cast<i32x4 *>(&Bs[stage * BSZ + (s * BN + col) * KP + 16 * h]) = rb[i];
        }
    };
    // lane holds 16 consecutive k of row (lane % 32): the same k slots for A and B
    const int kq = 16 * (lane >> 5);
    auto compute = [&](int stage) {
        const int8_t *as = As + stage * ASZ;
        const int8_t *bs = Bs + stage * BSZ;
        i32x4 fa[S][TM];
#pragma unroll
        for (int s = 0; s < S; ++s)
#pragma unroll
            for (int i = 0; i < TM; ++i)
                fa[s][i] = *reinterpret_cast<const i32x4 *>(&as[(s * BM + wm * WM + i * 32 + (lane & 31)) * KP + kq]);
#pragma unroll
        for (int sj = 0; sj < S; ++sj) {
            i32x4 fb[TN];
#pragma unroll
            for (int j = 0; j < TN; ++j)
                fb[j] = *reinterpret_cast<const i32x4 *>(&bs[(sj * BN + wn * WN + j * 32 + (lane & 31)) * KP + kq]);
            // pairs (si, sj) with si + sj <= S - 1 (0-based): diagonal si + sj
#pragma unroll
            for (int si = 0; si + sj < S; ++si)
#pragma unroll
                for (int i = 0; i < TM; ++i)
#pragma unroll
                    for (int j = 0; j < TN; ++j)
                        acc[si + sj][i][j] =
                            __builtin_amdgcn_mfma_i32_32x32x32_i8(fa[si][i], fb[j], acc[si + sj][i][j], 0, 0, 0);
        }
    };

    const int nk = Kp / BK;
    if (nk > 0) {
        load(0);
        store(0);
        load(min(1, nk - 1) * BK);
        __syncthreads();
        int cur = 0;
        for (int t = 0; t + 1 < nk; ++t) {
            // the stage written here was last read before the previous barrier
            store(cur ^ 1);
            load(min(t + 2, nk - 1) * BK);
            compute(cur);
            __syncthreads();
            cur ^= 1;
        }
        compute(cur);
    }

    // epilogue: combine the diagonals in fp64 (smallest first), scale by the
    // row and column exponents, C = alpha v + beta C
    const int *__restrict__ eA = a.eA + (long long)b * a.seA;
    const int *__restrict__ eB = a.eB + (long long)b * a.seB;
    double *__restrict__ C = static_cast<double *>(g.C) + (long long)b * g.sCb;
    const double alpha = g.alpha, beta = g.beta;
#pragma unroll
    for (int j = 0; j < TN; ++j) {
        const int n = n0 + wn * WN + j * 32 + (lane & 31);
        if (n >= g.N) continue;
        const long long cn = c_col_offset(g, n);
        const int fe = eB[n];
#pragma unroll
        for (int i = 0; i < TM; ++i) {
#pragma unroll
            for (int r = 0; r < 16; ++r) {
                const int m = m0 + wm * WM + i * 32 + (r >> 2) * 8 + (lane >> 5) * 4 + (r & 3);
                if (m >= g.M) continue;
                const int ee = eA[m];
                double v = 0.0;
#pragma unroll
                for (int d = S - 1; d >= 0; --d) v += ldexp((double)acc[d][i][j][r], -7 * (d + 2));
                v = (ee >= EXP_NONFINITE || fe >= EXP_NONFINITE) ? __builtin_nan("") : ldexp(v, ee + fe);
                const long long off = c_row_offset(g, m) + cn;
                C[off] = beta == 0.0 ? alpha * v : alpha * v + beta * C[off];
            }
        }
    }
}

// tile of the slice GEMM: 128 x 64, four waves of 32 x 64 (one per SIMD: the
// S accumulator sets take S x 32 registers per 32 x 32 block)
constexpr int I8_BM = 128, I8_BN = 64, I8_WM = 32, I8_WN = 64;

template <int S>
static int launch_slices(int batch, int R, int K, int Kp, const double *X, long long sXb, long long rsX,
                         long long csX, int8_t *out, int *exps, hipStream_t s) {
    dim3 grid((unsigned)cdiv(R, 64), batch);
    hipLaunchKernelGGL((slice_rows_kernel<S>), grid, dim3(256), 0, s, R, K, Kp, X, sXb, rsX, csX, out, exps);
    return check_launch();
}

template <int S>
static int launch_i8s(I8Args a, int batch, hipStream_t s) {
    a.g.tiles_m = (int)cdiv(a.g.M, I8_BM);
    a.g.tiles_n = (int)cdiv(a.g.N, I8_BN);
    dim3 grid(a.g.tiles_m * a.g.tiles_n, batch);
    hipLaunchKernelGGL((gemm_i8s_kernel<S, I8_BM, I8_BN, I8_WM, I8_WN>), grid,
                       dim3((I8_BM / I8_WM) * (I8_BN / I8_WN) * 64), 0, s, a);
    return check_launch();
}

static long long slice_pitch(int K) { return cdiv(K, SLICE_KB) * SLICE_KB; }

// largest K chunk whose diagonal sums stay inside int32: S 127^2 Kc < 2^31
static int k_chunk(int S) {
    const long long kc = ((1LL << 31) - 1) / (S * 127LL * 127LL);
    return (int)(kc / SLICE_KB * SLICE_KB);
}

}  // namespace trg

using namespace trg;

static int slices_ok(int S) { return S >= 4 && S <= 7; }

// bytes of workspace for trg_gemm_f64i8: slices of A and of B for one K chunk,
// 256-byte aligned, then the row exponents
extern "C" long long trg_gemm_f64i8_workspace(int slices, int batchA, int batchB, int M, int N, int K) {
    if (!slices_ok(slices) || batchA < 1 || batchB < 1 || M < 0 || N < 0 || K < 0) return -1;
    const long long kp = slice_pitch(std::min(K, k_chunk(slices)));
    auto al = [](long long v) { return (v + 255) / 256 * 256; };
    return al((long long)batchA * slices * M * kp) + al((long long)batchB * slices * N * kp) +
           al((long long)batchA * M * 4) + al((long long)batchB * N * 4);
}

extern "C" int trg_gemm_f64i8(int slices, int batch, int M, int N, int K, double alpha,
                              const void *A, long long sAb, long long rsA, long long csA,
                              const void *B, long long sBb, long long rsB, long long csB,
                              double beta, void *C, long long sCb, long long rsC, long long csC,
                              int cdm, long long csm, int cdn, long long csn,
                              void *workspace, long long ws_bytes, void *stream) {
    clear_hip_error();
    if (!slices_ok(slices)) return TRG_BAD_ARG;
    if (batch < 0 || M < 0 || N < 0 || K < 0) return TRG_BAD_ARG;
    if (batch == 0 || M == 0 || N == 0) return TRG_OK;
    if (!C || !workspace || (K > 0 && (!A || !B))) return TRG_BAD_ARG;   // K = 0: empty operands
    if (batch > 65535) return TRG_SHAPE;
    if (cdm < 0 || cdn < 0) return TRG_BAD_ARG;
    const int bA = sAb ? batch : 1, bB = sBb ? batch : 1;
    if (ws_bytes < trg_gemm_f64i8_workspace(slices, bA, bB, M, N, K)) return TRG_OOM;
    hipStream_t s = as_stream(stream);
    const int kc = k_chunk(slices);
    const long long kp = slice_pitch(std::min(K, kc));
    auto al = [](long long v) { return (v + 255) / 256 * 256; };
    int8_t *qa = static_cast<int8_t *>(workspace);
    int8_t *qb = qa + al((long long)bA * slices * M * kp);
    int *ea = reinterpret_cast<int *>(qb + al((long long)bB * slices * N * kp));
    int *eb = reinterpret_cast<int *>(reinterpret_cast<int8_t *>(ea) + al((long long)bA * M * 4));
    // K = 0 runs one pass with no slices (pitch 0, zero exponents): C = beta C
    const double *Ad = static_cast<const double *>(A);
    const double *Bd = static_cast<const double *>(B);
    for (int k0 = 0; k0 < std::max(K, 1); k0 += kc) {
        const int kn = std::max(0, std::min(kc, K - k0));
        const int kpn = (int)slice_pitch(kn);
        // the row exponents are those of the chunk's own rows
        int st;
        auto slice = [&](int nb, int R, const double *X, long long sXb, long long rsX, long long csX, int8_t *q,
                         int *e) {
            switch (slices) {
                case 4: return launch_slices<4>(nb, R, kn, kpn, X, sXb, rsX, csX, q, e, s);
                case 5: return launch_slices<5>(nb, R, kn, kpn, X, sXb, rsX, csX, q, e, s);
                case 6: return launch_slices<6>(nb, R, kn, kpn, X, sXb, rsX, csX, q, e, s);
                default: return launch_slices<7>(nb, R, kn, kpn, X, sXb, rsX, csX, q, e, s);
            }
        };
        if ((st = slice(bA, M, Ad + (long long)k0 * csA, sAb, rsA, csA, qa, ea)) != TRG_OK) return st;
        if ((st = slice(bB, N, Bd + (long long)k0 * rsB, sBb, csB, rsB, qb, eb)) != TRG_OK) return st;
        GemmArgs g{M, N, kpn, qa, sAb ? (long long)slices * M * kpn : 0, 0, 0,
                   qb, sBb ? (long long)slices * N * kpn : 0, 0, 0,
                   C, sCb, rsC, csC, alpha, k0 == 0 ? beta : 1.0, 0, 0};
        g.cdm = cdm;
        g.csm = csm;
        g.cdn = cdn;
        g.csn = csn;
        I8Args a{g, ea, sAb ? (long long)M : 0, eb, sBb ? (long long)N : 0};
        switch (slices) {
            case 4: st = launch_i8s<4>(a, batch, s); break;
            case 5: st = launch_i8s<5>(a, batch, s); break;
            case 6: st = launch_i8s<6>(a, batch, s); break;
            default: st = launch_i8s<7>(a, batch, s); break;
        }
        if (st != TRG_OK) return st;
    }
    return TRG_OK;
}
