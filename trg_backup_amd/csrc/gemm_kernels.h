// Batched strided GEMM kernels on CDNA4 matrix cores (gfx950): the kernel
// templates and their launch helpers.  Each compute family instantiates its
// tiles in a translation unit of its own (gemm_bf16.hip, gemm_x3.hip,
// gemm_x6.hip, gemm_fp.hip); gemm.hip holds the C entry points.
#pragma once

//
//   C[b] = alpha * A[b] * B[b] + beta * C[b]
//
// Three instantiation families share one tiling scheme:
//   * fp32 storage, fp32 MFMA   v_mfma_f32_32x32x2_f32   (32x32 block, k = 2)
//   * fp32 storage, bf16 MFMA   v_mfma_f32_32x32x16_bf16 (operands rounded to
//     bf16 while they are staged into LDS, fp32 accumulation)
//   * fp64 storage, fp64 MFMA   v_mfma_f64_16x16x4_f64   (16x16 block, k = 4)
//
// A workgroup owns a BM x BN tile of C; its waves each own a WM x WN sub-tile
// made of MFMA blocks.  K is walked in BK slices: the next slice is fetched
// from HBM into registers while the current one is consumed from LDS.
// Operand strides are arbitrary as long as one of the two strides of each
// operand is 1, so leg permutations of tensors become index maps here
// instead of separate copies whenever the contiguous leg allows it.
#include "common.h"

#include <type_traits>
#include <utility>

namespace trg {

typedef float f32x16 __attribute__((ext_vector_type(16)));
typedef double f64x4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

struct GemmArgs {
    int M, N, K;
    const void *A;
    long long sAb, rsA, csA;
    const void *B;
    long long sBb, rsB, csB;
    void *C;
    long long sCb, rsC, csC;
    double alpha, beta;
    int tiles_n;   // number of BN tiles along N
    int tiles_m;   // number of BM tiles along M
    // Two-level output index (0 = plain): row m of C sits at
    // (m / cdm) * csm + (m % cdm) * rsC, column n at (n / cdn) * csn + (n % cdn) * csC,
    // so a GEMM can write straight into a permuted tensor layout.
    int cdm, cdn;
    long long csm, csn;
    // set by the host when rows of C are contiguous (rsC == 1) and every
    // group of four consecutive rows starts on a 4-element boundary: the
    // epilogue then writes 4-row vectors
    int cvec;
};

__device__ inline long long c_row_offset(const GemmArgs &g, int m) {
    return g.cdm ? (long long)(m / g.cdm) * g.csm + (long long)(m % g.cdm) * g.rsC : (long long)m * g.rsC;
}
__device__ inline long long c_col_offset(const GemmArgs &g, int n) {
    return g.cdn ? (long long)(n / g.cdn) * g.csn + (long long)(n % g.cdn) * g.csC : (long long)n * g.csC;
}

// Epilogue C = alpha * acc + beta * C through the (possibly two-level) output
// index.  Column offsets are computed once per column block and row offsets
// once per accumulator row, so no index division sits in the inner loop.
// row(i, r) / col(j) give the C row of accumulator register r of block i and
// the C column of block j for this lane.
template <typename TA, typename TC, int TM, int TN, int NACC, typename Acc, typename RowF, typename ColF>
__device__ inline void store_c(const GemmArgs &g, TC *__restrict__ C, Acc (&acc)[TM][TN], RowF row, ColF col) {
    const TA alpha = (TA)g.alpha, beta = (TA)g.beta;
    long long on[TN];
#pragma unroll
    for (int j = 0; j < TN; ++j) {
        const int n = col(j);
        on[j] = n < g.N ? c_col_offset(g, n) : -1;
    }
    if constexpr (NACC == 16) {
        // 32x32 MFMA accumulators: registers 4q..4q+3 of a lane hold four
        // consecutive rows of one column.  With contiguous rows they leave as
        // one 4-element vector store (16 B float, 8 B bf16) instead of four
        // scalar ones; a group cut by the edge of M falls back to scalars.
        typedef TC vec4 __attribute__((ext_vector_type(4)));
        if (g.cvec) {
#pragma unroll
            for (int i = 0; i < TM; ++i)
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    const int m = row(i, 4 * q);
                    if (m >= g.M) continue;
                    const long long om = c_row_offset(g, m);
                    if (m + 3 < g.M) {
#pragma unroll
                        for (int j = 0; j < TN; ++j) {
                            if (on[j] < 0) continue;
                            vec4 *p = reinterpret_cast<vec4 *>(C + om + on[j]);
                            vec4 old;
                            if (g.beta != 0.0) old = *p;
                            vec4 out;
#pragma unroll
                            for (int t = 0; t < 4; ++t) {
                                TA v = alpha * acc[i][j][4 * q + t];
                                if (g.beta != 0.0) v += beta * (TA)old[t];
                                out[t] = (TC)v;
                            }
                            *p = out;
                        }
                    } else {
#pragma unroll
                        for (int t = 0; t < 4; ++t) {
                            if (m + t >= g.M) continue;
#pragma unroll
                            for (int j = 0; j < TN; ++j) {
                                if (on[j] < 0) continue;
                                const long long idx = om + t + on[j];
                                TA v = alpha * acc[i][j][4 * q + t];
                                if (g.beta != 0.0) v += beta * (TA)C[idx];
                                C[idx] = (TC)v;
                            }
                        }
                    }
                }
            return;
        }
    }
#pragma unroll
    for (int i = 0; i < TM; ++i)
#pragma unroll
        for (int r = 0; r < NACC; ++r) {
            const int m = row(i, r);
            if (m >= g.M) continue;
            const long long om = c_row_offset(g, m);
#pragma unroll
            for (int j = 0; j < TN; ++j) {
                if (on[j] < 0) continue;
                const long long idx = om + on[j];
                TA v = alpha * acc[i][j][r];
                if (g.beta != 0.0) v += beta * (TA)C[idx];
                C[idx] = (TC)v;
            }
        }
}

// ---------------------------------------------------------------------------
// MFMA traits: operand lane map and accumulator lane map.
// ---------------------------------------------------------------------------
template <typename T> struct Mfma;

template <> struct Mfma<float> {
    static constexpr int MB = 32, KB = 2, NACC = 16;
    typedef f32x16 acc_t;
    __device__ static inline acc_t zero() {
        acc_t a;
#pragma unroll
        for (int i = 0; i < 16; ++i) a[i] = 0.f;
        return a;
    }
    __device__ static inline void mma(acc_t &c, float a, float b) {
        c = __builtin_amdgcn_mfma_f32_32x32x2f32(a, b, c, 0, 0, 0);
    }
    // operand: lane holds A[m = lane % 32][k = lane / 32]
    __device__ static inline int op_mn(int lane) { return lane & 31; }
    __device__ static inline int op_k(int lane) { return lane >> 5; }
    // accumulator register r of lane -> (row, col) inside the 32x32 block
    __device__ static inline int c_row(int lane, int r) { return (r >> 2) * 8 + (lane >> 5) * 4 + (r & 3); }
    __device__ static inline int c_col(int lane) { return lane & 31; }
};

template <> struct Mfma<double> {
    static constexpr int MB = 16, KB = 4, NACC = 4;
    typedef f64x4 acc_t;
    __device__ static inline acc_t zero() {
        acc_t a;
#pragma unroll
        for (int i = 0; i < 4; ++i) a[i] = 0.0;
        return a;
    }
    __device__ static inline void mma(acc_t &c, double a, double b) {
        c = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
    __device__ static inline int op_mn(int lane) { return lane & 15; }
    __device__ static inline int op_k(int lane) { return lane >> 4; }
    // fp64 accumulators interleave rows: register r of lane holds row lane/16 + 4r
    __device__ static inline int c_row(int lane, int r) { return (lane >> 4) + 4 * r; }
    __device__ static inline int c_col(int lane) { return lane & 15; }
};

// Tile index of this workgroup.  Workgroups are dispatched round-robin over
// the 8 XCDs of the chip, each with its own L2; consecutive linear ids are
// re-mapped so that each XCD walks a contiguous band of tiles (tiles in a
// band share A rows / B columns, so the band stays L2-resident).
__device__ static inline void tile_coords(const GemmArgs &g, int &tm, int &tn) {
    const int total = g.tiles_m * g.tiles_n;
    int lin = blockIdx.x;
    if ((total & 7) == 0) {
        const int per = total >> 3;
        lin = (lin & 7) * per + (lin >> 3);
    }
    // group GROUP_M tile rows together so a band re-uses B columns
    constexpr int GROUP_M = 8;
    const int group_size = GROUP_M * g.tiles_n;
    const int group = lin / group_size;
    const int first_m = group * GROUP_M;
    const int gm = min(g.tiles_m - first_m, GROUP_M);
    const int in_group = lin - group * group_size;
    tm = first_m + in_group % gm;
    tn = in_group / gm;
}

// ---------------------------------------------------------------------------
// fp32 / fp64 kernel.  LDS holds As[k][m] and Bs[k][n] (m / n fastest) so a
// wave's operand fetch is one contiguous row segment per k.
// ---------------------------------------------------------------------------
template <typename T, int BM, int BN, int BK, int WM, int WN, bool A_MCONT, bool B_NCONT, int VEC>
__global__ void __launch_bounds__((BM / WM) * (BN / WN) * 64)
gemm_kernel(GemmArgs g) {
    using MF = Mfma<T>;
    using acc_t = typename MF::acc_t;
    constexpr int WAVES_M = BM / WM;
    constexpr int NT = (BM / WM) * (BN / WN) * 64;
    constexpr int TM = WM / MF::MB;
    constexpr int TN = WN / MF::MB;
    constexpr int PAD = 16 / sizeof(T);
    constexpr int AE = BM * BK / NT;
    constexpr int BE = BN * BK / NT;
    static_assert(AE * NT == BM * BK && BE * NT == BN * BK, "tile not divisible by threads");

    __shared__ T As[BK][BM + PAD];
    __shared__ T Bs[BK][BN + PAD];

    int tm, tn;
    tile_coords(g, tm, tn);
    const int b = blockIdx.y;
    const int m0 = tm * BM, n0 = tn * BN;
    const T *__restrict__ A = static_cast<const T *>(g.A) + (long long)b * g.sAb;
    const T *__restrict__ B = static_cast<const T *>(g.B) + (long long)b * g.sBb;
    T *__restrict__ C = static_cast<T *>(g.C) + (long long)b * g.sCb;

    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = wave % WAVES_M, wn = wave / WAVES_M;

    acc_t acc[TM][TN];
#pragma unroll
    for (int i = 0; i < TM; ++i)
#pragma unroll
        for (int j = 0; j < TN; ++j) acc[i][j] = MF::zero();

    // Global loads move VEC consecutive elements of an operand's contiguous
    // leg at once (VEC = 4 for 16-byte fp32 loads when sizes and strides
    // are multiples of 4, so a vector is either fully inside or outside).
    static_assert(AE % VEC == 0 && BE % VEC == 0, "vector width must divide the per-thread tile share");
    typedef T vec_t __attribute__((ext_vector_type(VEC)));
    T ra[AE], rb[BE];

    auto a_index = [&](int i, int &mm, int &kk) {
        const int e = tid + i * NT;
        if (A_MCONT) { kk = e / (BM / VEC); mm = VEC * (e % (BM / VEC)); }
        else { mm = e / (BK / VEC); kk = VEC * (e % (BK / VEC)); }
    };
    auto b_index = [&](int i, int &nn, int &kk) {
        const int e = tid + i * NT;
        if (B_NCONT) { kk = e / (BN / VEC); nn = VEC * (e % (BN / VEC)); }
        else { nn = e / (BK / VEC); kk = VEC * (e % (BK / VEC)); }
    };
    auto load = [&](int k0) {
#pragma unroll
        for (int i = 0; i < AE / VEC; ++i) {
            int mm, kk;
            a_index(i, mm, kk);
            const int m = m0 + mm, k = k0 + kk;
            if (m < g.M && k < g.K) {
                const vec_t v = *reinterpret_cast<const vec_t *>(A + (long long)m * g.rsA + (long long)k * g.csA);
#pragma unroll
                for (int t = 0; t < VEC; ++t) ra[i * VEC + t] = v[t];
            } else {
#pragma unroll
                for (int t = 0; t < VEC; ++t) ra[i * VEC + t] = T(0);
            }
        }
#pragma unroll
        for (int i = 0; i < BE / VEC; ++i) {
            int nn, kk;
            b_index(i, nn, kk);
            const int n = n0 + nn, k = k0 + kk;
            if (n < g.N && k < g.K) {
                const vec_t v = *reinterpret_cast<const vec_t *>(B + (long long)k * g.rsB + (long long)n * g.csB);
#pragma unroll
                for (int t = 0; t < VEC; ++t) rb[i * VEC + t] = v[t];
            } else {
#pragma unroll
                for (int t = 0; t < VEC; ++t) rb[i * VEC + t] = T(0);
            }
        }
    };
    auto store = [&]() {
#pragma unroll
        for (int i = 0; i < AE / VEC; ++i) {
            int mm, kk;
            a_index(i, mm, kk);
#pragma unroll
            for (int t = 0; t < VEC; ++t) {
                if (A_MCONT) As[kk][mm + t] = ra[i * VEC + t];
                else As[kk + t][mm] = ra[i * VEC + t];
            }
        }
#pragma unroll
        for (int i = 0; i < BE / VEC; ++i) {
            int nn, kk;
            b_index(i, nn, kk);
#pragma unroll
            for (int t = 0; t < VEC; ++t) {
                if (B_NCONT) Bs[kk][nn + t] = rb[i * VEC + t];
                else Bs[kk + t][nn] = rb[i * VEC + t];
            }
        }
    };

    load(0);
    store();
    __syncthreads();
    for (int k0 = 0; k0 < g.K; k0 += BK) {
        const bool more = k0 + BK < g.K;
        if (more) load(k0 + BK);
#pragma unroll
        for (int kk = 0; kk < BK; kk += MF::KB) {
            T a[TM], bb[TN];
            const int kr = kk + MF::op_k(lane);
#pragma unroll
            for (int i = 0; i < TM; ++i) a[i] = As[kr][wm * WM + i * MF::MB + MF::op_mn(lane)];
#pragma unroll
            for (int j = 0; j < TN; ++j) bb[j] = Bs[kr][wn * WN + j * MF::MB + MF::op_mn(lane)];
#pragma unroll
            for (int i = 0; i < TM; ++i)
#pragma unroll
                for (int j = 0; j < TN; ++j) MF::mma(acc[i][j], a[i], bb[j]);
        }
        __syncthreads();
        if (more) {
            store();
            __syncthreads();
        }
    }

    store_c<T, T, TM, TN, MF::NACC>(
        g, C, acc, [&](int i, int r) { return m0 + wm * WM + i * MF::MB + MF::c_row(lane, r); },
        [&](int j) { return n0 + wn * WN + j * MF::MB + MF::c_col(lane); });
}

// ---------------------------------------------------------------------------
// fp32 storage, bf16 MFMA.  LDS holds As[m][k] and Bs[n][k] in bf16 (k
// fastest) because v_mfma_f32_32x32x16_bf16 takes 8 consecutive k per lane:
// lane holds A[m = lane % 32][k = 8 * (lane / 32) + j], j < 8.
//
// SPLIT ("bf16x3"): every fp32 operand x is staged as hi = bf16(x) and
// lo = bf16(x - hi) (the subtraction is exact), and each product is formed
// as lo_a hi_b + hi_a lo_b + hi_a hi_b on the MFMA (the lo*lo term, below
// 2^-16 relative, is dropped).  That carries about 16 significant bits per
// operand instead of 8, at three bf16 MFMAs per block -- still far fewer
// matrix-core cycles than one fp32 MFMA chain over the same K.
// ---------------------------------------------------------------------------
__device__ inline void split_bf16(float x, __bf16 &hi, __bf16 &lo) {
    hi = (__bf16)x;
    lo = (__bf16)(x - (float)hi);
}

// SPLIT 2 ("bf16x6"): x = hi + mid + lo exactly (each residual is exact in
// fp32 and the last one fits bf16), 24 significant bits like fp32 itself.
// Products keep the six terms down to 2^-16 relative (lo lo, lo mid, mid lo
// are dropped, below fp32 rounding), six bf16 MFMAs per block.
__device__ inline void split3_bf16(float x, __bf16 &hi, __bf16 &mid, __bf16 &lo) {
    hi = (__bf16)x;
    const float r = x - (float)hi;
    mid = (__bf16)r;
    lo = (__bf16)(r - (float)mid);
}

// Stage one fp32 operand element as SPLIT + 1 bf16 pieces, piece p at S[p * stride].
template <int SPLIT>
__device__ inline void stage_pieces(float x, __bf16 *S, int stride) {
    if constexpr (SPLIT == 0) {
        S[0] = (__bf16)x;
    } else if constexpr (SPLIT == 1) {
        split_bf16(x, S[0], S[stride]);
    } else {
        split3_bf16(x, S[0], S[stride], S[2 * stride]);
    }
}

// One BK slice of MFMAs from the k-fastest LDS tiles (hi parts at offset 0,
// then lo (SPLIT 1) or mid and lo (SPLIT 2) parts, BM * KP / BN * KP
// elements apart).  Small terms are accumulated before large ones.
template <int BM, int BN, int BK, int KP, int WM, int WN, int SPLIT>
__device__ inline void bf16_slice(const __bf16 *As, const __bf16 *Bs, f32x16 (&acc)[WM / 32][WN / 32], int wm,
                                  int wn, int lane) {
    constexpr int TM = WM / 32, TN = WN / 32;
#pragma unroll
    for (int kk = 0; kk < BK; kk += 16) {
        bf16x8 a[TM], bb[TN];
        const int kr = kk + 8 * (lane >> 5);
        auto fa = [&](int piece, int i) {
            return *reinterpret_cast<const bf16x8 *>(&As[piece * BM * KP + (wm * WM + i * 32 + (lane & 31)) * KP + kr]);
        };
        auto fb = [&](int piece, int j) {
            return *reinterpret_cast<const bf16x8 *>(&Bs[piece * BN * KP + (wn * WN + j * 32 + (lane & 31)) * KP + kr]);
        };
#pragma unroll
        for (int i = 0; i < TM; ++i) a[i] = fa(0, i);
#pragma unroll
        for (int j = 0; j < TN; ++j) bb[j] = fb(0, j);
        if constexpr (SPLIT == 1) {
            bf16x8 al[TM], bl[TN];
#pragma unroll
            for (int i = 0; i < TM; ++i) al[i] = fa(1, i);
#pragma unroll
            for (int j = 0; j < TN; ++j) bl[j] = fb(1, j);
#pragma unroll
            for (int i = 0; i < TM; ++i)
#pragma unroll
                for (int j = 0; j < TN; ++j) {
                    acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(al[i], bb[j], acc[i][j], 0, 0, 0);
                    acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a[i], bl[j], acc[i][j], 0, 0, 0);
                }
        } else if constexpr (SPLIT == 2) {
            bf16x8 am[TM], bm[TN], al[TM], bl[TN];
#pragma unroll
            for (int i = 0; i < TM; ++i) { am[i] = fa(1, i); al[i] = fa(2, i); }
#pragma unroll
            for (int j = 0; j < TN; ++j) { bm[j] = fb(1, j); bl[j] = fb(2, j); }
#pragma unroll
            for (int i = 0; i < TM; ++i)
#pragma unroll
                for (int j = 0; j < TN; ++j) {
                    acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(al[i], bb[j], acc[i][j], 0, 0, 0);
                    acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a[i], bl[j], acc[i][j], 0, 0, 0);
                    acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(am[i], bm[j], acc[i][j], 0, 0, 0);
                    acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(am[i], bb[j], acc[i][j], 0, 0, 0);
                    acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a[i], bm[j], acc[i][j], 0, 0, 0);
                }
        }
#pragma unroll
        for (int i = 0; i < TM; ++i)
#pragma unroll
            for (int j = 0; j < TN; ++j)
                acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a[i], bb[j], acc[i][j], 0, 0, 0);
    }
}

template <int BM, int BN, int BK, int WM, int WN, bool A_MCONT, bool B_NCONT, int SPLIT, typename TC>
__global__ void __launch_bounds__((BM / WM) * (BN / WN) * 64)
gemm_bf16_kernel(GemmArgs g) {
    constexpr int WAVES_M = BM / WM;
    constexpr int NT = (BM / WM) * (BN / WN) * 64;
    constexpr int TM = WM / 32;
    constexpr int TN = WN / 32;
    constexpr int KP = BK + 8;  // row pitch in bf16 elements (16-byte aligned rows)
    constexpr int AE = BM * BK / NT;
    constexpr int BE = BN * BK / NT;
    constexpr int NS = SPLIT + 1;
    static_assert(BK % 16 == 0, "BK must be a multiple of 16");

    __shared__ __attribute__((aligned(16))) __bf16 As[NS * BM * KP];
    __shared__ __attribute__((aligned(16))) __bf16 Bs[NS * BN * KP];

    int tm, tn;
    tile_coords(g, tm, tn);
    const int b = blockIdx.y;
    const int m0 = tm * BM, n0 = tn * BN;
    const float *__restrict__ A = static_cast<const float *>(g.A) + (long long)b * g.sAb;
    const float *__restrict__ B = static_cast<const float *>(g.B) + (long long)b * g.sBb;
    TC *__restrict__ C = static_cast<TC *>(g.C) + (long long)b * g.sCb;

    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = wave % WAVES_M, wn = wave / WAVES_M;

    f32x16 acc[TM][TN];
#pragma unroll
    for (int i = 0; i < TM; ++i)
#pragma unroll
        for (int j = 0; j < TN; ++j)
#pragma unroll
            for (int r = 0; r < 16; ++r) acc[i][j][r] = 0.f;

    float ra[AE], rb[BE];
    auto load = [&](int k0) {
#pragma unroll
        for (int i = 0; i < AE; ++i) {
            const int e = tid + i * NT;
            int mm, kk;
            if (A_MCONT) { kk = e / BM; mm = e % BM; } else { mm = e / BK; kk = e % BK; }
            const int m = m0 + mm, k = k0 + kk;
            ra[i] = (m < g.M && k < g.K) ? A[(long long)m * g.rsA + (long long)k * g.csA] : 0.f;
        }
#pragma unroll
        for (int i = 0; i < BE; ++i) {
            const int e = tid + i * NT;
            int nn, kk;
            if (B_NCONT) { kk = e / BN; nn = e % BN; } else { nn = e / BK; kk = e % BK; }
            const int n = n0 + nn, k = k0 + kk;
            rb[i] = (n < g.N && k < g.K) ? B[(long long)k * g.rsB + (long long)n * g.csB] : 0.f;
        }
    };
    auto store = [&]() {
#pragma unroll
        for (int i = 0; i < AE; ++i) {
            const int e = tid + i * NT;
            int mm, kk;
            if (A_MCONT) { kk = e / BM; mm = e % BM; } else { mm = e / BK; kk = e % BK; }
            stage_pieces<SPLIT>(ra[i], As + mm * KP + kk, BM * KP);
        }
#pragma unroll
        for (int i = 0; i < BE; ++i) {
            const int e = tid + i * NT;
            int nn, kk;
            if (B_NCONT) { kk = e / BN; nn = e % BN; } else { nn = e / BK; kk = e % BK; }
            stage_pieces<SPLIT>(rb[i], Bs + nn * KP + kk, BN * KP);
        }
    };

    load(0);
    store();
    __syncthreads();
    for (int k0 = 0; k0 < g.K; k0 += BK) {
        const bool more = k0 + BK < g.K;
        if (more) load(k0 + BK);
        bf16_slice<BM, BN, BK, KP, WM, WN, SPLIT>(As, Bs, acc, wm, wn, lane);
        __syncthreads();
        if (more) {
            store();
            __syncthreads();
        }
    }

    store_c<float, TC, TM, TN, 16>(
        g, C, acc, [&](int i, int r) { return m0 + wm * WM + i * 32 + (r >> 2) * 8 + (lane >> 5) * 4 + (r & 3); },
        [&](int j) { return n0 + wn * WN + j * 32 + (lane & 31); });
}

// ---------------------------------------------------------------------------
// fp32 storage, bf16 MFMA, both operands K-contiguous: A(m,k) at m*rsA + k and
// B(k,n) at n*csB + k.  Each thread fetches 16-byte fp32 vectors along k,
// rounds them to bf16 and writes 8-byte runs into the k-fastest LDS rows that
// the MFMA operand fetch (8 consecutive k per lane, one 16-byte LDS read)
// expects, so no transposition happens on the way.  The row pitch BK + 8
// keeps the 16-byte reads of 16 consecutive rows on disjoint LDS banks.
// Requires 16-byte aligned bases, and K, rsA, csB, sAb, sBb multiples of 4
// (a vector is then either fully inside or fully outside K).  Same MFMA
// instruction and k order as gemm_bf16_kernel, so the two agree bitwise.
// ---------------------------------------------------------------------------

// DB: two LDS stages; the next slice is written into the stage the waves are
// not reading, so one barrier per K slice suffices instead of two.
template <int BM, int BN, int BK, int WM, int WN, int SPLIT, bool DB, typename TC>
__global__ void __launch_bounds__((BM / WM) * (BN / WN) * 64)
gemm_bf16k_kernel(GemmArgs g) {
    constexpr int WAVES_M = BM / WM;
    constexpr int NT = (BM / WM) * (BN / WN) * 64;
    constexpr int TM = WM / 32;
    constexpr int TN = WN / 32;
    constexpr int KP = BK + 8;
    constexpr int KQ = BK / 4;            // fp32 vectors per tile row
    constexpr int AV = BM * KQ / NT;
    constexpr int BV = BN * KQ / NT;
    constexpr int NS = SPLIT + 1;
    static_assert(AV * NT == BM * KQ && BV * NT == BN * KQ, "tile not divisible by threads");
    static_assert(BK % 16 == 0, "BK must be a multiple of 16");

    constexpr int NB = DB ? 2 : 1;
    constexpr int ASZ = NS * BM * KP, BSZ = NS * BN * KP;   // one stage
    __shared__ __attribute__((aligned(16))) __bf16 As[NB * ASZ];
    __shared__ __attribute__((aligned(16))) __bf16 Bs[NB * BSZ];

    int tm, tn;
    tile_coords(g, tm, tn);
    const int b = blockIdx.y;
    const int m0 = tm * BM, n0 = tn * BN;
    const float *__restrict__ A = static_cast<const float *>(g.A) + (long long)b * g.sAb;
    const float *__restrict__ B = static_cast<const float *>(g.B) + (long long)b * g.sBb;
    TC *__restrict__ C = static_cast<TC *>(g.C) + (long long)b * g.sCb;

    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = wave % WAVES_M, wn = wave / WAVES_M;

    f32x16 acc[TM][TN];
#pragma unroll
    for (int i = 0; i < TM; ++i)
#pragma unroll
        for (int j = 0; j < TN; ++j)
#pragma unroll
            for (int r = 0; r < 16; ++r) acc[i][j][r] = 0.f;

    f32x4 ra[AV], rb[BV];
    auto load = [&](int k0) {
#pragma unroll
        for (int i = 0; i < AV; ++i) {
            const int e = tid + i * NT;
            const int m = m0 + e / KQ, k = k0 + 4 * (e % KQ);
            if (m < g.M && k < g.K) ra[i] = *reinterpret_cast<const f32x4 *>(A + (long long)m * g.rsA + k);
            else ra[i] = f32x4{0.f, 0.f, 0.f, 0.f};
        }
#pragma unroll
        for (int i = 0; i < BV; ++i) {
            const int e = tid + i * NT;
            const int n = n0 + e / KQ, k = k0 + 4 * (e % KQ);
            if (n < g.N && k < g.K) rb[i] = *reinterpret_cast<const f32x4 *>(B + (long long)n * g.csB + k);
            else rb[i] = f32x4{0.f, 0.f, 0.f, 0.f};
        }
    };
    auto put = [&](__bf16 *S, int rows, int off, const f32x4 &x) {
        bf16x4 hi, mid, lo;
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            __bf16 h, m, l;
            if constexpr (SPLIT == 2) split3_bf16(x[t], h, m, l);
            else if constexpr (SPLIT == 1) split_bf16(x[t], h, l);
            else h = (__bf16)x[t];
            hi[t] = h;
            if constexpr (SPLIT == 2) mid[t] = m;
            if constexpr (SPLIT >= 1) lo[t] = l;
        }
        *reinterpret_cast<bf16x4 *>(&S[off]) = hi;
        if constexpr (SPLIT == 1) *reinterpret_cast<bf16x4 *>(&S[rows * KP + off]) = lo;
        if constexpr (SPLIT == 2) {
            *reinterpret_cast<bf16x4 *>(&S[rows * KP + off]) = mid;
            *reinterpret_cast<bf16x4 *>(&S[2 * rows * KP + off]) = lo;
        }
    };
    auto store = [&](int stage) {
#pragma unroll
        for (int i = 0; i < AV; ++i) {
            const int e = tid + i * NT;
            put(As + stage * ASZ, BM, (e / KQ) * KP + 4 * (e % KQ), ra[i]);
        }
#pragma unroll
        for (int i = 0; i < BV; ++i) {
            const int e = tid + i * NT;
            put(Bs + stage * BSZ, BN, (e / KQ) * KP + 4 * (e % KQ), rb[i]);
        }
    };

    load(0);
    store(0);
    __syncthreads();
    int cur = 0;
    for (int k0 = 0; k0 < g.K; k0 += BK) {
        const bool more = k0 + BK < g.K;
        if (more) load(k0 + BK);
        bf16_slice<BM, BN, BK, KP, WM, WN, SPLIT>(As + cur * ASZ, Bs + cur * BSZ, acc, wm, wn, lane);
        if constexpr (DB) {
            // the other stage was last read before the previous barrier
            if (more) store(cur ^ 1);
            __syncthreads();
            cur ^= 1;
        } else {
            __syncthreads();
            if (more) {
                store(0);
                __syncthreads();
            }
        }
    }

    store_c<float, TC, TM, TN, 16>(
        g, C, acc, [&](int i, int r) { return m0 + wm * WM + i * 32 + (r >> 2) * 8 + (lane >> 5) * 4 + (r & 3); },
        [&](int j) { return n0 + wn * WN + j * 32 + (lane & 31); });
}

// ---------------------------------------------------------------------------
// bf16 storage, bf16 MFMA, fp32 output: both operands bf16 and K-contiguous,
// A(m,k) at m*rsA + k and B(k,n) at n*csB + k.  Each thread moves 16-byte
// runs of 8 consecutive k straight from global memory into the k-fastest LDS
// rows, so an operand costs half the L2 traffic of the fp32-storage kernel
// and no conversion.  Requires 16-byte aligned bases and K, rsA, csB, sAb,
// sBb multiples of 8.  Bitwise equal to gemm_bf16k_kernel on operands that
// were rounded to bf16 beforehand (same MFMA, same k order).
// ---------------------------------------------------------------------------
typedef unsigned short u16x8 __attribute__((ext_vector_type(8)));

template <int BM, int BN, int BK, int WM, int WN, bool DB>
__global__ void __launch_bounds__((BM / WM) * (BN / WN) * 64)
gemm_bf16in_kernel(GemmArgs g) {
    constexpr int WAVES_M = BM / WM;
    constexpr int NT = (BM / WM) * (BN / WN) * 64;
    constexpr int TM = WM / 32;
    constexpr int TN = WN / 32;
    constexpr int KP = BK + 8;
    constexpr int KQ = BK / 8;            // bf16 vectors per tile row
    constexpr int AV = BM * KQ / NT;
    constexpr int BV = BN * KQ / NT;
    static_assert(AV * NT == BM * KQ && BV * NT == BN * KQ, "tile not divisible by threads");
    static_assert(BK % 16 == 0, "BK must be a multiple of 16");
    constexpr int NB = DB ? 2 : 1;
    constexpr int ASZ = BM * KP, BSZ = BN * KP;
    __shared__ __attribute__((aligned(16))) __bf16 As[NB * ASZ];
    __shared__ __attribute__((aligned(16))) __bf16 Bs[NB * BSZ];

    int tm, tn;
    tile_coords(g, tm, tn);
    const int b = blockIdx.y;
    const int m0 = tm * BM, n0 = tn * BN;
    const __bf16 *__restrict__ A = static_cast<const __bf16 *>(g.A) + (long long)b * g.sAb;
    const __bf16 *__restrict__ B = static_cast<const __bf16 *>(g.B) + (long long)b * g.sBb;
    float *__restrict__ C = static_cast<float *>(g.C) + (long long)b * g.sCb;

    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = wave % WAVES_M, wn = wave / WAVES_M;

    f32x16 acc[TM][TN];
#pragma unroll
    for (int i = 0; i < TM; ++i)
#pragma unroll
        for (int j = 0; j < TN; ++j)
#pragma unroll
            for (int r = 0; r < 16; ++r) acc[i][j][r] = 0.f;

    u16x8 ra[AV], rb[BV];
    const u16x8 zero = {0, 0, 0, 0, 0, 0, 0, 0};
    auto load = [&](int k0) {
#pragma unroll
        for (int i = 0; i < AV; ++i) {
            const int e = tid + i * NT;
            const int m = m0 + e / KQ, k = k0 + 8 * (e % KQ);
            ra[i] = (m < g.M && k < g.K) ? *reinterpret_cast<const u16x8 *>(A + (long long)m * g.rsA + k) : zero;
        }
#pragma unroll
        for (int i = 0; i < BV; ++i) {
            const int e = tid + i * NT;
            const int n = n0 + e / KQ, k = k0 + 8 * (e % KQ);
            rb[i] = (n < g.N && k < g.K) ? *reinterpret_cast<const u16x8 *>(B + (long long)n * g.csB + k) : zero;
        }
    };
    auto store = [&](int stage) {
#pragma unroll
        for (int i = 0; i < AV; ++i) {
            const int e = tid + i * NT;
            *reinterpret_cast<u16x8 *>(&As[stage * ASZ + (e / KQ) * KP + 8 * (e % KQ)]) = ra[i];
        }
#pragma unroll
        for (int i = 0; i < BV; ++i) {
            const int e = tid + i * NT;
            *reinterpret_cast<u16x8 *>(&Bs[stage * BSZ + (e / KQ) * KP + 8 * (e % KQ)]) = rb[i];
        }
    };

    load(0);
    store(0);
    __syncthreads();
    int cur = 0;
    for (int k0 = 0; k0 < g.K; k0 += BK) {
        const bool more = k0 + BK < g.K;
        if (more) load(k0 + BK);
        bf16_slice<BM, BN, BK, KP, WM, WN, false>(As + cur * ASZ, Bs + cur * BSZ, acc, wm, wn, lane);
        if constexpr (DB) {
            if (more) store(cur ^ 1);
            __syncthreads();
            cur ^= 1;
        } else {
            __syncthreads();
            if (more) {
                store(0);
                __syncthreads();
            }
        }
    }
    store_c<float, float, TM, TN, 16>(
        g, C, acc, [&](int i, int r) { return m0 + wm * WM + i * 32 + (r >> 2) * 8 + (lane >> 5) * 4 + (r & 3); },
        [&](int j) { return n0 + wn * WN + j * 32 + (lane & 31); });
}

// ---------------------------------------------------------------------------
// Software-pipelined K-contiguous kernel: fp32 storage staged as SPLIT + 1
// bf16 pieces, or bf16 storage (BIN) moved as is.  Two LDS stages and one
// register set.  Iteration t first converts and writes slice t + 1 (loaded
// into registers one iteration earlier) into the stage nobody reads, then
// issues the global loads of slice t + 2, then runs the MFMAs of slice t from
// the other stage, then one barrier: the global latency is hidden behind a
// whole iteration of MFMAs.  Loads are buffer loads whose out-of-range
// vectors come back as zeros, and the last slice is peeled off, so the loop
// body is a single basic block that the
// compiler can interleave; SCHED = 1 also asks its instruction-group
// scheduler for an explicit pattern (LDS reads of the first k step, then
// each MFMA followed by a share of the LDS reads, conversion VALU work, LDS
// writes and, halfway, the global loads).  Slices past the end are clamped
// to the last one (loaded again, never used).  Same MFMA instruction, k order
// and term order as gemm_bf16k_kernel and gemm_bf16in_kernel, so the results
// agree bit for bit.
// ---------------------------------------------------------------------------
template <int SPLIT>
__device__ inline void put_pieces(__bf16 *S, int piece_stride, int off, const f32x4 &x) {
    bf16x4 hi, mid, lo;
#pragma unroll
    for (int t = 0; t < 4; ++t) {
        __bf16 h, m, l;
        if constexpr (SPLIT == 2) split3_bf16(x[t], h, m, l);
        else if constexpr (SPLIT == 1) split_bf16(x[t], h, l);
        else h = (__bf16)x[t];
        hi[t] = h;
        if constexpr (SPLIT == 2) mid[t] = m;
        if constexpr (SPLIT >= 1) lo[t] = l;
    }
    *reinterpret_cast<bf16x4 *>(&S[off]) = hi;
    if constexpr (SPLIT == 1) *reinterpret_cast<bf16x4 *>(&S[piece_stride + off]) = lo;
    if constexpr (SPLIT == 2) {
        *reinterpret_cast<bf16x4 *>(&S[piece_stride + off]) = mid;
        *reinterpret_cast<bf16x4 *>(&S[2 * piece_stride + off]) = lo;
    }
}

// One slot of the interleave pattern: an MFMA followed by its share of the
// other instruction streams (sched_group_barrier masks: 0x008 MFMA, 0x100 LDS
// read, 0x002 VALU, 0x200 LDS write, 0x020 global read).
template <int DSR, int VA, int DSW, int VM>
__device__ inline void sched_slot() {
    __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);
    if constexpr (DSR > 0) __builtin_amdgcn_sched_group_barrier(0x100, DSR, 0);
    if constexpr (VA > 0) __builtin_amdgcn_sched_group_barrier(0x002, VA, 0);
    if constexpr (DSW > 0) __builtin_amdgcn_sched_group_barrier(0x200, DSW, 0);
    if constexpr (VM > 0) __builtin_amdgcn_sched_group_barrier(0x020, VM, 0);
}
// LDS reads spread over all NMF slots, VALU work and LDS writes over the
// first half, the global loads right after the last LDS write.
template <int NMF, int DSR, int VA, int DSW, int VM, int... I>
__device__ inline void sched_pattern(std::integer_sequence<int, I...>) {
    constexpr int H = NMF / 2;
    (sched_slot<(I * ((DSR + NMF - 1) / NMF) < DSR ? (DSR + NMF - 1) / NMF : 0),
                (I < H ? (VA + H - 1) / H : 0), (I < H ? (DSW + H - 1) / H : 0), (I == H - 1 ? VM : 0)>(),
     ...);
}

template <int BM, int BN, int BK, int WM, int WN, int SPLIT, bool BIN, int SCHED, typename TC>
__global__ void __launch_bounds__((BM / WM) * (BN / WN) * 64)
gemm_bf16p_kernel(GemmArgs g) {
    static_assert(!BIN || SPLIT == 0, "bf16 storage has no split pieces");
    constexpr int WAVES_M = BM / WM;
    constexpr int NT = (BM / WM) * (BN / WN) * 64;
    constexpr int TM = WM / 32, TN = WN / 32;
    constexpr int KP = BK + 8;
    constexpr int EV = BIN ? 8 : 4;           // elements per 16-byte global vector
    constexpr int KQ = BK / EV;
    constexpr int AV = BM * KQ / NT, BV = BN * KQ / NT;
    constexpr int NS = SPLIT + 1;
    static_assert(AV * NT == BM * KQ && BV * NT == BN * KQ, "tile not divisible by threads");
    static_assert(BK % 16 == 0, "BK must be a multiple of 16");
    constexpr int ASZ = NS * BM * KP, BSZ = NS * BN * KP;   // one stage
    __shared__ __attribute__((aligned(16))) __bf16 As[2 * ASZ];
    __shared__ __attribute__((aligned(16))) __bf16 Bs[2 * BSZ];
    typedef typename std::conditional<BIN, u16x8, f32x4>::type VT;
    typedef typename std::conditional<BIN, __bf16, float>::type ET;

    int tm, tn;
    tile_coords(g, tm, tn);
    const int b = blockIdx.y;
    const int m0 = tm * BM, n0 = tn * BN;
    const ET *__restrict__ A = static_cast<const ET *>(g.A) + (long long)b * g.sAb;
    const ET *__restrict__ B = static_cast<const ET *>(g.B) + (long long)b * g.sBb;
    TC *__restrict__ C = static_cast<TC *>(g.C) + (long long)b * g.sCb;

    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = wave % WAVES_M, wn = wave / WAVES_M;

    f32x16 acc[TM][TN];
#pragma unroll
    for (int i = 0; i < TM; ++i)
#pragma unroll
        for (int j = 0; j < TN; ++j)
#pragma unroll
            for (int r = 0; r < 16; ++r) acc[i][j][r] = 0.f;

    VT ra[AV], rb[BV];
    // Buffer loads: a vector outside the operand gets an offset past the
    // resource's extent, which the hardware returns as zeros, so no load sits
    // behind a branch.  The host guarantees both extents below 2^31 bytes.
    constexpr int ES = (int)sizeof(ET);
    constexpr unsigned OOB = 0x80000000u;
    const __amdgpu_buffer_rsrc_t rsA = __builtin_amdgcn_make_buffer_rsrc(
        (void *)A, (short)0, (int)(((long long)(g.M - 1) * g.rsA + g.K) * ES), 0x00020000);
    const __amdgpu_buffer_rsrc_t rsB = __builtin_amdgcn_make_buffer_rsrc(
        (void *)B, (short)0, (int)(((long long)(g.N - 1) * g.csB + g.K) * ES), 0x00020000);
    auto load = [&](int k0) {
#pragma unroll
        for (int i = 0; i < AV; ++i) {
            const int e = tid + i * NT;
            const int m = m0 + e / KQ, k = k0 + EV * (e % KQ);
            const unsigned off = (m < g.M && k < g.K) ? (unsigned)((m * (int)g.rsA + k) * ES) : OOB;
            ra[i] = __builtin_bit_cast(VT, __builtin_amdgcn_raw_buffer_load_b128(rsA, off, 0, 0));
        }
#pragma unroll
        for (int i = 0; i < BV; ++i) {
            const int e = tid + i * NT;
            const int n = n0 + e / KQ, k = k0 + EV * (e % KQ);
            const unsigned off = (n < g.N && k < g.K) ? (unsigned)((n * (int)g.csB + k) * ES) : OOB;
            rb[i] = __builtin_bit_cast(VT, __builtin_amdgcn_raw_buffer_load_b128(rsB, off, 0, 0));
        }
    };
    auto store = [&](int stage) {
#pragma unroll
        for (int i = 0; i < AV; ++i) {
            const int e = tid + i * NT;
            const int off = stage * ASZ + (e / KQ) * KP + EV * (e % KQ);
            if constexpr (BIN) *reinterpret_cast<u16x8 *>(&As[off]) = ra[i];
            else put_pieces<SPLIT>(As, BM * KP, off, ra[i]);
        }
#pragma unroll
        for (int i = 0; i < BV; ++i) {
            const int e = tid + i * NT;
            const int off = stage * BSZ + (e / KQ) * KP + EV * (e % KQ);
            if constexpr (BIN) *reinterpret_cast<u16x8 *>(&Bs[off]) = rb[i];
            else put_pieces<SPLIT>(Bs, BN * KP, off, rb[i]);
        }
    };

    const int nk = (g.K + BK - 1) / BK;
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
            // SCHED 2: keep the conversion, LDS writes and global loads ahead
            // of every MFMA of the slice (the loads then have a whole slice of
            // MFMAs to land)
            if constexpr (SCHED == 2) __builtin_amdgcn_sched_barrier(0);
            bf16_slice<BM, BN, BK, KP, WM, WN, SPLIT>(As + cur * ASZ, Bs + cur * BSZ, acc, wm, wn, lane);
            if constexpr (SCHED == 1) {
                constexpr int MFP = SPLIT == 0 ? 1 : (SPLIT == 1 ? 3 : 6);   // MFMAs per block product
                constexpr int NMF = TM * TN * (BK / 16) * MFP;
                constexpr int DSR0 = (TM + TN) * NS;                        // LDS reads of one k step
                constexpr int DSR = DSR0 * (BK / 16 - 1);
                constexpr int DSW = (AV + BV) * NS;
                constexpr int VM = AV + BV;
                constexpr int VA = BIN ? 0 : (AV + BV) * 4 * (SPLIT == 0 ? 1 : (SPLIT == 1 ? 3 : 5));
                __builtin_amdgcn_sched_group_barrier(0x100, DSR0, 0);
                sched_pattern<NMF, DSR, VA, DSW, VM>(std::make_integer_sequence<int, NMF>{});
            }
            __syncthreads();
            cur ^= 1;
        }
        bf16_slice<BM, BN, BK, KP, WM, WN, SPLIT>(As + cur * ASZ, Bs + cur * BSZ, acc, wm, wn, lane);
    }

    store_c<float, TC, TM, TN, 16>(
        g, C, acc, [&](int i, int r) { return m0 + wm * WM + i * 32 + (r >> 2) * 8 + (lane >> 5) * 4 + (r & 3); },
        [&](int j) { return n0 + wn * WN + j * 32 + (lane & 31); });
}

// ---------------------------------------------------------------------------
// Deep-prefetch form of gemm_bf16p_kernel: two register sets, so the global
// loads of a slice are issued two iterations before the iteration that
// writes it into LDS (twice the latency budget of the single-set kernel, at
// the price of one more register set).  Slice s lives in set s % 2; the loop
// is unrolled by two so that the set index is a compile-time constant.  Same
// MFMA instruction, k order and term order, so the results agree bit for bit
// with gemm_bf16p_kernel.  It pays where a slice is short next to the load
// latency: the three-piece split's 16-wide K slices (4 % faster on the chi^7
// merge GEMM); the bf16 and split-bf16 tiles gained nothing from it (the
// split-bf16 tile spills with the second register set).
// ---------------------------------------------------------------------------
template <int BM, int BN, int BK, int WM, int WN, int SPLIT, bool BIN, typename TC>
__global__ void __launch_bounds__((BM / WM) * (BN / WN) * 64)
gemm_bf16d_kernel(GemmArgs g) {
    static_assert(!BIN || SPLIT == 0, "bf16 storage has no split pieces");
    constexpr int WAVES_M = BM / WM;
    constexpr int NT = (BM / WM) * (BN / WN) * 64;
    constexpr int TM = WM / 32, TN = WN / 32;
    constexpr int KP = BK + 8;
    constexpr int EV = BIN ? 8 : 4;           // elements per 16-byte global vector
    constexpr int KQ = BK / EV;
    constexpr int AV = BM * KQ / NT, BV = BN * KQ / NT;
    constexpr int NS = SPLIT + 1;
    static_assert(AV * NT == BM * KQ && BV * NT == BN * KQ, "tile not divisible by threads");
    static_assert(BK % 16 == 0, "BK must be a multiple of 16");
    constexpr int ASZ = NS * BM * KP, BSZ = NS * BN * KP;   // one stage
    __shared__ __attribute__((aligned(16))) __bf16 As[2 * ASZ];
    __shared__ __attribute__((aligned(16))) __bf16 Bs[2 * BSZ];
    typedef typename std::conditional<BIN, u16x8, f32x4>::type VT;
    typedef typename std::conditional<BIN, __bf16, float>::type ET;

    int tm, tn;
    tile_coords(g, tm, tn);
    const int b = blockIdx.y;
    const int m0 = tm * BM, n0 = tn * BN;
    const ET *__restrict__ A = static_cast<const ET *>(g.A) + (long long)b * g.sAb;
    const ET *__restrict__ B = static_cast<const ET *>(g.B) + (long long)b * g.sBb;
    TC *__restrict__ C = static_cast<TC *>(g.C) + (long long)b * g.sCb;

    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = wave % WAVES_M, wn = wave / WAVES_M;

    f32x16 acc[TM][TN];
#pragma unroll
    for (int i = 0; i < TM; ++i)
#pragma unroll
        for (int j = 0; j < TN; ++j)
#pragma unroll
            for (int r = 0; r < 16; ++r) acc[i][j][r] = 0.f;

    VT ra0[AV], rb0[BV], ra1[AV], rb1[BV];
    constexpr int ES = (int)sizeof(ET);
    constexpr unsigned OOB = 0x80000000u;
    const __amdgpu_buffer_rsrc_t rsA = __builtin_amdgcn_make_buffer_rsrc(
        (void *)A, (short)0, (int)(((long long)(g.M - 1) * g.rsA + g.K) * ES), 0x00020000);
    const __amdgpu_buffer_rsrc_t rsB = __builtin_amdgcn_make_buffer_rsrc(
        (void *)B, (short)0, (int)(((long long)(g.N - 1) * g.csB + g.K) * ES), 0x00020000);
    auto load = [&](VT (&ra)[AV], VT (&rb)[BV], int k0) {
#pragma unroll
        for (int i = 0; i < AV; ++i) {
            const int e = tid + i * NT;
            const int m = m0 + e / KQ, k = k0 + EV * (e % KQ);
            const unsigned off = (m < g.M && k < g.K) ? (unsigned)((m * (int)g.rsA + k) * ES) : OOB;
            ra[i] = __builtin_bit_cast(VT, __builtin_amdgcn_raw_buffer_load_b128(rsA, off, 0, 0));
        }
#pragma unroll
        for (int i = 0; i < BV; ++i) {
            const int e = tid + i * NT;
            const int n = n0 + e / KQ, k = k0 + EV * (e % KQ);
            const unsigned off = (n < g.N && k < g.K) ? (unsigned)((n * (int)g.csB + k) * ES) : OOB;
            rb[i] = __builtin_bit_cast(VT, __builtin_amdgcn_raw_buffer_load_b128(rsB, off, 0, 0));
        }
    };
    auto store = [&](const VT (&ra)[AV], const VT (&rb)[BV], int stage) {
#pragma unroll
        for (int i = 0; i < AV; ++i) {
            const int e = tid + i * NT;
            const int off = stage * ASZ + (e / KQ) * KP + EV * (e % KQ);
            if constexpr (BIN) *reinterpret_cast<u16x8 *>(&As[off]) = ra[i];
            else put_pieces<SPLIT>(As, BM * KP, off, ra[i]);
        }
#pragma unroll
        for (int i = 0; i < BV; ++i) {
            const int e = tid + i * NT;
            const int off = stage * BSZ + (e / KQ) * KP + EV * (e % KQ);
            if constexpr (BIN) *reinterpret_cast<u16x8 *>(&Bs[off]) = rb[i];
            else put_pieces<SPLIT>(Bs, BN * KP, off, rb[i]);
        }
    };
    auto mfma = [&](int stage) {
        bf16_slice<BM, BN, BK, KP, WM, WN, SPLIT>(As + stage * ASZ, Bs + stage * BSZ, acc, wm, wn, lane);
    };

    const int nk = (g.K + BK - 1) / BK;
    if (nk > 0) {
        load(ra0, rb0, 0);
        load(ra1, rb1, min(1, nk - 1) * BK);
        store(ra0, rb0, 0);
        load(ra0, rb0, min(2, nk - 1) * BK);
        __syncthreads();
        int cur = 0, t = 0;
        for (; t + 2 < nk; t += 2) {
            // iteration t (even): slice t + 1 sits in set 1, slice t + 3 goes there
            store(ra1, rb1, cur ^ 1);
            load(ra1, rb1, min(t + 3, nk - 1) * BK);
            __builtin_amdgcn_sched_barrier(0);
            mfma(cur);
            __syncthreads();
            cur ^= 1;
            // iteration t + 1: slice t + 2 sits in set 0, slice t + 4 goes there
            store(ra0, rb0, cur ^ 1);
            load(ra0, rb0, min(t + 4, nk - 1) * BK);
            __builtin_amdgcn_sched_barrier(0);
            mfma(cur);
            __syncthreads();
            cur ^= 1;
        }
        if (t + 1 < nk) {
            store(ra1, rb1, cur ^ 1);
            __builtin_amdgcn_sched_barrier(0);
            mfma(cur);
            __syncthreads();
            cur ^= 1;
        }
        mfma(cur);
    }

    store_c<float, TC, TM, TN, 16>(
        g, C, acc, [&](int i, int r) { return m0 + wm * WM + i * 32 + (r >> 2) * 8 + (lane >> 5) * 4 + (r & 3); },
        [&](int j) { return n0 + wn * WN + j * 32 + (lane & 31); });
}

// ---------------------------------------------------------------------------
// Host dispatch.
// ---------------------------------------------------------------------------
template <typename T, int BM, int BN, int BK, int WM, int WN, int VEC>
static int launch_v(GemmArgs g, int batch, bool amc, bool bnc, hipStream_t s) {
    g.tiles_m = (int)cdiv(g.M, BM);
    g.tiles_n = (int)cdiv(g.N, BN);
    dim3 grid(g.tiles_m * g.tiles_n, batch);
    dim3 block((BM / WM) * (BN / WN) * 64);
    if (amc && bnc) hipLaunchKernelGGL((gemm_kernel<T, BM, BN, BK, WM, WN, true, true, VEC>), grid, block, 0, s, g);
    else if (amc) hipLaunchKernelGGL((gemm_kernel<T, BM, BN, BK, WM, WN, true, false, VEC>), grid, block, 0, s, g);
    else if (bnc) hipLaunchKernelGGL((gemm_kernel<T, BM, BN, BK, WM, WN, false, true, VEC>), grid, block, 0, s, g);
    else hipLaunchKernelGGL((gemm_kernel<T, BM, BN, BK, WM, WN, false, false, VEC>), grid, block, 0, s, g);
    return check_launch();
}

// 16-byte vector loads are legal when every vector starts on a 16-byte
// boundary and never straddles the operand's edge.
static inline bool vec_ok(const GemmArgs &g, bool amc, bool bnc, int esize, int vec) {
    auto al = [&](const void *p, long long sb) {
        return (reinterpret_cast<uintptr_t>(p) % (esize * vec) == 0) && (sb % vec == 0);
    };
    const bool a = al(g.A, g.sAb) && (amc ? (g.M % vec == 0 && g.csA % vec == 0) : (g.K % vec == 0 && g.rsA % vec == 0));
    const bool b = al(g.B, g.sBb) && (bnc ? (g.N % vec == 0 && g.rsB % vec == 0) : (g.K % vec == 0 && g.csB % vec == 0));
    return a && b;
}

template <typename T, int BM, int BN, int BK, int WM, int WN>
static int launch_t(GemmArgs g, int batch, bool amc, bool bnc, hipStream_t s) {
    constexpr int V = 16 / (int)sizeof(T);
    if (vec_ok(g, amc, bnc, sizeof(T), V)) return launch_v<T, BM, BN, BK, WM, WN, V>(g, batch, amc, bnc, s);
    return launch_v<T, BM, BN, BK, WM, WN, 1>(g, batch, amc, bnc, s);
}

// Both operands K-contiguous with 16-byte vectors that never straddle K.
static inline bool kvec_ok(const GemmArgs &g, bool amc, bool bnc) {
    if (amc || bnc) return false;
    auto al = [](const void *p) { return reinterpret_cast<uintptr_t>(p) % 16 == 0; };
    return al(g.A) && al(g.B) && g.K % 4 == 0 && g.rsA % 4 == 0 && g.csB % 4 == 0 && g.sAb % 4 == 0 &&
           g.sBb % 4 == 0;
}

// KONLY: the tile exists for the K-contiguous kernel only (its generic
// counterpart would spill registers); other layouts run tile 256x128x32 w64x64
// (128x128x32 w64x64 when SPLIT), which gives the same bits.
template <int BM, int BN, int BK, int WM, int WN, int SPLIT, bool KONLY = false, bool DB = false,
          typename TC = float>
static int launch_bf16(GemmArgs g, int batch, bool amc, bool bnc, hipStream_t s) {
    static_assert(!DB || KONLY, "double-buffered tiles exist for the K-contiguous kernel only");
    if constexpr (KONLY) {
        if (!kvec_ok(g, amc, bnc)) {
            if constexpr (SPLIT > 0) return launch_bf16<128, 128, 32, 64, 64, SPLIT, false, false, TC>(g, batch, amc, bnc, s);
            else return launch_bf16<256, 128, 32, 64, 64, 0, false, false, TC>(g, batch, amc, bnc, s);
        }
    }
    g.tiles_m = (int)cdiv(g.M, BM);
    g.tiles_n = (int)cdiv(g.N, BN);
    dim3 grid(g.tiles_m * g.tiles_n, batch);
    dim3 block((BM / WM) * (BN / WN) * 64);
    if (KONLY || kvec_ok(g, amc, bnc)) {
        hipLaunchKernelGGL((gemm_bf16k_kernel<BM, BN, BK, WM, WN, SPLIT, DB, TC>), grid, block, 0, s, g);
        return check_launch();
    }
    if constexpr (!KONLY) {
        if (amc && bnc)
            hipLaunchKernelGGL((gemm_bf16_kernel<BM, BN, BK, WM, WN, true, true, SPLIT, TC>), grid, block, 0, s, g);
        else if (amc)
            hipLaunchKernelGGL((gemm_bf16_kernel<BM, BN, BK, WM, WN, true, false, SPLIT, TC>), grid, block, 0, s, g);
        else if (bnc)
            hipLaunchKernelGGL((gemm_bf16_kernel<BM, BN, BK, WM, WN, false, true, SPLIT, TC>), grid, block, 0, s, g);
        else
            hipLaunchKernelGGL((gemm_bf16_kernel<BM, BN, BK, WM, WN, false, false, SPLIT, TC>), grid, block, 0, s, g);
    }
    return check_launch();
}

// bf16 operands in memory: both must be K-contiguous with 16-byte runs of 8
// elements that never straddle K.
static inline bool kvec8_ok(const GemmArgs &g, bool amc, bool bnc) {
    if (amc || bnc) return false;
    auto al = [](const void *p) { return reinterpret_cast<uintptr_t>(p) % 16 == 0; };
    return al(g.A) && al(g.B) && g.K % 8 == 0 && g.rsA % 8 == 0 && g.csB % 8 == 0 && g.sAb % 8 == 0 &&
           g.sBb % 8 == 0;
}

template <int BM, int BN, int BK, int WM, int WN, bool DB = false>
static int launch_bf16in(GemmArgs g, int batch, bool amc, bool bnc, hipStream_t s) {
    if (!kvec8_ok(g, amc, bnc)) return TRG_UNSUPPORTED;
    g.tiles_m = (int)cdiv(g.M, BM);
    g.tiles_n = (int)cdiv(g.N, BN);
    dim3 grid(g.tiles_m * g.tiles_n, batch);
    dim3 block((BM / WM) * (BN / WN) * 64);
    hipLaunchKernelGGL((gemm_bf16in_kernel<BM, BN, BK, WM, WN, DB>), grid, block, 0, s, g);
    return check_launch();
}

// Pipelined kernel: fp32 storage needs the K-contiguous layout of
// gemm_bf16k_kernel (other layouts run the generic fallback tile, same bits);
// bf16 storage needs that of gemm_bf16in_kernel.
template <int BM, int BN, int BK, int WM, int WN, int SPLIT, bool BIN, int SCHED, typename TC = float, bool DEEP = false>
static int launch_bf16p(GemmArgs g, int batch, bool amc, bool bnc, hipStream_t s) {
    // the kernel's buffer loads address each operand instance with 32-bit
    // byte offsets and use 2^31 as the out-of-range offset
    const long long es = BIN ? 2 : 4;
    const bool fits = (((long long)(g.M - 1) * g.rsA + g.K) * es < (1LL << 31)) &&
                      (((long long)(g.N - 1) * g.csB + g.K) * es < (1LL << 31));
    if constexpr (BIN) {
        if (!kvec8_ok(g, amc, bnc)) return TRG_UNSUPPORTED;
        if (!fits) return launch_bf16in<128, 128, 32, 64, 64>(g, batch, amc, bnc, s);
    } else {
        if (!kvec_ok(g, amc, bnc) || !fits) {
            if constexpr (SPLIT > 0) return launch_bf16<128, 128, 32, 64, 64, SPLIT, false, false, TC>(g, batch, amc, bnc, s);
            else return launch_bf16<256, 128, 32, 64, 64, 0, false, false, TC>(g, batch, amc, bnc, s);
        }
    }
    g.tiles_m = (int)cdiv(g.M, BM);
    g.tiles_n = (int)cdiv(g.N, BN);
    dim3 grid(g.tiles_m * g.tiles_n, batch);
    dim3 block((BM / WM) * (BN / WN) * 64);
    if constexpr (DEEP)
        hipLaunchKernelGGL((gemm_bf16d_kernel<BM, BN, BK, WM, WN, SPLIT, BIN, TC>), grid, block, 0, s, g);
    else
        hipLaunchKernelGGL((gemm_bf16p_kernel<BM, BN, BK, WM, WN, SPLIT, BIN, SCHED, TC>), grid, block, 0, s, g);
    return check_launch();
}

// Tile tables of the compute families (config -1 selects by shape; every
// configuration of one family gives the same bits).
int gemm_dispatch_bin(int config, GemmArgs g, int batch, bool amc, bool bnc, hipStream_t s);
int gemm_dispatch_bf16out(int config, GemmArgs g, int batch, bool amc, bool bnc, hipStream_t s);
int gemm_dispatch_f32(int config, GemmArgs g, int batch, bool amc, bool bnc, hipStream_t s);
int gemm_dispatch_bf16(int config, GemmArgs g, int batch, bool amc, bool bnc, hipStream_t s);
int gemm_dispatch_x3(int config, GemmArgs g, int batch, bool amc, bool bnc, hipStream_t s);
int gemm_dispatch_x6(int config, GemmArgs g, int batch, bool amc, bool bnc, hipStream_t s);
int gemm_dispatch_f64(int config, GemmArgs g, int batch, bool amc, bool bnc, hipStream_t s);

}  // namespace trg
