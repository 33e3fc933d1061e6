// Batched strided GEMM on CDNA4 matrix cores (gfx950): C entry points and the
// dispatch to the tile tables of each (storage, output, compute) family.
#include "gemm_kernels.h"

namespace trg {

// 4-row vector stores in the epilogue: rows of C contiguous and every place a
// group of four consecutive rows (m a multiple of 4) can start a multiple of
// 4 elements, on a base aligned to the vector.
static int cvec_ok(const GemmArgs &g, int c_dtype) {
    const int esize = c_dtype == TRG_F64 ? 8 : (c_dtype == TRG_BF16 ? 2 : 4);
    if (g.rsC != 1) return 0;
    if (g.cdm && (g.cdm % 4 || g.csm % 4)) return 0;
    if (g.csC % 4 || (g.cdn && g.csn % 4) || g.sCb % 4) return 0;
    return reinterpret_cast<uintptr_t>(g.C) % (4 * esize) == 0 ? 1 : 0;
}

// Tile configuration table.  config -1 selects by shape (the default
// heuristic: the pipelined 256x256 tiles once a launch holds at least 512 of
// them, otherwise the generic tiles).  Every fp32 configuration uses the same MFMA instruction and
// walks K in the same order, so the choice never changes the result bits.
static int gemm_dispatch(int dtype, int c_dtype, int compute, int config, GemmArgs g, int batch, bool amc,
                         bool bnc, hipStream_t s) {
    if (dtype == TRG_BF16 && c_dtype == TRG_F32 && compute == TRG_BF16) return gemm_dispatch_bin(config, g, batch, amc, bnc, s);
    if (dtype == TRG_F32 && c_dtype == TRG_BF16 && compute == TRG_BF16) return gemm_dispatch_bf16out(config, g, batch, amc, bnc, s);
    if (c_dtype != dtype) return TRG_UNSUPPORTED;
    if (dtype == TRG_F32 && compute == TRG_F32) return gemm_dispatch_f32(config, g, batch, amc, bnc, s);
    if (dtype == TRG_F32 && compute == TRG_BF16) return gemm_dispatch_bf16(config, g, batch, amc, bnc, s);
    if (dtype == TRG_F32 && compute == TRG_BF16X3) return gemm_dispatch_x3(config, g, batch, amc, bnc, s);
    if (dtype == TRG_F32 && compute == TRG_BF16X6) return gemm_dispatch_x6(config, g, batch, amc, bnc, s);
    if (dtype == TRG_F64 && compute == TRG_F64) return gemm_dispatch_f64(config, g, batch, amc, bnc, s);
    return TRG_UNSUPPORTED;
}

}  // namespace trg

using namespace trg;

extern "C" int trg_gemm_config_count(int dtype, int c_dtype, int compute);

extern "C" int trg_gemm(int dtype, int compute, int batch, int M, int N, int K, double alpha,
                        const void *A, long long sAb, long long rsA, long long csA,
                        const void *B, long long sBb, long long rsB, long long csB,
                        double beta, void *C, long long sCb, long long rsC, long long csC,
                        void *stream) {
    clear_hip_error();
    if (batch < 0 || M < 0 || N < 0 || K < 0) return TRG_BAD_ARG;
    if (batch == 0 || M == 0 || N == 0) return TRG_OK;
    if (!A || !B || !C) return TRG_BAD_ARG;
    if (batch > 65535) return TRG_SHAPE;
    const bool amc = (rsA == 1);
    const bool bnc = (csB == 1);
    if (!amc && csA != 1) return TRG_BAD_ARG;
    if (!bnc && rsB != 1) return TRG_BAD_ARG;
    hipStream_t s = as_stream(stream);
    if (K == 0) {
        // C = beta * C: a GEMM with an empty sum.  Use the kernel with K = 0 (loop skipped).
    }
    GemmArgs g{M, N, K, A, sAb, rsA, csA, B, sBb, rsB, csB, C, sCb, rsC, csC, alpha, beta, 0, 0};
    g.cvec = cvec_ok(g, dtype);
    if (dtype == TRG_BF16) return TRG_UNSUPPORTED;   // bf16 storage: trg_gemm_ex
    return gemm_dispatch(dtype, dtype, compute, -1, g, batch, amc, bnc, s);
}

extern "C" int trg_gemm_config_count(int dtype, int c_dtype, int compute) {
    if (dtype == TRG_BF16 && c_dtype == TRG_F32 && compute == TRG_BF16) return 17;
    if (dtype == TRG_F32 && c_dtype == TRG_BF16 && compute == TRG_BF16) return 3;
    if (c_dtype != dtype) return 0;
    if (dtype == TRG_F32 && compute == TRG_F32) return 8;
    if (dtype == TRG_F32 && compute == TRG_BF16) return 15;
    if (dtype == TRG_F32 && compute == TRG_BF16X3) return 15;
    if (dtype == TRG_F32 && compute == TRG_BF16X6) return 13;
    if (dtype == TRG_F64 && compute == TRG_F64) return 4;
    return 0;
}

extern "C" const char *trg_gemm_config_name(int dtype, int c_dtype, int compute, int config) {
    static const char *f32[] = {"128x128x32 w64x64", "128x128x16 w64x64", "64x64x16 w32x32", "256x64x16 w64x64",
                                "64x256x16 w64x64", "256x128x32 w64x64", "128x256x32 w64x64", "256x128x16 w128x64"};
    static const char *bf[] = {"128x128x32 w64x64", "64x64x32 w32x32", "256x128x32 w64x64", "256x128x32 w128x64",
                               "128x256x32 w64x128", "128x128x64 w64x64", "256x128x64 w128x64",
                               "256x256x32 w128x64", "256x256x16 w128x64", "128x256x32 w64x64",
                               "256x256x32 w128x64 db", "256x128x32 w64x64 db", "256x256x32 w128x64 pipe",
                               "256x256x32 w128x64 pipe sched", "64x256x32 w64x64 pipe phased"};
    static const char *bx3[] = {"128x128x32 w64x64", "64x64x32 w32x32", "256x128x32 w64x64", "128x256x32 w64x64",
                                "256x256x32 w128x64", "256x128x32 w128x64", "256x128x32 w64x64 db",
                                "128x256x32 w64x64 db", "256x256x32 w128x64 pipe", "256x256x32 w128x64 pipe sched",
                                "256x128x32 w64x64 pipe sched", "256x256x32 w128x64 pipe phased",
                                "256x256x32 w128x128 pipe", "256x256x32 w128x128 pipe phased",
                                "64x256x32 w64x64 pipe phased"};
    static const char *bin[] = {"128x128x32 w64x64", "64x64x32 w32x32", "256x128x32 w64x64", "256x256x32 w128x64",
                                "256x256x64 w128x64", "256x256x32 w128x64 db", "256x128x64 w64x64 db",
                                "128x256x32 w64x64", "256x256x64 w128x64 db", "256x256x128 w128x64",
                                "128x256x64 w64x64 db", "256x256x64 w128x64 pipe", "256x256x64 w128x64 pipe sched",
                                "256x256x32 w128x64 pipe sched", "256x256x64 w128x64 pipe phased",
                                "256x256x64 w128x128 pipe", "256x256x64 w128x128 pipe phased"};
    static const char *bx6[] = {"128x128x32 w64x64", "64x64x32 w32x32", "256x128x32 w64x64", "128x256x32 w64x64",
                                "256x256x32 w128x64", "256x128x32 w128x64", "128x128x32 w64x64 pipe sched",
                                "256x256x16 w128x64 pipe", "256x256x16 w128x64 pipe sched",
                                "256x256x16 w128x64 pipe phased", "256x256x16 w128x128 pipe",
                                "64x256x16 w64x64 pipe phased", "256x256x16 w128x64 pipe deep"};
    static const char *f64[] = {"128x128x16 w64x64", "64x64x16 w32x32", "128x64x16 w64x32", "64x128x16 w32x64"};
    if (config < 0 || config >= trg_gemm_config_count(dtype, c_dtype, compute)) return "";
    if (dtype == TRG_BF16) return bin[config];
    if (c_dtype == TRG_BF16) return bf[config];
    if (dtype == TRG_F32 && compute == TRG_F32) return f32[config];
    if (dtype == TRG_F32 && compute == TRG_BF16) return bf[config];
    if (dtype == TRG_F32 && compute == TRG_BF16X3) return bx3[config];
    if (dtype == TRG_F32 && compute == TRG_BF16X6) return bx6[config];
    return f64[config];
}

extern "C" int trg_gemm_ex(int config, int dtype, int c_dtype, int compute, int batch, int M, int N, int K,
                           double alpha, const void *A, long long sAb, long long rsA, long long csA,
                           const void *B, long long sBb, long long rsB, long long csB,
                           double beta, void *C, long long sCb, long long rsC, long long csC,
                           int cdm, long long csm, int cdn, long long csn, void *stream) {
    clear_hip_error();
    if (batch < 0 || M < 0 || N < 0 || K < 0 || cdm < 0 || cdn < 0) return TRG_BAD_ARG;
    if (batch == 0 || M == 0 || N == 0) return TRG_OK;
    if (!A || !B || !C) return TRG_BAD_ARG;
    if (batch > 65535) return TRG_SHAPE;
    if (trg_gemm_config_count(dtype, c_dtype, compute) == 0) return TRG_UNSUPPORTED;
    if (config < -1 || config >= trg_gemm_config_count(dtype, c_dtype, compute)) return TRG_BAD_ARG;
    const bool amc = (rsA == 1);
    const bool bnc = (csB == 1);
    if (!amc && csA != 1) return TRG_BAD_ARG;
    if (!bnc && rsB != 1) return TRG_BAD_ARG;
    GemmArgs g{M, N, K, A, sAb, rsA, csA, B, sBb, rsB, csB, C, sCb, rsC, csC, alpha, beta, 0, 0,
               cdm, cdn, csm, csn};
    g.cvec = cvec_ok(g, c_dtype);
    return gemm_dispatch(dtype, c_dtype, compute, config, g, batch, amc, bnc, as_stream(stream));
}
