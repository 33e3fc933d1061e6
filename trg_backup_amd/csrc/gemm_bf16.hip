// bf16 MFMA tile tables: bf16 operands in memory, bf16 output, fp32 operands rounded to bf16.
#include "gemm_kernels.h"

namespace trg {

int gemm_dispatch_bin(int config, GemmArgs g, int batch, bool amc, bool bnc, hipStream_t s) {
    const long long work = (long long)g.M * g.N;
    (void)work;
    if (config < 0) config = work * batch >= 256LL * 256 * 512 ? 14 : (work >= 256LL * 256 * 64 ? 3 : 1);
    switch (config) {
        case 0: return launch_bf16in<128, 128, 32, 64, 64>(g, batch, amc, bnc, s);
        case 1: return launch_bf16in<64, 64, 32, 32, 32>(g, batch, amc, bnc, s);
        case 2: return launch_bf16in<256, 128, 32, 64, 64>(g, batch, amc, bnc, s);
        case 3: return launch_bf16in<256, 256, 32, 128, 64>(g, batch, amc, bnc, s);
        case 4: return launch_bf16in<256, 256, 64, 128, 64>(g, batch, amc, bnc, s);
        case 5: return launch_bf16in<256, 256, 32, 128, 64, true>(g, batch, amc, bnc, s);
        case 6: return launch_bf16in<256, 128, 64, 64, 64, true>(g, batch, amc, bnc, s);
        case 7: return launch_bf16in<128, 256, 32, 64, 64>(g, batch, amc, bnc, s);
        case 8: return launch_bf16in<256, 256, 64, 128, 64, true>(g, batch, amc, bnc, s);
        case 9: return launch_bf16in<256, 256, 128, 128, 64>(g, batch, amc, bnc, s);
        case 10: return launch_bf16in<128, 256, 64, 64, 64, true>(g, batch, amc, bnc, s);
        case 11: return launch_bf16p<256, 256, 64, 128, 64, 0, true, 0>(g, batch, amc, bnc, s);
        case 12: return launch_bf16p<256, 256, 64, 128, 64, 0, true, 1>(g, batch, amc, bnc, s);
        case 13: return launch_bf16p<256, 256, 32, 128, 64, 0, true, 1>(g, batch, amc, bnc, s);
        case 14: return launch_bf16p<256, 256, 64, 128, 64, 0, true, 2>(g, batch, amc, bnc, s);
        case 15: return launch_bf16p<256, 256, 64, 128, 128, 0, true, 0>(g, batch, amc, bnc, s);
        case 16: return launch_bf16p<256, 256, 64, 128, 128, 0, true, 2>(g, batch, amc, bnc, s);
        default: return TRG_BAD_ARG;
    }
}

int gemm_dispatch_bf16out(int config, GemmArgs g, int batch, bool amc, bool bnc, hipStream_t s) {
    const long long work = (long long)g.M * g.N;
    (void)work;
    if (config < 0) config = work >= 256LL * 128 * 128 ? 0 : 1;
    switch (config) {
        case 0: return launch_bf16<128, 128, 32, 64, 64, false, false, false, __bf16>(g, batch, amc, bnc, s);
        case 1: return launch_bf16<64, 64, 32, 32, 32, false, false, false, __bf16>(g, batch, amc, bnc, s);
        case 2: return launch_bf16<256, 128, 32, 64, 64, false, false, false, __bf16>(g, batch, amc, bnc, s);
        default: return TRG_BAD_ARG;
    }
}

int gemm_dispatch_bf16(int config, GemmArgs g, int batch, bool amc, bool bnc, hipStream_t s) {
    const long long work = (long long)g.M * g.N;
    (void)work;
    if (config < 0) config = work >= 256LL * 128 * 128 ? 0 : 1;
    switch (config) {
        case 0: return launch_bf16<128, 128, 32, 64, 64, false>(g, batch, amc, bnc, s);
        case 1: return launch_bf16<64, 64, 32, 32, 32, false>(g, batch, amc, bnc, s);
        case 2: return launch_bf16<256, 128, 32, 64, 64, false>(g, batch, amc, bnc, s);
        case 3: return launch_bf16<256, 128, 32, 128, 64, false>(g, batch, amc, bnc, s);
        case 4: return launch_bf16<128, 256, 32, 64, 128, false>(g, batch, amc, bnc, s);
        case 5: return launch_bf16<128, 128, 64, 64, 64, false>(g, batch, amc, bnc, s);
        case 6: return launch_bf16<256, 128, 64, 128, 64, false, true>(g, batch, amc, bnc, s);
        case 7: return launch_bf16<256, 256, 32, 128, 64, false, true>(g, batch, amc, bnc, s);
        case 8: return launch_bf16<256, 256, 16, 128, 64, false>(g, batch, amc, bnc, s);
        case 9: return launch_bf16<128, 256, 32, 64, 64, false>(g, batch, amc, bnc, s);
        case 10: return launch_bf16<256, 256, 32, 128, 64, false, true, true>(g, batch, amc, bnc, s);
        case 11: return launch_bf16<256, 128, 32, 64, 64, false, true, true>(g, batch, amc, bnc, s);
        case 12: return launch_bf16p<256, 256, 32, 128, 64, 0, false, 0>(g, batch, amc, bnc, s);
        case 13: return launch_bf16p<256, 256, 32, 128, 64, 0, false, 1>(g, batch, amc, bnc, s);
        case 14: return launch_bf16p<64, 256, 32, 64, 64, 0, false, 2>(g, batch, amc, bnc, s);
        default: return TRG_BAD_ARG;
    }
}

}  // namespace trg
