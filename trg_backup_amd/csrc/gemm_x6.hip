// Three-piece split-bf16 (bf16x6) tile table.
#include "gemm_kernels.h"

namespace trg {

int gemm_dispatch_x6(int config, GemmArgs g, int batch, bool amc, bool bnc, hipStream_t s) {
    const long long work = (long long)g.M * g.N;
    (void)work;
    if (config < 0) config = work * batch >= 256LL * 256 * 512 ? 12 : (work >= 256LL * 128 * 128 ? 0 : 1);
    switch (config) {
        case 0: return launch_bf16<128, 128, 32, 64, 64, 2>(g, batch, amc, bnc, s);
        case 1: return launch_bf16<64, 64, 32, 32, 32, 2>(g, batch, amc, bnc, s);
        case 2: return launch_bf16<256, 128, 32, 64, 64, 2>(g, batch, amc, bnc, s);
        case 3: return launch_bf16<128, 256, 32, 64, 64, 2>(g, batch, amc, bnc, s);
        case 4: return launch_bf16<256, 256, 32, 128, 64, 2, true>(g, batch, amc, bnc, s);
        case 5: return launch_bf16<256, 128, 32, 128, 64, 2, true>(g, batch, amc, bnc, s);
        case 6: return launch_bf16p<128, 128, 32, 64, 64, 2, false, 1>(g, batch, amc, bnc, s);
        case 7: return launch_bf16p<256, 256, 16, 128, 64, 2, false, 0>(g, batch, amc, bnc, s);
        case 8: return launch_bf16p<256, 256, 16, 128, 64, 2, false, 1>(g, batch, amc, bnc, s);
        case 9: return launch_bf16p<256, 256, 16, 128, 64, 2, false, 2>(g, batch, amc, bnc, s);
        case 10: return launch_bf16p<256, 256, 16, 128, 128, 2, false, 0>(g, batch, amc, bnc, s);
        case 11: return launch_bf16p<64, 256, 16, 64, 64, 2, false, 2>(g, batch, amc, bnc, s);
        case 12: return launch_bf16p<256, 256, 16, 128, 64, 2, false, 2, float, true>(g, batch, amc, bnc, s);
        default: return TRG_BAD_ARG;
    }
}

}  // namespace trg
