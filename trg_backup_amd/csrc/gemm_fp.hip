// fp32 and fp64 MFMA tile tables.
#include "gemm_kernels.h"

namespace trg {

int gemm_dispatch_f32(int config, GemmArgs g, int batch, bool amc, bool bnc, hipStream_t s) {
    const long long work = (long long)g.M * g.N;
    (void)work;
    if (config < 0) {
        if (g.N <= 64 && g.M >= 4096) config = 3;
        else if (g.M <= 64 && g.N >= 4096) config = 4;
        else if (work >= 256LL * 128 * 128) config = 0;
        else config = 2;
    }
    switch (config) {
        case 0: return launch_t<float, 128, 128, 32, 64, 64>(g, batch, amc, bnc, s);
        case 1: return launch_t<float, 128, 128, 16, 64, 64>(g, batch, amc, bnc, s);
        case 2: return launch_t<float, 64, 64, 16, 32, 32>(g, batch, amc, bnc, s);
        case 3: return launch_t<float, 256, 64, 16, 64, 64>(g, batch, amc, bnc, s);
        case 4: return launch_t<float, 64, 256, 16, 64, 64>(g, batch, amc, bnc, s);
        case 5: return launch_t<float, 256, 128, 32, 64, 64>(g, batch, amc, bnc, s);
        case 6: return launch_t<float, 128, 256, 32, 64, 64>(g, batch, amc, bnc, s);
        case 7: return launch_t<float, 256, 128, 16, 128, 64>(g, batch, amc, bnc, s);
        default: return TRG_BAD_ARG;
    }
}

int gemm_dispatch_f64(int config, GemmArgs g, int batch, bool amc, bool bnc, hipStream_t s) {
    const long long work = (long long)g.M * g.N;
    (void)work;
    if (config < 0) config = work >= 256LL * 128 * 128 ? 0 : 1;
    switch (config) {
        case 0: return launch_t<double, 128, 128, 16, 64, 64>(g, batch, amc, bnc, s);
        case 1: return launch_t<double, 64, 64, 16, 32, 32>(g, batch, amc, bnc, s);
        case 2: return launch_t<double, 128, 64, 16, 64, 32>(g, batch, amc, bnc, s);
        case 3: return launch_t<double, 64, 128, 16, 32, 64>(g, batch, amc, bnc, s);
        default: return TRG_BAD_ARG;
    }
}

}  // namespace trg
