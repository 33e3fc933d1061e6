// Memory-bound tensor kernels: strided copy / permutation with optional
// fp32<->fp64 conversion, fill, max-|T| and trace reductions, normalisation
// with the per-step log-norm, leg scaling, isometry selection and the Philox
// Gaussian sketch.  All of them are HBM-bound; they are written for
// coalesced stores (the output index is the thread index) and keep every
// reduction order fixed so results do not depend on the batch a tensor sits in.
#include "common.h"

namespace trg {

thread_local int g_last_hip_error = 0;

struct StridedDesc {
    int rank;
    long long dims[8];
    long long si[8];
    long long so[8];
};

template <typename Ti, typename To>
__global__ void copy_strided_kernel(StridedDesc d, const Ti *__restrict__ in, To *__restrict__ out, long long total) {
    for (long long lin = blockIdx.x * (long long)blockDim.x + threadIdx.x; lin < total;
         lin += (long long)gridDim.x * blockDim.x) {
        long long r = lin, oi = 0, oo = 0;
#pragma unroll
        for (int k = 7; k >= 0; --k) {
            if (k < d.rank) {
                const long long ik = r % d.dims[k];
                r /= d.dims[k];
                oi += ik * d.si[k];
                oo += ik * d.so[k];
            }
        }
        out[oo] = (To)in[oi];
    }
}

// Tiled transpose-copy for a permutation whose input-contiguous leg (ii)
// differs from the output-contiguous leg (oi): a 32 x 32 tile is read
// coalesced along ii, staged in LDS and written coalesced along oi.  The
// remaining legs are enumerated by the workgroup index (decomposed once per
// workgroup).  Output and input strides are arbitrary.
constexpr int TT = 32;
struct TransDesc {
    int nouter;
    long long odims[6];
    long long osi[6], oso[6];  // strides of the outer legs
    long long ni, nj;          // extents of leg ii (input-contiguous) and oi (output-contiguous)
    long long si_j, so_i;      // input stride of oi, output stride of ii
    long long tiles_i, tiles_j;
    long long nblk;            // number of tiles in the launch
};

template <typename Ti, typename To>
__global__ void __launch_bounds__(256) transpose_kernel(TransDesc d, const Ti *__restrict__ in, To *__restrict__ out) {
    __shared__ To tile[TT][TT + 1];
    long long bid = blockIdx.x + (long long)blockIdx.y * gridDim.x;
    if (bid >= d.nblk) return;
    const long long tiles = d.tiles_i * d.tiles_j;
    long long outer = bid / tiles;
    const long long t = bid - outer * tiles;
    const long long ti = t % d.tiles_i, tj = t / d.tiles_i;
    long long offi = 0, offo = 0;
    for (int k = d.nouter - 1; k >= 0; --k) {
        const long long ik = outer % d.odims[k];
        outer /= d.odims[k];
        offi += ik * d.osi[k];
        offo += ik * d.oso[k];
    }
    const int tx = threadIdx.x & (TT - 1), ty = threadIdx.x / TT;   // 32 x 8
    const long long i0 = ti * TT, j0 = tj * TT;
#pragma unroll
    for (int r = 0; r < TT; r += 8) {
        const long long i = i0 + tx, j = j0 + ty + r;
        if (i < d.ni && j < d.nj) tile[ty + r][tx] = (To)in[offi + i + j * d.si_j];
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < TT; r += 8) {
        const long long j = j0 + tx, i = i0 + ty + r;
        if (i < d.ni && j < d.nj) out[offo + j + i * d.so_i] = tile[tx][ty + r];
    }
}

// Row copy when input and output share their contiguous leg: each
// workgroup copies ROWS rows of length n (coalesced on both sides).
struct RowDesc {
    int nouter;
    long long odims[7];
    long long osi[7], oso[7];
    long long n;        // row length
    long long rows;
};

// Each workgroup copies a tile of RCH elements: either R = RCH / n whole rows
// (short rows, every thread busy) or one RCH-long segment of a long row.
// Row base offsets are computed once per row into LDS.
constexpr int RCH = 4096;

template <typename Ti, typename To>
__global__ void __launch_bounds__(256) rowcopy_kernel(RowDesc d, const Ti *__restrict__ in, To *__restrict__ out,
                                                      long long nblk) {
    __shared__ long long offi_s[256], offo_s[256];
    const long long bid = blockIdx.x + (long long)blockIdx.y * gridDim.x;
    if (bid >= nblk) return;
    const long long R = d.n >= RCH ? 1 : min(256LL, RCH / d.n);   // rows in this tile
    const long long segs = d.n >= RCH ? (d.n + RCH - 1) / RCH : 1;  // segments per row
    const long long r0 = (bid / segs) * R;
    const long long c0 = (bid % segs) * RCH;
    const long long ncol = min((long long)RCH, d.n - c0);
    const int nr = (int)min(R, d.rows - r0);
    for (int rr = threadIdx.x; rr < nr; rr += blockDim.x) {
        long long o = r0 + rr, oi = 0, oo = 0;
        for (int k = d.nouter - 1; k >= 0; --k) {
            const long long ik = o % d.odims[k];
            o /= d.odims[k];
            oi += ik * d.osi[k];
            oo += ik * d.oso[k];
        }
        offi_s[rr] = oi + c0;
        offo_s[rr] = oo + c0;
    }
    __syncthreads();
    const int tot = (int)(nr * ncol);
    for (int e = threadIdx.x; e < tot; e += blockDim.x) {
        const int rr = e / (int)ncol, c = e - rr * (int)ncol;
        out[offo_s[rr] + c] = (To)in[offi_s[rr] + c];
    }
}

template <typename T>
__global__ void fill_kernel(T *out, long long n, T v) {
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
        out[i] = v;
}

// Block max of |T| per instance; deterministic (max is order independent).
// One workgroup per (chunk, instance); per-chunk maxima go to `part`,
// the final pass reduces them.
template <typename T>
__global__ void maxabs_partial(const T *__restrict__ X, long long n, long long sb, double *part, int nchunk) {
    __shared__ double red[256];
    const int b = blockIdx.y;
    const T *x = X + (long long)b * sb;
    const long long per = (n + nchunk - 1) / nchunk;
    const long long beg = (long long)blockIdx.x * per;
    const long long end = min(n, beg + per);
    double m = 0.0;
    for (long long i = beg + threadIdx.x; i < end; i += blockDim.x) {
        const double v = fabs((double)x[i]);
        m = (v > m || v != v) ? v : m;   // NaN propagates
    }
    red[threadIdx.x] = m;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s) {
            const double o = red[threadIdx.x + s];
            if (o > red[threadIdx.x] || o != o) red[threadIdx.x] = o;
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) part[(long long)b * nchunk + blockIdx.x] = red[0];
}

__global__ void maxabs_final(const double *part, int nchunk, double *c) {
    __shared__ double red[256];
    const int b = blockIdx.x;
    double m = 0.0;
    for (int i = threadIdx.x; i < nchunk; i += blockDim.x) {
        const double v = part[(long long)b * nchunk + i];
        m = (v > m || v != v) ? v : m;
    }
    red[threadIdx.x] = m;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s) {
            const double o = red[threadIdx.x + s];
            if (o > red[threadIdx.x] || o != o) red[threadIdx.x] = o;
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) c[b] = red[0];
}

// Generalised trace in fp64 with a fixed tree order (one workgroup per instance).
struct TraceDesc {
    int rank;
    long long dims[4];
    long long strides[4];
};

template <typename T>
__global__ void trace_kernel(TraceDesc d, const T *__restrict__ X, long long sb, double *t) {
    __shared__ double red[256];
    const int b = blockIdx.x;
    const T *x = X + (long long)b * sb;
    long long total = 1;
    for (int k = 0; k < d.rank; ++k) total *= d.dims[k];
    double acc = 0.0;
    for (long long lin = threadIdx.x; lin < total; lin += blockDim.x) {
        long long r = lin, off = 0;
        for (int k = d.rank - 1; k >= 0; --k) {
            const long long ik = r % d.dims[k];
            r /= d.dims[k];
            off += ik * d.strides[k];
        }
        acc += (double)x[off];
    }
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) t[b] = red[0];
}

// lnf[b, step] = ln|c_b| / 2^step (one thread per instance); the step index
// comes from the device when step_ptr is given so captured graphs advance it.
__global__ void lognorm_kernel(int batch, const double *c, double *lnf, long long lnf_stride, int step,
                               int *step_ptr) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    const int st = step_ptr ? *step_ptr : step;
    if (b < batch && lnf) lnf[(long long)b * lnf_stride + st] = log(fabs(c[b])) / ldexp(1.0, st);
}

__global__ void bump_step_kernel(int *step_ptr) {
    if (threadIdx.x == 0 && blockIdx.x == 0) *step_ptr += 1;
}

template <typename T>
__global__ void scale_by_inv_kernel(T *__restrict__ X, long long n, long long sb, const double *c) {
    const int b = blockIdx.y;
    const double cb = fabs(c[b]);
    const T inv = (T)(1.0 / cb);
    T *x = X + (long long)b * sb;
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
        x[i] *= inv;
}

// Y[b][i] *= f(x[b][(i / inner) % period]); mode 0: x, 1: sqrt(x), 2: 1/sqrt(x), 3: 1/x.
template <typename T>
__global__ void scale_rows_kernel(T *__restrict__ Y, long long n, long long sY, long long period, long long inner,
                                  const T *__restrict__ x, long long sx, int mode) {
    const int b = blockIdx.y;
    T *y = Y + (long long)b * sY;
    const T *xb = x + (long long)b * sx;
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
        const T v = xb[(i / inner) % period];
        T f;
        if (mode == 0) f = v;
        else if (mode == 1) f = sqrt(v < (T)0 ? (T)0 : v);   // eigenvalues a rounding below 0 -> 0; NaN stays
        else if (mode == 2) f = (T)1 / sqrt(v);
        else f = (T)1 / v;
        y[i] *= f;
    }
}

template <typename T>
__global__ void select_kernel(long long nk, const double *eps0, const double *eps1, const double *V0,
                              const double *V1, T *U, int *flag) {
    const int b = blockIdx.y;
    const bool side1 = !(eps0[b] < eps1[b]);
    const double *src = (side1 ? V1 : V0) + (long long)b * nk;
    T *dst = U + (long long)b * nk;
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < nk; i += (long long)gridDim.x * blockDim.x)
        dst[i] = (T)src[i];
    if (flag && blockIdx.x == 0 && threadIdx.x == 0) flag[b] = side1 ? 1 : 0;
}

// Philox4x32-10 (Salmon et al., SC'11) -> Box-Muller, identical to the
// NumPy generator used by the tests.
__device__ static inline void philox10(unsigned int c[4], unsigned int k0, unsigned int k1) {
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        const unsigned long long p0 = 0xD2511F53ull * (unsigned long long)c[0];
        const unsigned long long p1 = 0xCD9E8D57ull * (unsigned long long)c[2];
        const unsigned int hi0 = (unsigned int)(p0 >> 32), lo0 = (unsigned int)p0;
        const unsigned int hi1 = (unsigned int)(p1 >> 32), lo1 = (unsigned int)p1;
        const unsigned int n0 = hi1 ^ c[1] ^ k0, n2 = hi0 ^ c[3] ^ k1;
        c[0] = n0; c[1] = lo1; c[2] = n2; c[3] = lo0;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
}

template <typename T>
__global__ void gaussian_kernel(long long n, unsigned long long seed, unsigned int stream_id, unsigned int stride,
                                unsigned long long offset, T *out, long long sb) {
    const int b = blockIdx.y;
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
        const unsigned long long e = offset + (unsigned long long)i;
        unsigned int c[4] = {(unsigned int)e, (unsigned int)(e >> 32), stream_id + (unsigned int)b * stride, 0u};
        philox10(c, (unsigned int)seed, (unsigned int)(seed >> 32));
        const double u0 = ((double)c[0] + 0.5) * (1.0 / 4294967296.0);
        const double u1 = ((double)c[1] + 0.5) * (1.0 / 4294967296.0);
        out[(long long)b * sb + i] = (T)(sqrt(-2.0 * log(u0)) * cos(6.283185307179586 * u1));
    }
}

// lnz[b] = sum_i lnf[b, i] + ln|tr_b| * inv_volume (fixed summation order)
// health[0] |= 1 if any info[i] != 0, |= 2 if any vals[i] is NaN, infinite or zero
// (one workgroup; a wavefront OR-reduction, then one lane per wave into LDS).
__global__ void __launch_bounds__(256) health_or_kernel(long long n_info, const int *info, long long n_vals,
                                                        const double *vals, int *health) {
    __shared__ int part[4];
    int bits = 0;
    for (long long i = threadIdx.x; i < n_info; i += blockDim.x) bits |= info[i] != 0 ? 1 : 0;
    for (long long i = threadIdx.x; i < n_vals; i += blockDim.x) {
        const double v = vals[i];
        bits |= (v != v || fabs(v) > 1.7976931348623157e308 || v == 0.0) ? 2 : 0;
    }
    for (int off = 32; off > 0; off >>= 1) bits |= __shfl_xor(bits, off);
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = bits;
    __syncthreads();
    if (threadIdx.x == 0) {
        const int all = part[0] | part[1] | part[2] | part[3];
        if (all) health[0] = health[0] | all;
    }
}

__global__ void lnz_kernel(int batch, int nterms, const double *lnf, long long lnf_stride, const double *tr,
                           double inv_volume, double *out) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= batch) return;
    double s = 0.0;
    for (int i = 0; i < nterms; ++i) s += lnf[(long long)b * lnf_stride + i];
    if (tr) s += log(fabs(tr[b])) * inv_volume;
    out[b] = s;
}

static inline unsigned grid_for(long long n, int block) {
    long long g = cdiv(n, block);
    if (g > 65536) g = 65536;
    if (g < 1) g = 1;
    return (unsigned)g;
}

}  // namespace trg

using namespace trg;

extern "C" int trg_last_hip_error(void) { return g_last_hip_error; }

extern "C" int trg_copy_strided(int dtype_in, int dtype_out, int rank, const long long *dims, const void *in,
                                const long long *si, void *out, const long long *so, void *stream) {
    clear_hip_error();
    if (rank < 0 || rank > 8 || !dims || !si || !so) return TRG_BAD_ARG;
    StridedDesc d;
    d.rank = rank;
    long long total = 1;
    for (int k = 0; k < rank; ++k) {
        if (dims[k] < 0) return TRG_BAD_ARG;
        d.dims[k] = dims[k];
        d.si[k] = si[k];
        d.so[k] = so[k];
        total *= dims[k];
    }
    for (int k = rank; k < 8; ++k) { d.dims[k] = 1; d.si[k] = 0; d.so[k] = 0; }
    if (total == 0) return TRG_OK;
    if (!in || !out) return TRG_BAD_ARG;
    hipStream_t s = as_stream(stream);
    {
        // fast paths: the leg contiguous in the output (oi) and in the input (ii)
        int oi = -1, ii = -1;
        for (int k = 0; k < rank; ++k) {
            if (so[k] == 1 && dims[k] > 1) oi = k;
            if (si[k] == 1 && dims[k] > 1) ii = k;
        }
        if (oi >= 0 && ii >= 0 && oi == ii && dims[oi] >= 16) {
            RowDesc r;
            r.nouter = 0;
            for (int k = 0; k < rank; ++k) {
                if (k == oi) continue;
                r.odims[r.nouter] = dims[k];
                r.osi[r.nouter] = si[k];
                r.oso[r.nouter] = so[k];
                ++r.nouter;
            }
            r.n = dims[oi];
            r.rows = total / dims[oi];
            const long long R = r.n >= RCH ? 1 : std::min(256LL, RCH / r.n);
            const long long segs = r.n >= RCH ? cdiv(r.n, RCH) : 1;
            const long long nblk = cdiv(r.rows, R) * segs;
            const dim3 grid((unsigned)std::min<long long>(nblk, 65535), (unsigned)cdiv(nblk, 65535));
            if (dtype_in == TRG_F32 && dtype_out == TRG_F32) hipLaunchKernelGGL((rowcopy_kernel<float, float>), grid, 256, 0, s, r, (const float *)in, (float *)out, nblk);
            else if (dtype_in == TRG_F64 && dtype_out == TRG_F64) hipLaunchKernelGGL((rowcopy_kernel<double, double>), grid, 256, 0, s, r, (const double *)in, (double *)out, nblk);
            else if (dtype_in == TRG_F32 && dtype_out == TRG_F64) hipLaunchKernelGGL((rowcopy_kernel<float, double>), grid, 256, 0, s, r, (const float *)in, (double *)out, nblk);
            else if (dtype_in == TRG_F64 && dtype_out == TRG_F32) hipLaunchKernelGGL((rowcopy_kernel<double, float>), grid, 256, 0, s, r, (const double *)in, (float *)out, nblk);
            else if (dtype_in == TRG_F32 && dtype_out == TRG_BF16) hipLaunchKernelGGL((rowcopy_kernel<float, __bf16>), grid, 256, 0, s, r, (const float *)in, (__bf16 *)out, nblk);
            else if (dtype_in == TRG_BF16 && dtype_out == TRG_F32) hipLaunchKernelGGL((rowcopy_kernel<__bf16, float>), grid, 256, 0, s, r, (const __bf16 *)in, (float *)out, nblk);
            else if (dtype_in == TRG_BF16 && dtype_out == TRG_BF16) hipLaunchKernelGGL((rowcopy_kernel<__bf16, __bf16>), grid, 256, 0, s, r, (const __bf16 *)in, (__bf16 *)out, nblk);
            else return TRG_UNSUPPORTED;
            return check_launch();
        }
        if (oi >= 0 && ii >= 0 && oi != ii && dims[oi] >= 8 && dims[ii] >= 8 && rank <= 8) {
            TransDesc t;
            t.nouter = 0;
            for (int k = 0; k < rank; ++k) {
                if (k == oi || k == ii) continue;
                t.odims[t.nouter] = dims[k];
                t.osi[t.nouter] = si[k];
                t.oso[t.nouter] = so[k];
                ++t.nouter;
            }
            t.ni = dims[ii];
            t.nj = dims[oi];
            t.si_j = si[oi];
            t.so_i = so[ii];
            t.tiles_i = cdiv(t.ni, TT);
            t.tiles_j = cdiv(t.nj, TT);
            const long long nblk = t.tiles_i * t.tiles_j * (total / (t.ni * t.nj));
            t.nblk = nblk;
            const dim3 grid((unsigned)std::min<long long>(nblk, 65535), (unsigned)cdiv(nblk, 65535));
            if (dtype_in == TRG_F32 && dtype_out == TRG_F32) hipLaunchKernelGGL((transpose_kernel<float, float>), grid, 256, 0, s, t, (const float *)in, (float *)out);
            else if (dtype_in == TRG_F64 && dtype_out == TRG_F64) hipLaunchKernelGGL((transpose_kernel<double, double>), grid, 256, 0, s, t, (const double *)in, (double *)out);
            else if (dtype_in == TRG_F32 && dtype_out == TRG_F64) hipLaunchKernelGGL((transpose_kernel<float, double>), grid, 256, 0, s, t, (const float *)in, (double *)out);
            else if (dtype_in == TRG_F64 && dtype_out == TRG_F32) hipLaunchKernelGGL((transpose_kernel<double, float>), grid, 256, 0, s, t, (const double *)in, (float *)out);
            else if (dtype_in == TRG_F32 && dtype_out == TRG_BF16) hipLaunchKernelGGL((transpose_kernel<float, __bf16>), grid, 256, 0, s, t, (const float *)in, (__bf16 *)out);
            else if (dtype_in == TRG_BF16 && dtype_out == TRG_F32) hipLaunchKernelGGL((transpose_kernel<__bf16, float>), grid, 256, 0, s, t, (const __bf16 *)in, (float *)out);
            else if (dtype_in == TRG_BF16 && dtype_out == TRG_BF16) hipLaunchKernelGGL((transpose_kernel<__bf16, __bf16>), grid, 256, 0, s, t, (const __bf16 *)in, (__bf16 *)out);
            else return TRG_UNSUPPORTED;
            return check_launch();
        }
    }
    const unsigned g = grid_for(total, 256);
    if (dtype_in == TRG_F32 && dtype_out == TRG_F32)
        hipLaunchKernelGGL((copy_strided_kernel<float, float>), g, 256, 0, s, d, (const float *)in, (float *)out, total);
    else if (dtype_in == TRG_F64 && dtype_out == TRG_F64)
        hipLaunchKernelGGL((copy_strided_kernel<double, double>), g, 256, 0, s, d, (const double *)in, (double *)out, total);
    else if (dtype_in == TRG_F32 && dtype_out == TRG_F64)
        hipLaunchKernelGGL((copy_strided_kernel<float, double>), g, 256, 0, s, d, (const float *)in, (double *)out, total);
    else if (dtype_in == TRG_F64 && dtype_out == TRG_F32)
        hipLaunchKernelGGL((copy_strided_kernel<double, float>), g, 256, 0, s, d, (const double *)in, (float *)out, total);
    else if (dtype_in == TRG_F32 && dtype_out == TRG_BF16)
        hipLaunchKernelGGL((copy_strided_kernel<float, __bf16>), g, 256, 0, s, d, (const float *)in, (__bf16 *)out, total);
    else if (dtype_in == TRG_BF16 && dtype_out == TRG_F32)
        hipLaunchKernelGGL((copy_strided_kernel<__bf16, float>), g, 256, 0, s, d, (const __bf16 *)in, (float *)out, total);
    else if (dtype_in == TRG_BF16 && dtype_out == TRG_BF16)
        hipLaunchKernelGGL((copy_strided_kernel<__bf16, __bf16>), g, 256, 0, s, d, (const __bf16 *)in, (__bf16 *)out, total);
    else
        return TRG_UNSUPPORTED;
    return check_launch();
}

extern "C" int trg_fill(int dtype, long long n, double value, void *out, void *stream) {
    clear_hip_error();
    if (n < 0) return TRG_BAD_ARG;
    if (n == 0) return TRG_OK;
    if (!out) return TRG_BAD_ARG;
    hipStream_t s = as_stream(stream);
    if (dtype == TRG_F32) hipLaunchKernelGGL(fill_kernel<float>, grid_for(n, 256), 256, 0, s, (float *)out, n, (float)value);
    else if (dtype == TRG_F64) hipLaunchKernelGGL(fill_kernel<double>, grid_for(n, 256), 256, 0, s, (double *)out, n, value);
    else if (dtype == TRG_I32) hipLaunchKernelGGL(fill_kernel<int>, grid_for(n, 256), 256, 0, s, (int *)out, n, (int)value);
    else return TRG_UNSUPPORTED;
    return check_launch();
}

extern "C" int trg_maxabs(int dtype, int batch, long long n, const void *T, long long sb, double *c,
                          double *part, int nchunk, void *stream) {
    clear_hip_error();
    if (batch < 0 || n < 0 || nchunk < 1) return TRG_BAD_ARG;
    if (batch == 0) return TRG_OK;
    if (!T || !c || !part) return TRG_BAD_ARG;
    if (batch > 65535 || nchunk > 65535) return TRG_SHAPE;
    hipStream_t s = as_stream(stream);
    if (dtype == TRG_F32) hipLaunchKernelGGL(maxabs_partial<float>, dim3(nchunk, batch), 256, 0, s, (const float *)T, n, sb, part, nchunk);
    else if (dtype == TRG_F64) hipLaunchKernelGGL(maxabs_partial<double>, dim3(nchunk, batch), 256, 0, s, (const double *)T, n, sb, part, nchunk);
    else return TRG_UNSUPPORTED;
    TRG_TRY(check_launch());
    hipLaunchKernelGGL(maxabs_final, batch, 256, 0, s, (const double *)part, nchunk, c);
    return check_launch();
}

extern "C" int trg_trace(int dtype, int batch, int rank, const long long *dims, const long long *strides,
                         const void *T, long long sb, double *t, void *stream) {
    clear_hip_error();
    if (batch < 0 || rank < 1 || rank > 4 || !dims || !strides) return TRG_BAD_ARG;
    if (batch == 0) return TRG_OK;
    if (!T || !t) return TRG_BAD_ARG;
    TraceDesc d;
    d.rank = rank;
    for (int k = 0; k < 4; ++k) {
        d.dims[k] = k < rank ? dims[k] : 1;
        d.strides[k] = k < rank ? strides[k] : 0;
    }
    hipStream_t s = as_stream(stream);
    if (dtype == TRG_F32) hipLaunchKernelGGL(trace_kernel<float>, batch, 256, 0, s, d, (const float *)T, sb, t);
    else if (dtype == TRG_F64) hipLaunchKernelGGL(trace_kernel<double>, batch, 256, 0, s, d, (const double *)T, sb, t);
    else return TRG_UNSUPPORTED;
    return check_launch();
}

extern "C" int trg_normalize(int dtype, int batch, long long n, void *T, long long sb, const double *c_in,
                             double *lnf, long long lnf_stride, int step, int *step_ptr, void *stream) {
    clear_hip_error();
    if (batch < 0 || n < 0 || step < 0) return TRG_BAD_ARG;
    if (batch == 0) return TRG_OK;
    if (!c_in || (n > 0 && !T)) return TRG_BAD_ARG;
    if (batch > 65535) return TRG_SHAPE;
    hipStream_t s = as_stream(stream);
    if (lnf) {
        hipLaunchKernelGGL(lognorm_kernel, (unsigned)cdiv(batch, 256), 256, 0, s, batch, c_in, lnf, lnf_stride, step, step_ptr);
        TRG_TRY(check_launch());
    }
    if (n > 0) {
        const dim3 grid(grid_for(n, 256) > 4096 ? 4096 : grid_for(n, 256), batch);
        if (dtype == TRG_F32) hipLaunchKernelGGL(scale_by_inv_kernel<float>, grid, 256, 0, s, (float *)T, n, sb, c_in);
        else if (dtype == TRG_F64) hipLaunchKernelGGL(scale_by_inv_kernel<double>, grid, 256, 0, s, (double *)T, n, sb, c_in);
        else return TRG_UNSUPPORTED;
        TRG_TRY(check_launch());
    }
    if (step_ptr) {
        hipLaunchKernelGGL(bump_step_kernel, 1, 64, 0, s, step_ptr);
        TRG_TRY(check_launch());
    }
    return TRG_OK;
}

extern "C" int trg_scale_rows(int dtype, int batch, long long n, long long period, long long inner, void *Y,
                              long long sY, const void *x, long long sx, int mode, void *stream) {
    clear_hip_error();
    if (batch < 0 || n < 0 || period < 1 || inner < 1 || mode < 0 || mode > 3) return TRG_BAD_ARG;
    if (batch == 0 || n == 0) return TRG_OK;
    if (!Y || !x) return TRG_BAD_ARG;
    if (batch > 65535) return TRG_SHAPE;
    hipStream_t s = as_stream(stream);
    const dim3 grid(grid_for(n, 256) > 4096 ? 4096 : grid_for(n, 256), batch);
    if (dtype == TRG_F32) hipLaunchKernelGGL(scale_rows_kernel<float>, grid, 256, 0, s, (float *)Y, n, sY, period, inner, (const float *)x, sx, mode);
    else if (dtype == TRG_F64) hipLaunchKernelGGL(scale_rows_kernel<double>, grid, 256, 0, s, (double *)Y, n, sY, period, inner, (const double *)x, sx, mode);
    else return TRG_UNSUPPORTED;
    return check_launch();
}

extern "C" int trg_select_isometry(int dtype, int batch, long long nk, const double *eps0, const double *eps1,
                                   const double *V0, const double *V1, void *U, int *flag, void *stream) {
    clear_hip_error();
    if (batch < 0 || nk < 0) return TRG_BAD_ARG;
    if (batch == 0) return TRG_OK;
    if (!eps0 || !eps1 || !V0 || !V1 || !U) return TRG_BAD_ARG;
    if (batch > 65535) return TRG_SHAPE;
    hipStream_t s = as_stream(stream);
    const dim3 grid(grid_for(nk, 256) > 1024 ? 1024 : grid_for(nk, 256), batch);
    if (dtype == TRG_F32) hipLaunchKernelGGL(select_kernel<float>, grid, 256, 0, s, nk, eps0, eps1, V0, V1, (float *)U, flag);
    else if (dtype == TRG_F64) hipLaunchKernelGGL(select_kernel<double>, grid, 256, 0, s, nk, eps0, eps1, V0, V1, (double *)U, flag);
    else return TRG_UNSUPPORTED;
    return check_launch();
}

extern "C" int trg_gaussian(int dtype, int batch, long long n, unsigned long long seed, unsigned int stream_id,
                            unsigned int stream_stride, unsigned long long offset, void *out, long long sb,
                            void *stream) {
    clear_hip_error();
    if (batch < 0 || n < 0) return TRG_BAD_ARG;
    if (batch == 0 || n == 0) return TRG_OK;
    if (!out) return TRG_BAD_ARG;
    if (batch > 65535) return TRG_SHAPE;
    hipStream_t s = as_stream(stream);
    const dim3 grid(grid_for(n, 256) > 4096 ? 4096 : grid_for(n, 256), batch);
    if (dtype == TRG_F32) hipLaunchKernelGGL(gaussian_kernel<float>, grid, 256, 0, s, n, seed, stream_id, stream_stride, offset, (float *)out, sb);
    else if (dtype == TRG_F64) hipLaunchKernelGGL(gaussian_kernel<double>, grid, 256, 0, s, n, seed, stream_id, stream_stride, offset, (double *)out, sb);
    else return TRG_UNSUPPORTED;
    return check_launch();
}

extern "C" int trg_lnz(int batch, int nterms, const double *lnf, long long lnf_stride, const double *tr,
                       double inv_volume, double *out, void *stream) {
    clear_hip_error();
    if (batch < 0 || nterms < 0) return TRG_BAD_ARG;
    if (batch == 0) return TRG_OK;
    if (!lnf || !out) return TRG_BAD_ARG;
    hipLaunchKernelGGL(lnz_kernel, (unsigned)cdiv(batch, 256), 256, 0, as_stream(stream), batch, nterms, lnf,
                       lnf_stride, tr, inv_volume, out);
    return check_launch();
}

extern "C" int trg_health_or(long long n_info, const int *info, long long n_vals, const double *vals, int *health,
                             void *stream) {
    clear_hip_error();
    if (n_info < 0 || n_vals < 0) return TRG_BAD_ARG;
    if (!health || (n_info > 0 && !info) || (n_vals > 0 && !vals)) return TRG_BAD_ARG;
    if (n_info == 0 && n_vals == 0) return TRG_OK;
    hipLaunchKernelGGL(health_or_kernel, 1, 256, 0, as_stream(stream), n_info, info, n_vals, vals, health);
    return check_launch();
}
