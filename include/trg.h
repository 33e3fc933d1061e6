/*
 * libtrg -- MI355X (gfx950) kernels for tensor-renormalisation-group coarse
 * graining: HOTRG 2D/3D and ATRG 3D.  Plain C ABI: device pointers, sizes,
 * element strides and a hipStream_t passed as void*.  Every call is
 * asynchronous on the given stream and returns a trg_status.
 *
 * Each entry point names the reference operation it replaces (file:line in
 * LUOXIAO92/TRG_backup).  The reference has no native layer: it calls
 * cuquantum/opt_einsum contractions and cupy.linalg decompositions from
 * Python, so the "interface replaced" is that call site.
 */
#ifndef TRG_H
#define TRG_H

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
    TRG_OK = 0,
    TRG_BAD_ARG = 1,      /* null pointer, negative size, unsupported stride pattern */
    TRG_SHAPE = 2,        /* sizes beyond a kernel's supported maximum */
    TRG_OOM = 3,          /* workspace too small */
    TRG_NUMERIC = 4,      /* NaN/Inf, non-positive pivot */
    TRG_UNSUPPORTED = 5,  /* dtype / device not supported */
    TRG_HIP_ERROR = 6     /* a HIP runtime call failed (see trg_last_hip_error) */
} trg_status;

/* Storage types (F32, F64, BF16 for GEMM operands / products and copies;
 * I32 for trg_fill only) and GEMM compute types (F32, F64, BF16, BF16X3 =
 * fp32 operands split into two bf16 parts, three bf16 MFMAs per product,
 * about 16 significant bits per operand; BF16X6 = three bf16 parts, six
 * MFMAs per product, the 24 bits of fp32). */
typedef enum { TRG_F32 = 0, TRG_F64 = 1, TRG_BF16 = 2, TRG_I32 = 3, TRG_BF16X3 = 4, TRG_BF16X6 = 5 } trg_dtype;

const char *trg_version(void);
const char *trg_status_string(int status);
int trg_last_hip_error(void);
/* Device properties the host side uses for tiling decisions. */
int trg_device_info(int device, int *num_cus, int *lds_per_block, int *warp_size);

/*
 * Batched strided GEMM:  C[b] = alpha * A[b] B[b] + beta * C[b].
 * Element (m,k) of A[b] is at A + b*sAb + m*rsA + k*csA (elements, not
 * bytes); likewise B (k,n) and C (m,n).  One of rsA/csA and one of rsB/csB
 * must be 1.  `dtype` is the storage type of A, B and C (F32 or F64);
 * `compute` is F32, BF16, BF16X3 or BF16X6 for F32 storage (BF16: operands
 * rounded to bf16, fp32 accumulation; BF16X3 / BF16X6: split-bf16 products)
 * and F64 for F64 storage.
 * Replaces every pairwise step of the cuquantum/opt_einsum contractions on
 * the path: trg/HOTRG_2d.py:26-27,65-67; trg/HOTRG_3d.py:42-45,107-116;
 * trg/ATRG_3d_new.py:99-122,232-233,261-265,379-405,424-433.
 */
int trg_gemm(int dtype, int compute, int batch, int M, int N, int K, double alpha,
             const void *A, long long sAb, long long rsA, long long csA,
             const void *B, long long sBb, long long rsB, long long csB,
             double beta, void *C, long long sCb, long long rsC, long long csC,
             void *stream);

/*
 * trg_gemm with separate storage types for the operands (dtype, A and B) and
 * the product (c_dtype), an explicit tile configuration (0 <= config <
 * trg_gemm_config_count(dtype, c_dtype, compute); -1 = the shape heuristic)
 * and an optional two-level output index.  Storage combinations:
 * (F32, F32, F32|BF16|BF16X3|BF16X6), (F64, F64, F64), (BF16, F32, BF16) -- bf16
 * operands, both K-contiguous with K and strides multiples of 8 -- and
 * (F32, BF16, BF16), a product rounded to bf16 on the way out.  With cdm > 0 row m of
 * C[b] sits at (m / cdm) * csm + (m % cdm) * rsC, with cdn > 0 column n at
 * (n / cdn) * csn + (n % cdn) * csC (0 = the plain strides), so the product
 * lands directly in a permuted tensor layout.  All configurations of one
 * (dtype, compute) pair issue the same MFMA instruction over K in the same
 * order, so results are bitwise equal across configurations; only speed
 * differs.  Used by the autotuner that stands in for cuTENSOR's plan
 * selection behind the reference's cuquantum.contract calls
 * (trg/HOTRG_2d.py:26-27,65-67, trg/HOTRG_3d.py:42-45).
 */
int trg_gemm_ex(int config, int dtype, int c_dtype, int compute, int batch, int M, int N, int K,
                double alpha, const void *A, long long sAb, long long rsA, long long csA,
                const void *B, long long sBb, long long rsB, long long csB,
                double beta, void *C, long long sCb, long long rsC, long long csC,
                int cdm, long long csm, int cdn, long long csn, void *stream);
int trg_gemm_config_count(int dtype, int c_dtype, int compute);

/*
 * Float64 GEMM on the int8 matrix cores, same operands, strides and output
 * layout as trg_gemm_ex with dtype F64 (sAb / sBb = 0: one operand shared by
 * the batch).  Every row of A and column of B becomes 2^e times `slices`
 * (4..7) signed 7-bit integer slices; the slice products are summed exactly
 * in int32 on v_mfma_i32_32x32x32_i8 and the slices(slices+1)/2 products
 * whose weight reaches 2^-7(slices+1) are combined in fp64.  The product is
 * within about 2^-7 slices of (row max of |A|) x (column max of |B|) x K per
 * entry (slices = 6: 21 int8 MFMAs per product, 2^-42).  A non-finite entry
 * makes its row or column of C NaN.  The workspace holds the slices and the
 * row exponents, trg_gemm_f64i8_workspace bytes; K beyond the int32 range of
 * the sums (slices 127^2 K < 2^31) runs in chunks added in fp64.  The result
 * depends on nothing but the inputs and the K chunking: no tile configuration
 * or batch enters it.  Replaces the same contraction steps as trg_gemm in
 * its float64 form: trg/HOTRG_3d.py:107-116, trg/HOTRG_2d.py:65-67.
 */
long long trg_gemm_f64i8_workspace(int slices, int batchA, int batchB, int M, int N, int K);
int trg_gemm_f64i8(int slices, int batch, int M, int N, int K, double alpha,
                   const void *A, long long sAb, long long rsA, long long csA,
                   const void *B, long long sBb, long long rsB, long long csB,
                   double beta, void *C, long long sCb, long long rsC, long long csC,
                   int cdm, long long csm, int cdn, long long csn,
                   void *workspace, long long ws_bytes, void *stream);
const char *trg_gemm_config_name(int dtype, int c_dtype, int compute, int config);

/*
 * Strided copy with optional F32<->F64 or F32<->BF16 conversion (round to
 * nearest even) over a rank-`rank` index
 * space (rank <= 8):  out[sum i_k so[k]] = in[sum i_k si[k]].  Replaces the
 * leg permutations trg/HOTRG_2d.py:8-20, trg/HOTRG_3d.py:18-36,
 * trg/ATRG_3d_new.py:18-65 and the transpose-before-reshape of
 * utility/truncated_svd.py:117-118.
 */
int trg_copy_strided(int dtype_in, int dtype_out, int rank, const long long *dims,
                     const void *in, const long long *si, void *out, const long long *so,
                     void *stream);

/* out[i] = value for i < n (n elements of dtype: F32, F64 or I32). */
int trg_fill(int dtype, long long n, double value, void *out, void *stream);

/*
 * c_b = max |T_b| over n elements of each of `batch` tensors at stride sb,
 * written as float64 to c[b] (trg/HOTRG_2d.py:181-185).  Two passes:
 * `nchunk` workgroups per instance write partial maxima to part
 * (batch * nchunk doubles, caller-owned), one more pass reduces them.
 * NaN propagates.
 */
int trg_maxabs(int dtype, int batch, long long n, const void *T, long long sb, double *c,
               double *part, int nchunk, void *stream);

/*
 * Generalised trace: t_b = sum over (i_0..i_{r-1}) < dims of
 * T_b[sum i_k strides[k]] accumulated in float64 (r <= 4).  With strides
 * (s_x + s_x', s_y + s_y', ...) this is Tr T over paired legs:
 * "aabb" / "xxyytt" (trg/HOTRG_3d.py:234, measurement/Ising_model_3d.py:112).
 */
int trg_trace(int dtype, int batch, int rank, const long long *dims, const long long *strides,
              const void *T, long long sb, double *t, void *stream);

/*
 * Normalise and accumulate the log-norm:  for each b,
 *   c = |c_in[b]| ; T_b /= c ; lnf[b*lnf_stride + step] = ln(c) / 2^step.
 * `step_ptr` (optional, device int) overrides `step` and is incremented by
 * one (so a captured graph can be replayed across RG steps).
 * trg/HOTRG_2d.py:181-185,193,202,212; trg/HOTRG_3d.py:232-238,270;
 * trg/ATRG_3d_new.py:954-958,967-972.
 */
int trg_normalize(int dtype, int batch, long long n, void *T, long long sb, const double *c_in,
                  double *lnf, long long lnf_stride, int step, int *step_ptr, void *stream);

/*
 * Batched symmetric eigensolver (cyclic Jacobi, float64), n <= 256.
 * A: batch x (n x n) at stride sA (row-major, only symmetric input is
 * meaningful; A is overwritten).  Outputs eigenvalues ascending in w
 * (stride n) and eigenvectors as columns of V (row-major n x n, stride
 * n*n).  Replaces cupy.linalg.eigh (trg/HOTRG_2d.py:33-34,
 * trg/HOTRG_3d.py:55-58, utility/truncated_svd.py:188) for small n and the
 * Rayleigh-Ritz step of the large-n eigensolver.
 */
int trg_syevj_f64(int batch, int n, double *A, long long sA, double *w, double *V,
                  int max_sweeps, int *info, void *stream);

/*
 * Batched R^{-1} from a shifted Cholesky:  R^T R = G + shift_rel*tr(G)*I,
 * Rinv upper triangular (row-major l x l, stride l*l).  G (l x l, stride
 * sG) is overwritten.  l <= 256.  info[b] != 0 flags a non-positive pivot.
 * Building block of CholeskyQR2 (replaces cupy.linalg.qr in
 * trg/ATRG_3d_new.py:104,111,117 and utility/randomized_svd.py:59,79,84-91).
 */
int trg_chol_inv_f64(int batch, int l, double *G, long long sG, double shift_rel,
                     double *Rinv, int *info, void *stream);

/*
 * Gaussian sketch: out[b][i] = N(0,1) from Philox4x32-10 with key `seed`
 * and counter (offset + i, stream_id + b*stream_stride).  dtype F32 or F64.
 * Replaces cp.random.RandomState(seed).standard_normal
 * (trg/ATRG_3d_new.py:90-91, utility/randomized_svd.py:14,19).
 */
int trg_gaussian(int dtype, int batch, long long n, unsigned long long seed, unsigned int stream_id,
                 unsigned int stream_stride, unsigned long long offset, void *out, long long sb,
                 void *stream);

/*
 * Top-k eigenpairs of batch symmetric PSD n x n float64 matrices by
 * block subspace iteration (block l, q iterations, shift-regularised
 * CholeskyQR3) and Rayleigh-Ritz with trg_syevj_f64.  Outputs the k
 * largest eigenvalues ascending in w (stride k), their eigenvectors as the
 * columns of V (row-major n x k, stride n*k), and the discarded mass
 * eps[b] = tr(A_b) - sum(w_b).  Needs ws_bytes >= trg_eigh_topk_workspace.
 * Replaces cupy.linalg.eigh + the [:, D-Dcut:] slice + the discarded-sum
 * of trg/HOTRG_2d.py:33-50 and trg/HOTRG_3d.py:55-88.
 */
long long trg_eigh_topk_workspace(int batch, int n, int l);
int trg_eigh_topk_f64(int batch, int n, int k, int l, int q, const double *A, long long sA,
                      double *w, double *V, double *eps, unsigned long long seed,
                      void *ws, long long ws_bytes, int *info, void *stream);

/*
 * Allocation-free form of trg_eigh_topk_f64 (same kernels, same results bit
 * for bit).  No counterpart in the reference: it is for callers that cannot
 * allocate device memory -- code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc.  The workspace is a file-scope
 * device array of trg_eigh_topk_noalloc_max_bytes() bytes (64 MiB) that each
 * GPU holds until the process ends; a call whose trg_eigh_topk_workspace
 * exceeds it returns TRG_OOM.  All calls on one GPU share that array, so calls
 * that overlap on one GPU (different streams or threads) race.
 */
long long trg_eigh_topk_noalloc_max_bytes(void);
int trg_eigh_topk_f64_noalloc(int batch, int n, int k, int l, int q, const double *A, long long sA,
                              double *w, double *V, double *eps, unsigned long long seed, int *info,
                              void *stream);

/*
 * HOTRG isometry choice on the device: for each b pick side 0 if
 * eps0[b] < eps1[b] else side 1 (trg/HOTRG_2d.py:45, trg/HOTRG_3d.py:76,83),
 * copy the chosen n x k block into U (stride n*k) and write the flag
 * (0 = "LD"/"minus", 1 = "RU"/"plus") to flag[b].
 */
int trg_select_isometry(int dtype, int batch, long long nk, const double *eps0, const double *eps1,
                        const double *V0, const double *V1, void *U, int *flag, void *stream);

/*
 * Diagonal scaling of one leg:  Y_b[i] *= f(x_b[(i / inner) % period]) with
 * f = x, sqrt(max(x, 0)), 1/sqrt(x), 1/x for mode 0..3 (f32 or f64; mode 1
 * maps eigenvalues a rounding below zero to 0, NaN stays NaN).  Replaces the
 * s_i leg products of trg/ATRG_3d_new.py:180,188,278-279,285-286,451-452 and
 * tensor_class/tensor_class.py:66-86.
 */
int trg_scale_rows(int dtype, int batch, long long n, long long period, long long inner,
                   void *Y, long long sY, const void *x, long long sx, int mode, void *stream);

/*
 * Batched one-sided Jacobi SVD of m x n float64 matrices (n <= m <= 256,
 * m*(n|1)*8 + 16 KiB within the LDS of one workgroup).  A (stride sA,
 * row-major) is read only; s: singular values descending (stride n), U:
 * m x n (stride m*n), V: n x n (stride n*n, right singular vectors as
 * columns).  Replaces cupy.linalg.svd(full_matrices=False) of the small
 * factors in utility/truncated_svd.py:119, trg/ATRG_3d_new.py:124,416-427.
 */
int trg_gesvj_f64(int batch, int m, int n, const double *A, long long sA, double *s, double *U,
                  double *V, int max_sweeps, int *info, void *stream);

/*
 * Rank selection of utility/truncated_svd.py:17-78 (`truncate`) on a
 * descending vector s of length n: writes chi (int) and the tail-norm
 * error (double) to device memory.
 */
int trg_truncate_rank(int n, const double *s, int k, double degeneracy_eps, double truncate_eps,
                      int *chi, double *err, void *stream);

/*
 * Positive-semidefinite Cholesky of l x l blocks (l <= 256): R^T R = G with
 * rows of R zeroed where the pivot falls below psd_tol * scale[b] (scale
 * NULL: trace of G); R (optional) and a generalised inverse Rinv are
 * written row-major (stride l*l).  G is overwritten.  The diagonal-block
 * kernel of the blocked square roots R^T R = XdagX, Rx Rx^T = xxdag that
 * replace the full SVDs of trg/ATRG_3d_new.py:271-279 (any square root
 * gives the same squeezer).
 */
int trg_potrf_psd_f64(int batch, int l, double *G, long long sG, double psd_tol, const double *scale,
                      double *R, double *Rinv, void *stream);

/*
 * Orthonormal basis of the column span of batch m x l float64 matrices
 * (l <= 256, l <= m) by shifted CholeskyQR3 (passes = 3), or a
 * well-conditioned basis of the same span from its shifted first pass
 * (passes = 1, for the inner steps of a power iteration); X (stride sX) is
 * clobbered, Q is written with stride m*l.  Tall blocks with few output
 * tiles split the Gram over row slabs (count fixed by m and l) summed in
 * slab order.  Replaces cupy.linalg.qr in the range finders
 * (trg/ATRG_3d_new.py:104,111,117, utility/randomized_svd.py:59,84-91).
 */
long long trg_orthonormalize_workspace(int batch, int m, int l);
int trg_orthonormalize_f64(int batch, int m, int l, double *X, long long sX, double *Q, void *ws,
                           long long ws_bytes, int passes, int *info, void *stream);

/* Y[b][i] += alpha_rel * scal[b] * X[b][i] (float64, device-side scale). */
int trg_axpy_dev_f64(int batch, long long n, double alpha_rel, const double *scal, const double *X,
                     long long sX, double *Y, long long sY, void *stream);

/*
 * ln Z / V assembly on the device: out[b] = sum_{i<nterms} lnf[b*lnf_stride+i]
 * + ln|tr[b]| * inv_volume (tr may be NULL).  Replaces the callers'
 * cp.sum(ln_normfact) + cp.log(trace)/V (measurement/Ising_model_3d.py:137,
 * measurement/Potts3_model_3d.py:54, measurement/SU2_pcm_measurement.py:32-36).
 */
int trg_lnz(int batch, int nterms, const double *lnf, long long lnf_stride, const double *tr,
            double inv_volume, double *out, void *stream);

/*
 * Solve health word, written without a host sync (graph-capturable):
 * health[0] |= 1 if any of the n_info flags is non-zero (eigensolver,
 * Jacobi or Cholesky breakdown flags of the calls above) and |= 2 if any of
 * the n_vals normalisation constants is NaN, infinite or zero.  The host
 * reads the word once at the end of a solve.  Replaces the Python
 * exceptions a failed cupy.linalg.eigh / svd raises inside
 * trg/HOTRG_2d.py:33-34, trg/HOTRG_3d.py:55-58, trg/ATRG_3d_new.py:175-188,
 * and the division by a zero norm in trg/HOTRG_2d.py:181-185.
 */
int trg_health_or(long long n_info, const int *info, long long n_vals, const double *vals, int *health,
                  void *stream);

#ifdef __cplusplus
}
#endif
#endif /* TRG_H */
