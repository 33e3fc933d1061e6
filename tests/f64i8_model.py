"""CPU model of the float64-on-int8 GEMM (trg_gemm_f64i8), for the tests.

Every row of A and column of B is 2^e times a value in (-1, 1), cut into S
signed 7-bit slices by q_i = trunc(128 r_{i-1}), r_i = 128 r_{i-1} - q_i.
The slice products of the diagonals d = i + j <= S + 1 are integer sums,
here in int64 (exact); they are combined in float64, smallest diagonal
weight first, then scaled by 2^(e_m + f_n) and alpha.  With beta = 0 and K
inside one chunk this is the device computation operation for operation, so
the two agree bit for bit.
"""
import numpy as np

EXP_NONFINITE = 1 << 20


def k_chunk(S):
    return ((2 ** 31 - 1) // (S * 127 * 127)) // 64 * 64


def slices(X, S):
    """X (R, K) float64 -> ([q_1 .. q_S] int64 (R, K) each, e (R,) int)."""
    R, K = X.shape
    fin = np.all(np.isfinite(X), axis=1) if K else np.ones(R, bool)
    mx = np.max(np.abs(np.where(np.isfinite(X), X, 0.0)), axis=1) if K else np.zeros(R)
    _, e = np.frexp(mx)
    e = np.where(mx > 0, e, 0).astype(np.int64)
    e = np.where(fin, e, EXP_NONFINITE)
    r = np.ldexp(np.where(fin[:, None], X, 0.0), np.where(fin, -e, 0)[:, None])
    out = []
    for _ in range(S):
        t = r * 128.0
        q = np.trunc(t)
        r = t - q
        out.append(q.astype(np.int64))
    return out, e


def gemm_f64i8(A, B, S=6, alpha=1.0):
    """alpha A B for A (M, K), B (K, N) float64, one K chunk."""
    M, K = A.shape
    N = B.shape[1]
    assert K <= k_chunk(S)
    qa, ea = slices(A, S)
    qb, eb = slices(np.ascontiguousarray(B.T), S)
    P = [np.zeros((M, N), np.int64) for _ in range(S)]
    for i in range(S):
        for j in range(S - i):
            P[i + j] += qa[i] @ qb[j].T
    v = np.zeros((M, N))
    for d in reversed(range(S)):
        v = v + np.ldexp(P[d].astype(np.float64), -7 * (d + 2))
    bad = (ea[:, None] >= EXP_NONFINITE) | (eb[None, :] >= EXP_NONFINITE)
    sc = np.where(bad, 0, ea[:, None] + eb[None, :])
    v = np.where(bad, np.nan, np.ldexp(v, sc))
    return alpha * v
