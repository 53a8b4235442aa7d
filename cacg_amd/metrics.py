"""Algorithmic bytes and flops of the s-step Chebyshev (block) CG kernels.

These are the counts a kernel cannot avoid (each operand read once from HBM,
each result written once), used to turn measured durations into achieved
GB/s or TFLOP/s against MI355X's published peaks.  w_a / w_v are value bytes
of the matrix and of the basis vectors (8 for FP64); indices are int32.
"""

HBM_PEAK_GBPS = 8000.0          # MI355X HBM3E, published
FP64_PEAK_TFLOPS = 78.6         # MI355X dense FP64 (vector and MFMA), published


def basis_bytes(n, nnz, m, s, w_a=8, w_v=8):
    """One Chebyshev basis generation (s - 1 sparse products): step 1 reads
    S_0 and writes S_1, each later step reads S_{j-1}, S_{j-2} and writes S_j;
    the matrix (values + column ids + row pointers) is read once per step."""
    if s < 2:
        return 0
    return (s - 1) * (nnz * (w_a + 4) + 4 * (n + 1)) + n * m * w_v * (2 + 3 * (s - 2))


def aq_bytes(n, nnz, sm, w_a=8, w_v=8):
    """A Q with k = s m columns: matrix once, Q read once, A Q written once."""
    return nnz * (w_a + 4) + 4 * (n + 1) + 2 * n * sm * w_v


def gram_flops(n, kl, kr):
    return 2.0 * n * kl * kr


def gram_bytes(n, kl, kr, w=8):
    return n * (kl + kr) * w


def qupdate_flops(n, sm):
    return 2.0 * n * sm * sm


def qupdate_bytes(n, sm, w=8):
    return 3 * n * sm * w         # read S, read Q, write Q


def xrupdate_bytes(n, sm, m, w=8):
    return 2 * n * sm * w + 4 * n * m * w   # read Q, AQ; read+write X, R


# fast order with the X update deferred (cacg_k_qxupdate / cacg_k_xrupdate_split):
# the Q update also forms X += Q alpha, the X/R update only R -= AQ alpha
def qxupdate_flops(n, sm, m):
    return 2.0 * n * sm * sm + 2.0 * n * sm * m


def rupdate_bytes(n, sm, m, w=8):
    return n * sm * w + 2 * n * m * w       # read AQ; read+write R


def lu_flops(sm, k):
    return 2.0 * (2.0 / 3.0) * sm ** 3 + 2.0 * sm * sm * k


def iteration_flops(n, nnz, m, s, first=False):
    """Algorithmic flops of one outer iteration (the first skips Q^T A S,
    Q beta and the beta solve)."""
    sm = s * m
    f = 2.0 * nnz * m * (s - 1) + 2.0 * nnz * sm                 # basis + A Q
    f += 3.0 * n * m * max(s - 1, 0) + 4.0 * n * m               # recurrence + X/R updates
    f += gram_flops(n, sm, sm + m) + 2.0 * 2 * n * sm * m        # Q^T[AQ|R], Q alpha, AQ alpha
    f += lu_flops(sm, m)
    if not first:
        f += gram_flops(n, sm, sm) + qupdate_flops(n, sm) + n * sm + lu_flops(sm, sm)
    return f


def solve_flops(n, nnz, m, s, iters):
    if iters <= 0:
        return 0.0
    return iteration_flops(n, nnz, m, s, first=True) + (iters - 1) * iteration_flops(n, nnz, m, s)
