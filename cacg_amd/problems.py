"""Seeded synthetic inputs for the s-step Chebyshev (block) CG path.

The reference reads a Matrix Market file and a CSV of right-hand sides
(src/readMTX.c:8-136, src/GetRHS.c:124-208) and seeds nothing
(src/GetRHS.c:74 calls srand(time(NULL))).  The workloads here are generated
instead, reproducibly from a seed:

* 2D 5-point and 3D 7-point Dirichlet Laplacians in natural lexicographic
  order (x fastest), diagonal 4 / 6, off-diagonal -1, columns ascending in
  each row.  nnz = 5L^2 - 4L (2D) and 7N^3 - 6N^2 (3D).
* A random SPD variant with the same pattern: off-diagonals -U(0.5, 1.5),
  drawn once per (i < j) edge in CSR order of the upper triangle and
  mirrored; diagonal = 1.01 x absolute off-diagonal row sum.
* Dense right-hand-side blocks, n x m row-major (src/GetRHS.c:9 layout),
  filled row by row.

Random numbers come from SplitMix64 (Steele, Lea, Flood, OOPSLA 2014):
output i (0-based) of the stream with seed `seed` is
mix(seed + (i + 1) * 0x9E3779B97F4A7C15), where mix is SplitMix64's
finaliser.  A uniform double is (u >> 11) * 2^-53 in [0, 1); a normal pair is
Box-Muller on (1 - U_2k, U_2k+1).  The stream is counter-based, so it is
vectorised here with numpy and could be generated in parallel anywhere.
"""

from __future__ import annotations

import dataclasses

import numpy as np

_GOLDEN = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


def splitmix64(seed: int, count: int, offset: int = 0) -> np.ndarray:
    """Outputs offset .. offset+count-1 of the SplitMix64 stream for `seed`."""
    with np.errstate(over="ignore"):
        idx = np.arange(offset + 1, offset + count + 1, dtype=np.uint64)
        z = np.uint64(seed & 0xFFFFFFFFFFFFFFFF) + idx * _GOLDEN
        z = (z ^ (z >> np.uint64(30))) * _M1
        z = (z ^ (z >> np.uint64(27))) * _M2
        z = z ^ (z >> np.uint64(31))
    return z


def uniform(seed: int, count: int, offset: int = 0) -> np.ndarray:
    """`count` doubles uniform in [0, 1)."""
    u = splitmix64(seed, count, offset) >> np.uint64(11)
    return u.astype(np.float64) * (1.0 / 9007199254740992.0)


def normal(seed: int, count: int) -> np.ndarray:
    """`count` standard normal doubles (Box-Muller on consecutive pairs)."""
    npair = (count + 1) // 2
    u = uniform(seed, 2 * npair)
    u1 = 1.0 - u[0::2]          # (0, 1], safe for log
    u2 = u[1::2]
    r = np.sqrt(-2.0 * np.log(u1))
    t = 2.0 * np.pi * u2
    out = np.empty(2 * npair)
    out[0::2] = r * np.cos(t)
    out[1::2] = r * np.sin(t)
    return out[:count]


@dataclasses.dataclass
class CSR:
    """A square CSR matrix with int64 row pointers and column indices.

    The field meaning follows csrType_local (src/DataTypes.h:22-42):
    rowptr = row_start (n + 1 entries), col = col_idx, val = csrdata.
    """
    n: int
    rowptr: np.ndarray   # int64, n + 1
    col: np.ndarray      # int64, nnz
    val: np.ndarray      # float64, nnz

    @property
    def nnz(self) -> int:
        return int(self.rowptr[-1])

    def to_dense(self) -> np.ndarray:
        a = np.zeros((self.n, self.n))
        rows = np.repeat(np.arange(self.n), np.diff(self.rowptr))
        np.add.at(a, (rows, self.col), self.val)
        return a


def _stencil_csr(dims: tuple[int, ...], diag: float) -> CSR:
    """(2d+1)-point Dirichlet Laplacian on a grid of shape dims (x fastest)."""
    shape = tuple(int(d) for d in dims)
    n = int(np.prod(shape))
    coords = np.indices(shape[::-1]).reshape(len(shape), -1)[::-1]  # coords[0] = x
    strides = [int(np.prod(shape[:a])) for a in range(len(shape))]
    idx = np.arange(n, dtype=np.int64)
    # neighbour offsets in ascending column order: -stride_{d-1} .. -1, 0, +1 .. +stride_{d-1}
    offs = []
    for a in reversed(range(len(shape))):
        offs.append((a, -1))
    offs.append((None, 0))
    for a in range(len(shape)):
        offs.append((a, +1))
    cols, vals, valid = [], [], []
    for axis, sgn in offs:
        if axis is None:
            cols.append(idx)
            vals.append(np.full(n, diag))
            valid.append(np.ones(n, dtype=bool))
        else:
            c = coords[axis] + sgn
            ok = (c >= 0) & (c < shape[axis])
            cols.append(idx + sgn * strides[axis])
            vals.append(np.full(n, -1.0))
            valid.append(ok)
    cols = np.stack(cols, axis=1)
    vals = np.stack(vals, axis=1)
    valid = np.stack(valid, axis=1)
    counts = valid.sum(axis=1)
    rowptr = np.zeros(n + 1, dtype=np.int64)
    np.cumsum(counts, out=rowptr[1:])
    return CSR(n, rowptr, cols[valid].astype(np.int64), vals[valid].astype(np.float64))


def laplacian_2d(L: int, shift: float = 0.0) -> CSR:
    """5-point Laplacian on an L x L grid, plus `shift` on the diagonal."""
    return _stencil_csr((L, L), 4.0 + shift)


def laplacian_3d(N: int, shift: float = 0.0) -> CSR:
    """7-point Laplacian on an N x N x N grid, plus `shift` on the diagonal."""
    return _stencil_csr((N, N, N), 6.0 + shift)


def random_spd_like(A: CSR, seed: int) -> CSR:
    """Same pattern as A; off-diagonals -U(0.5,1.5) mirrored; diag 1.01*sum|off|."""
    n = A.n
    rows = np.repeat(np.arange(n, dtype=np.int64), np.diff(A.rowptr))
    upper = A.col > rows
    nup = int(upper.sum())
    draws = 0.5 + uniform(seed, nup)           # CSR order of the upper triangle
    val = np.zeros_like(A.val)
    val[upper] = -draws
    # mirror: entry (j, i) of the lower triangle gets the draw of (i, j)
    key_up = rows[upper] * n + A.col[upper]
    lower = A.col < rows
    key_lo = A.col[lower] * n + rows[lower]
    order = np.argsort(key_up)
    pos = np.searchsorted(key_up[order], key_lo)
    if not np.array_equal(key_up[order][pos], key_lo):
        raise ValueError("random_spd_like: pattern is not structurally symmetric")
    val[lower] = -draws[order][pos]
    diag = rows == A.col
    offsum = np.zeros(n)
    np.add.at(offsum, rows[~diag], np.abs(val[~diag]))
    val[diag] = 1.01 * offsum[rows[diag]]
    return CSR(n, A.rowptr.copy(), A.col.copy(), val)


def power_law_spd(L: int, hubs: int, dmax: int, seed: int, alpha: float = 1.0) -> CSR:
    """An irregular SPD matrix: the 5-point pattern of an L x L grid plus `hubs`
    hub vertices, hub h (vertex (h * n) // hubs) joined to d_h = max(1,
    floor(dmax / (h + 1)^alpha)) other vertices, so the row lengths follow a
    power law from dmax + 5 down to the grid's 5.  Hub neighbours: stream
    outputs of SplitMix64(seed) taken mod n in order (duplicates and the hub
    itself dropped).  Values as random_spd_like: off-diagonals -U(0.5, 1.5)
    drawn per upper-triangle edge in CSR order with seed + 1 and mirrored,
    diagonal 1.01 x the absolute off-diagonal row sum (strictly diagonally
    dominant, so SPD)."""
    G = laplacian_2d(L)
    n = G.n
    rows = np.repeat(np.arange(n, dtype=np.int64), np.diff(G.rowptr))
    up = G.col > rows
    keys = [rows[up] * n + G.col[up]]
    off = 0
    for h in range(hubs):
        v = (h * n) // hubs
        d = max(1, int(dmax / (h + 1) ** alpha))
        nb = (splitmix64(seed, d, off) % np.uint64(n)).astype(np.int64)
        off += d
        nb = nb[nb != v]
        lo, hi = np.minimum(nb, v), np.maximum(nb, v)
        keys.append(lo * n + hi)
    ukeys = np.unique(np.concatenate(keys))            # upper edges, CSR order
    draws = 0.5 + uniform(seed + 1, len(ukeys))
    ui, uj = ukeys // n, ukeys % n
    # both triangles, then the diagonal
    r = np.concatenate([ui, uj, np.arange(n, dtype=np.int64)])
    c = np.concatenate([uj, ui, np.arange(n, dtype=np.int64)])
    v = np.concatenate([-draws, -draws, np.zeros(n)])
    order = np.lexsort((c, r))
    r, c, v = r[order], c[order], v[order]
    offsum = np.zeros(n)
    np.add.at(offsum, r[r != c], np.abs(v[r != c]))
    v[r == c] = 1.01 * offsum[r[r == c]]
    rowptr = np.zeros(n + 1, dtype=np.int64)
    np.cumsum(np.bincount(r, minlength=n), out=rowptr[1:])
    return CSR(n, rowptr, c, v)


def rhs_uniform(n: int, m: int, seed: int) -> np.ndarray:
    """n x m row-major block, entries U[0,1), filled row by row."""
    return uniform(seed, n * m).reshape(n, m)


def rhs_normal(n: int, m: int, seed: int) -> np.ndarray:
    """n x m row-major block, entries N(0,1), filled row by row."""
    return normal(seed, n * m).reshape(n, m)


def sample_batch_systems(count: int, seed: int, lmin: int = 32, lmax: int = 256,
                         ms=(1, 2, 4, 8, 16, 32)):
    """Draw (L, shift, m) for a batch of shifted 2D Laplacians.

    L is log-uniform in [lmin, lmax], shift ~ U(0, 0.1), m uniform over `ms`;
    three consecutive stream outputs per system.
    """
    u = uniform(seed, 3 * count).reshape(count, 3)
    L = np.floor(np.exp(np.log(lmin) + u[:, 0] * (np.log(lmax + 1) - np.log(lmin)))).astype(int)
    L = np.clip(L, lmin, lmax)
    shift = 0.1 * u[:, 1]
    m = np.asarray(ms)[np.minimum((u[:, 2] * len(ms)).astype(int), len(ms) - 1)]
    return [(int(a), float(b), int(c)) for a, b, c in zip(L, shift, m)]
