"""Seeded synthetic problem generators for the sparse solve path.

Every matrix comes back as CSR arrays with the reference's layout: uint32
``row_ptr``/``col_idx`` and float64 ``values``, columns ascending within each
row, exactly what the reference's COO constructor produces
(ode-solver/src/math/array.cpp:2072-2182).

Generators:
  * ``poisson2d(N)`` – 5-point Laplacian, diagonal 4, off-diagonals -1,
    Dirichlet-eliminated; the matrix of ode-solver/test/test_SMatrix_pcg.cpp:66-102.
  * ``poisson3d(N)`` – the 7-point analogue (diagonal 6).
  * ``convdiff2d(N, peclet)`` – first-order upwind convection-diffusion with
    velocity ``peclet*(cos 2*pi*y, sin 2*pi*x)`` at cell centres; reduces to
    ode-solver/test/test_SMatrix_bicgstab.cpp:78-106 for positive velocity.
  * ``vardiff3d(N, seed)`` – 7-point diffusion with coefficient ``exp(g)``, g a
    Gaussian random field (sigma 1, Gaussian covariance with correlation length
    8 cells); face coefficients are harmonic means, boundary faces use the
    cell's own coefficient.
  * ``block_test_scalar(N)`` – node-major 3-component test matrix of
    ode-solver/test/test_BlockMatrix_mpi.cpp:29-59.
  * ``bsr7(t, seed)`` – SPD block-CSR (3x3 blocks) 7-point coupling on a t^3
    grid: symmetric off-diagonal blocks U[-1,1], diagonal blocks shifted to
    strict diagonal dominance.
  * ``batch_c5(seed)`` – the 512-system batch (256 Poisson-2D, 256 BSR).
  * ``rhs(A, seed)`` – x_true ~ U[-1,1] (numpy PCG64 seeded), b = A x_true.
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np


@dataclass
class CSR:
    n: int
    m: int
    row_ptr: np.ndarray  # uint32 [n+1]
    col_idx: np.ndarray  # uint32 [nnz]
    values: np.ndarray   # float64 [nnz]

    @property
    def nnz(self) -> int:
        return int(self.values.size)

    def matvec(self, x: np.ndarray) -> np.ndarray:
        """y = A x with a per-row sequential sum in stored order (x may hold
        several vectors as columns)."""
        return self.to_scipy() @ x

    def to_scipy(self):
        import scipy.sparse as sp
        return sp.csr_matrix((self.values, self.col_idx.astype(np.int64), self.row_ptr.astype(np.int64)),
                             shape=(self.n, self.m))


def _assemble(n: int, rows: np.ndarray, cols: np.ndarray, vals: np.ndarray, m: int | None = None) -> CSR:
    """Assemble CSR from entries whose (row, col) pairs are unique; sorts by
    (row, col) and drops exact zeros like the reference COO constructor."""
    m = n if m is None else m
    keep = vals != 0.0
    rows, cols, vals = rows[keep], cols[keep], vals[keep]
    order = np.lexsort((cols, rows))
    rows, cols, vals = rows[order], cols[order], vals[order]
    counts = np.bincount(rows, minlength=n)
    rp = np.zeros(n + 1, dtype=np.uint64)
    np.cumsum(counts, out=rp[1:])
    if rp[-1] >= 2**32:
        raise ValueError("nnz exceeds the uint32 index range")
    return CSR(n, m, rp.astype(np.uint32), cols.astype(np.uint32), vals.astype(np.float64))


def _stencil(shape, diag, offs):
    """Grid stencil: ``offs`` = list of (axis, step, value-array-or-scalar)."""
    dims = len(shape)
    n = int(np.prod(shape))
    idx = np.arange(n, dtype=np.int64).reshape(shape)
    rows = [idx.ravel()]
    cols = [idx.ravel()]
    vals = [np.broadcast_to(np.asarray(diag, dtype=np.float64), shape).ravel()]
    for axis, step, val in offs:
        sl_src = [slice(None)] * dims
        sl_dst = [slice(None)] * dims
        if step > 0:
            sl_src[axis] = slice(0, shape[axis] - step)
            sl_dst[axis] = slice(step, None)
        else:
            sl_src[axis] = slice(-step, None)
            sl_dst[axis] = slice(0, shape[axis] + step)
        v = np.broadcast_to(np.asarray(val, dtype=np.float64), shape)
        rows.append(idx[tuple(sl_src)].ravel())
        cols.append(idx[tuple(sl_dst)].ravel())
        vals.append(v[tuple(sl_src)].ravel())
    return _assemble(n, np.concatenate(rows), np.concatenate(cols), np.concatenate(vals))


def poisson2d(N: int) -> CSR:
    """5-point Poisson on an N x N interior grid, row index i*N + j."""
    return _stencil((N, N), 4.0, [(0, -1, -1.0), (0, 1, -1.0), (1, -1, -1.0), (1, 1, -1.0)])


def poisson3d(N: int) -> CSR:
    """7-point Poisson on an N^3 interior grid, row index (k*N + i)*N + j."""
    offs = [(a, s, -1.0) for a in range(3) for s in (-1, 1)]
    return _stencil((N, N, N), 6.0, offs)


def aniso2d(N: int, eps: float = 0.01) -> CSR:
    """Anisotropic 5-point Laplacian -u_xx - eps u_yy on an N x N interior
    grid, row index i*N + j (x along j): diagonal 2 + 2 eps, x neighbours -1,
    y neighbours -eps. Interior rows sum to zero; for eps below an AMG
    strength threshold the y couplings are weak."""
    return _stencil((N, N), 2.0 + 2.0 * eps, [(0, -1, -eps), (0, 1, -eps), (1, -1, -1.0), (1, 1, -1.0)])


def convdiff2d(N: int, peclet: float = 10.0) -> CSR:
    """Upwind convection-diffusion on N x N cells, D = 1, h = 1.

    Row index i*N + j with y = (i+1/2)/N, x = (j+1/2)/N and velocity
    v = peclet * (cos 2 pi y, sin 2 pi x).
    diag = 4 + |vx| + |vy|; west -1 - max(vx,0); east -1 - max(-vx,0);
    south -1 - max(vy,0); north -1 - max(-vy,0).
    """
    i = (np.arange(N) + 0.5) / N
    y, x = np.meshgrid(i, i, indexing="ij")
    vx = peclet * np.cos(2 * np.pi * y)
    vy = peclet * np.sin(2 * np.pi * x)
    diag = 4.0 + np.abs(vx) + np.abs(vy)
    west = -1.0 - np.maximum(vx, 0.0)
    east = -1.0 - np.maximum(-vx, 0.0)
    south = -1.0 - np.maximum(vy, 0.0)
    north = -1.0 - np.maximum(-vy, 0.0)
    return _stencil((N, N), diag, [(0, -1, south), (0, 1, north), (1, -1, west), (1, 1, east)])


def gaussian_random_field(N: int, seed: int, sigma: float = 1.0, ell: float = 8.0) -> np.ndarray:
    """Periodic Gaussian random field on N^3 cells with covariance
    sigma^2 exp(-r^2 / (2 ell^2)): white noise (PCG64, ``seed``) convolved in
    Fourier space with a Gaussian of standard deviation ell/sqrt(2), then
    normalised to zero mean and standard deviation sigma."""
    rng = np.random.default_rng(seed)
    w = rng.standard_normal((N, N, N))
    k = np.fft.fftfreq(N) * 2 * np.pi
    kx, ky, kz = np.meshgrid(k, k, k, indexing="ij")
    s = ell / math.sqrt(2.0)
    filt = np.exp(-0.5 * s * s * (kx * kx + ky * ky + kz * kz))
    g = np.real(np.fft.ifftn(np.fft.fftn(w) * filt))
    g -= g.mean()
    g *= sigma / g.std()
    return g


def vardiff3d(N: int, seed: int, sigma: float = 1.0, ell: float = 8.0) -> tuple[CSR, np.ndarray]:
    """Variable-coefficient 7-point diffusion on N^3 cells; returns (A, kappa).

    kappa = exp(g). Interior face coefficient = harmonic mean of the two cells;
    a boundary face uses the cell's own kappa (Dirichlet at the face).
    Row index (k*N + i)*N + j.
    """
    kappa = np.exp(gaussian_random_field(N, seed, sigma, ell))
    diag = np.zeros_like(kappa)
    offs = []
    for a in range(3):
        lo = [slice(None)] * 3
        hi = [slice(None)] * 3
        lo[a] = slice(0, N - 1)
        hi[a] = slice(1, N)
        k1, k2 = kappa[tuple(lo)], kappa[tuple(hi)]
        face = 2.0 * k1 * k2 / (k1 + k2)
        # coefficient array indexed at the source cell of each direction
        plus = np.zeros_like(kappa)
        minus = np.zeros_like(kappa)
        plus[tuple(lo)] = face
        minus[tuple(hi)] = face
        # boundary faces
        b0 = [slice(None)] * 3
        b1 = [slice(None)] * 3
        b0[a] = 0
        b1[a] = N - 1
        bnd = np.zeros_like(kappa)
        bnd[tuple(b0)] += kappa[tuple(b0)]
        bnd[tuple(b1)] += kappa[tuple(b1)]
        diag += plus + minus + bnd
        offs.append((a, 1, -plus))
        offs.append((a, -1, -minus))
    return _stencil((N, N, N), diag, offs), kappa


def drand(i: np.ndarray | int) -> np.ndarray:
    """Deterministic pseudo-random of the reference tests:
    frac(sin(0.371 i + 1.7) * 43758.5453) (test_BlockMatrix_mpi.cpp:22-23)."""
    i = np.asarray(i, dtype=np.float64)
    s = np.sin(i * 0.371 + 1.7) * 43758.5453
    return s - np.floor(s)


def block_test_scalar(N: int, BS: int = 3) -> CSR:
    """Node-major BS-component test matrix on an N x N node grid
    (test_BlockMatrix_mpi.cpp:29-59): diagonal block (4+BS) I with 0.1
    off-diagonal couplings, neighbour blocks -I."""
    nodes = N * N
    n = nodes * BS
    gi, gj = np.meshgrid(np.arange(N), np.arange(N), indexing="ij")
    idx = (gi * N + gj).ravel()
    rows, cols, vals = [], [], []
    for r in range(BS):
        row = idx * BS + r
        for c in range(BS):
            rows.append(row)
            cols.append(idx * BS + c)
            vals.append(np.full(nodes, 4.0 + BS if r == c else 0.1))
        for cond, d in (((gj > 0).ravel(), -1), ((gj < N - 1).ravel(), 1), ((gi > 0).ravel(), -N), ((gi < N - 1).ravel(), N)):
            rows.append(row[cond])
            cols.append((idx[cond] + d) * BS + r)
            vals.append(np.full(int(cond.sum()), -1.0))
    return _assemble(n, np.concatenate(rows), np.concatenate(cols), np.concatenate(vals))


def stokes3d(N: int, alpha: float = 0.1) -> dict:
    """Periodic 3-D MAC Stokes test system of test_BlockMatrix_schur3d.cpp:39-129
    on an N^3 grid (point (i, j, k) -> i + j N + k N^2): velocity block A =
    three periodic 7-point Laplacians plus alpha I (diagonal 6 + alpha),
    divergence B (pressure p, velocity component c: +1 at p, -1 at the point
    before it along c), B^T, the pressure Laplacian Lp (diagonal 6 + 1e-8)
    and K = [[A, B^T], [B, 0]]. Repeated (row, column) pairs are summed and
    zeros dropped, as the reference's COO constructor does. Returns CSR
    matrices A, B, Bt, Lp, K and the sizes nv, np."""
    import scipy.sparse as sp
    N3 = N ** 3
    nv, npres = 3 * N3, N3
    i, j, k = np.meshgrid(np.arange(N), np.arange(N), np.arange(N), indexing="ij")
    i, j, k = i.ravel(), j.ravel(), k.ravel()
    p = i + j * N + k * N * N
    nbrs = [((i + 1) % N, j, k), ((i - 1) % N, j, k), (i, (j + 1) % N, k), (i, (j - 1) % N, k),
            (i, j, (k + 1) % N), (i, j, (k - 1) % N)]
    nidx = [a + b * N + c * N * N for a, b, c in nbrs]

    def lap(n_off, diag):
        rows = [p + n_off] + [p + n_off for _ in nidx]
        cols = [p + n_off] + [q + n_off for q in nidx]
        vals = [np.full(N3, diag)] + [np.full(N3, -1.0) for _ in nidx]
        return np.concatenate(rows), np.concatenate(cols), np.concatenate(vals)

    rA, cA, vA = zip(*[lap(c * N3, 6.0 + alpha) for c in range(3)])
    rA, cA, vA = np.concatenate(rA), np.concatenate(cA), np.concatenate(vA)
    rL, cL, vL = lap(0, 6.0 + 1e-8)
    prev = [((i - 1) % N) + j * N + k * N * N, i + ((j - 1) % N) * N + k * N * N, i + j * N + ((k - 1) % N) * N * N]
    rB = np.concatenate([np.concatenate([p, p]) for _ in range(3)])
    cB = np.concatenate([np.concatenate([c * N3 + p, c * N3 + prev[c]]) for c in range(3)])
    vB = np.concatenate([np.concatenate([np.ones(N3), -np.ones(N3)]) for _ in range(3)])

    def csr(r, c, v, shape):
        M = sp.coo_matrix((v, (r, c)), shape=shape).tocsr()
        M.sum_duplicates()
        M.eliminate_zeros()
        M.sort_indices()
        return CSR(shape[0], shape[1], M.indptr.astype(np.uint32), M.indices.astype(np.uint32), M.data.copy())

    A = csr(rA, cA, vA, (nv, nv))
    B = csr(rB, cB, vB, (npres, nv))
    Bt = csr(cB, rB, vB, (nv, npres))
    Lp = csr(rL, cL, vL, (npres, npres))
    K = csr(np.concatenate([rA, rB + nv, cB]), np.concatenate([cA, cB, rB + nv]), np.concatenate([vA, vB, vB]),
            (nv + npres, nv + npres))
    return dict(A=A, B=B, Bt=Bt, Lp=Lp, K=K, nv=nv, np=npres)


@dataclass
class BSR:
    bs: int
    nbr: int
    nbc: int
    row_ptr: np.ndarray  # uint32 [nbr+1]
    col_idx: np.ndarray  # uint32 [nnzb]
    values: np.ndarray   # float64 [nnzb*bs*bs], row-major blocks

    @property
    def nnzb(self) -> int:
        return int(self.col_idx.size)

    def to_scipy(self):
        import scipy.sparse as sp
        bs = self.bs
        return sp.bsr_matrix((self.values.reshape(-1, bs, bs), self.col_idx.astype(np.int64),
                              self.row_ptr.astype(np.int64)), shape=(self.nbr * bs, self.nbc * bs))

    def to_csr(self) -> CSR:
        """Flatten to scalar CSR in node-major numbering (block row b, row r
        -> scalar row b*bs + r)."""
        bs = self.bs
        rp = self.row_ptr.astype(np.int64)
        nb = self.nbr
        cnt = np.diff(rp)
        brow = np.repeat(np.arange(nb), cnt)
        blocks = self.values.reshape(-1, bs, bs)
        rows = (brow[:, None, None] * bs + np.arange(bs)[None, :, None]) + 0 * np.arange(bs)[None, None, :]
        cols = self.col_idx.astype(np.int64)[:, None, None] * bs + np.arange(bs)[None, None, :] + 0 * np.arange(bs)[None, :, None]
        return _assemble(nb * bs, rows.ravel(), cols.ravel(), blocks.ravel().copy(), m=self.nbc * bs)


def bsr7(t: int, seed: int, bs: int = 3) -> BSR:
    """SPD block 7-point system on a t^3 node grid with bs x bs blocks.

    Off-diagonal blocks: for each grid edge (i < j) B ~ U[-1,1]^{bs x bs},
    A_ij = B, A_ji = B^T. Diagonal block: symmetric U[-1,1] part plus a shift
    making every scalar row strictly diagonally dominant (margin 1), so A is
    symmetric with positive diagonal and strictly diagonally dominant: SPD.
    """
    rng = np.random.default_rng(seed)
    nb = t ** 3
    idx = np.arange(nb, dtype=np.int64).reshape(t, t, t)
    src, dst = [], []
    for a in range(3):
        lo = [slice(None)] * 3
        hi = [slice(None)] * 3
        lo[a] = slice(0, t - 1)
        hi[a] = slice(1, t)
        src.append(idx[tuple(lo)].ravel())
        dst.append(idx[tuple(hi)].ravel())
    src = np.concatenate(src)
    dst = np.concatenate(dst)
    ne = src.size
    B = rng.uniform(-1.0, 1.0, size=(ne, bs, bs))
    R = rng.uniform(-1.0, 1.0, size=(nb, bs, bs))
    D = 0.5 * (R + R.transpose(0, 2, 1))
    rowabs = np.zeros((nb, bs))
    np.add.at(rowabs, src, np.abs(B).sum(axis=2))
    np.add.at(rowabs, dst, np.abs(B).sum(axis=1))
    off_d = np.abs(D).sum(axis=2) - np.abs(np.einsum("bii->bi", D))
    need = rowabs + off_d + 1.0 - np.einsum("bii->bi", D)
    shift = need.max(axis=1)
    D = D + shift[:, None, None] * np.eye(bs)[None]
    brows = np.concatenate([np.arange(nb), src, dst])
    bcols = np.concatenate([np.arange(nb), dst, src])
    blks = np.concatenate([D, B, B.transpose(0, 2, 1)])
    order = np.lexsort((bcols, brows))
    brows, bcols, blks = brows[order], bcols[order], blks[order]
    rp = np.zeros(nb + 1, dtype=np.uint32)
    np.cumsum(np.bincount(brows, minlength=nb), out=rp[1:])
    return BSR(bs, nb, nb, rp, bcols.astype(np.uint32), np.ascontiguousarray(blks).ravel())


def rhs(A: CSR, seed: int, nrhs: int | None = None) -> tuple[np.ndarray, np.ndarray]:
    """x_true ~ U[-1,1] from PCG64(seed); b = A x_true in float64."""
    rng = np.random.default_rng(seed)
    if nrhs is None:
        x = rng.uniform(-1.0, 1.0, A.n)
        return x, A.matvec(x)
    X = rng.uniform(-1.0, 1.0, (nrhs, A.n))
    return X, np.ascontiguousarray((A.to_scipy() @ X.T).T)


def batch_c5(seed: int = 0, count: int = 512, poisson_sides=(32, 256), bsr_sides=(16, 48), nrhs: int = 4):
    """The batch of independent SPD systems: the first half 2-D Poisson with
    side ~ U{32..256}, the second half block-CSR (3x3 blocks, 7-point) with
    side ~ U{16..48}. Returns a list of dicts with keys kind, side, A (CSR or
    BSR), X (true solutions, nrhs x n), B (right-hand sides)."""
    rng = np.random.default_rng(seed)
    half = count // 2
    ps = rng.integers(poisson_sides[0], poisson_sides[1] + 1, size=half)
    ts = rng.integers(bsr_sides[0], bsr_sides[1] + 1, size=count - half)
    out = []
    for k, s in enumerate(ps):
        A = poisson2d(int(s))
        X, B = rhs(A, seed * 100003 + k, nrhs)
        out.append(dict(kind="poisson2d", side=int(s), A=A, X=X, B=B))
    for k, t in enumerate(ts):
        Ab = bsr7(int(t), seed * 100003 + half + k)
        X = np.random.default_rng(seed * 100003 + half + k + 7).uniform(-1.0, 1.0, (nrhs, Ab.nbr * Ab.bs))
        B = np.ascontiguousarray((Ab.to_scipy() @ X.T).T)
        out.append(dict(kind="bsr3", side=int(t), A=Ab, X=X, B=B))
    return out
