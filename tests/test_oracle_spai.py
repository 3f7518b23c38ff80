"""CPU checks that pin the learned-SPAI oracle (oracle/spai_oracle.cpp):
feature direction mapping, the network's layout, the SPD assembly in closed
form, known-answer losses and the gradient against central differences."""
import math

import numpy as np
import pytest
import scipy.sparse as sp

import oracle_api as O
from femgpu import gen

NP = 6 * 64 + 64 + 64 * 64 + 64 + 64 * 7 + 7
OFF = {"W1": 0, "b1": 384, "W2": 448, "b2": 4544, "W3": 4608, "b3": 5056}


def test_nparam():
    assert O.spai_nparam() == NP


def _axis_matrix(nx, ny, nz, cx=0.1, cy=0.2, cz=0.3, diag=2.0):
    """7-point matrix with coupling -cx/-cy/-cz along x/y/z."""
    n = nx * ny * nz
    rows, cols, vals = [], [], []
    for z in range(nz):
        for y in range(ny):
            for x in range(nx):
                i = (z * ny + y) * nx + x
                rows.append(i); cols.append(i); vals.append(diag)
                for (dx, dy, dz, c) in ((1, 0, 0, cx), (-1, 0, 0, cx), (0, 1, 0, cy), (0, -1, 0, cy),
                                        (0, 0, 1, cz), (0, 0, -1, cz)):
                    X, Y, Z = x + dx, y + dy, z + dz
                    if 0 <= X < nx and 0 <= Y < ny and 0 <= Z < nz:
                        rows.append(i); cols.append((Z * ny + Y) * nx + X); vals.append(-c)
    S = sp.csr_matrix((vals, (rows, cols)), shape=(n, n))
    S.sort_indices()
    return gen.CSR(n, n, S.indptr.astype(np.uint32), S.indices.astype(np.uint32), S.data.astype(np.float64))


def test_features_direction_mapping():
    G = _axis_matrix(3, 4, 5)
    feat, dis = O.spai_features(G, [3, 4, 5])
    assert np.allclose(dis, 1 / math.sqrt(2.0))
    # interior cell: directions -nxny, -nx, -1, +1, +nx, +nxny = z, y, x, x, y, z
    i = (2 * 4 + 1) * 3 + 1
    np.testing.assert_allclose(feat[i], 6 * np.array([0.3, 0.2, 0.1, 0.1, 0.2, 0.3]) / 2.0, rtol=1e-15)
    # corner cell 0 has only the + neighbours
    np.testing.assert_allclose(feat[0], 6 * np.array([0, 0, 0, 0.1, 0.2, 0.3]) / 2.0, rtol=1e-15)


def test_features_poisson_interior_are_one():
    G = gen.poisson3d(4)
    feat, dis = O.spai_features(G, [4, 4, 4])
    i = (1 * 4 + 1) * 4 + 1
    np.testing.assert_allclose(feat[i], np.ones(6), rtol=1e-15)
    assert np.allclose(dis, 1 / math.sqrt(6.0))


def test_features_reject_non_stencil():
    G = gen.poisson2d(4)      # offsets +-1, +-4 on a 4x4 grid
    with pytest.raises(O.OracleError):
        O.spai_features(G, [3, 3, 1])


def _rand_params(seed, scale=0.3):
    return (np.random.default_rng(seed).standard_normal(NP) * scale).astype(np.float32)


def test_mlp_forward_matches_numpy():
    P = _rand_params(0)
    X = np.random.default_rng(1).uniform(0, 1.5, (50, 6))
    W1 = P[:384].reshape(6, 64); b1 = P[384:448]
    W2 = P[448:4544].reshape(64, 64); b2 = P[4544:4608]
    W3 = P[4608:5056].reshape(64, 7); b3 = P[5056:]
    W1, b1, W2, b2, W3, b3 = (a.astype(np.float64) for a in (W1, b1, W2, b2, W3, b3))
    h1 = np.maximum(X @ W1 + b1, 0)
    h2 = np.maximum(h1 @ W2 + b2, 0)
    ref = h2 @ W3 + b3
    np.testing.assert_allclose(O.mlp_forward(P, X), ref, rtol=1e-12, atol=1e-12)
    # bfloat16 operand rounding moves the result by about 2^-8 relative
    b = O.mlp_forward(P, X, bf16=True)
    assert np.abs(b - ref).max() < 0.05 * np.abs(ref).max()


def _const_output_params(b3):
    P = np.zeros(NP, np.float32)
    P[5056:] = b3
    return P


def test_assembly_closed_form_and_spd():
    """Constant network output o = b3: G_ii = softplus(b3[0]) / sqrt(a_ii),
    G_ij = b3[1+d] / sqrt(a_jj) for the lower neighbour j in direction d
    (d = 0, 1, 2: offsets -16, -4, -1 on the 4^3 grid), zero above the
    diagonal; M = G^T G is SPD for any output."""
    G = gen.vardiff3d(4, seed=2, ell=1.5)[0]
    b3 = np.array([0.3, 0.1, -0.2, 0.05, 0.4, -0.1, 0.2], np.float32)
    g = O.spai_assemble(G, [4, 4, 4], _const_output_params(b3))
    A = G.to_scipy().tocsr()
    d = A.diagonal()
    off = [-16, -4, -1]
    ref = np.zeros_like(g)
    for i in range(G.n):
        for p in range(A.indptr[i], A.indptr[i + 1]):
            j = A.indices[p]
            if j == i:
                ref[p] = math.log1p(math.exp(float(b3[0]))) / math.sqrt(d[i])
            elif j - i in off:
                ref[p] = float(b3[1 + off.index(j - i)]) / math.sqrt(d[j])
    np.testing.assert_allclose(g, ref, rtol=1e-14, atol=0)
    # random network: G lower triangular with a positive diagonal, M SPD
    g = O.spai_assemble(G, [4, 4, 4], _rand_params(3))
    Gm = sp.csr_matrix((g, A.indices, A.indptr), shape=A.shape).toarray()
    assert np.all(np.triu(Gm, 1) == 0) and np.all(np.diag(Gm) > 0)
    M = Gm.T @ Gm
    assert np.linalg.eigvalsh(0.5 * (M + M.T)).min() > 0


def test_loss_known_answers():
    """A diagonal (no neighbours): G = softplus(b3[0]) D^{-1/2}, so
    G^T G A v = c^2 v with c = softplus(b3[0]) and L = (c^2 - 1)^2 mean(v^2)."""
    n = 27
    G = gen.CSR(n, n, np.arange(n + 1, dtype=np.uint32), np.arange(n, dtype=np.uint32),
                np.random.default_rng(0).uniform(1, 3, n))
    V = np.random.default_rng(1).standard_normal((3, n))
    for c in (1.0, 1.5):
        b0 = math.log(math.expm1(c))
        L, _ = O.spai_loss_grad(G, [3, 3, 3], _const_output_params([b0, 0, 0, 0, 0, 0, 0]), V)
        assert abs(L - (c * c - 1) ** 2 * np.mean(V ** 2)) < 1e-6 * max(1.0, L)
    # G ~ 0 (softplus(-80) ~ 2e-35): M A v ~ 0 and L = mean(v^2)
    P0 = _const_output_params([-80.0, 0, 0, 0, 0, 0, 0])
    L0, _ = O.spai_loss_grad(G, [3, 3, 3], P0, V)
    assert abs(L0 - np.mean(V ** 2)) < 1e-9 * np.mean(V ** 2)


def test_gradient_central_differences():
    G = gen.vardiff3d(4, seed=1, ell=1.5)[0]
    grid = [4, 4, 4]
    P = _rand_params(5, 0.2)
    V = np.random.default_rng(2).choice([-1.0, 1.0], (3, G.n))
    L, g = O.spai_loss_grad(G, grid, P, V)
    rng = np.random.default_rng(7)
    idx = []
    for name, (a, b) in {"W1": (0, 384), "b1": (384, 448), "W2": (448, 4544), "b2": (4544, 4608),
                         "W3": (4608, 5056), "b3": (5056, NP)}.items():
        idx += list(rng.choice(np.arange(a, b), size=6, replace=False))
    errs = []
    for t in idx:
        h = np.float32(max(1e-3, 1e-3 * abs(float(P[t]))))
        Pp, Pm = P.copy(), P.copy()
        Pp[t] = P[t] + h
        Pm[t] = P[t] - h
        Lp, _ = O.spai_loss_grad(G, grid, Pp, V)
        Lm, _ = O.spai_loss_grad(G, grid, Pm, V)
        fd = (Lp - Lm) / (float(Pp[t]) - float(Pm[t]))
        errs.append(abs(fd - g[t]) / max(abs(fd), 1e-3 * np.abs(g).max()))
    assert max(errs) < 2e-3, (max(errs), idx[int(np.argmax(errs))])
