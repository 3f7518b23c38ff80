"""GPU parity of the Schur-complement preconditioner
(make_schur_preconditioner, blockMatrix.hpp:625-685) against the oracle on
the reference's 3-D Stokes test system (test_BlockMatrix_schur3d.cpp:39-193):
ILU(0) velocity block and the scaled SA-AMG of the pressure Laplacian, or the
identity, composed on the device."""
import numpy as np
import pytest

import femgpu as F
import oracle_api as O
from femgpu import gen

pytestmark = pytest.mark.gpu


def _systems(N):
    S3 = gen.stokes3d(N)
    dev = {k: F.CsrMatrix.from_gen(S3[k]) for k in ("K", "A", "B", "Bt", "Lp")}
    orc = {k: O.Csr.from_gen(S3[k]) for k in ("K", "A", "B", "Bt", "Lp")}
    return S3, dev, orc


@pytest.mark.parametrize("spc", ["amg", "identity"])
def test_schur_apply_and_gmres_match_oracle(spc):
    S3, D, Or = _systems(10)
    pa = F.Preconditioner(D["A"], "ilu0")
    pao = O.Pc(Or["A"], O.K_ILU0)
    if spc == "amg":
        pl = F.Preconditioner(D["Lp"], "saamg")
        plo = O.Pc(Or["Lp"], O.K_SAAMG, [0.25, 0.0, 1, 1, 20, 4, 50, 0.0, 1])
        M, Mo = F.SchurPreconditioner(D["B"], D["Bt"], pa, pl, 6.1), O.Pc.schur(Or["B"], Or["Bt"], pao, plo, 6.1)
    else:
        M, Mo = F.SchurPreconditioner(D["B"], D["Bt"], pa, None, 1.0), O.Pc.schur(Or["B"], Or["Bt"], pao, None, 1.0)
    r = np.random.default_rng(3).standard_normal(D["K"].rows)
    zo = Mo.apply(r)
    np.testing.assert_allclose(M.apply(r), zo, rtol=1e-10, atol=1e-10 * np.abs(zo).max())
    xe = gen.drand(np.arange(D["K"].rows))
    b = S3["K"].matvec(xe)
    x, st = F.solve(D["K"], M, b, solver="gmres", restart=100, max_iter=500)
    xo, ito = O.solve(Or["K"], 2, Mo, b, tol=1e-10, max_iter=500, restart=100)
    print(f"{spc}: GMRES(100) steps gpu {st['iterations'] + 1} oracle {ito + 1}", flush=True)
    # GMRES stops on its Givens estimate |g_j+1| / ||b|| < 1e-10 (array.cpp:
    # 4361); the true residual of x follows it to rounding, as the oracle's
    res_o = np.linalg.norm(b - S3["K"].matvec(xo)) / np.linalg.norm(b)
    assert st["true_relres"] < max(1.5e-10, 1.5 * res_o)
    # K is symmetric with the one null vector (0, constant pressure), which
    # GMRES leaves where its Krylov space puts it: off that direction two
    # iterates differ by at most cond'(K) (r1 + r2), cond' = |lambda|_max /
    # the smallest nonzero |lambda| (130 for N = 10)
    lam = np.sort(np.abs(np.linalg.eigvalsh(S3["K"].to_scipy().toarray())))
    assert lam[0] < 1e-12 * lam[-1] < lam[1]
    nv = S3["nv"]
    d = x - xo
    d[nv:] -= d[nv:].mean()
    assert np.linalg.norm(d) / np.linalg.norm(xo) < max(1e-8, lam[-1] / lam[1] * (st["true_relres"] + res_o))
    assert abs(st["iterations"] - ito) <= 2
    # K's null vector is (0, constant pressure): the drand solution is met up
    # to a pressure constant (the reference test's 1e-6 bar)
    d = x - xe
    assert np.linalg.norm(d[:nv]) < 1e-6 and np.linalg.norm(d[nv:] - d[nv:].mean()) < 1e-6


def test_schur_fp32_refinement_and_sizes():
    S3, D, Or = _systems(8)
    pa = F.Preconditioner(D["A"], "ilu0", fp32=1)
    pl = F.Preconditioner(D["Lp"], "saamg", fp32=1)
    M = F.SchurPreconditioner(D["B"], D["Bt"], pa, pl, 6.1)
    xe = gen.drand(np.arange(D["K"].rows))
    b = S3["K"].matvec(xe)
    x, st = F.solve(D["K"], M, b, solver="gmres", restart=100, max_iter=1000, precision=F.FP32_IR)
    assert st["converged"] == 1 and st["true_relres"] < 1e-10
    # size checks: the preconditioners must match n_vel / n_pres
    with pytest.raises(F.InvalidArgument):
        F.SchurPreconditioner(D["B"], D["Bt"], pl, pa, 1.0)
