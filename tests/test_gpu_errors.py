"""GPU: error paths of the C ABI.  A failed call reports its error code and
leaves the context usable with the geometry it had; a refinement that does
not reach the FP64 level is never accepted because the caller's tolerance is
loose."""
import numpy as np
import pytest

from optimet_amd.problem import drude_cluster

pytestmark = pytest.mark.gpu


def rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def test_failed_dense_solve_restores_the_geometry(ora):
    """omx_dense_solve on an exactly singular matrix (a zero column) raises
    OMX_ERR_SINGULAR after both pivot searches; the context keeps the
    geometry's size, solver kind and second-harmonic flag, so the next
    update() and solve() size their buffers from the geometry and match the
    oracle (PreconditionedMatrixSolver.h:45-100)."""
    from optimet_amd import solver
    c = drude_cluster(3, 3, 700e-9, 40, 80, 500, 10, seed=4, sh=True)
    ctx = solver.Context()
    try:
        s = solver.Solver(c, kind="lu32ir", context=ctx)
        rng = np.random.default_rng(5)
        n = 200  # larger than the cluster's N = 90: a failed solve must not leave it behind
        A = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
        A[:, 5] = 0.0
        b = rng.standard_normal(n) + 0j
        with pytest.raises(solver.OmxError) as e:
            ctx.dense_solve(A, b, "lu64")
        assert e.value.code == -4
        assert s.scattering_size() == c.N
        s.update()
        out = s.solve()
        ref = ora.solve_all(c)
        assert rel(out["x_sca"], ref["x_sca"]) < 1e-9
        assert rel(out["x_sca_sh"], ref["x_sca_sh"]) < 1e-9
    finally:
        ctx.close()


def test_loose_tolerance_does_not_accept_fp32_quality():
    """tol = 1e-10 and a refinement capped at one iteration, on cond(A) = 1e6
    with a random right-hand side: the relative residual stays far above
    1e3 tol, and the backward error of the FP32-factor answer (about 1e-11)
    is far above the FP64 level.  The solve falls back to partial pivoting and
    then reports no convergence, instead of returning an FP32-quality x."""
    from optimet_amd import solver
    rng = np.random.default_rng(91)
    N = 512
    U, _ = np.linalg.qr(rng.standard_normal((N, N)) + 1j * rng.standard_normal((N, N)))
    V, _ = np.linalg.qr(rng.standard_normal((N, N)) + 1j * rng.standard_normal((N, N)))
    A = (U * np.logspace(0, -6, N)) @ V.conj().T
    b = rng.standard_normal(N) + 1j * rng.standard_normal(N)
    xr = np.linalg.solve(A, b)
    ctx = solver.Context()
    try:
        ctx.set_solver("lu32ir", tol=1e-10, maxit=1, restart=40)
        before = ctx.stats()["pivot_fallbacks"]
        try:
            x = ctx.dense_solve(A, b, "lu32ir")
        except solver.OmxError as e:
            assert e.code == -6, e
        else:  # accepted only at FP64 quality
            assert rel(x, xr) < 1e-9
        assert ctx.stats()["pivot_fallbacks"] == before + 1
    finally:
        ctx.close()
