"""Continuous batching across requests (omx_solve_group): clusters of
different sizes (P, nmax) and wavelengths, each in a context of its own,
solved as one batch bordered by the identity to the largest padded order.

Each member's answer is held to the oracle at the path's bar (FP64 1e-10,
mixed 1e-9: x_sca, x_sca_SH and the cross sections) and to the same member
solved alone with omx_solve (1e-10: two converged refinements of one system);
x_int is held to x_sca (.) Iaux with the oracle's Iaux (1e-12), and x_int_SH
enters C_abs,SH, which is held to the oracle.  The leader is reused for a second, smaller group."""
import numpy as np
import pytest

from optimet_amd.problem import drude_cluster

pytestmark = pytest.mark.gpu

KEYS = ("cext", "csca", "cabs", "csca_sh", "cabs_sh")
# (P, nmax, wavelength nm, seed): N from 48 to 400, padded orders 64 to 448
MEMBERS = [(3, 4, 520, 31), (5, 5, 610, 32), (2, 6, 700, 33), (4, 3, 760, 34), (6, 4, 830, 35)]


def rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def clusters():
    return [drude_cluster(P, nmax, lam * 1e-9, 40, 80, 600, 10, seed=seed) for P, nmax, lam, seed in MEMBERS]


@pytest.fixture(scope="module")
def ora():
    import oracle
    return oracle


@pytest.mark.parametrize("kind,tol", [("lu32ir", 1e-9), ("lu64", 1e-10)])
def test_group_matches_alone_and_oracle(ora, kind, tol):
    from optimet_amd import solver
    cs = clusters()
    print(f"group of {len(cs)}: N = {[c.N for c in cs]}, {kind}", flush=True)
    members = [solver.Solver(c, kind=kind) for c in cs]
    for s in members:
        s.update()
    leader = solver.Context()
    outs = solver.solve_group(members, leader=leader)
    xs_group = [s.cross_sections() for s in members]
    assert leader.stats()["sweep_batch"] == len(cs) and leader.stats()["pivot_fallbacks"] == 0
    for c, s, out, xs in zip(cs, members, outs, xs_group):
        st = s.stats()
        assert st["residual"] <= 1e-13 and st["residual_sh"] <= 1e-13, st
        ref = ora.solve_all(c)
        for k in ("x_sca", "x_sca_sh"):
            assert rel(out[k], ref[k]) < tol, (c.N, k, rel(out[k], ref[k]))
        for k in KEYS:
            assert abs(xs[k] - ref[k]) <= tol * abs(ref[k]), (c.N, k)
        alone = solver.Solver(c, kind=kind)
        alone.update()
        a = alone.solve()
        for k in ("x_sca", "x_sca_sh"):
            assert rel(out[k], a[k]) < 1e-10, (c.N, k, rel(out[k], a[k]))
        # each member's own unpreconditioning of its batch solution: x_int =
        # x_sca (.) Iaux (Solver.cpp:57-77) with the oracle's Iaux
        iaux = np.concatenate([ora.object_vector(c, j, ora.OBJ_IAUX) for j in range(c.nobj)])
        assert rel(out["x_int"], out["x_sca"] * iaux) < 1e-12
        alone.ctx.close()
    # the same leader, a smaller group of the first two members
    for s in members[:2]:
        s.update()
    again = solver.solve_group(members[:2], leader=leader)
    for o1, o2 in zip(outs[:2], again):
        assert rel(o2["x_sca"], o1["x_sca"]) < 1e-10
    leader.close()
    for s in members:
        s.ctx.close()


def test_group_argument_checks():
    from optimet_amd import solver
    cs = clusters()[:2]
    a = solver.Solver(cs[0], kind="lu32ir")
    b = solver.Solver(cs[1], kind="lu64")
    a.update()
    b.update()
    leader = solver.Context()
    with pytest.raises(solver.OmxError):  # two LU kinds
        solver.solve_group([a, b], leader=leader)
    c = solver.Solver(cs[1], kind="lu32ir")
    with pytest.raises(solver.OmxError):  # not updated
        solver.solve_group([a, c], leader=leader)
    with pytest.raises(ValueError):  # one context twice
        solver.solve_group([a, a], leader=leader)
    c.update()
    out = solver.solve_group([a, c], leader=leader)  # and then it works
    assert len(out) == 2 and np.all(np.isfinite(out[1]["x_sca"]))
    for x in (a, b, c):
        x.ctx.close()
    leader.close()
