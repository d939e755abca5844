"""Factorisations captured in HIP graphs (omx_set_graphs): every replay gives
the eager path's answer.

One context with graphs on solves several wavelengths one after another (one
capture, then replays reading each new matrix); a second context solves the
same systems eagerly in between, so eager work and synchronisations separate
the replays.  Every replay is compared with its eager twin.  The graphs hold
the same kernels in the same order, so the two agree to rounding of the
(non-associative) parallel sums at most; the bar is 1e-14 relative.  With
second-harmonic generation S's graph is cut at V's start mark and V's
factorisation replays on the second stream pair beside S's."""
import numpy as np
import pytest

from optimet_amd.problem import drude_cluster

pytestmark = pytest.mark.gpu

KEYS = ("x_sca", "x_int", "x_sca_sh", "x_int_sh")


def rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


@pytest.mark.parametrize("kind,sh", [("lu32ir", True), ("lu64", True), ("lu32ir", False)])
def test_single_solve_replays(kind, sh):
    from optimet_amd import solver
    c = drude_cluster(12, 8, 700e-9, 40, 80, 900, 10, seed=21, sh=sh)  # N = 1920
    lams = [520e-9, 610e-9, 700e-9, 760e-9, 610e-9]
    print(f"graphs: {kind}, sh {sh}, N {c.N}, {len(lams)} solves", flush=True)
    sg = solver.Solver(c, kind=kind)
    sg.ctx.set_graphs(True)
    se = solver.Solver(c, kind=kind)
    worst = 0.0
    for lam in lams:
        cl = c.with_wavelength(lam)
        sg.set_cluster(cl)
        sg.update()
        og = sg.solve()
        se.set_cluster(cl)
        se.update()
        oe = se.solve()
        for k in KEYS:
            if k in oe:
                worst = max(worst, rel(og[k], oe[k]))
                assert rel(og[k], oe[k]) < 1e-14, (lam, k)
        cg, ce = sg.cross_sections(), se.cross_sections()
        assert cg["cext"] == pytest.approx(ce["cext"], rel=1e-14)
    st = sg.stats()
    print(f"worst rel. difference {worst:.2e}; captures {st['graph_captures']}, launches {st['graph_launches']}",
          flush=True)
    assert st["graph_captures"] == 1 and st["graph_launches"] == len(lams)
    assert se.stats()["graph_launches"] == 0
    sg.ctx.close()
    se.ctx.close()


@pytest.mark.parametrize("kind", ["lu32ir", "lu64"])
def test_sweep_replays(kind):
    """Three chunks of eight wavelengths: captured at the first, replayed at
    the next two; the same sweep eagerly in another context."""
    from optimet_amd import solver
    c = drude_cluster(12, 8, 700e-9, 40, 80, 900, 10, seed=21)
    lams = np.linspace(480e-9, 860e-9, 24)
    print(f"graph sweep: {kind}, N {c.N}, 24 wavelengths in batches of 8", flush=True)
    g = solver.Solver(c, kind=kind)
    g.ctx.set_graphs(True)
    og = g.sweep(lams, batch=8)
    e = solver.Solver(c, kind=kind)
    oe = e.sweep(lams, batch=8)
    for k in ("x_sca", "x_sca_sh", "x_int", "x_int_sh"):
        for i in range(lams.size):
            assert rel(og[k][i], oe[k][i]) < 1e-14, (k, i)
    np.testing.assert_allclose(og["xs"], oe["xs"], rtol=1e-13)
    st = g.stats()
    print(f"captures {st['graph_captures']}, launches {st['graph_launches']}", flush=True)
    assert st["graph_captures"] == 1 and st["graph_launches"] == 3
    g.ctx.set_graphs(False)  # releases the graphs; eager again
    again = g.sweep(lams[:8], batch=8)
    assert g.stats()["graph_launches"] == 3
    assert rel(again["x_sca"][0], oe["x_sca"][0]) < 1e-14
    g.ctx.close()
    e.ctx.close()
