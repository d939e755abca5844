"""The automatic solver choice (OMX_SOLVER_AUTO) and the mixed Krylov path
(OMX_SOLVER_GMRES32_IR) at a size where the choice matters.

At N >= 16384 the automatic kind starts with the mixed Krylov; with an
iteration budget too small to converge it must fall back to the mixed LU on
the same matrices and return the LU's answer.  Below that, with the second
harmonic, S is tried with FP64 GMRES while V's mixed LU factorises.  Every path is held to the
mixed bar (1e-9) against the mixed LU of the same system, whose own parity
with the oracle is tested at smaller sizes (tests/test_gpu_parity.py) and at
config 4 by residual (tests/test_gpu_configs.py)."""
import numpy as np
import pytest

from optimet_amd.problem import drude_cluster

pytestmark = pytest.mark.gpu


def rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


@pytest.fixture(scope="module")
def big():
    """171 Drude spheres at nmax 6: N = 16416 (a 4.3 GB FP64 matrix)."""
    return drude_cluster(171, 6, 800e-9, 30, 60, 2000, 10, seed=77, sh=False)


@pytest.fixture(scope="module")
def reference(big):
    from optimet_amd import solver
    s = solver.Solver(big, kind="lu32ir")
    s.update()
    out = s.solve()
    cs = s.cross_sections()
    s.ctx.close()
    return out, cs


def run(c, kind, budget=0, **kw):
    from optimet_amd import solver
    s = solver.Solver(c, kind=kind, **kw)
    s.ctx.set_auto_budget(budget)
    s.update()
    out = s.solve()
    cs = s.cross_sections()
    st = s.stats()
    s.ctx.close()
    return out, cs, st


def test_auto_picks_the_mixed_krylov(big, reference):
    print(f"auto: N {big.N}", flush=True)
    out, cs, st = run(big, "auto")
    assert st["pivot_fallbacks"] == 0 and st["matvec_calls"] > 0
    assert rel(out["x_sca"], reference[0]["x_sca"]) < 1e-9
    assert cs["cext"] == pytest.approx(reference[1]["cext"], rel=1e-9)


def test_auto_falls_back_to_the_mixed_lu(big, reference):
    """One outer step cannot reach FP64 accuracy: the LU takes over."""
    print(f"auto with one outer step: N {big.N}", flush=True)
    out, cs, st = run(big, "auto", budget=1)
    assert st["pivot_fallbacks"] >= 1
    assert rel(out["x_sca"], reference[0]["x_sca"]) < 1e-12


@pytest.mark.parametrize("kind", ["gmres32ir", "gmres64"])
def test_krylov_paths_at_size(big, reference, kind):
    print(f"{kind}: N {big.N}", flush=True)
    out, cs, st = run(big, kind)
    assert rel(out["x_sca"], reference[0]["x_sca"]) < 1e-9
    assert st["residual"] <= 1e-13 or kind == "gmres64"


def test_auto_small_system_uses_the_lu():
    """Below N = 16384, without the second harmonic, the automatic kind is the
    mixed LU: the LU's iterations and its answer to rounding."""
    c = drude_cluster(8, 5, 700e-9, 40, 80, 600, 10, seed=11, sh=False)
    out_a, cs_a, st_a = run(c, "auto")
    out_l, cs_l, st_l = run(c, "lu32ir")
    assert st_a["iterations"] == st_l["iterations"]
    assert rel(out_a["x_sca"], out_l["x_sca"]) < 1e-14


def test_auto_small_sh_system_tries_gmres_for_s():
    """With the second harmonic, V's mixed LU starts at once and S is solved by
    FP64 GMRES beside it: S takes exactly the iterations of the FP64 GMRES kind
    (same restart, tolerance and zero start), V about those of the mixed LU, and both
    answers agree with the mixed LU's to the mixed bar."""
    c = drude_cluster(8, 5, 700e-9, 40, 80, 600, 10, seed=11, sh=True)
    print(f"auto with SH: N {c.N}", flush=True)
    out_a, cs_a, st_a = run(c, "auto")
    out_l, cs_l, st_l = run(c, "lu32ir")
    out_g, cs_g, st_g = run(c, "gmres64")
    assert st_a["pivot_fallbacks"] == 0
    assert st_a["iterations"] == st_g["iterations"]
    # V's right-hand side comes from the fundamental solution, which differs
    # from the LU's at rounding level: its refinement at tol 1e-16 may take
    # one step more or less
    assert abs(st_a["iterations_sh"] - st_l["iterations_sh"]) <= 1
    for k in ("x_sca", "x_int", "x_sca_sh", "x_int_sh"):
        assert rel(out_a[k], out_l[k]) < 1e-10, k
    assert rel(out_a["x_sca"], out_g["x_sca"]) < 1e-14
    for k in ("cext", "csca", "cabs", "csca_sh", "cabs_sh"):
        assert cs_a[k] == pytest.approx(cs_l[k], rel=1e-10), k


def test_auto_small_sh_system_falls_back_to_the_lu():
    """One GMRES step cannot reach FP64 accuracy: S is factored after all and
    the answer is the mixed LU's."""
    c = drude_cluster(8, 5, 700e-9, 40, 80, 600, 10, seed=11, sh=True)
    out_a, cs_a, st_a = run(c, "auto", budget=1)
    out_l, cs_l, st_l = run(c, "lu32ir")
    assert st_a["pivot_fallbacks"] == 1
    assert st_a["iterations"] == st_l["iterations"] and st_a["iterations_sh"] == st_l["iterations_sh"]
    for k in ("x_sca", "x_sca_sh"):
        assert rel(out_a[k], out_l[k]) < 1e-13, k


@pytest.fixture(scope="module")
def ora():
    import oracle
    return oracle


@pytest.fixture(scope="module")
def sweep_case():
    """14 Drude spheres at nmax 8 (N = 2240), eight wavelengths, with SH: the
    automatic sweep solves S by lockstep FP64 GMRES beside V's batched LU."""
    c = drude_cluster(14, 8, 700e-9, 40, 80, 1000, 10, seed=91)
    return c, np.linspace(450e-9, 850e-9, 8)


def sweep(c, lams, kind, budget=0):
    from optimet_amd import solver
    s = solver.Solver(c, kind=kind)
    s.ctx.set_auto_budget(budget)
    out = s.sweep(lams, batch=4)
    st = s.stats()
    s.ctx.close()
    return out, st


def test_auto_sweep_gmres_for_s(ora, sweep_case):
    c, lams = sweep_case
    print(f"auto sweep: N {c.N}, {lams.size} wavelengths", flush=True)
    out_a, st_a = sweep(c, lams, "auto")
    out_l, st_l = sweep(c, lams, "lu32ir")
    assert st_a["pivot_fallbacks"] == 0
    assert out_a["iterations"][:, 0].max() > out_l["iterations"][:, 0].max()  # GMRES steps, not refinement steps
    for k in ("x_sca", "x_sca_sh"):
        for i in range(lams.size):
            assert rel(out_a[k][i], out_l[k][i]) < 1e-10, (k, i)
    # against the oracle: x_sca at the mixed bar; x_sca_SH no farther from it
    # than the mixed LU's (at nmax 8 the SH solution's conditioning spreads
    # FP64 solvers over about 1e-9; the spread-aware bars are the full-size
    # tests')
    ref = ora.solve_all(c.with_wavelength(lams[3]))
    assert rel(out_a["x_sca"][3], ref["x_sca"]) < 1e-9
    assert rel(out_a["x_sca_sh"][3], ref["x_sca_sh"]) <= max(1e-9, 2 * rel(out_l["x_sca_sh"][3], ref["x_sca_sh"]))


def test_auto_sweep_falls_back_to_the_batched_lu(sweep_case):
    """One GMRES step per system: every chunk factors S after all and refines
    from the GMRES iterate; the answers are the mixed LU sweep's."""
    c, lams = sweep_case
    out_a, st_a = sweep(c, lams, "auto", budget=1)
    out_l, st_l = sweep(c, lams, "lu32ir")
    assert st_a["pivot_fallbacks"] == 2  # both chunks of four
    for k in ("x_sca", "x_sca_sh"):
        for i in range(lams.size):
            assert rel(out_a[k][i], out_l[k][i]) < 1e-11, (k, i)
