"""GPU parity: every stage of the HIP path against the CPU oracle on the same
seeded inputs.  Bars: FP64 path <= 1e-10 relative L2 (coefficients and cross
sections), mixed precision <= 1e-9; assembled matrices <= 1e-12."""
import numpy as np
import pytest

from optimet_amd.problem import Cluster, config1, drude_cluster

pytestmark = pytest.mark.gpu

FP64_TOL = 1e-10
MIXED_TOL = 1e-9


def rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


@pytest.fixture(scope="module")
def ctx():
    from optimet_amd import solver
    return solver.Context()


def small_drude(P=5, nmax=4, sh=True, seed=11):
    return drude_cluster(P, nmax, 700e-9, 40, 80, 600, 10, seed=seed, sh=sh)


def check_internal(ora, c, out, ref, tol):
    """x_int = x_sca (.) Iaux (Solver.cpp:57-77).  The unpreconditioning step
    itself is pinned: x_int equals x_sca (.) Iaux_oracle to 1e-12.  x_int
    against the oracle then obeys the bound that the elementwise product
    fixes, ||dx_int|| / ||x_int|| <= amp ||dx_sca|| / ||x_sca|| (+ the 1e-12 of
    the Iaux pin), amp = max|Iaux| ||x_sca|| / ||x_int||: Iaux grows like
    (2n+1)!!/x^n with the order, so amp reaches 1e15 for plasmonic spheres at
    nmax 8, where two FP64 solvers of the same system differ by O(1) on x_int
    (test_oracle_pins.py::test_internal_coefficients_are_ill_conditioned)."""
    iaux = np.concatenate([ora.object_vector(c, j, ora.OBJ_IAUX) for j in range(c.nobj)])
    assert rel(out["x_int"], out["x_sca"] * iaux) < 1e-12
    amp = np.max(np.abs(iaux)) * np.linalg.norm(ref["x_sca"]) / np.linalg.norm(ref["x_int"])
    e = rel(out["x_sca"], ref["x_sca"])
    assert e < tol
    assert rel(out["x_int"], ref["x_int"]) <= amp * e * (1 + 1e-6) + 1e-12


def make(ctx, c, kind="lu64", pivoting="block"):
    from optimet_amd import solver
    s = solver.Solver(c, kind=kind, context=ctx, pivoting=pivoting)
    s.update()
    return s


CASES = {
    "config1": config1,
    "drude5": lambda: small_drude(),
    "offaxis": lambda: Cluster(xyz=np.array([[0, 0, 0], [210e-9, 40e-9, -30e-9], [-50e-9, 190e-9, 120e-9]]),
                               radius=np.array([90e-9, 70e-9, 60e-9]), eps_r=np.array([4 + 0.2j, 12 + 0.1j, 2.25]),
                               eps_r_sh=np.array([4.1 + 0.2j, 12.5 + 0.3j, 2.3]), nmax=5, wavelength=650e-9,
                               theta_k=0.6, phi_k=-1.1, etheta=0.8, ephi=0.6j),
}


@pytest.mark.parametrize("name", sorted(CASES))
def test_source_and_incident(ctx, ora, name):
    c = CASES[name]()
    s = make(ctx, c)
    assert rel(s.vector("Q"), ora.source_vector(c)) < 1e-12
    inc = np.concatenate([ora.inc_local(c, j) for j in range(c.nobj)])
    assert rel(s.vector("inc"), inc) < 1e-12


@pytest.mark.parametrize("name", sorted(CASES))
def test_scattering_matrix(ctx, ora, name):
    c = CASES[name]()
    s = make(ctx, c)
    S = s.matrix()
    So = ora.scattering_matrix(c)
    assert np.max(np.abs(S - So)) / np.max(np.abs(So)) < 1e-12
    assert rel(S, So) < 1e-12


def test_sh_matrix(ctx, ora):
    c = small_drude()
    s = make(ctx, c)
    V = s.matrix(sh=True)
    Vo = ora.scattering_matrix(c, sh=True)
    assert rel(V, Vo) < 1e-12


@pytest.mark.parametrize("kind,tol", [("lu64", FP64_TOL), ("lu32ir", MIXED_TOL), ("gmres64", MIXED_TOL),
                                      ("gmres32ir", MIXED_TOL)])
@pytest.mark.parametrize("name", ["config1", "offaxis"])
def test_fundamental_solve(ctx, ora, name, kind, tol):
    c = CASES[name]()
    c.sh = False
    ref = ora.solve_all(c)
    s = make(ctx, c, kind)
    out = s.solve()
    assert rel(out["x_sca"], ref["x_sca"]) < tol
    check_internal(ora, c, out, ref, tol)
    cs = s.cross_sections()
    for k in ("cext", "csca", "cabs"):
        assert abs(cs[k] - ref[k]) <= tol * abs(ref[k]), k


@pytest.mark.parametrize("kind,tol", [("lu64", FP64_TOL), ("lu32ir", MIXED_TOL), ("gmres32ir", MIXED_TOL),
                                      ("auto", MIXED_TOL)])
def test_second_harmonic_chain(ctx, ora, kind, tol):
    c = small_drude(P=4, nmax=3)
    ref = ora.solve_all(c)
    s = make(ctx, c, kind)
    out = s.solve()
    assert rel(out["x_sca"], ref["x_sca"]) < tol
    K, K1 = ora.source_sh(c, ora.cg_tables(c.nmax), ref["x_int"])
    assert rel(s.vector("K"), K) < max(tol, 1e-11)
    assert rel(out["x_sca_sh"], ref["x_sca_sh"]) < tol
    assert rel(out["x_int_sh"], ref["x_int_sh"]) < tol
    cs = s.cross_sections()
    assert abs(cs["csca_sh"] - ref["csca_sh"]) <= tol * abs(ref["csca_sh"])
    assert abs(cs["cabs_sh"] - ref["cabs_sh"]) <= tol * abs(ref["cabs_sh"])
    # C_abs,SH evaluated from the device's own internal coefficients
    own = ora.absorption_sh(c, ora.cg_tables(c.nmax), out["x_int"], out["x_int_sh"])
    assert abs(cs["cabs_sh"] - own) <= 1e-12 * abs(own)


@pytest.mark.parametrize("kind,tol", [("lu64", FP64_TOL), ("lu32ir", MIXED_TOL), ("auto", MIXED_TOL)])
def test_nmax9_cluster(ctx, ora, kind, tol):
    """nmax 9 (pair_kernel<9>): three Drude spheres, omega and 2 omega; S, V
    and Q to 1e-12, x_sca, x_sca_sh and the cross sections at the bar."""
    c = drude_cluster(3, 9, 650e-9, 50, 80, 500, 10, seed=9)
    s = make(ctx, c, kind)
    assert rel(s.vector("Q"), ora.source_vector(c)) < 1e-12
    assert rel(s.matrix(), ora.scattering_matrix(c)) < 1e-12
    assert rel(s.matrix(sh=True), ora.scattering_matrix(c, sh=True)) < 1e-12
    ref = ora.solve_all(c)
    out = s.solve()
    cs = s.cross_sections()
    assert rel(out["x_sca"], ref["x_sca"]) < tol
    assert rel(out["x_sca_sh"], ref["x_sca_sh"]) < tol
    for key in ("cext", "csca", "cabs", "csca_sh", "cabs_sh"):
        assert abs(cs[key] - ref[key]) <= tol * abs(ref[key]), key


def test_caller_supplied_cg_tables(ctx, ora):
    c = small_drude(P=3, nmax=3, seed=5)
    ref = ora.solve_all(c)
    from optimet_amd import solver
    s = solver.Solver(c, kind="lu64", context=ctx)
    s.set_cg_tables(ora.cg_tables(c.nmax))
    s.update()
    out = s.solve()
    assert rel(out["x_sca_sh"], ref["x_sca_sh"]) < FP64_TOL


@pytest.mark.parametrize("N", [37, 200, 700])
@pytest.mark.parametrize("kind,tol", [("lu64", 1e-11), ("lu32ir", 1e-10)])
def test_dense_solve_random(ctx, N, kind, tol):
    rng = np.random.default_rng(N)
    A = rng.standard_normal((N, N)) + 1j * rng.standard_normal((N, N)) + 2 * np.sqrt(N) * np.eye(N)
    b = rng.standard_normal(N) + 1j * rng.standard_normal(N)
    x = ctx.dense_solve(A, b, kind)
    xr = np.linalg.solve(A, b)
    assert rel(x, xr) < tol


@pytest.mark.parametrize("N", [300, 700, 1500])
def test_lu32_gemm_modes(ctx, N):
    """The mixed path's bulk updates on the bf16 matrix cores (three-term
    splits, six products) and on the FP32 ones: the same solution, and an
    FP32-accurate factorisation -- on a matrix with no diagonal boost (so the
    refinement's contraction is set by the factors' accuracy) the refinement
    takes as many iterations as with FP32 products, give or take one."""
    rng = np.random.default_rng(100 + N)
    A = rng.standard_normal((N, N)) + 1j * rng.standard_normal((N, N))
    b = rng.standard_normal(N) + 1j * rng.standard_normal(N)
    xr = np.linalg.solve(A, b)
    its = {}
    try:
        for mode in ("fp32", "bf16x3"):
            ctx.set_lu32_gemm(mode)
            x = ctx.dense_solve(A, b, "lu32ir")
            its[mode] = ctx.stats()["iterations"]
            assert rel(x, xr) < 1e-9, mode
    finally:
        ctx.set_lu32_gemm("bf16x3")
    assert abs(its["bf16x3"] - its["fp32"]) <= 1, its


def test_lu32_gemm_modes_cluster(ctx, ora):
    """N = 768 cluster, mixed path with either product mode, against the oracle
    (x_sca and x_sca_sh at the mixed bar)."""
    c = drude_cluster(8, 6, 750e-9, 40, 80, 800, 10, seed=21, sh=True)
    ref = ora.solve_all(c)
    try:
        for mode in ("fp32", "bf16x3"):
            ctx.set_lu32_gemm(mode)
            s = make(ctx, c, "lu32ir")
            out = s.solve()
            assert rel(out["x_sca"], ref["x_sca"]) < MIXED_TOL, mode
            assert rel(out["x_sca_sh"], ref["x_sca_sh"]) < MIXED_TOL, mode
    finally:
        ctx.set_lu32_gemm("bf16x3")


def test_mixed_refinement_ill_conditioned(ctx):
    """cond(A) = 1e6: the FP32 factors are a weak preconditioner
    (||I - M^{-1} A|| about cond * 6e-8, near 0.1), and the flexible refinement
    still reaches the FP64 answer (to cond * 1e-16 relative), as the FP64 LU
    does.  b = A x for a random x, so that ||A|| ||x|| / ||b|| is of order
    one: the refinement's relative residual ||b - A x|| / ||b|| then has an
    FP64 floor near 1e-16 (for a b with ||x|| ~ cond ||b|| the floor is
    cond * 1e-16, above the acceptance band of tol = 1e-16).  Beyond
    cond * 6e-8 ~ 1 the FP32 factors stop being a preconditioner."""
    rng = np.random.default_rng(77)
    N = 512
    U, _ = np.linalg.qr(rng.standard_normal((N, N)) + 1j * rng.standard_normal((N, N)))
    V, _ = np.linalg.qr(rng.standard_normal((N, N)) + 1j * rng.standard_normal((N, N)))
    A = (U * np.logspace(0, -6, N)) @ V.conj().T
    xt = rng.standard_normal(N) + 1j * rng.standard_normal(N)
    b = A @ xt
    xr = np.linalg.solve(A, b)
    for kind in ("lu64", "lu32ir"):
        x = ctx.dense_solve(A, b, kind)
        assert rel(x, xr) < 1e-9, kind
    assert ctx.stats()["iterations"] > 5  # the mixed refinement had real work to do


def test_refinement_accepts_backward_stable_large_solution(ctx):
    """cond(A) = 1e6 and a random b, so ||x|| ~ 1e6 ||b||: the relative
    residual cannot go below about u ||A|| ||x|| / ||b|| ~ 1e-10, far above the
    acceptance band of tol = 1e-16, but the normwise backward error reaches the
    FP64 unit roundoff and the solvers return x (to cond * 1e-15 relative)
    instead of reporting no convergence."""
    rng = np.random.default_rng(78)
    N = 512
    U, _ = np.linalg.qr(rng.standard_normal((N, N)) + 1j * rng.standard_normal((N, N)))
    V, _ = np.linalg.qr(rng.standard_normal((N, N)) + 1j * rng.standard_normal((N, N)))
    A = (U * np.logspace(0, -6, N)) @ V.conj().T
    b = rng.standard_normal(N) + 1j * rng.standard_normal(N)
    xr = np.linalg.solve(A, b)
    for kind in ("lu64", "lu32ir"):
        x = ctx.dense_solve(A, b, kind)
        assert rel(x, xr) < 1e-8, kind


def test_dense_solve_needs_pivoting(ctx):
    """Zero leading diagonal: only a pivoting LU survives."""
    rng = np.random.default_rng(3)
    N = 300
    A = rng.standard_normal((N, N)) + 1j * rng.standard_normal((N, N))
    A[np.arange(N), np.arange(N)] = 0.0
    b = rng.standard_normal(N) + 0j
    x = ctx.dense_solve(A, b, "lu64")
    assert rel(x, np.linalg.solve(A, b)) < 1e-9


def test_medium_cluster_multi_panel(ctx, ora):
    """N = 768: twelve LU panels, several panel workgroups."""
    c = drude_cluster(8, 6, 750e-9, 40, 80, 800, 10, seed=21, sh=False)
    ref = ora.solve_all(c)
    s = make(ctx, c, "lu64")
    out = s.solve()
    assert rel(out["x_sca"], ref["x_sca"]) < FP64_TOL


@pytest.mark.parametrize("kind,tol", [("lu64", FP64_TOL), ("lu32ir", MIXED_TOL)])
def test_partial_pivoting_mode(ctx, ora, kind, tol):
    """Whole-column (tournament) pivoting: the same answers as block pivoting."""
    c = small_drude(P=6, nmax=5, seed=5)
    ref = ora.solve_all(c)
    s = make(ctx, c, kind, pivoting="partial")
    out = s.solve()
    cs = s.cross_sections()
    assert rel(out["x_sca"], ref["x_sca"]) < tol
    assert rel(out["x_sca_sh"], ref["x_sca_sh"]) < tol
    assert abs(cs["cext"] - ref["cext"]) / abs(ref["cext"]) < tol
    ctx.set_pivoting("block")


@pytest.mark.parametrize("kind", ["lu64", "lu32ir"])
def test_block_pivoting_falls_back(ctx, kind):
    """Zero diagonal blocks: block pivoting meets a zero pivot, the solver
    refactors with partial pivoting and still returns the solution."""
    rng = np.random.default_rng(8)
    N = 256
    A = np.zeros((N, N), complex)
    A[:128, 128:] = rng.standard_normal((128, 128)) + 1j * rng.standard_normal((128, 128)) + 12 * np.eye(128)
    A[128:, :128] = rng.standard_normal((128, 128)) + 1j * rng.standard_normal((128, 128)) + 12 * np.eye(128)
    b = rng.standard_normal(N) + 1j * rng.standard_normal(N)
    before = ctx.stats()["pivot_fallbacks"]
    x = ctx.dense_solve(A, b, kind)
    assert ctx.stats()["pivot_fallbacks"] == before + 1
    assert rel(x, np.linalg.solve(A, b)) < 1e-10


@pytest.mark.parametrize("kind", ["lu64", "lu32ir"])
def test_dense_solve_partial_mode(ctx, kind):
    rng = np.random.default_rng(12)
    N = 500
    A = rng.standard_normal((N, N)) + 1j * rng.standard_normal((N, N))
    b = rng.standard_normal(N) + 1j * rng.standard_normal(N)
    ctx.set_pivoting("partial")
    try:
        x = ctx.dense_solve(A, b, kind)
    finally:
        ctx.set_pivoting("block")
    assert rel(x, np.linalg.solve(A, b)) < 1e-9


@pytest.mark.parametrize("kind,tol", [("lu64", FP64_TOL), ("lu32ir", MIXED_TOL)])
def test_golden_fixtures_gpu(ctx, kind, tol):
    """The HIP path against the committed fixtures (tests/golden/make_golden.py)."""
    import os
    from tests.golden.make_golden import CASES
    here = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
    for name, build in CASES.items():
        ref = np.load(os.path.join(here, name + ".npz"))
        s = make(ctx, build(), kind)
        out = s.solve()
        cs = s.cross_sections()
        assert rel(out["x_sca"], ref["x_sca"]) < tol, name
        for key in ("cext", "csca", "cabs"):
            assert abs(cs[key] - ref[key]) <= tol * abs(ref[key]), (name, key)
        if "x_sca_sh" in ref.files:
            assert rel(out["x_sca_sh"], ref["x_sca_sh"]) < tol, name
            assert abs(cs["csca_sh"] - ref["csca_sh"]) <= tol * abs(ref["csca_sh"]), name


@pytest.mark.parametrize("kind,tol", [("lu64", FP64_TOL), ("lu32ir", MIXED_TOL), ("auto", MIXED_TOL)])
@pytest.mark.parametrize("P,nmax,batch", [(5, 4, 3), (8, 6, 4)])
def test_wavelength_sweep_batched(ctx, ora, kind, tol, P, nmax, batch):
    """omx_solve_sweep: several wavelengths solved as batches (chunks of
    `batch` systems, the last one partial) against the oracle at each
    wavelength, Drude permittivities re-evaluated per wavelength.  "auto"
    solves S by lockstep FP64 GMRES beside V's batched mixed LU."""
    from optimet_amd import solver
    c = drude_cluster(P, nmax, 600e-9, 40, 80, 600 if P < 8 else 800, 10, seed=31 + P)
    lams = np.linspace(450e-9, 850e-9, 7)
    s = solver.Solver(c, kind=kind, context=ctx)
    out = s.sweep(lams, batch=batch)
    assert s.stats()["sweep_batch"] == batch
    for i, lam in enumerate(lams):
        ref = ora.solve_all(c.with_wavelength(lam))
        assert rel(out["x_sca"][i], ref["x_sca"]) < tol, (i, lam)
        assert rel(out["x_sca_sh"][i], ref["x_sca_sh"]) < tol, (i, lam)
        xs = out["xs"][i]
        for j, key in enumerate(("cext", "csca", "cabs", "csca_sh", "cabs_sh")):
            assert abs(xs[j] - ref[key]) <= tol * abs(ref[key]), (i, key)
        assert out["iterations"][i, 0] > 0


def test_sweep_matches_single_solves(ctx):
    """A batched sweep and the same wavelengths solved one at a time agree."""
    from optimet_amd import solver
    c = drude_cluster(6, 5, 700e-9, 40, 80, 700, 10, seed=77)
    lams = np.array([520e-9, 610e-9, 760e-9])
    s = solver.Solver(c, kind="lu32ir", context=ctx)
    out = s.sweep(lams, batch=0)
    for i, lam in enumerate(lams):
        s1 = solver.Solver(c.with_wavelength(lam), kind="lu32ir", context=ctx)
        s1.update()
        one = s1.solve()
        assert rel(out["x_sca"][i], one["x_sca"]) < 1e-12
        assert rel(out["x_sca_sh"][i], one["x_sca_sh"]) < 1e-10


@pytest.mark.parametrize("kind", ["lu64", "lu32ir"])
def test_field_maps(ctx, ora, kind):
    """omx_fields (Result::getEHFields, Result.cpp:74-301) against the oracle
    at points outside and inside the spheres (the SH bulk term of
    Geometry::COEFFpartSH included), on an object's z axis (the reference's
    1e-6 polar shift): E, H at omega and 2 omega to 1e-10 of each field's
    largest component.  On an object's z axis the reference forms the point
    relative to the object through a spherical -> Cartesian round trip
    (Tools::toPoint, as the oracle does), which leaves an arbitrary azimuth,
    and its 1e-6 shift of the polar angle makes the field depend on that
    azimuth at O(1e-6): those points are held to 1e-6.  (At a sphere centre or
    the origin the reference's N waves divide by k r = 0; no point is placed
    there.)"""
    c = small_drude(P=3, nmax=4, seed=13)
    s = make(ctx, c, kind)
    out = s.solve()
    rng = np.random.default_rng(2)
    lo, hi = c.xyz.min(axis=0) - 150e-9, c.xyz.max(axis=0) + 150e-9
    pts = [lo + (hi - lo) * rng.random(3) for _ in range(40)]
    for j in range(c.nobj):  # inside each sphere
        pts.append(c.xyz[j] + 0.5 * c.radius[j] * rng.standard_normal(3) / np.sqrt(3))
    axis = [c.xyz[j] + np.array([0, 0, 1.7 * c.radius[j]]) for j in range(c.nobj)]
    pts = np.array(pts + axis)
    on_axis = np.arange(len(pts)) >= len(pts) - len(axis)
    f = s.fields(pts)
    ref = ora.fields(c, ora.cg_tables(c.nmax), out, pts)
    for key in ("E", "H", "E_sh", "H_sh"):
        scale = np.max(np.abs(ref[key]))
        err = np.max(np.abs(f[key] - ref[key]), axis=1) / scale
        assert np.max(err[~on_axis]) < 1e-10, (key, np.max(err[~on_axis]))
        assert np.max(err[on_axis]) < 1e-6, (key, np.max(err[on_axis]))
