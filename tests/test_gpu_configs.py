"""GPU parity at the benchmark configurations' real shapes.

The oracle assembles the reference's matrices (PreconditionedMatrix.cpp:350-400,
555-610) and right-hand sides (:1327-1436) on all host cores; numpy's LAPACK
solves them where that is affordable (the reference solves with a dense
factorisation, PreconditionedMatrixSolver.h:58, 75).  Larger systems are
checked through properties that do not need a host factorisation: sampled
matrix blocks, the residual of the device's solution on the oracle's matrix
(assembled one block row at a time), and agreement of two device solvers.

Bars: assembled matrices and right-hand sides 1e-12; residuals of the
device's x on the oracle's matrix 1e-10 of the right-hand side; scattering
coefficients and cross sections 1e-10 (lu64) and 1e-9 (lu32ir), unless the
reference's own chain does not determine them that well.  Each reference
solve is repeated on the oracle's matrices perturbed entrywise by 1e-15
relative (random phases; about the device-oracle assembly difference, and
ten units of FP64 roundoff), and an output is held to max(bar, 10 x the
change that perturbation causes); C_abs and C_sca,SH are pinned through the
device's own coefficients instead (check_solution).  The second-harmonic source is built from
conj(x_int) = conj(x_sca (.) Iaux) (PreconditionedMatrixSolver.h:65-66), and
Iaux reaches 1e15-1e17 at high orders, so x_sca_sh of large plasmonic
clusters moves by up to 1e-3 under such a perturbation: two FP64 solvers of
the reference's own chain disagree that much.
"""
import numpy as np
import pytest

from optimet_amd import problem

pytestmark = pytest.mark.gpu

TOL = {"lu64": 1e-10, "lu32ir": 1e-9}
ETA = 1e-15  # entrywise relative perturbation of the oracle's matrices
MARGIN = 10.0


def rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def relabs(a, b):
    return abs(a - b) / abs(b)


def iaux(ora, c):
    return np.concatenate([ora.object_vector(c, j, ora.OBJ_IAUX) for j in range(c.nobj)])


def perturbed_solve(A, b, rng):
    """LAPACK solve of A (.) (1 + ETA e^{i phi}) x = b, random phases phi."""
    Ap = A * (1.0 + ETA * np.exp(2j * np.pi * rng.random(A.shape)))
    return np.linalg.solve(Ap, b)


class Chain:
    """The reference's update() + solve() chain at one wavelength, with the
    dense solves done by LAPACK: x_sca = S^-1 Q, x_int = x_sca (.) Iaux,
    K = source_vectorSH(conj x_int), x_sca_sh = V^-1 K, cross sections; and
    the same chain on the perturbed matrices, whose distance to the first is
    each output's spread."""

    def __init__(self, ora, c, tables, keep=False, seed=0):
        rng = np.random.default_rng(seed)
        ia = iaux(ora, c)
        self.Q = ora.source_vector(c)
        S = ora.scattering_matrix(c)
        self.x = np.linalg.solve(S, self.Q)
        xp = perturbed_solve(S, self.Q, rng)
        self.S = S if keep else None
        del S
        self.xint = self.x * ia
        self.K, self.K1 = ora.source_sh(c, tables, self.xint)
        Kp, _ = ora.source_sh(c, tables, xp * ia)
        V = ora.scattering_matrix(c, sh=True)
        self.xsh = np.linalg.solve(V, self.K)
        yp = perturbed_solve(V, Kp, rng)
        self.V = V if keep else None
        del V
        self.cs = ora.cross_sections(c, self.x, self.xsh)
        csp = ora.cross_sections(c, xp, yp)
        self.spread = {"x_sca": rel(xp, self.x), "x_sca_sh": rel(yp, self.xsh)}
        for k, key in enumerate(("cext", "csca", "cabs", "csca_sh")):
            self.spread[key] = relabs(csp[k], self.cs[k])

    def tol(self, key, bar):
        return max(bar, MARGIN * self.spread[key])


def check_solution(name, out, cs, ref, bar, ora, c):
    """Against a Chain: x_sca, x_sca_sh, C_ext and C_sca at max(bar, 10 x
    spread).  The device's four cross sections equal the oracle's cross
    sections (Result.cpp:557-761) of the device's own coefficients to 1e-12,
    which pins their evaluation.  C_abs and C_sca,SH then differ from the
    reference's only through the coefficients: C_abs weights the high orders
    as Iaux does, so its sampled spread varies thirtyfold between samples and
    makes no bar; both are printed beside their spread."""
    keys = ("cext", "csca", "cabs", "csca_sh")
    csv = [cs[k] for k in keys] if isinstance(cs, dict) else list(cs)
    errs = {"x_sca": rel(out["x_sca"], ref.x), "x_sca_sh": rel(out["x_sca_sh"], ref.xsh)}
    for k, key in enumerate(keys):
        errs[key] = relabs(csv[k], ref.cs[k])
    print(f"[{name}] " + ", ".join(f"{k} {errs[k]:.1e} (spread {ref.spread[k]:.1e})" for k in errs), flush=True)
    for key in ("x_sca", "x_sca_sh", "cext", "csca"):
        assert errs[key] < ref.tol(key, bar), (name, key, errs[key], ref.spread[key])
    own = ora.cross_sections(c, out["x_sca"], out["x_sca_sh"])
    for k, key in enumerate(keys):
        assert relabs(csv[k], own[k]) < 1e-12, (name, key, "evaluation", relabs(csv[k], own[k]))


# ------------------------------------------------------------------ config 2
_TABLES = {}
_CHAINS = {}


def ora_tables(ora, nmax):
    if nmax not in _TABLES:
        _TABLES[nmax] = ora.cg_tables(nmax)
    return _TABLES[nmax]


@pytest.fixture(scope="module")
def config2_ref(ora):
    c = problem.config2()
    print(f"\n[config2] P={c.nobj} nmax={c.nmax} N={c.N}: oracle chain", flush=True)
    return c, Chain(ora, c, ora_tables(ora, c.nmax), keep=True)


@pytest.mark.parametrize("kind", ["lu64", "lu32ir", "auto"])
def test_config2_full_size(ora, config2_ref, kind):
    """64 Drude spheres, nmax 8, N = 10240, omega and 2 omega: the device's
    whole S, V and Q against the oracle's, then x_sca, x_sca_sh and the cross
    sections against the oracle's matrices solved by LAPACK.  "auto" (the
    benchmark's path: S by FP64 GMRES beside V's mixed LU) is held to the
    mixed bar and must have taken the GMRES branch."""
    from optimet_amd import solver
    c, ref = config2_ref
    s = solver.Solver(c, kind=kind)
    try:
        s.update()
        print(f"[config2 {kind}] assembled; comparing S, V, Q", flush=True)
        assert rel(s.vector("Q"), ref.Q) < 1e-12
        if kind == "lu64":
            assert rel(s.matrix(), ref.S) < 1e-12
            assert rel(s.matrix(sh=True), ref.V) < 1e-12
        out = s.solve()
        cs = s.cross_sections()
        if kind == "auto":
            assert s.stats()["pivot_fallbacks"] == 0
        check_solution(f"config2 {kind}", out, cs, ref, TOL["lu32ir" if kind == "auto" else kind], ora, c)
        assert rel(out["x_int"], out["x_sca"] * iaux(ora, c)) < 1e-12
        # C_abs,SH of the device's own internal coefficients (Result.cpp:763-794)
        cabs_sh = ora.absorption_sh(c, ora_tables(ora, c.nmax), out["x_int"], out["x_int_sh"])
        assert relabs(cs["cabs_sh"], cabs_sh) < 1e-12, relabs(cs["cabs_sh"], cabs_sh)
        # the second-harmonic source of the device's own x_int (Geometry.cpp:250-310)
        K, _ = ora.source_sh(c, ora_tables(ora, c.nmax), out["x_int"])
        assert rel(s.vector("K"), K) < 1e-11, rel(s.vector("K"), K)
    finally:
        s.ctx.close()


# ------------------------------------------------------------------ config 3
@pytest.mark.parametrize("kind", ["lu32ir", "lu64", "auto"])
def test_config3_sweep_full(ora, kind):
    """16 Drude spheres, nmax 10, N = 3840: the whole 256-wavelength batched
    sweep at omega and 2 omega; six wavelengths (both ends, the ones with the
    most refinement iterations at omega and at 2 omega, two inside) against
    the oracle and LAPACK.  "auto" (the benchmark's path: S by lockstep FP64
    GMRES beside V's batched mixed LU) is held to the mixed bar."""
    from optimet_amd import solver
    c = problem.config3_cluster()
    lams = problem.config3_wavelengths()
    s = solver.Solver(c, kind=kind)
    try:
        out = s.sweep(lams, outputs=("x_sca", "x_sca_sh"))
    finally:
        s.ctx.close()
    its = out["iterations"]
    pick = sorted({0, len(lams) - 1, int(np.argmax(its[:, 0])), int(np.argmax(its[:, 1])), 97, 181})
    print(f"\n[config3 {kind}] swept 256 wavelengths, max iterations {its.max(axis=0)}; checking {pick}", flush=True)
    for i in pick:
        if ("config3", i) not in _CHAINS:
            _CHAINS[("config3", i)] = Chain(ora, c.with_wavelength(lams[i]), ora_tables(ora, c.nmax))
        ref = _CHAINS[("config3", i)]
        res = {"x_sca": out["x_sca"][i], "x_sca_sh": out["x_sca_sh"][i]}
        check_solution(f"config3 {kind} lambda[{i}]", res, out["xs"][i], ref, TOL["lu32ir" if kind == "auto" else kind],
                       ora, c.with_wavelength(lams[i]))


# ------------------------------------------------------------------ config 4
def test_config4_full_size_properties(ora):
    """512 Drude spheres, nmax 6, N = 49152 (38.7 GB FP64 matrix): 64 sampled
    blocks of S and the whole Q against the oracle; the mixed solution's
    residual on the oracle's matrix; the mixed LU + refinement and
    unpreconditioned FP64 GMRES agree; C_ext from the device equals the
    oracle's C_ext of the device's solution."""
    from optimet_amd import solver
    c = problem.config4()
    ctx = solver.Context()
    try:
        s = solver.Solver(c, kind="lu32ir", context=ctx)
        s.update()
        Q = ora.source_vector(c)
        assert rel(s.vector("Q"), Q) < 1e-12
        rng = np.random.default_rng(4)
        pairs = [(int(i), int(j)) for i, j in rng.integers(0, c.nobj, size=(64, 2))]
        pairs[0] = (7, 7)  # a diagonal (identity) block
        blocks = ora.scattering_blocks(c, pairs)
        for (i, j), b in zip(pairs, blocks):
            assert rel(s.matrix_block(i, j), b) < 1e-12, (i, j)
        print(f"\n[config4] N={c.N}: Q and 64 blocks match; solving (lu32ir)", flush=True)
        x = s.solve()["x_sca"]
        cs = s.cross_sections()
        rr = np.linalg.norm(ora.residual(c, x, Q)) / np.linalg.norm(Q)
        print(f"[config4] residual of x on the oracle's S {rr:.2e}", flush=True)
        assert rr < 1e-10
        assert relabs(cs["cext"], ora.cross_sections(c, x)[0]) < 1e-10
        g = solver.Solver(c, kind="gmres64", tol=1e-14, context=ctx)
        g.update()
        xg = g.solve()["x_sca"]
        print(f"[config4] gmres64 vs lu32ir {rel(xg, x):.2e}", flush=True)
        assert rel(xg, x) < 1e-9
    finally:
        ctx.close()


# ------------------------------------------------------------------ config 5
def _config5_sample():
    """From the seeded 2000-request stream: for each nmax in 4..10 the first
    request with N <= 4000 (full oracle + LAPACK comparison), and three large
    requests (P 83-128, nmax 8-10, N 17440-20480) checked by residuals; each
    truncated to its first and last wavelength."""
    reqs = problem.serving_requests()
    small, big = {}, [3, 4, 32]
    for r in reqs:
        c = r["cluster"]
        if c.N <= 4000 and c.nmax not in small:
            small[c.nmax] = r["id"]
        if len(small) == 7:
            break
    ids = sorted(small.values()) + big
    out = []
    for rid in ids:
        r = reqs[rid]
        lams = r["wavelengths"]
        out.append((rid, r["cluster"], np.array([lams[0], lams[-1]]) if len(lams) > 1 else lams))
    return out


def test_config5_served_sample(ora):
    """Requests served concurrently (three slots, footprint admission, the
    automatic choice per request) give the oracle's answer: full chain where
    N <= 4000 (batched mixed-LU sweeps), residuals of x_sca on S
    and of x_sca_sh on V (with the oracle's SH source from the served x_sca)
    otherwise."""
    from optimet_amd.serving import Request, Server
    sample = _config5_sample()
    assert sorted({c.nmax for _, c, _ in sample}) == list(range(4, 11))
    assert max(c.nobj for _, c, _ in sample) == 128
    reqs = [Request(rid, 0.0, c, lams) for rid, c, lams in sample]
    srv = Server(slots=3, kind="auto")  # the large requests go through the mixed Krylov
    try:
        results = srv.run(reqs, realtime=False)
    finally:
        srv.close()
    for (rid, c, lams), res in zip(sample, results):
        assert res.error is None, (rid, res.error)
        tables = ora_tables(ora, c.nmax)
        for l, lam in enumerate(lams):
            cl = c.with_wavelength(lam)
            x, xsh = res.out["x_sca"][l], res.out["x_sca_sh"][l]
            if c.N <= 4000:
                ref = Chain(ora, cl, tables)
                check_solution(f"request {rid} lambda {lam:.3e}", {"x_sca": x, "x_sca_sh": xsh}, res.out["xs"][l], ref,
                               TOL["lu32ir"], ora, cl)
            else:
                Q = ora.source_vector(cl)
                rs = np.linalg.norm(ora.residual(cl, x, Q)) / np.linalg.norm(Q)
                K, _ = ora.source_sh(cl, tables, x * iaux(ora, cl))
                rv = np.linalg.norm(ora.residual(cl, xsh, K, sh=True)) / np.linalg.norm(K)
                print(f"\n[config5] request {rid} P={c.nobj} nmax={c.nmax} N={c.N}: residuals {rs:.1e} {rv:.1e}",
                      flush=True)
                assert rs < 1e-10 and rv < 1e-10, (rid, rs, rv)
