"""Python mirror of the solver boundary, over the C ABI of liboptimet_mi355x.so.

``Solver`` follows optimet::solver::AbstractSolver (srcAna/Solver.h:62-144):
construct it from a cluster (Geometry + Excitation), call ``update()``
(assembly of Q, S and, for the second harmonic, V) and then ``solve()``, which
returns the scattered and internal coefficients at ω and 2ω in the reference's
layout.  The HIP library is the only compute path: if it is missing or no GPU
is present the calls raise, they never fall back to a CPU implementation.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liboptimet_mi355x.so")

LU64, LU32_IR, GMRES64, ACA_GMRES, GMRES32_IR, AUTO = 0, 1, 2, 3, 4, 5
PIVOT_BLOCK, PIVOT_PARTIAL = 0, 1
PIVOTS = {"block": PIVOT_BLOCK, "partial": PIVOT_PARTIAL}
GEMMS = {"bf16x3": 0, "fp32": 1}
KINDS = {"lu64": LU64, "lu32ir": LU32_IR, "gmres64": GMRES64, "aca": ACA_GMRES, "gmres32ir": GMRES32_IR,
         "auto": AUTO}
# Gmres_Zcomp settings of the reference's ACA path: tol 1e-6, 240 steps per
# cycle, 2 cycles (PreconditionedMatrixSolver.h:47-56)
ACA_DEFAULTS = {"tol": 1e-6, "maxit": 480, "restart": 240}

ERRORS = {-1: "argument", -2: "device", -3: "memory", -4: "singular", -5: "state", -6: "no convergence",
          -7: "overlap"}

EXPORTS = ["omx_last_error", "omx_create", "omx_destroy", "omx_set_geometry", "omx_set_excitation",
           "omx_set_solver", "omx_set_cg_tables", "omx_build_cg_tables", "omx_update", "omx_solve",
           "omx_cross_sections", "omx_scattering_size", "omx_get_stats", "omx_read_matrix", "omx_read_vector",
           "omx_dense_solve", "omx_set_profile", "omx_set_pivoting", "omx_set_lu32_gemm", "omx_solve_sweep", "omx_mem_info", "omx_trim",
           "omx_dense_solve_noalloc", "omx_read_matrix_block", "omx_absorption_cross_section_sh",
           "omx_fields", "omx_aca_info", "omx_aca_block", "omx_set_auto_budget", "omx_set_autotune",
           "omx_tune_query", "omx_tune_reset", "omx_set_graphs", "omx_solve_group", "omx_read_solution"]


class OmxError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"[{ERRORS.get(code, code)}] {msg}")
        self.code = code


class Stats(ctypes.Structure):
    _fields_ = [("N", ctypes.c_int), ("Np", ctypes.c_int), ("iterations", ctypes.c_int),
                ("iterations_sh", ctypes.c_int), ("residual", ctypes.c_double), ("residual_sh", ctypes.c_double),
                ("t_update_ms", ctypes.c_double), ("t_factor_ms", ctypes.c_double), ("t_solve_ms", ctypes.c_double),
                ("t_sh_source_ms", ctypes.c_double), ("t_assemble_ms", ctypes.c_double),
                ("device_bytes", ctypes.c_size_t), ("t_lu_gemm_ms", ctypes.c_double),
                ("lu_gemm_flop", ctypes.c_double), ("lu_gemm_launches", ctypes.c_int),
                ("t_lu_panel_ms", ctypes.c_double), ("t_factor_total_ms", ctypes.c_double),
                ("pivot_fallbacks", ctypes.c_int), ("sweep_batch", ctypes.c_int), ("matvec_calls", ctypes.c_int),
                ("t_matvec_ms", ctypes.c_double), ("matvec_bytes", ctypes.c_double),
                ("graph_captures", ctypes.c_int), ("graph_launches", ctypes.c_int)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


_lib = None


def load_library(path=LIB_PATH):
    """Loads the HIP library; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise RuntimeError(f"{path} is missing: build it with `python -m optimet_amd.build`")
    L = ctypes.CDLL(path)
    P, D, I, V = ctypes.c_void_p, ctypes.POINTER(ctypes.c_double), ctypes.c_int, ctypes.c_void_p
    L.omx_last_error.restype = ctypes.c_char_p
    L.omx_create.argtypes = [ctypes.POINTER(ctypes.c_void_p), I]
    L.omx_destroy.argtypes = [P]
    L.omx_destroy.restype = None
    L.omx_set_geometry.argtypes = [P, I, I, D, D, D, D, D, D, D, D, D, D]
    L.omx_set_excitation.argtypes = [P, ctypes.c_double, ctypes.c_double, ctypes.c_double, D, D, I]
    L.omx_set_solver.argtypes = [P, I, ctypes.c_double, I, I]
    L.omx_set_cg_tables.argtypes = [P, I, ctypes.POINTER(D)]
    L.omx_build_cg_tables.argtypes = [P, I]
    L.omx_update.argtypes = [P]
    L.omx_solve.argtypes = [P, D, D, D, D]
    L.omx_cross_sections.argtypes = [P, D]
    L.omx_scattering_size.argtypes = [P]
    L.omx_get_stats.argtypes = [P, ctypes.POINTER(Stats)]
    L.omx_read_matrix.argtypes = [P, I, D]
    L.omx_read_vector.argtypes = [P, I, D]
    L.omx_dense_solve.argtypes = [P, I, D, D, D, I]
    L.omx_set_profile.argtypes = [P, I]
    L.omx_set_pivoting.argtypes = [P, I]
    L.omx_set_lu32_gemm.argtypes = [P, I]
    L.omx_solve_sweep.argtypes = [P, I, D, D, D, I, D, D, D, D, D, ctypes.POINTER(ctypes.c_int)]
    SZ = ctypes.POINTER(ctypes.c_size_t)
    L.omx_mem_info.argtypes = [P, SZ, SZ, SZ]
    L.omx_trim.argtypes = [P]
    L.omx_dense_solve_noalloc.argtypes = [P, I, V, ctypes.c_size_t, V, V]
    L.omx_read_matrix_block.argtypes = [P, I, I, I, I, I, D]
    L.omx_absorption_cross_section_sh.argtypes = [P, D]
    L.omx_fields.argtypes = [P, I, D, D, D, D, D]
    L.omx_aca_info.argtypes = [P, I, ctypes.POINTER(ctypes.c_int), D]
    L.omx_aca_block.argtypes = [P, I, I, I, ctypes.POINTER(ctypes.c_int), D, D]
    L.omx_set_auto_budget.argtypes = [P, I]
    L.omx_set_autotune.argtypes = [P, I]
    L.omx_tune_query.argtypes = [I, I, I, I, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int), D]
    L.omx_tune_reset.argtypes = []
    L.omx_set_graphs.argtypes = [P, I]
    L.omx_solve_group.argtypes = [P, I, ctypes.POINTER(ctypes.c_void_p)]
    L.omx_read_solution.argtypes = [P, D, D, D, D]
    _lib = L
    return L


def _d(a):
    return None if a is None else a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _c(a):
    return np.ascontiguousarray(np.asarray(a, dtype=np.complex128))


def _check(rc):
    if rc != 0:
        raise OmxError(rc, load_library().omx_last_error().decode())


class Context:
    """One device context (one HIP stream, one set of device buffers)."""

    def __init__(self, device=0):
        self.L = load_library()
        h = ctypes.c_void_p()
        _check(self.L.omx_create(ctypes.byref(h), device))
        self.h = h

    def close(self):
        if self.h:
            self.L.omx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_profile(self, on=True):
        _check(self.L.omx_set_profile(self.h, 1 if on else 0))

    def stats(self):
        s = Stats()
        _check(self.L.omx_get_stats(self.h, ctypes.byref(s)))
        return s.as_dict()

    def mem_info(self):
        """(free, total, held) device bytes."""
        f, t, h = ctypes.c_size_t(), ctypes.c_size_t(), ctypes.c_size_t()
        _check(self.L.omx_mem_info(self.h, ctypes.byref(f), ctypes.byref(t), ctypes.byref(h)))
        return f.value, t.value, h.value

    def trim(self):
        """Release the matrix-sized device buffers."""
        _check(self.L.omx_trim(self.h))

    def set_graphs(self, on):
        """Factorisations captured in HIP graphs once per state and replayed (default off)."""
        _check(self.L.omx_set_graphs(self.h, 1 if on else 0))

    def set_autotune(self, on):
        """Batched sweeps tune their products and triangular solves per shape (default on)."""
        _check(self.L.omx_set_autotune(self.h, 1 if on else 0))

    def set_auto_budget(self, steps):
        """Outer Krylov steps the "auto" kind tries before the LU (0: the cost model)."""
        _check(self.L.omx_set_auto_budget(self.h, int(steps)))

    def set_pivoting(self, mode):
        """"block" (default) or "partial" (OMX_PIVOT_*)."""
        _check(self.L.omx_set_pivoting(self.h, PIVOTS[mode]))

    def set_lu32_gemm(self, mode):
        """"bf16x3" (default: FP32 operands split in three bf16 terms on the
        bf16 matrix cores) or "fp32" (the FP32 matrix cores): the products of
        the mixed-precision factorisation's bulk updates (OMX_GEMM_*)."""
        _check(self.L.omx_set_lu32_gemm(self.h, GEMMS[mode]))

    def set_solver(self, kind="lu64", tol=1e-16, maxit=300, restart=40):
        """omx_set_solver: solver kind and the iterative parts' settings."""
        _check(self.L.omx_set_solver(self.h, KINDS[kind], tol, maxit, restart))

    def scattering_size(self):
        return self.L.omx_scattering_size(self.h)

    def dense_solve(self, A, b, kind="lu64"):
        A = np.asfortranarray(A, dtype=np.complex128)
        N = A.shape[0]
        flat = A.reshape(-1, order="F").copy()
        bb = _c(b)
        x = np.zeros(N, dtype=np.complex128)
        _check(self.L.omx_dense_solve(self.h, N, _d(flat.view(np.float64)), _d(bb.view(np.float64)),
                                      _d(x.view(np.float64)), KINDS[kind]))
        return x

    def dense_solve_noalloc(self, N, A_dev, lda, b_dev, x_dev):
        """omx_dense_solve_noalloc on device addresses (ints): A (lda x Np,
        column-major, overwritten by its factors), b and x (N complex each)."""
        _check(self.L.omx_dense_solve_noalloc(self.h, int(N), ctypes.c_void_p(A_dev), int(lda),
                                              ctypes.c_void_p(b_dev), ctypes.c_void_p(x_dev)))


class Solver:
    """MI355X backend behind the AbstractSolver interface (Solver.h:62-144).

    kind: "lu64" (FP64 LU), "lu32ir" (complex-FP32 MFMA LU + FP64 GMRES-IR),
    "gmres64" (GMRES on the FP64 matrix), "gmres32ir" (FP64 flexible GMRES
    preconditioned by GMRES cycles on the complex-FP32 copy), "auto" (by size:
    "gmres32ir" from N = 16384 on, within the mixed LU's cost, else or on
    failure "lu32ir") or "aca" (<ACA compression="yes"/>: the admissible
    blocks ACA-compressed, GMRES on the compressed operator; tol, maxit and
    restart default to the reference's 1e-6, 480 and 240).
    pivoting: "block" (pivots within each 64 x 64 diagonal block, FP64
    refinement, automatic fallback) or "partial" (whole-column partial
    pivoting; with "lu64" the plain FP64 LU).
    """

    def __init__(self, cluster, kind="lu64", tol=None, maxit=None, restart=None, context=None, pivoting="block"):
        self.ctx = context or Context()
        self.kind = kind
        d = ACA_DEFAULTS if kind == "aca" else {"tol": 1e-16, "maxit": 300, "restart": 40}
        tol = d["tol"] if tol is None else tol
        maxit = d["maxit"] if maxit is None else maxit
        restart = d["restart"] if restart is None else restart
        _check(self.ctx.L.omx_set_solver(self.ctx.h, KINDS[kind], tol, maxit, restart))
        self.ctx.set_pivoting(pivoting)
        self.set_cluster(cluster)

    def set_cluster(self, c):
        """AbstractSolver::update(geometry, incWave) inputs (Solver.h:94-99)."""
        self.cluster = c
        L, h = self.ctx.L, self.ctx.h
        self._keep = [np.ascontiguousarray(c.xyz.reshape(-1)), np.ascontiguousarray(c.radius),
                      _c(c.eps_r), _c(c.mu_r), _c(c.eps_r_sh), _c(c.ksippp), _c(c.ksiparppar), _c(c.gamma),
                      _c([c.bg_eps_r]), _c([c.bg_mu_r])]
        k = self._keep
        args = [_d(k[0]), _d(k[1])] + [_d(a.view(np.float64)) for a in k[2:]]
        _check(L.omx_set_geometry(h, c.nobj, c.nmax, *args))
        et, ep = _c([c.etheta]), _c([c.ephi])
        _check(L.omx_set_excitation(h, c.wavelength, c.theta_k, c.phi_k, _d(et.view(np.float64)),
                                    _d(ep.view(np.float64)), 1 if c.sh else 0))

    def set_cg_tables(self, tables):
        """The CGcoeff argument of solve() (9 tables [k][p][q])."""
        t = [np.ascontiguousarray(tables[i], dtype=np.float64).reshape(-1) for i in range(9)]
        arr = (ctypes.POINTER(ctypes.c_double) * 9)(*[_d(x) for x in t])
        _check(self.ctx.L.omx_set_cg_tables(self.ctx.h, self.cluster.nmax, arr))

    def update(self):
        _check(self.ctx.L.omx_update(self.ctx.h))

    def solve(self):
        """Returns dict x_sca, x_int (and x_sca_sh, x_int_sh with SH)."""
        N = self.cluster.N
        xs, xi = np.zeros(N, np.complex128), np.zeros(N, np.complex128)
        xss, xis = (np.zeros(N, np.complex128), np.zeros(N, np.complex128)) if self.cluster.sh else (None, None)
        _check(self.ctx.L.omx_solve(self.ctx.h, _d(xs.view(np.float64)), _d(xi.view(np.float64)),
                                    None if xss is None else _d(xss.view(np.float64)),
                                    None if xis is None else _d(xis.view(np.float64))))
        out = {"x_sca": xs, "x_int": xi}
        if self.cluster.sh:
            out.update({"x_sca_sh": xss, "x_int_sh": xis})
        return out

    def read_solution(self):
        """The solution of the last solve or group solve, as solve() returns it."""
        N = self.cluster.N
        out = {k: np.zeros(N, np.complex128) for k in ("x_sca", "x_int")}
        if self.cluster.sh:
            out.update({k: np.zeros(N, np.complex128) for k in ("x_sca_sh", "x_int_sh")})
        p = [_d(out[k].view(np.float64)) if k in out else None for k in ("x_sca", "x_int", "x_sca_sh", "x_int_sh")]
        _check(self.ctx.L.omx_read_solution(self.ctx.h, *p))
        return out

    def sweep(self, wavelengths, batch=0, outputs=("x_sca", "x_int", "x_sca_sh", "x_int_sh"), eps=None,
              eps_sh=None):
        """Batched wavelength sweep (omx_solve_sweep): the cluster at every
        wavelength, Drude spheres re-evaluated per wavelength (or the given
        (n_lambda, P) permittivities eps, eps_sh).  Returns a dict with the
        requested (n_lambda, N) coefficient arrays, "xs" (n_lambda, 5: C_ext,
        C_sca, C_abs, C_sca_SH, C_abs_SH) and "iterations" (n_lambda, 2)."""
        c = self.cluster
        lams = np.ascontiguousarray(np.asarray(wavelengths, dtype=np.float64))
        nl, N = lams.size, c.N
        if eps is None:
            eps = np.empty((nl, c.nobj), np.complex128)
            eps_sh = np.empty((nl, c.nobj), np.complex128)
            for i, lam in enumerate(lams):
                ci = c.with_wavelength(lam)
                eps[i], eps_sh[i] = ci.eps_r, ci.eps_r_sh
        eps = np.ascontiguousarray(np.asarray(eps, np.complex128).reshape(nl, c.nobj))
        eps_sh = np.ascontiguousarray(np.asarray(eps_sh, np.complex128).reshape(nl, c.nobj))
        out = {}
        ptrs = []
        for name in ("x_sca", "x_int", "x_sca_sh", "x_int_sh"):
            if name in outputs and (c.sh or not name.endswith("_sh")):
                out[name] = np.zeros((nl, N), np.complex128)
                ptrs.append(_d(out[name].view(np.float64)))
            else:
                ptrs.append(None)
        xs = np.zeros((nl, 5))
        its = np.zeros((nl, 2), np.int32)
        _check(self.ctx.L.omx_solve_sweep(self.ctx.h, nl, _d(lams), _d(eps.view(np.float64)),
                                          _d(eps_sh.view(np.float64)), int(batch), *ptrs, _d(xs),
                                          its.ctypes.data_as(ctypes.POINTER(ctypes.c_int))))
        out["xs"] = xs
        out["iterations"] = its
        return out

    def cross_sections(self):
        """C_ext, C_sca, C_abs, C_sca_SH and, with the second harmonic, C_abs_SH."""
        cs = np.zeros(5)
        _check(self.ctx.L.omx_cross_sections(self.ctx.h, _d(cs)))
        if self.cluster.sh:
            _check(self.ctx.L.omx_absorption_cross_section_sh(self.ctx.h, _d(cs[4:])))
        return {"cext": cs[0], "csca": cs[1], "cabs": cs[2], "csca_sh": cs[3], "cabs_sh": cs[4]}

    def fields(self, xyz, sh=None):
        """E and H (Cartesian components) at points xyz (n, 3) in metres, at
        omega and, for a second-harmonic solve (or sh=True), at 2 omega:
        dict E, H (and E_sh, H_sh), each (n, 3) complex (omx_fields)."""
        pts = np.ascontiguousarray(np.asarray(xyz, dtype=np.float64).reshape(-1, 3))
        n = pts.shape[0]
        sh = self.cluster.sh if sh is None else sh
        out = {k: np.zeros((n, 3), np.complex128) for k in (("E", "H", "E_sh", "H_sh") if sh else ("E", "H"))}
        ptr = lambda k: _d(out[k].view(np.float64)) if k in out else None
        _check(self.ctx.L.omx_fields(self.ctx.h, n, _d(pts), ptr("E"), ptr("H"), ptr("E_sh"), ptr("H_sh")))
        return out

    def matrix(self, sh=False):
        N = self.cluster.N
        out = np.zeros(N * N, np.complex128)
        _check(self.ctx.L.omx_read_matrix(self.ctx.h, 1 if sh else 0, _d(out.view(np.float64))))
        return out.reshape(N, N).T

    def aca_info(self, sh=False):
        """The compressed operator of kind "aca" after update(): (P, P) ranks
        (r, -1 dense, 0 diagonal) and a dict of its sizes."""
        P = self.cluster.nobj
        ranks = np.zeros(P * P, np.int32)
        info = np.zeros(7)
        _check(self.ctx.L.omx_aca_info(self.ctx.h, 1 if sh else 0, ranks.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
                                       _d(info)))
        keys = ("bytes", "dense_bytes", "reference_mb", "compressed_blocks", "dense_blocks", "max_rank", "mean_rank")
        return ranks.reshape(P, P), dict(zip(keys, info.tolist()))

    def aca_block(self, i, j, sh=False):
        """Block (i, j) of the compressed operator: (U (n, r), V (r, n)) of a
        compressed block, (block, None) of a dense one, (None, None) on the
        diagonal."""
        n = 2 * self.cluster.nf
        U, V = np.zeros(n * n, np.complex128), np.zeros(n * n, np.complex128)
        r = ctypes.c_int(0)
        _check(self.ctx.L.omx_aca_block(self.ctx.h, 1 if sh else 0, int(i), int(j), ctypes.byref(r),
                                        _d(U.view(np.float64)), _d(V.view(np.float64))))
        k = r.value
        if k < 0:
            return U.reshape(n, n).T.copy(), None
        if k == 0:
            return None, None
        return U[:n * k].reshape(k, n).T.copy(), V[:n * k].reshape(k, n).copy()

    def matrix_block(self, i, j, sh=False):
        """Block (row object i, column object j) of S (or V): 2nf x 2nf."""
        n2 = 2 * self.cluster.nf
        out = np.zeros(n2 * n2, np.complex128)
        _check(self.ctx.L.omx_read_matrix_block(self.ctx.h, 1 if sh else 0, i * n2, j * n2, n2, n2,
                                                _d(out.view(np.float64))))
        return out.reshape(n2, n2).T

    def vector(self, which):
        """which: "Q", "K", "K1", "inc"."""
        N = self.cluster.N
        out = np.zeros(N, np.complex128)
        idx = {"Q": 0, "K": 1, "K1": 2, "inc": 3}[which]
        _check(self.ctx.L.omx_read_vector(self.ctx.h, idx, _d(out.view(np.float64))))
        return out

    def stats(self):
        return self.ctx.stats()

    def scattering_size(self):
        return self.ctx.L.omx_scattering_size(self.ctx.h)


def tune_query(Np, nb, fp64=False, device=0):
    """The sweep autotuner's choice for a shape, or None when not tuned:
    dict split (bool), trsv (variant), times_ms (product probes split-bf16
    and FP32 MFMA, then the triangular-solve variants)."""
    L = load_library()
    sp, tv = ctypes.c_int(0), ctypes.c_int(0)
    t = np.zeros(5)
    rc = L.omx_tune_query(device, Np, nb, 1 if fp64 else 0, ctypes.byref(sp), ctypes.byref(tv), _d(t))
    if rc == -5:
        return None
    _check(rc)
    return {"split": bool(sp.value), "trsv": tv.value, "times_ms": t.tolist()}


def tune_reset():
    """Forget the tuned choices of this process."""
    load_library().omx_tune_reset()


def solve_group(solvers, leader=None):
    """omx_solve_group: the solvers (each with its own context, updated, one
    LU kind, one SH flag; clusters of any size) solved as one batch in the
    leader context (a new one when None).  Returns each member's solution."""
    solvers = list(solvers)
    if not solvers:
        return []
    ctxs = [s.ctx for s in solvers]
    if len({id(c) for c in ctxs}) != len(ctxs):
        raise ValueError("solve_group: every member needs a context of its own")
    L = leader or Context()
    try:
        arr = (ctypes.c_void_p * len(ctxs))(*[c.h.value for c in ctxs])
        _check(L.L.omx_solve_group(L.h, len(ctxs), arr))
    finally:
        if leader is None:
            L.close()
    return [s.read_solution() for s in solvers]
