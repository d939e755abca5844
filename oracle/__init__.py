"""TEST INFRASTRUCTURE ONLY: ctypes front end of the CPU FP64 oracle.

The oracle (``optimet_oracle.cpp``) restates the reference's srcAna hot path on
the CPU.  Only ``tests/``, ``__graft_entry__.smoke()`` and the CPU-baseline leg
of ``bench.py`` may import this module; it is the checker, never the product.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboptimet_oracle.so")
_lib = None


def build():
    """Compile the oracle library in place."""
    subprocess.check_call(["make", "-s", "-C", _HERE, "liboptimet_oracle.so"])


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        _lib = ctypes.CDLL(_LIB_PATH)
        _declare(_lib)
    return _lib


class _Problem(ctypes.Structure):
    _fields_ = [
        ("nobj", ctypes.c_int),
        ("nmax", ctypes.c_int),
        ("sh", ctypes.c_int),
        ("xyz", ctypes.c_void_p),
        ("radius", ctypes.c_void_p),
        ("eps_r", ctypes.c_void_p),
        ("mu_r", ctypes.c_void_p),
        ("eps_r_sh", ctypes.c_void_p),
        ("ksippp", ctypes.c_void_p),
        ("ksiparppar", ctypes.c_void_p),
        ("gamma", ctypes.c_void_p),
        ("bg_eps_r", ctypes.c_double * 2),
        ("bg_mu_r", ctypes.c_double * 2),
        ("lam", ctypes.c_double),
        ("theta_k", ctypes.c_double),
        ("phi_k", ctypes.c_double),
        ("etheta", ctypes.c_double * 2),
        ("ephi", ctypes.c_double * 2),
    ]


_P = ctypes.POINTER(_Problem)
_D = ctypes.POINTER(ctypes.c_double)
_I = ctypes.c_int
_Dbl = ctypes.c_double


def _declare(L):
    L.ora_last_error.restype = ctypes.c_char_p
    L.ora_bessel.argtypes = [_I, _Dbl, _Dbl, _I, _D, _D]
    L.ora_ynm.argtypes = [_I, _I, _Dbl, _Dbl, _D]
    L.ora_tac.argtypes = [_D, _Dbl, _Dbl, _I, _I, _I, _I, _I, _D]
    L.ora_coupling.argtypes = [_D, _Dbl, _Dbl, _I, _I, _D, _D]
    L.ora_object_vector.argtypes = [_P, _I, _I, _D]
    L.ora_incident.argtypes = [_P, _D, _D]
    L.ora_inc_local.argtypes = [_P, _I, _D]
    L.ora_source_vector.argtypes = [_P, _D]
    L.ora_scattering_matrix.argtypes = [_P, _I, _D]
    L.ora_dense_solve.argtypes = [_I, _D, _D, _D, _I]
    L.ora_scattering_blocks.argtypes = [_P, _I, _I, ctypes.POINTER(ctypes.c_int), _D]
    L.ora_residual.argtypes = [_P, _I, _D, _D, _D]
    L.ora_cg_tables.argtypes = [_I, _I, _D]
    L.ora_wigner3j.argtypes = [_I] * 6
    L.ora_wigner3j.restype = _Dbl
    L.ora_wigner6j.argtypes = [_I] * 6
    L.ora_wigner6j.restype = _Dbl
    L.ora_wigner9j.argtypes = [_I] * 9
    L.ora_wigner9j.restype = _Dbl
    L.ora_source_sh.argtypes = [_P, _D, _D, _D, _D]
    L.ora_solve_all.argtypes = [_P, _I, _D, _D, _D, _D, _D]
    L.ora_cross_sections.argtypes = [_P, _D, _D, _D]
    L.ora_absorption_sh.argtypes = [_P, _D, _D, _D, _D]
    L.ora_fields.argtypes = [_P, _D, _D, _D, _D, _D, _I, _D, _D, _D, _D, _D]
    L.ora_time_sample.argtypes = [_P, _I, _I, _I, _D]
    L.ora_aca_block.argtypes = [_I, _D, _Dbl, ctypes.POINTER(ctypes.c_int), _D, _D, _D]
    L.ora_aca_operator.argtypes = [_P, _I, _Dbl, _D, ctypes.POINTER(ctypes.c_int)]
    L.ora_gmres_zcomp.argtypes = [_I, _D, _D, _Dbl, _I, _I, _D, _D]


def _ptr(a):
    return a.ctypes.data_as(_D)


def _check(rc):
    if rc != 0:
        raise RuntimeError("oracle: " + lib().ora_last_error().decode())


def _cz(n):
    return np.zeros(n, dtype=np.complex128)


class _Pinned:
    """Keeps the numpy buffers of a problem alive while the struct is in use."""

    def __init__(self, prob):
        P = prob.nobj
        self.keep = []

        def arr(x, dtype, n):
            a = np.ascontiguousarray(np.broadcast_to(np.asarray(x, dtype=dtype), (n,)).copy())
            self.keep.append(a)
            return a.ctypes.data

        s = _Problem()
        s.nobj = P
        s.nmax = prob.nmax
        s.sh = 1 if prob.sh else 0
        s.xyz = arr(np.asarray(prob.xyz, dtype=np.float64).reshape(-1), np.float64, 3 * P)
        s.radius = arr(prob.radius, np.float64, P)
        s.eps_r = arr(prob.eps_r, np.complex128, P)
        s.mu_r = arr(prob.mu_r, np.complex128, P)
        s.eps_r_sh = arr(prob.eps_r_sh, np.complex128, P)
        s.ksippp = arr(prob.ksippp, np.complex128, P)
        s.ksiparppar = arr(prob.ksiparppar, np.complex128, P)
        s.gamma = arr(prob.gamma, np.complex128, P)
        s.bg_eps_r[0], s.bg_eps_r[1] = prob.bg_eps_r.real, prob.bg_eps_r.imag
        s.bg_mu_r[0], s.bg_mu_r[1] = prob.bg_mu_r.real, prob.bg_mu_r.imag
        s.lam = prob.wavelength
        s.theta_k = prob.theta_k
        s.phi_k = prob.phi_k
        s.etheta[0], s.etheta[1] = prob.etheta.real, prob.etheta.imag
        s.ephi[0], s.ephi[1] = prob.ephi.real, prob.ephi.imag
        self.s = s

    def ref(self):
        return ctypes.byref(self.s)


def bessel(kind, z, nmax):
    """kind 0: j_n, 1: h_n^(1); returns (data, ddata) for orders 0..nmax."""
    d, dd = _cz(nmax + 1), _cz(nmax + 1)
    _check(lib().ora_bessel(kind, float(np.real(z)), float(np.imag(z)), nmax,
                            _ptr(d.view(np.float64)), _ptr(dd.view(np.float64))))
    return d, dd


def ynm(n, m, theta, phi):
    out = _cz(1)
    _check(lib().ora_ynm(n, m, theta, phi, _ptr(out.view(np.float64))))
    return out[0]


def tac(R_sph, k, regular, n, m, l, kk):
    R = np.asarray(R_sph, dtype=np.float64)
    out = _cz(1)
    _check(lib().ora_tac(_ptr(R), k.real, k.imag, int(regular), n, m, l, kk, _ptr(out.view(np.float64))))
    return out[0]


def coupling(R_sph, k, nmax, hankel):
    """Coupling(R, k, nmax, hankel) -> (A, B), each nf x nf with A[p, q].

    ``hankel`` is Coupling's own ``regular`` argument (Coupling.cpp:80-87):
    True builds the irregular (Hankel) coefficients used in S, False the
    regular (Bessel) ones used for the incident field.
    """
    nf = nmax * (nmax + 2)
    R = np.asarray(R_sph, dtype=np.float64)
    A, B = _cz(nf * nf), _cz(nf * nf)
    k = complex(k)
    _check(lib().ora_coupling(_ptr(R), k.real, k.imag, nmax, int(hankel),
                              _ptr(A.view(np.float64)), _ptr(B.view(np.float64))))
    return A.reshape(nf, nf), B.reshape(nf, nf)


OBJ_T, OBJ_TSH, OBJ_IAUX, OBJ_IAUXSH1, OBJ_IAUXSH2, OBJ_TSH1_OUTER, OBJ_TSH2_OUTER = range(7)


def object_vector(prob, obj, which):
    nf = prob.nmax * (prob.nmax + 2)
    out = _cz(2 * nf)
    p = _Pinned(prob)
    _check(lib().ora_object_vector(p.ref(), obj, which, _ptr(out.view(np.float64))))
    return out


def incident(prob):
    nf = prob.nmax * (prob.nmax + 2)
    a, b = _cz(nf), _cz(nf)
    p = _Pinned(prob)
    _check(lib().ora_incident(p.ref(), _ptr(a.view(np.float64)), _ptr(b.view(np.float64))))
    return a, b


def inc_local(prob, obj):
    nf = prob.nmax * (prob.nmax + 2)
    out = _cz(2 * nf)
    p = _Pinned(prob)
    _check(lib().ora_inc_local(p.ref(), obj, _ptr(out.view(np.float64))))
    return out


def source_vector(prob):
    out = _cz(prob.N)
    p = _Pinned(prob)
    _check(lib().ora_source_vector(p.ref(), _ptr(out.view(np.float64))))
    return out


def scattering_matrix(prob, sh=False):
    """S (sh=False) or V (sh=True) as an N x N array (row index first)."""
    N = prob.N
    out = _cz(N * N)
    p = _Pinned(prob)
    _check(lib().ora_scattering_matrix(p.ref(), int(sh), _ptr(out.view(np.float64))))
    return out.reshape(N, N).T  # stored column-major


def scattering_blocks(prob, pairs, sh=False):
    """Blocks (i, j) of S (or V) for the given object pairs: array (n, 2nf, 2nf),
    row index first, as scattering_matrix(prob)[i*2nf:(i+1)*2nf, j*2nf:(j+1)*2nf]."""
    nf = prob.nmax * (prob.nmax + 2)
    ij = np.ascontiguousarray(np.asarray(pairs, dtype=np.int32).reshape(-1, 2))
    n = ij.shape[0]
    out = _cz(n * 4 * nf * nf)
    p = _Pinned(prob)
    _check(lib().ora_scattering_blocks(p.ref(), int(sh), n, ij.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
                                       _ptr(out.view(np.float64))))
    return out.reshape(n, 2 * nf, 2 * nf).transpose(0, 2, 1)  # each block stored column-major


def residual(prob, x, b, sh=False):
    """b - S x (or b - V x) with the oracle's matrix assembled block row by block row."""
    N = prob.N
    xx = np.ascontiguousarray(x, dtype=np.complex128)
    bb = np.ascontiguousarray(b, dtype=np.complex128)
    r = _cz(N)
    p = _Pinned(prob)
    _check(lib().ora_residual(p.ref(), int(sh), _ptr(xx.view(np.float64)), _ptr(bb.view(np.float64)),
                              _ptr(r.view(np.float64))))
    return r


def dense_solve(S, b, method=0):
    N = S.shape[0]
    Sc = np.asfortranarray(S, dtype=np.complex128)
    flat = Sc.reshape(-1, order="F").copy()
    bb = np.ascontiguousarray(b, dtype=np.complex128)
    x = _cz(N)
    _check(lib().ora_dense_solve(N, _ptr(flat.view(np.float64)), _ptr(bb.view(np.float64)),
                                 _ptr(x.view(np.float64)), method))
    return x


def aca_block(A, eps=1e-3, trace=False):
    """ACA_compression of the square block A: (U (n, r), V (r, n)) and, with
    trace, the (r, 3) per-step ratios: runner-up / largest modulus of the
    pivot column's choice and of the next row's choice, and the stopping
    ratio (>= 1 stops)."""
    n = A.shape[0]
    flat = np.asfortranarray(A, dtype=np.complex128).reshape(-1, order="F").copy()
    U, V = _cz(n * n), _cz(n * n)
    tr = np.zeros(3 * n)
    r = ctypes.c_int(0)
    _check(lib().ora_aca_block(n, _ptr(flat.view(np.float64)), eps, ctypes.byref(r), _ptr(U.view(np.float64)),
                               _ptr(V.view(np.float64)), _ptr(tr)))
    k = r.value
    out = (U[:n * k].reshape(k, n).T.copy(), V[:n * k].reshape(k, n).copy())
    return out + (tr[:3 * k].reshape(k, 3),) if trace else out


def aca_operator(prob, sh=False, eps=1e-3):
    """The compressed operator (N, N) and the (P, P) ranks (-1 dense, 0 diagonal)."""
    N, P = prob.N, prob.nobj
    S = _cz(N * N)
    rank = np.zeros(P * P, np.int32)
    p = _Pinned(prob)
    _check(lib().ora_aca_operator(p.ref(), int(sh), eps, _ptr(S.view(np.float64)),
                                  rank.ctypes.data_as(ctypes.POINTER(ctypes.c_int))))
    return S.reshape(N, N).T, rank.reshape(P, P)


def gmres_zcomp(A, b, tol=1e-6, maxit=240, restarts=2):
    """Gmres_Zcomp from x = 0: (x, steps, final residual estimate)."""
    N = A.shape[0]
    flat = np.asfortranarray(A, dtype=np.complex128).reshape(-1, order="F").copy()
    bb = np.ascontiguousarray(b, dtype=np.complex128)
    x = _cz(N)
    info = np.zeros(2)
    _check(lib().ora_gmres_zcomp(N, _ptr(flat.view(np.float64)), _ptr(bb.view(np.float64)), tol, maxit, restarts,
                                 _ptr(x.view(np.float64)), _ptr(info)))
    return x, int(info[0]), float(info[1])


def cg_tables(nmax, nmaxs=None):
    nmaxs = nmax if nmaxs is None else nmaxs
    nf, nfs = nmax * (nmax + 2), nmaxs * (nmaxs + 2)
    out = np.zeros(9 * nfs * nf * nf)
    _check(lib().ora_cg_tables(nmax, nmaxs, _ptr(out)))
    return out.reshape(9, nfs, nf, nf)


def wigner3j(*a):
    return lib().ora_wigner3j(*a)


def wigner6j(*a):
    return lib().ora_wigner6j(*a)


def wigner9j(*a):
    return lib().ora_wigner9j(*a)


def source_sh(prob, tables, x_int):
    N = prob.N
    K, K1 = _cz(N), _cz(N)
    t = np.ascontiguousarray(tables, dtype=np.float64).reshape(-1)
    xi = np.ascontiguousarray(x_int, dtype=np.complex128)
    p = _Pinned(prob)
    _check(lib().ora_source_sh(p.ref(), _ptr(t), _ptr(xi.view(np.float64)),
                               _ptr(K.view(np.float64)), _ptr(K1.view(np.float64))))
    return K, K1


def solve_all(prob, method=0):
    """Full update()+solve(): dict with x_sca, x_int, x_sca_sh, x_int_sh, xs."""
    N = prob.N
    xs_, xi_, xss, xis = _cz(N), _cz(N), _cz(N), _cz(N)
    cs = np.zeros(5)
    p = _Pinned(prob)
    _check(lib().ora_solve_all(p.ref(), method, _ptr(xs_.view(np.float64)), _ptr(xi_.view(np.float64)),
                               _ptr(xss.view(np.float64)), _ptr(xis.view(np.float64)), _ptr(cs)))
    out = {"x_sca": xs_, "x_int": xi_, "cext": cs[0], "csca": cs[1], "cabs": cs[2]}
    if prob.sh:
        out.update({"x_sca_sh": xss, "x_int_sh": xis, "csca_sh": cs[3], "cabs_sh": cs[4]})
    return out


def cross_sections(prob, x_sca, x_sca_sh=None):
    cs = np.zeros(4)
    xs = np.ascontiguousarray(x_sca, dtype=np.complex128)
    p = _Pinned(prob)
    if x_sca_sh is None:
        _check(lib().ora_cross_sections(p.ref(), _ptr(xs.view(np.float64)), None, _ptr(cs)))
    else:
        xh = np.ascontiguousarray(x_sca_sh, dtype=np.complex128)
        _check(lib().ora_cross_sections(p.ref(), _ptr(xs.view(np.float64)),
                                        _ptr(xh.view(np.float64)), _ptr(cs)))
    return cs


def absorption_sh(prob, tables, x_int, x_int_sh):
    """Result::getAbsorptionCrossSection_SH from x_int and x_int_sh."""
    t = np.ascontiguousarray(tables, dtype=np.float64).reshape(-1)
    a = np.ascontiguousarray(x_int, dtype=np.complex128)
    b = np.ascontiguousarray(x_int_sh, dtype=np.complex128)
    out = np.zeros(1)
    p = _Pinned(prob)
    _check(lib().ora_absorption_sh(p.ref(), _ptr(t), _ptr(a.view(np.float64)), _ptr(b.view(np.float64)), _ptr(out)))
    return float(out[0])


def fields(prob, tables, sol, xyz):
    """Result::getEHFields (Cartesian components) at points xyz (n, 3):
    dict E, H, E_sh, H_sh, each (n, 3) complex; sol holds x_sca, x_int and,
    with the second harmonic, x_sca_sh, x_int_sh."""
    pts = np.ascontiguousarray(np.asarray(xyz, dtype=np.float64).reshape(-1, 3))
    n = pts.shape[0]
    outs = [_cz(3 * n) for _ in range(4)]
    t = np.ascontiguousarray(tables, dtype=np.float64).reshape(-1)
    keep = {k: np.ascontiguousarray(sol[k], dtype=np.complex128) for k in sol if k.startswith("x_")}
    ptr = lambda k: _ptr(keep[k].view(np.float64)) if k in keep else None
    p = _Pinned(prob)
    _check(lib().ora_fields(p.ref(), _ptr(t), ptr("x_sca"), ptr("x_int"), ptr("x_sca_sh"), ptr("x_int_sh"), n,
                            _ptr(pts), *[_ptr(o.view(np.float64)) for o in outs]))
    return {k: o.reshape(n, 3) for k, o in zip(("E", "H", "E_sh", "H_sh"), outs)}


def time_sample(prob, npairs, nsolve, method=1):
    """(s per pair block, s per dense solve of order nsolve, s per object SH source)."""
    out = np.zeros(3)
    p = _Pinned(prob)
    _check(lib().ora_time_sample(p.ref(), npairs, nsolve, method, _ptr(out)))
    return tuple(out)
