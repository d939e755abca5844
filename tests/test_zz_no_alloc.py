"""GPU: omx_dense_solve_noalloc (no device allocation; the library's scratch
array holds every work buffer) against omx_dense_solve with the FP64 LU and
partial pivoting, on two calls in a row, the second with a smaller system.

Device buffers come from the HIP runtime the library itself links, through
ctypes (the test is the caller that owns the inputs and outputs)."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2


def _hip():
    from optimet_amd import solver
    solver.load_library()  # loads libamdhip64 as the library's dependency
    hip = ctypes.CDLL("libamdhip64.so.7")
    hip.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    hip.hipFree.argtypes = [ctypes.c_void_p]
    hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    return hip


class _Dev:
    def __init__(self, hip, nbytes):
        self.hip, self.p = hip, ctypes.c_void_p()
        assert hip.hipMalloc(ctypes.byref(self.p), nbytes) == 0
        self.nbytes = nbytes

    def put(self, a):
        a = np.ascontiguousarray(a)
        assert a.nbytes <= self.nbytes
        assert self.hip.hipMemcpy(self.p, a.ctypes.data, a.nbytes, H2D) == 0

    def get(self, shape, dtype):
        out = np.empty(shape, dtype)
        assert self.hip.hipMemcpy(out.ctypes.data, self.p, out.nbytes, D2H) == 0
        return out

    def free(self):
        if self.p:
            self.hip.hipFree(self.p)
            self.p = None


def _system(N, seed):
    rng = np.random.default_rng(seed)
    A = rng.standard_normal((N, N)) + 1j * rng.standard_normal((N, N))
    b = rng.standard_normal(N) + 1j * rng.standard_normal(N)
    return A, b


def rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def test_noalloc_matches_dense_solve_twice_second_smaller():
    """Two calls in a row (N = 1000, then N = 300 in the same scratch, whose
    status words, pivots and staged vectors still hold the first call's values): each x equals the
    allocating entry point's FP64 partial-pivoting solution to 1e-12 (the
    same kernels run on the same data) and LAPACK's to 1e-10 (the FP64 bar)."""
    from optimet_amd import solver
    hip = _hip()
    ctx = solver.Context()
    bufs = []
    try:
        ctx.set_pivoting("partial")
        for N, seed, lda_extra in ((1000, 1, 64), (300, 2, 0)):
            A, b = _system(N, seed)
            ref = ctx.dense_solve(A, b, "lu64")
            Np = -(-N // 64) * 64
            lda = Np + lda_extra
            colmajor = np.zeros((Np, lda), np.complex128)  # row c of this array is column c of A
            colmajor[:N, :N] = A.T
            dA, db, dx = _Dev(hip, colmajor.nbytes), _Dev(hip, 16 * N), _Dev(hip, 16 * N)
            bufs += [dA, db, dx]
            dA.put(colmajor)
            db.put(b)
            ctx.dense_solve_noalloc(N, dA.p.value, lda, db.p.value, dx.p.value)
            x = dx.get(N, np.complex128)
            assert rel(x, ref) < 1e-12, (N, rel(x, ref))
            assert rel(x, np.linalg.solve(A, b)) < 1e-10, N
    finally:
        ctx.set_pivoting("block")
        for d in bufs:
            d.free()
        ctx.close()


def test_noalloc_rejects_sizes_beyond_its_scratch():
    """N above 4096 (the scratch array's limit) and lda below N rounded up to
    64 are argument errors, reported before any device work."""
    from optimet_amd import solver
    hip = _hip()
    ctx = solver.Context()
    d = _Dev(hip, 256)
    try:
        with pytest.raises(solver.OmxError) as e:
            ctx.dense_solve_noalloc(4097, d.p.value, 4160, d.p.value, d.p.value)
        assert e.value.code == -1 and "4096" in str(e.value)
        with pytest.raises(solver.OmxError) as e:
            ctx.dense_solve_noalloc(100, d.p.value, 100, d.p.value, d.p.value)
        assert e.value.code == -1 and "lda" in str(e.value)
    finally:
        d.free()
        ctx.close()
