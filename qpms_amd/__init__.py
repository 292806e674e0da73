"""ctypes bindings to libqpms_hip.so, the gfx950 multiple-scattering hot path.

The C library (include/qpms_hip.h) is the product; this module only loads it
and wraps a few entry points for Python callers and tests. It never computes
anything itself and raises if the shared library is missing: there is no CPU
fallback.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.join(_HERE, "libqpms_hip.so")

# qpms_types.h:128-210
NORM_KRISTENSSON = 65536 | 8
NORM_DEFAULT = NORM_KRISTENSSON
BESSEL_REGULAR, BESSEL_SINGULAR, HANKEL_PLUS, HANKEL_MINUS = 1, 2, 3, 4


class c_cplx(ctypes.Structure):
    """Layout of C99 `double complex` (two doubles; same SysV classification)."""
    _fields_ = [("re", ctypes.c_double), ("im", ctypes.c_double)]

    @classmethod
    def of(cls, z):
        z = complex(z)
        return cls(z.real, z.imag)

    def value(self):
        return complex(self.re, self.im)


class csph_t(ctypes.Structure):
    _fields_ = [("r", c_cplx), ("theta", ctypes.c_double), ("phi", ctypes.c_double)]


class cart3_t(ctypes.Structure):
    _fields_ = [("x", ctypes.c_double), ("y", ctypes.c_double), ("z", ctypes.c_double)]


class cart2_t(ctypes.Structure):
    _fields_ = [("x", ctypes.c_double), ("y", ctypes.c_double)]


class trans_calculator(ctypes.Structure):
    """translations.h:64-84"""
    _fields_ = [("normalisation", ctypes.c_int), ("lMax", ctypes.c_int), ("nelem", ctypes.c_size_t),
                ("A_multipliers", ctypes.POINTER(ctypes.POINTER(c_cplx))),
                ("B_multipliers", ctypes.POINTER(ctypes.POINTER(c_cplx))),
                ("e3c", ctypes.c_void_p), ("legendre0", ctypes.POINTER(ctypes.c_double))]


_lib = None


def lib():
    """Load libqpms_hip.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIBPATH):
            raise RuntimeError(f"{LIBPATH} is missing: build it first (python -c 'import __graft_entry__ as g; g.build()')")
        L = ctypes.CDLL(LIBPATH)
        vp, sz, dbl, i = ctypes.c_void_p, ctypes.c_size_t, ctypes.c_double, ctypes.c_int
        L.qpms_trans_calculator_init.restype = ctypes.POINTER(trans_calculator)
        L.qpms_trans_calculator_init.argtypes = [i, i]
        L.qpms_trans_calculator_free.argtypes = [ctypes.POINTER(trans_calculator)]
        L.qpms_trans_calculator_get_AB_arrays.argtypes = [ctypes.POINTER(trans_calculator), vp, vp, sz, sz,
                                                          csph_t, ctypes.c_bool, i]
        L.qpms_trans_calculator_get_AB_arrays_batch_dev.argtypes = [ctypes.POINTER(trans_calculator), sz, vp, vp,
                                                                    vp, i, i, vp, vp, vp]
        L.qpms_hip_build_translation_matrix_dev.argtypes = [ctypes.POINTER(trans_calculator), sz, vp, c_cplx, i,
                                                            vp, vp]
        L.qpms_hip_build_translation_matrix_rows_dev.argtypes = [ctypes.POINTER(trans_calculator), sz, vp, c_cplx,
                                                                 i, sz, sz, vp, sz, vp]
        L.qpms_hip_build_modeproblem_matrix_dev.argtypes = [ctypes.POINTER(trans_calculator), sz, vp, vp, vp,
                                                            c_cplx, vp, vp]
        L.qpms_hip_init.restype = i
        L.qpms_hip_malloc.restype = vp
        L.qpms_hip_malloc.argtypes = [sz]
        L.qpms_hip_free.argtypes = [vp]
        L.qpms_hip_memcpy_h2d.argtypes = [vp, vp, sz]
        L.qpms_hip_memcpy_d2h.argtypes = [vp, vp, sz]
        L.qpms_hip_zgetrf_dev.argtypes = [sz, vp, vp, vp]
        L.qpms_hip_zgetrs_dev.argtypes = [sz, sz, vp, vp, vp, vp]
        L.qpms_hip_solve_mixed_dev.argtypes = [sz, sz, vp, vp, vp, dbl, i, vp]
        L.qpms_hip_version.restype = ctypes.c_char_p
        L.qpms_hip_scattered_E_dev.argtypes = [i, sz, vp, c_cplx, i, vp, sz, vp, vp, i, vp]
        L.qpms_hip_scattered_E_dev.restype = i
        _lib = L
    return _lib


class DeviceBuffer:
    """Device allocation owned by libqpms_hip (hipMalloc)."""

    def __init__(self, nbytes):
        self.nbytes = int(nbytes)
        self.ptr = lib().qpms_hip_malloc(self.nbytes)
        if not self.ptr:
            raise MemoryError(f"device allocation of {nbytes} bytes failed")

    @classmethod
    def from_array(cls, a):
        a = np.ascontiguousarray(a)
        b = cls(a.nbytes)
        lib().qpms_hip_memcpy_h2d(b.ptr, a.ctypes.data, a.nbytes)
        return b

    def to_array(self, dtype, shape):
        out = np.empty(shape, dtype=dtype)
        assert out.nbytes <= self.nbytes
        lib().qpms_hip_memcpy_d2h(out.ctypes.data, self.ptr, out.nbytes)
        return out

    def free(self):
        if self.ptr:
            lib().qpms_hip_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class TransCalculator:
    """qpms_trans_calculator (translations.h:64-106) backed by the GPU."""

    def __init__(self, lMax, norm=NORM_DEFAULT):
        self.c = lib().qpms_trans_calculator_init(int(lMax), int(norm))
        self.lMax = int(lMax)
        self.nelem = self.c.contents.nelem

    def multipliers(self, y, yu):
        """A and B multiplier arrays of (y <- yu) from the calculator struct."""
        c = self.c.contents
        i = y * self.nelem + yu
        out = []
        for arr in (c.A_multipliers, c.B_multipliers):
            a0 = ctypes.cast(arr[i], ctypes.c_void_p).value
            a1 = ctypes.cast(arr[i + 1], ctypes.c_void_p).value
            n = (a1 - a0) // 16
            buf = (ctypes.c_double * (2 * n)).from_address(a0)
            v = np.frombuffer(buf, dtype=np.float64).copy().view(np.complex128)
            out.append(v)
        return out

    def get_AB_arrays(self, kr, theta, phi, r_ge_d=False, J=HANKEL_PLUS):
        ne = self.nelem
        A = np.empty((ne, ne), np.complex128)
        B = np.empty((ne, ne), np.complex128)
        lib().qpms_trans_calculator_get_AB_arrays(self.c, A.ctypes.data, B.ctypes.data, ne, 1,
                                                  csph_t(c_cplx.of(kr), theta, phi), bool(r_ge_d), int(J))
        return A, B

    def get_AB_arrays_batch(self, kr, theta, phi, J=HANKEL_PLUS, r_ge_d=False):
        kr = np.ascontiguousarray(kr, dtype=np.complex128).ravel()
        th = np.ascontiguousarray(theta, dtype=np.float64).ravel()
        ph = np.ascontiguousarray(phi, dtype=np.float64).ravel()
        n, ne = kr.size, self.nelem
        dk, dt, dp = DeviceBuffer.from_array(kr), DeviceBuffer.from_array(th), DeviceBuffer.from_array(ph)
        dA, dB = DeviceBuffer(n * ne * ne * 16), DeviceBuffer(n * ne * ne * 16)
        lib().qpms_trans_calculator_get_AB_arrays_batch_dev(self.c, n, dk.ptr, dt.ptr, dp.ptr, int(r_ge_d), int(J),
                                                            dA.ptr, dB.ptr, None)
        lib().qpms_hip_sync()
        return dA.to_array(np.complex128, (n, ne, ne)), dB.to_array(np.complex128, (n, ne, ne))

    def build_W(self, pos, k, J=HANKEL_PLUS):
        pos = np.ascontiguousarray(pos, dtype=np.float64)
        npart = pos.shape[0]
        L = npart * 2 * self.nelem
        dpos = DeviceBuffer.from_array(pos)
        dW = DeviceBuffer(L * L * 16)
        lib().qpms_hip_build_translation_matrix_dev(self.c, npart, dpos.ptr, c_cplx.of(k), int(J), dW.ptr, None)
        lib().qpms_hip_sync()
        return dW.to_array(np.complex128, (L, L))

    def build_M(self, pos, ttype, tdiag, k):
        pos = np.ascontiguousarray(pos, dtype=np.float64)
        npart = pos.shape[0]
        L = npart * 2 * self.nelem
        dpos = DeviceBuffer.from_array(pos)
        dty = DeviceBuffer.from_array(np.ascontiguousarray(ttype, dtype=np.int32))
        dtd = DeviceBuffer.from_array(np.ascontiguousarray(tdiag, dtype=np.complex128))
        dM = DeviceBuffer(L * L * 16)
        lib().qpms_hip_build_modeproblem_matrix_dev(self.c, npart, dpos.ptr, dty.ptr, dtd.ptr, c_cplx.of(k), dM.ptr,
                                                    None)
        lib().qpms_hip_sync()
        return dM.to_array(np.complex128, (L, L))

    def __del__(self):
        try:
            lib().qpms_trans_calculator_free(self.c)
        except Exception:
            pass


def scattered_E(lMax, pos, k, cvf, where, J=HANKEL_PLUS, parity=False):
    """Field of the coefficient vectors cvf (npart x 2 nelem, default basis:
    electric block, then magnetic) at the points where (npts x 3), waves of
    kind J at wavenumber k: complex (npts, 3) cartesian components
    (qpms_hip_scattered_E_dev)."""
    pos = np.ascontiguousarray(pos, np.float64).reshape(-1, 3)
    where = np.ascontiguousarray(where, np.float64).reshape(-1, 3)
    cvf = np.ascontiguousarray(cvf, np.complex128).reshape(len(pos), -1)
    assert cvf.shape[1] == 2 * lMax * (lMax + 2)
    if len(where) == 0:   # nothing to launch; the entry point is still called with no buffers
        assert lib().qpms_hip_scattered_E_dev(lMax, len(pos), None, c_cplx.of(k), J, None, 0, None, None,
                                              int(parity), None) == 0
        return np.zeros((0, 3), np.complex128)
    dp, dc, dw = DeviceBuffer.from_array(pos), DeviceBuffer.from_array(cvf), DeviceBuffer.from_array(where)
    do = DeviceBuffer(len(where) * 48)
    rc = lib().qpms_hip_scattered_E_dev(lMax, len(pos), dp.ptr, c_cplx.of(k), J, dc.ptr, len(where), dw.ptr, do.ptr,
                                        int(parity), None)
    assert rc == 0
    out = do.to_array(np.complex128, (len(where), 3))
    for b in (dp, dc, dw, do):
        b.free()
    return out


def zgetrf(A):
    """LU on the GPU: returns (LU, ipiv 1-based, info)."""
    A = np.ascontiguousarray(A, dtype=np.complex128)
    n = A.shape[0]
    dA = DeviceBuffer.from_array(A)
    dp = DeviceBuffer(4 * n)
    info = lib().qpms_hip_zgetrf_dev(n, dA.ptr, dp.ptr, None)
    return dA.to_array(np.complex128, (n, n)), dp.to_array(np.int32, (n,)), info


def zgetrs(LU, ipiv, B):
    LU = np.ascontiguousarray(LU, dtype=np.complex128)
    B = np.ascontiguousarray(B, dtype=np.complex128)
    n = LU.shape[0]
    nrhs = 1 if B.ndim == 1 else B.shape[1]
    dA, dp, dB = DeviceBuffer.from_array(LU), DeviceBuffer.from_array(np.asarray(ipiv, np.int32)), DeviceBuffer.from_array(B)
    lib().qpms_hip_zgetrs_dev(n, nrhs, dA.ptr, dp.ptr, dB.ptr, None)
    return dB.to_array(np.complex128, B.shape)


def solve_mixed(A, B, tol=1e-14, maxit=30):
    """FP32 MFMA LU + FP64 refinement; returns (X, refinement steps or -1)."""
    A = np.ascontiguousarray(A, dtype=np.complex128)
    B = np.ascontiguousarray(B, dtype=np.complex128)
    n = A.shape[0]
    nrhs = 1 if B.ndim == 1 else B.shape[1]
    dA, dB = DeviceBuffer.from_array(A), DeviceBuffer.from_array(B)
    dX = DeviceBuffer(B.nbytes)
    it = lib().qpms_hip_solve_mixed_dev(n, nrhs, dA.ptr, dB.ptr, dX.ptr, float(tol), int(maxit), None)
    return dX.to_array(np.complex128, B.shape), it


# ---------------------------------------------------------------------------
# Beyn contour eigensolver (beyn.h:11-126)
# ---------------------------------------------------------------------------
class beyn_result_t(ctypes.Structure):
    """beyn.h:82-94"""
    _fields_ = [("neig", ctypes.c_size_t), ("vlen", ctypes.c_size_t),
                ("eigval", ctypes.POINTER(c_cplx)), ("eigval_err", ctypes.POINTER(c_cplx)),
                ("residuals", ctypes.POINTER(ctypes.c_double)), ("eigvec", ctypes.POINTER(c_cplx)),
                ("ranktest_SV", ctypes.POINTER(ctypes.c_double)), ("gsl", ctypes.c_void_p)]


# int (*)(complex double *target_M, size_t m, complex double z, void *params), beyn.h:14-16
beyn_function_M_t = ctypes.CFUNCTYPE(ctypes.c_int, ctypes.c_void_p, ctypes.c_size_t, c_cplx, ctypes.c_void_p)


def _beyn_bind():
    L = lib()
    if getattr(L, "_beyn_bound", False):
        return L
    vp, sz, dbl = ctypes.c_void_p, ctypes.c_size_t, ctypes.c_double
    L.beyn_contour_ellipse.restype = vp
    L.beyn_contour_ellipse.argtypes = [c_cplx, dbl, dbl, sz]
    L.beyn_solve_v.restype = ctypes.POINTER(beyn_result_t)
    L.beyn_solve_v.argtypes = [sz, sz, beyn_function_M_t, vp, vp, vp, dbl, sz, dbl, vp]
    L.beyn_solve.restype = ctypes.POINTER(beyn_result_t)
    L.beyn_solve.argtypes = [sz, sz, beyn_function_M_t, vp, vp, vp, dbl, sz, dbl]
    L.beyn_result_free.argtypes = [ctypes.POINTER(beyn_result_t)]
    L._beyn_bound = True
    return L


def _libc_free(p):
    ctypes.CDLL(None).free(ctypes.c_void_p(p))


class BeynResult:
    """Host copy of a beyn_result_t: eigval, eigval_err, residuals, eigvec
    (neig x vlen) and ranktest_SV (l singular values of the fine A0)."""

    def __init__(self, r, l):
        c = r.contents
        n, m = c.neig, c.vlen
        self.neig, self.vlen = n, m

        def carr(p, k):
            return np.ctypeslib.as_array(ctypes.cast(p, ctypes.POINTER(ctypes.c_double)), (2 * k,)).copy().view(
                np.complex128) if k else np.zeros(0, np.complex128)
        self.eigval = carr(c.eigval, n)
        self.eigval_err = carr(c.eigval_err, n)
        self.residuals = np.ctypeslib.as_array(c.residuals, (n,)).copy() if n else np.zeros(0)
        self.eigvec = carr(c.eigvec, n * m).reshape(n, m)
        self.ranktest_SV = np.ctypeslib.as_array(c.ranktest_SV, (l,)).copy() if l else np.zeros(0)


def beyn_solve(Mfun, m, l, centre, halfax_re, halfax_im, npoints, rank_tol=1e-4, rank_sel_min=1, res_tol=0.0,
               Vhat=None):
    """Run beyn_solve (seeded probe) or beyn_solve_v (given probe Vhat, m x l)
    with M(z) supplied by the Python callable Mfun(z) -> (m, m) complex array."""
    L = _beyn_bind()

    def cb(target, mm, z, params):
        A = np.ascontiguousarray(Mfun(complex(z.re, z.im)), dtype=np.complex128)
        if A.shape != (mm, mm):
            return 1
        ctypes.memmove(target, A.ctypes.data, A.nbytes)
        return 0
    fn = beyn_function_M_t(cb)
    cont = L.beyn_contour_ellipse(c_cplx.of(centre), float(halfax_re), float(halfax_im), int(npoints))
    try:
        if Vhat is None:
            r = L.beyn_solve(m, l, fn, None, None, cont, float(rank_tol), int(rank_sel_min), float(res_tol))
        else:
            V = np.ascontiguousarray(Vhat, dtype=np.complex128)
            assert V.shape == (m, l)
            r = L.beyn_solve_v(m, l, fn, None, None, cont, float(rank_tol), int(rank_sel_min), float(res_tol),
                               V.ctypes.data)
        out = BeynResult(r, l)
        L.beyn_result_free(r)
    finally:
        _libc_free(cont)
    return out
