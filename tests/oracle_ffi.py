"""ctypes access to the CPU oracle (oracle/_build/liboracle.so) and to the
reference's own GSL-free sources compiled by oracle/build_ref.sh. Test
infrastructure only."""
import ctypes
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_SO = os.path.join(ROOT, "oracle", "_build", "liboracle.so")
REF_SO = os.path.join(ROOT, "oracle", "_ref", "libqpms_ref.so")

vp, i32, dbl = ctypes.c_void_p, ctypes.c_int, ctypes.c_double


class c_cplx(ctypes.Structure):
    _fields_ = [("re", ctypes.c_double), ("im", ctypes.c_double)]

    @classmethod
    def of(cls, z):
        z = complex(z)
        return cls(z.real, z.imag)


_o = None


def oracle():
    global _o
    if _o is None:
        if not os.path.exists(ORACLE_SO):
            import subprocess
            subprocess.check_call(["make", "-s", "-C", os.path.join(ROOT, "oracle")])
        L = ctypes.CDLL(ORACLE_SO)
        L.or_calc_init.restype = vp
        L.or_calc_init.argtypes = [i32, i32]
        L.or_calc_free.argtypes = [vp]
        L.or_get_AB_arrays.argtypes = [vp, c_cplx, dbl, dbl, i32, i32, vp, vp]
        L.or_trans_single_A.restype = c_cplx
        L.or_trans_single_A.argtypes = [i32, i32, i32, i32, i32, c_cplx, dbl, dbl, i32]
        L.or_trans_single_B.restype = c_cplx
        L.or_trans_single_B.argtypes = [i32, i32, i32, i32, i32, c_cplx, dbl, dbl, i32]
        L.or_build_W_rows.argtypes = [vp, i32, vp, c_cplx, i32, i32, i32, vp]
        L.or_build_M.argtypes = [vp, i32, vp, vp, vp, c_cplx, vp]
        L.or_mie.argtypes = [i32, dbl, c_cplx, c_cplx, c_cplx, c_cplx, vp, vp]
        L.or_zgetrf.argtypes = [i32, vp, vp, i32]
        L.or_zgetrf.restype = i32
        L.or_zgetrs.argtypes = [i32, i32, vp, vp, vp]
        L.or_gaunt.restype = dbl
        L.or_gaunt.argtypes = [i32] * 5
        L.or_wigner3j.restype = dbl
        L.or_wigner3j.argtypes = [i32] * 6
        L.or_legendre_array.argtypes = [i32, dbl, i32, vp]
        L.or_sph_bessel.argtypes = [i32, i32, c_cplx, vp]
        L.or_calc_multipliers.argtypes = [vp, i32, i32, vp, vp, vp, vp]
        L.or_set_threads.argtypes = [i32]
        L.or_scattered_E.argtypes = [i32, i32, vp, c_cplx, i32, vp, i32, vp, vp]
        L.or_scattered_E_alt.argtypes = [vp, i32, vp, c_cplx, i32, vp, i32, vp, vp]
        _o = L
    return _o


def ref():
    """Reference gaunt.c / latticegens.c / lattices2d.c, or None if not built."""
    if not os.path.exists(REF_SO):
        return None
    L = ctypes.CDLL(REF_SO, mode=os.RTLD_LAZY)
    L.gaunt_xu.argtypes = [i32, i32, i32, i32, i32, vp, vp]
    return L


class OracleCalc:
    def __init__(self, lMax, norm=65536 | 8):
        self.o = oracle()
        self.c = self.o.or_calc_init(lMax, norm)
        self.lMax, self.nelem, self.norm = lMax, lMax * (lMax + 2), norm

    def __del__(self):
        try:
            self.o.or_calc_free(self.c)
        except Exception:
            pass

    def get_AB_arrays(self, kr, theta, phi, J=3, r_ge_d=False):
        ne = self.nelem
        A = np.empty((ne, ne), np.complex128)
        B = np.empty((ne, ne), np.complex128)
        self.o.or_get_AB_arrays(self.c, c_cplx.of(kr), theta, phi, int(r_ge_d), J, A.ctypes.data, B.ctypes.data)
        return A, B

    def multipliers(self, y, yu):
        A = np.zeros(self.lMax + 1, np.complex128)
        B = np.zeros(self.lMax + 1, np.complex128)
        nA, nB = ctypes.c_int(), ctypes.c_int()
        self.o.or_calc_multipliers(self.c, y, yu, A.ctypes.data, ctypes.byref(nA), B.ctypes.data, ctypes.byref(nB))
        return A[:nA.value], B[:nB.value]

    def build_W_rows(self, pos, k, row0, nrows, J=3):
        pos = np.ascontiguousarray(pos, np.float64)
        npart = pos.shape[0]
        nb = 2 * self.nelem
        W = np.empty((nrows * nb, npart * nb), np.complex128)
        self.o.or_build_W_rows(self.c, npart, pos.ctypes.data, c_cplx.of(k), J, row0, nrows, W.ctypes.data)
        return W

    def build_M(self, pos, ttype, tdiag, k):
        pos = np.ascontiguousarray(pos, np.float64)
        npart = pos.shape[0]
        L = npart * 2 * self.nelem
        M = np.empty((L, L), np.complex128)
        ttype = np.ascontiguousarray(ttype, np.int32)
        tdiag = np.ascontiguousarray(tdiag, np.complex128)
        self.o.or_build_M(self.c, npart, pos.ctypes.data, ttype.ctypes.data, tdiag.ctypes.data, c_cplx.of(k),
                          M.ctypes.data)
        return M


def sph_bessel(J, lmax, x):
    out = np.empty(lmax + 1, np.complex128)
    oracle().or_sph_bessel(J, lmax, c_cplx.of(x), out.ctypes.data)
    return out


def legendre_array(lmax, x, csphase=-1):
    out = np.empty((lmax + 1) * (lmax + 2) // 2)
    oracle().or_legendre_array(lmax, x, csphase, out.ctypes.data)
    return out


def mie(lMax, a, ki, ke, mui=1, mue=1):
    RV = np.empty(lMax + 1, np.complex128)
    RH = np.empty(lMax + 1, np.complex128)
    oracle().or_mie(lMax, a, c_cplx.of(ki), c_cplx.of(ke), c_cplx.of(mui), c_cplx.of(mue), RV.ctypes.data,
                    RH.ctypes.data)
    return RV, RH


def zgetrf(A, nthreads=1):
    A = np.ascontiguousarray(A, np.complex128).copy()
    n = A.shape[0]
    ipiv = np.empty(n, np.int32)
    info = oracle().or_zgetrf(n, A.ctypes.data, ipiv.ctypes.data, nthreads)
    return A, ipiv, info


def zgetrs(LU, ipiv, B):
    B = np.ascontiguousarray(B, np.complex128).copy()
    n = LU.shape[0]
    nrhs = 1 if B.ndim == 1 else B.shape[1]
    oracle().or_zgetrs(n, nrhs, np.ascontiguousarray(LU).ctypes.data, np.ascontiguousarray(ipiv, np.int32).ctypes.data,
                       B.ctypes.data)
    return B


EV = 1.602176634e-19 / 1.054571817e-34   # rad/s per eV
C0 = 2.99792458e8


def planewave(lMax, pos, kvec, E):
    """Plane-wave coefficients (npart, 2*nelem): electric block then magnetic."""
    o = oracle()
    o.or_planewave.argtypes = [i32, i32, vp, vp, vp, vp]
    pos = np.ascontiguousarray(pos, np.float64).reshape(-1, 3)
    kvec = np.ascontiguousarray(kvec, np.float64)
    E = np.ascontiguousarray(E, np.complex128)
    ne = lMax * (lMax + 2)
    out = np.empty((len(pos), 2 * ne), np.complex128)
    o.or_planewave(lMax, len(pos), pos.ctypes.data, kvec.ctypes.data, E.ctypes.data, out.ctypes.data)
    return out


def mie_tdiag(lMax, radius, eps_in, eps_bg, omega):
    """Diagonal of the sphere T-matrix in the default basis (tmatrices.c:500-591)."""
    ke = np.sqrt(complex(eps_bg)) * omega / C0
    ki = np.sqrt(complex(eps_in)) * omega / C0
    RV, RH = mie(lMax, radius, ki, ke)
    el = np.concatenate([[RV[l]] * (2 * l + 1) for l in range(1, lMax + 1)])
    mg = np.concatenate([[RH[l]] * (2 * l + 1) for l in range(1, lMax + 1)])
    return np.concatenate([el, mg]), ke


def finite_cluster(pos, lMax, radius, eps_in, eps_bg, eV, kdir, E):
    """Oracle pipeline of a plane-wave scattering solve on a finite cluster of
    identical spheres: W, M = I - T W, a, T a, f = M^-1 T a, and the
    extinction cross-section -Re(a^H f)/k^2 (misc/finiterectlat-scatter.py:104-109)."""
    omega = eV * EV
    tdiag, k = mie_tdiag(lMax, radius, eps_in, eps_bg, omega)
    oc = OracleCalc(lMax)
    pos = np.ascontiguousarray(pos, np.float64)
    n = len(pos)
    W = oc.build_W_rows(pos, k, 0, n)
    M = oc.build_M(pos, np.zeros(n, np.int32), tdiag, k)
    a = planewave(lMax, pos, k.real * np.asarray(kdir, float), E).ravel()
    Ta = np.tile(tdiag, n) * a
    LU, ipiv, info = zgetrf(M)
    assert info == 0
    f = zgetrs(LU, ipiv, Ta)
    sigma = -np.vdot(a, f).real / (k * k.conjugate()).real
    return dict(W=W, M=M, a=a, Ta=Ta, f=f, tdiag=tdiag, k=k, sigma_ext=sigma)


def scattered_E(lMax, pos, k, cvf, where, J=3):
    """Field of the coefficient vector cvf (npart x 2 nelem, default basis) at
    the points where (npts x 3): complex (npts, 3) cartesian components."""
    o = oracle()
    pos = np.ascontiguousarray(pos, np.float64).reshape(-1, 3)
    where = np.ascontiguousarray(where, np.float64).reshape(-1, 3)
    cvf = np.ascontiguousarray(cvf, np.complex128)
    out = np.empty((len(where), 3), np.complex128)
    o.or_scattered_E(lMax, len(pos), pos.ctypes.data, c_cplx.of(k), J, cvf.ctypes.data, len(where),
                     where.ctypes.data, out.ctypes.data)
    return out


def scattered_E_alt(lMax, pos, k, cvf, where, J=3):
    """The same field through translation arrays to regular dipoles, as the
    reference's __alt variant forms it."""
    o = oracle()
    c = OracleCalc(lMax)
    pos = np.ascontiguousarray(pos, np.float64).reshape(-1, 3)
    where = np.ascontiguousarray(where, np.float64).reshape(-1, 3)
    cvf = np.ascontiguousarray(cvf, np.complex128)
    out = np.empty((len(where), 3), np.complex128)
    o.or_scattered_E_alt(c.c, len(pos), pos.ctypes.data, c_cplx.of(k), J, cvf.ctypes.data, len(where),
                         where.ctypes.data, out.ctypes.data)
    return out


# ---------------------------------------------------------------------------
# Host LAPACK (the reference's own dependency for the LU: LAPACKE_zgetrf /
# zgetrs at scatsystem.c:2214, 2270). Intel MKL's single dynamic library when
# present (this image ships it), else scipy's bundled LAPACK.
# ---------------------------------------------------------------------------
MKL_CANDIDATES = ["/opt/conda/lib/libmkl_rt.so.1", "/opt/conda/lib/libmkl_rt.so", "libmkl_rt.so.1", "libmkl_rt.so"]
_lap = None


class HostLapack:
    """zgetrf / zgetrs on the host, row-major, with a settable thread count."""

    def __init__(self):
        self.name, self.mkl = None, None
        for p in MKL_CANDIDATES:
            try:
                L = ctypes.CDLL(p)
            except OSError:
                continue
            if hasattr(L, "LAPACKE_zgetrf") and hasattr(L, "LAPACKE_zgetrs"):
                L.LAPACKE_zgetrf.restype = ctypes.c_int
                L.LAPACKE_zgetrf.argtypes = [i32, i32, i32, vp, i32, vp]
                L.LAPACKE_zgetrs.restype = ctypes.c_int
                L.LAPACKE_zgetrs.argtypes = [i32, ctypes.c_char, i32, i32, vp, i32, vp, vp, i32]
                L.MKL_Set_Num_Threads.argtypes = [i32]
                L.MKL_Get_Max_Threads.restype = i32
                self.mkl, self.name = L, f"Intel MKL ({p}), LAPACKE_zgetrf/zgetrs"
                break
        if self.mkl is None:
            import scipy.linalg.lapack as sl
            self.sl = sl
            self.name = "scipy.linalg.lapack zgetrf/zgetrs (bundled OpenBLAS)"

    def set_threads(self, n):
        if self.mkl is not None:
            self.mkl.MKL_Set_Num_Threads(int(n))
        else:   # OpenBLAS reads its thread count at load; nothing to do per call
            pass

    def zgetrf(self, A):
        """LU of A (complex128, C-contiguous), in place with MKL; returns
        (LU, ipiv 1-based, info)."""
        n = A.shape[0]
        if self.mkl is not None:
            ipiv = np.zeros(n, np.int32)
            info = self.mkl.LAPACKE_zgetrf(101, n, n, A.ctypes.data, n, ipiv.ctypes.data)   # LAPACK_ROW_MAJOR
            return A, ipiv, info
        lu, piv, info = self.sl.zgetrf(A)
        return lu, piv.astype(np.int32) + 1, info

    def zgetrs(self, LU, ipiv, b):
        """Solve with one right-hand side (row-major, ldb = 1 as the reference passes it)."""
        n = LU.shape[0]
        x = np.ascontiguousarray(b, np.complex128).copy()
        if self.mkl is not None:
            info = self.mkl.LAPACKE_zgetrs(101, b"N", n, 1, LU.ctypes.data, n, ipiv.ctypes.data, x.ctypes.data, 1)
        else:
            x, info = self.sl.zgetrs(LU, ipiv - 1, x)
        assert info == 0
        return x


def host_lapack():
    global _lap
    if _lap is None:
        _lap = HostLapack()
    return _lap


def reference_cpu_solve(pos, lMax, radius, eps_in, eps_bg, omega, kdir, E, lap):
    """The reference's CPU path of one plane-wave solve, as a caller of the
    reference library runs it (misc/finiterectlat-scatter.py:80-111 through
    scatsystem.c:1283-1343, 2202-2273): T from Mie, M = I - T W built by the
    oracle's pair loop (threads as set by or_set_threads), LU and solve by host
    LAPACK, extinction -Re(a^H f)/k^2. Returns sigma_ext."""
    tdiag, k = mie_tdiag(lMax, radius, eps_in, eps_bg, omega)
    oc = OracleCalc(lMax)
    pos = np.ascontiguousarray(pos, np.float64)
    n = len(pos)
    M = oc.build_M(pos, np.zeros(n, np.int32), tdiag, k)
    a = planewave(lMax, pos, k.real * np.asarray(kdir, float), E).ravel()
    Ta = np.tile(tdiag, n) * a
    LU, ipiv, info = lap.zgetrf(M)
    assert info == 0
    f = lap.zgetrs(LU, ipiv, Ta)
    return -np.vdot(a, f).real / (k * k.conjugate()).real
