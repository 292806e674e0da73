"""GPU parity of the 2D-periodic builders through the public C ABI
(qpms_scatsyswk_build_translation_matrix_full, ..._modeproblem_matrix_full,
the FP64 LU of M), against the oracle's Ewald sums and e32 combination.

Lattice: the 2D-periodic configuration's hexagonal lattice (a = 375 nm, two
particles at (0,0) and (a/sqrt3, 0)), silver-like spheres r = 50 nm,
eps = -12+1i in n = 1.52, lMax 2 and 3, Bloch vectors inside the zone. The
library sums over maxR = 64 sqrt(A), maxK = 2048 2pi/maxR
(scatsystem.c:1226-1227); the oracle over 12.5 rings and |K| < 30 2pi/a at
the same eta (the adjusted eta the library reports): the terms beyond are
below exp(-400) relative. Tolerance 1e-10 relative."""
import ctypes
import os
import subprocess

import numpy as np
import pytest

import ewald_common as ec
import oracle_ffi as of

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "tests", "c", "_build", "ss_periodic")


def run_driver(tmp_path, lMax, radius, eps_in, eps_bg, eV, beta, points=None):
    out = tmp_path / "per.bin"
    args = [DRIVER, str(out), repr(ec.A_LAT), str(lMax), repr(radius), repr(eps_in.real), repr(eps_in.imag),
            repr(eps_bg), repr(eV), repr(beta[0]), repr(beta[1])]
    if points is not None:
        pf = tmp_path / "points.bin"
        pf.write_bytes(np.ascontiguousarray(points, np.float64).tobytes())
        args.append(str(pf))
    subprocess.run(args, check=True, timeout=300)
    raw = out.read_bytes()
    L = int(np.frombuffer(raw[:8], np.int64)[0])
    hdr = np.frombuffer(raw[8:80], np.float64)
    o = 80
    c = np.frombuffer(raw[o:o + 16 * 2 * L * L], np.complex128)
    W, M = c[:L * L].reshape(L, L), c[L * L:].reshape(L, L)
    o += 32 * L * L
    nb = L // 2
    td = np.frombuffer(raw[o:o + 16 * nb], np.complex128)
    o += 16 * nb
    LU = np.frombuffer(raw[o:o + 16 * L * L], np.complex128).reshape(L, L)
    o += 16 * L * L
    ipiv = np.frombuffer(raw[o:o + 4 * L], np.int32)
    o += 4 * L
    r = dict(L=L, eta=hdr[0], k=complex(hdr[1], hdr[2]), basis=hdr[3:].reshape(2, 3), W=W, M=M, tdiag=td,
             LU=LU, ipiv=ipiv)
    if points is not None:
        npts = int(np.frombuffer(raw[o:o + 8], np.int64)[0])
        r["E"] = np.frombuffer(raw[o + 8:], np.complex128).reshape(npts, 3)
    return r


def sigma_total(LE, k, beta, sh, eta):
    o = of.oracle()
    o.or_sigma_short.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_double, of.c_cplx, ctypes.c_int,
                                 ctypes.c_void_p] + [ctypes.c_double] * 4 + [ctypes.c_void_p]
    s0 = sh[0] == 0 and sh[1] == 0
    R = np.ascontiguousarray(ec.xyweb(ec.B1, ec.B2, 0, not s0, 12.5 * ec.H_R, False))
    out = np.empty((LE + 1) ** 2, np.complex128)
    o.or_sigma_short(LE, -1, eta, of.c_cplx.of(k), len(R), R.ctypes.data, beta[0], beta[1], sh[0], sh[1],
                     out.ctypes.data)
    out += ec.oracle_sigma_long(LE, k, beta, sh, ec.K_points(), eta=eta)
    if s0:
        o.or_sigma0.argtypes = [ctypes.c_int, ctypes.c_double, of.c_cplx, ctypes.c_void_p]
        z = np.empty(1, np.complex128)
        o.or_sigma0(LE, eta, of.c_cplx.of(k), z.ctypes.data)
        out[0] += z[0]
    return out


def oracle_periodic_W(lMax, k, beta, eta):
    pos = np.array([[0.0, 0.0], [ec.A_LAT / np.sqrt(3), 0.0]])
    oc = of.OracleCalc(lMax)
    o = of.oracle()
    o.or_e32_combine.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p]
    ne = lMax * (lMax + 2)
    nb = 2 * ne
    W = np.zeros((2 * nb, 2 * nb), np.complex128)
    cache = {}
    for d in range(2):
        for s in range(2):
            sh = tuple(pos[d] - pos[s])
            if sh not in cache:
                sig = sigma_total(2 * lMax + 1, k, beta, sh, eta)
                A = np.empty((ne, ne), np.complex128)
                B = np.empty((ne, ne), np.complex128)
                o.or_e32_combine(oc.c, sig.ctypes.data, 0, A.ctypes.data, B.ctypes.data)
                cache[sh] = np.block([[A, B], [B, A]])
            W[d * nb:(d + 1) * nb, s * nb:(s + 1) * nb] = cache[sh]
    return W


def rel(a, b):
    return np.max(np.abs(a - b)) / np.max(np.abs(b))


@pytest.mark.parametrize("lMax,beta", [(2, (0.0, 0.0)), (2, (0.25, 0.1)), (3, (0.1, -0.3))])
def test_periodic_modeproblem_abi(gpu, tmp_path, lMax, beta):
    b = (beta[0] * 2 * np.pi / ec.A_LAT, beta[1] * 2 * np.pi / ec.A_LAT)
    g = run_driver(tmp_path, lMax, 50e-9, -12 + 1j, 1.52 ** 2, 2.0, b)
    Wo = oracle_periodic_W(lMax, g["k"], b, g["eta"])
    assert rel(g["W"], Wo) < 1e-10
    td, _ = of.mie_tdiag(lMax, 50e-9, -12 + 1j, 1.52 ** 2, 2.0 * of.EV)
    assert rel(g["tdiag"], td) < 1e-12
    Mo = np.eye(len(Wo)) - np.tile(td, 2)[:, None] * Wo
    assert rel(g["M"], Mo) < 1e-10
    # LU: P A = L U reconstructs M, and log|det| agrees with the oracle matrix's
    n = g["L"]
    LU = g["LU"]
    Lm = np.tril(LU, -1) + np.eye(n)
    Um = np.triu(LU)
    PA = g["M"].copy()
    for i, p in enumerate(g["ipiv"]):
        PA[[i, p - 1]] = PA[[p - 1, i]]
    assert rel(Lm @ Um, PA) < 1e-12
    ld = np.sum(np.log(np.abs(np.diag(Um))))
    assert abs(ld - np.linalg.slogdet(Mo)[1]) < 1e-9 * max(1, abs(ld))


@pytest.mark.parametrize("lMax,beta", [(2, (0.25, 0.1)), (3, (0.0, 0.0))])
def test_periodic_scattered_field_abi(gpu, tmp_path, lMax, beta):
    """qpms_scatsyswk_scattered_E at in-plane points against the oracle's
    lattice-summed translation arrays (destination: the regular electric
    dipoles at the point; source: each particle's basis; shift: point -
    particle) contracted with the coefficients and the dipoles' value at the
    origin (scatsystem.c:2113-2188)."""
    b = (beta[0] * 2 * np.pi / ec.A_LAT, beta[1] * 2 * np.pi / ec.A_LAT)
    pts = np.array([[60e-9, 140e-9, 0.0], [-150e-9, -40e-9, 0.0], [ec.A_LAT / np.sqrt(3) + 90e-9, 70e-9, 0.0]])
    g = run_driver(tmp_path, lMax, 50e-9, -12 + 1j, 1.52 ** 2, 2.0, b, points=pts)
    k, eta = g["k"], g["eta"]
    ne = lMax * (lMax + 2)
    L = g["L"]
    cv = np.exp(0.7j * np.arange(L)) / (1.0 + np.arange(L))
    dip = np.array([of.scattered_E(1, [[0, 0, 0]], k, np.eye(6)[j], [[0, 0, 0]], J=1)[0] for j in range(3)])
    pos = np.array([[0.0, 0.0], [ec.A_LAT / np.sqrt(3), 0.0]])
    o = of.oracle()
    o.or_e32_combine.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p]
    oc = of.OracleCalc(lMax)
    ref = np.zeros((len(pts), 3), np.complex128)
    for t, w in enumerate(pts):
        for p in range(2):
            sig = sigma_total(2 * lMax + 1, k, b, tuple(w[:2] - pos[p]), eta)
            A = np.empty((ne, ne), np.complex128)
            B = np.empty((ne, ne), np.complex128)
            o.or_e32_combine(oc.c, sig.ctypes.data, 0, A.ctypes.data, B.ctypes.data)
            c = cv[p * 2 * ne:(p + 1) * 2 * ne]
            for j in range(3):
                ref[t] += (A[j] @ c[:ne] + B[j] @ c[ne:]) * dip[j]
    assert rel(g["E"], ref) < 1e-10
