"""GPU parity of the plane-wave scattering solve on a finite cluster, through
the public C ABI (qpms_incfield_planewave, qpms_scatsys_incident_field_vector_full,
qpms_scatsysw_apply_Tmatrices_full, the W / M builders, LU and
qpms_scatsys_scatter_solve), against the oracle pipeline.

The 4x4 cluster is the plain finite-cluster configuration: pitch 500 nm,
spheres r = 50 nm, eps = -12+1i in eps = 2.31, lMax = 2 (L = 256), 2.0 eV,
x-polarised plane wave at normal incidence. Tolerances: 1e-10 relative for
builds and incident coefficients, 1e-9 for the solve and the extinction."""
import ctypes
import os
import subprocess

import numpy as np
import pytest

import oracle_ffi as of

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "tests", "c", "_build", "ss_finite")
MIXED = os.path.join(ROOT, "tests", "c", "_build", "ss_field_mixed")


class ccart3(ctypes.Structure):
    _fields_ = [("x", ctypes.c_double * 2), ("y", ctypes.c_double * 2), ("z", ctypes.c_double * 2)]


class pw_params(ctypes.Structure):
    """vswf.h:93-103 (both unions' largest member is ccart3_t)."""
    _fields_ = [("use_cartesian", ctypes.c_bool), ("k", ccart3), ("E", ccart3)]


def _c3(v):
    v = np.asarray(v, complex)
    return ccart3((ctypes.c_double * 2)(v[0].real, v[0].imag), (ctypes.c_double * 2)(v[1].real, v[1].imag),
                  (ctypes.c_double * 2)(v[2].real, v[2].imag))


DIRS = [((0, 0, 1.0), (1, 0, 0)), ((0, 0, -1.0), (0, 1j, 0)), ((0.3, 0.4, np.sqrt(0.75)), (0.8, -0.6, 0)),
        ((1.0, 0, 0), (0, 0.6, 0.8j))]


@pytest.mark.parametrize("lMax", [1, 2, 3, 4])
@pytest.mark.parametrize("case", range(len(DIRS)))
def test_incfield_planewave(gpu, lMax, case):
    import qpms_amd
    L = qpms_amd.lib()
    L.qpms_vswf_set_spec_from_lMax.restype = ctypes.c_void_p
    L.qpms_vswf_set_spec_from_lMax.argtypes = [ctypes.c_int, ctypes.c_int]
    L.qpms_vswf_set_spec_free.argtypes = [ctypes.c_void_p]
    L.qpms_incfield_planewave.argtypes = [ctypes.c_void_p, ctypes.c_void_p, qpms_amd.cart3_t, ctypes.c_void_p,
                                          ctypes.c_bool]
    kd, E = DIRS[case]
    k = 1.1e7 * np.asarray(kd)
    bspec = L.qpms_vswf_set_spec_from_lMax(lMax, qpms_amd.NORM_DEFAULT)
    ne = lMax * (lMax + 2)
    pos = np.array([[120e-9, -40e-9, 75e-9]])
    p = pw_params(True, _c3(k), _c3(E))
    out = np.full(2 * ne, 1 + 1j)
    L.qpms_incfield_planewave(out.ctypes.data, bspec, qpms_amd.cart3_t(*pos[0]), ctypes.byref(p), False)
    ref = of.planewave(lMax, pos, k, np.asarray(E, complex))[0]
    assert np.max(np.abs(out - ref)) <= 1e-12 * np.max(np.abs(ref))
    # add = true accumulates
    L.qpms_incfield_planewave(out.ctypes.data, bspec, qpms_amd.cart3_t(*pos[0]), ctypes.byref(p), True)
    assert np.max(np.abs(out - 2 * ref)) <= 2e-12 * np.max(np.abs(ref))
    L.qpms_vswf_set_spec_free(bspec)


def run_driver(tmp_path, nx, ny, pitch, radius, eps_in, eps_bg, eV, lMax, kd, E, points=None):
    out = tmp_path / "ss.bin"
    args = [DRIVER, str(out), str(nx), str(ny), repr(pitch), repr(radius), repr(eps_in.real), repr(eps_in.imag),
            repr(eps_bg), repr(eV), str(lMax)] + [repr(float(x)) for x in kd]
    for e in E:
        args += [repr(complex(e).real), repr(complex(e).imag)]
    if points is not None:
        pf = tmp_path / "points.bin"
        pf.write_bytes(np.ascontiguousarray(points, np.float64).tobytes())
        args.append(str(pf))
    subprocess.run(args, check=True, timeout=120)
    raw = out.read_bytes()
    L = int(np.frombuffer(raw[:8], np.int64)[0])
    nb = L // (nx * ny)
    nc = 2 * L * L + 3 * L + nb
    c = np.frombuffer(raw[8:8 + 16 * nc], np.complex128)
    tail = np.frombuffer(raw[8 + 16 * nc:8 + 16 * nc + 24], np.float64)
    W = c[:L * L].reshape(L, L)
    M = c[L * L:2 * L * L].reshape(L, L)
    a, Ta, f = c[2 * L * L:2 * L * L + L], c[2 * L * L + L:2 * L * L + 2 * L], c[2 * L * L + 2 * L:2 * L * L + 3 * L]
    td = c[2 * L * L + 3 * L:]
    r = dict(W=W, M=M, a=a, Ta=Ta, f=f, tdiag=td, sigma_ext=tail[0], k=complex(tail[1], tail[2]))
    if points is not None:
        o = 8 + 16 * nc + 24
        npts = int(np.frombuffer(raw[o:o + 8], np.int64)[0])
        e = np.frombuffer(raw[o + 8:], np.complex128).reshape(4, npts, 3)
        r.update(E_single=e[0], E_w=e[1], E_alt=e[2], E_points=e[3])
    return r


def rel(a, b):
    return np.max(np.abs(a - b)) / np.max(np.abs(b))


@pytest.mark.parametrize("shape,lMax,kd,E", [
    ((4, 4), 2, (0, 0, 1.0), (1, 0, 0)),                       # plain finite-cluster configuration
    ((3, 2), 3, (0.3, 0.4, np.sqrt(0.75)), (0.8, -0.6, 0)),    # oblique incidence, lMax 3
    ((2, 2), 4, (0, 0, 1.0), (0, 1, 0)),
])
def test_finite_cluster_abi(gpu, tmp_path, shape, lMax, kd, E):
    nx, ny = shape
    pitch, radius, eps_in, eps_bg, eV = 500e-9, 50e-9, -12 + 1j, 2.31, 2.0
    g = run_driver(tmp_path, nx, ny, pitch, radius, eps_in, eps_bg, eV, lMax, kd, E)
    pos = np.array([[(i - 0.5 * (nx - 1)) * pitch, (j - 0.5 * (ny - 1)) * pitch, 0.0]
                    for i in range(nx) for j in range(ny)])
    o = of.finite_cluster(pos, lMax, radius, eps_in, eps_bg, eV, kd, np.asarray(E, complex))
    assert abs(g["k"] - o["k"]) <= 1e-15 * abs(o["k"])
    assert rel(g["tdiag"], o["tdiag"]) < 1e-12
    assert rel(g["W"], o["W"]) < 1e-10
    assert rel(g["M"], o["M"]) < 1e-10
    assert rel(g["a"], o["a"]) < 1e-12
    assert rel(g["Ta"], o["Ta"]) < 1e-12
    assert rel(g["f"], o["f"]) < 1e-9
    assert abs(g["sigma_ext"] - o["sigma_ext"]) < 1e-9 * abs(o["sigma_ext"])


def _field_points(pos, rng, n=40):
    """Points near and between the particles, on a particle's z axis (both
    poles) and far away."""
    lo, hi = pos.min(axis=0) - 300e-9, pos.max(axis=0) + 300e-9
    pts = [lo + (hi - lo) * rng.uniform(size=3) for _ in range(n)]
    pts += [pos[0] + [0, 0, 120e-9], pos[-1] - [0, 0, 90e-9], pos[0] + [70e-9, 0, 0]]
    pts += [pos.mean(axis=0) + 1e-3 * np.array([0.3, -0.5, 0.8]), pos.mean(axis=0) + [0, 0, 2e-6]]
    return np.array(pts)


@pytest.mark.parametrize("shape,lMax,kd,E", [
    ((4, 4), 2, (0, 0, 1.0), (1, 0, 0)),
    ((3, 2), 3, (0.3, 0.4, np.sqrt(0.75)), (0.8, -0.6, 0)),
    ((2, 2), 4, (0, 0, 1.0), (0, 1, 0)),
])
def test_scattered_field_abi(gpu, tmp_path, shape, lMax, kd, E):
    """qpms_scatsys_scattered_E and its variants through the C ABI, on the
    solved cluster's own f, against the oracle's field of that f (1e-10
    relative to the largest field in the sample)."""
    nx, ny = shape
    pitch, radius, eps_in, eps_bg, eV = 500e-9, 50e-9, -12 + 1j, 2.31, 2.0
    pos = np.array([[(i - 0.5 * (nx - 1)) * pitch, (j - 0.5 * (ny - 1)) * pitch, 0.0]
                    for i in range(nx) for j in range(ny)])
    pts = _field_points(pos, np.random.default_rng(lMax))
    g = run_driver(tmp_path, nx, ny, pitch, radius, eps_in, eps_bg, eV, lMax, kd, E, points=pts)
    ref = of.scattered_E(lMax, pos, g["k"], g["f"], pts)
    alt = of.scattered_E_alt(lMax, pos, g["k"], g["f"], pts)
    near = slice(0, len(pts) - 2)   # far points compared separately (1/r smaller)
    for key in ("E_single", "E_w", "E_points"):
        assert rel(g[key][near], ref[near]) < 1e-10, key
        for t in (-2, -1):
            assert np.max(np.abs(g[key][t] - ref[t])) < 1e-10 * np.max(np.abs(ref[t])), (key, t)
    assert rel(g["E_alt"][near], alt[near]) < 1e-10


def test_scattered_field_mixed_bases(gpu, tmp_path):
    """Particles with bases of lMax 2 and 3 in one system: the library pads
    each particle's coefficients into the calculator's lMax; compared with the
    oracle's field of the same padded coefficients."""
    pos = np.array([[-400e-9, 0, 0], [0, 350e-9, 40e-9], [300e-9, -200e-9, 0]])
    lms = [2, 3, 2]
    pts = np.concatenate([_field_points(pos, np.random.default_rng(5), n=20), pos[1:2] + [[0, 0, 90e-9]]])
    pf, out = tmp_path / "pts.bin", tmp_path / "mixed.bin"
    pf.write_bytes(np.ascontiguousarray(pts, np.float64).tobytes())
    subprocess.run([MIXED, str(out), str(pf)], check=True, timeout=120)
    raw = out.read_bytes()
    L = int(np.frombuffer(raw[:8], np.int64)[0])
    k = complex(*np.frombuffer(raw[8:24], np.float64))
    e = np.frombuffer(raw[24:], np.complex128).reshape(2, len(pts), 3)
    assert L == sum(2 * l * (l + 2) for l in lms)
    cv = np.exp(0.37j * np.arange(L)) / (1.0 + 0.1 * np.arange(L))
    pad = np.zeros((3, 30), np.complex128)
    off = 0
    for p, l in enumerate(lms):
        ne = l * (l + 2)
        pad[p, :ne] = cv[off:off + ne]
        pad[p, 15:15 + ne] = cv[off + ne:off + 2 * ne]
        off += 2 * ne
    ref = of.scattered_E(3, pos, k, pad.ravel(), pts)
    near = slice(0, len(pts) - 3)
    assert rel(e[0][near], ref[near]) < 1e-10
    assert rel(e[1][near], ref[near]) < 1e-10
    for t in range(len(pts) - 3, len(pts)):
        assert np.max(np.abs(e[0][t] - ref[t])) < 1e-10 * np.max(np.abs(ref[t]))
