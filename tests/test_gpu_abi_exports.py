"""GPU parity of the translation and eigenmode entry points that only a C
caller reaches (tests/c/trans_arrays.c drives them through the public ABI):

* qpms_trans_calculator_get_trans_array (row- and column-major strides,
  a reversed basis and an electric-only basis) and _lc3p: the oracle's A/B
  arrays rearranged into the two bases by the rule of
  translations_inlines.h:8-35 (same wave type takes A, the other B), 1e-10;
* get_trans_array_e32 equals the same rearrangement of get_AB_arrays_e32
  (bit for bit), and its _e parts (short + long + sigma0) add up to it;
* qpms_scatsys_build_translation_matrix_full on a cluster whose particles
  alternate lMax 2 and 3 (the general-basis builder): every block against
  the oracle's A/B at k (r_R - r_C), diagonal blocks zero, 1e-10;
* qpms_scatsys_finite_find_eigenmodes on a 2 x 2 silver-like cluster: the
  eigenfrequencies against the oracle's Beyn solver on the oracle's mode
  problem (same contour, full-rank probe of its own), 1e-9 relative, and
  M(lambda) numerically singular at each.
"""
import os
import subprocess

import numpy as np
import pytest

import oracle_beyn as ob
import oracle_ffi as of

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "tests", "c", "_build", "trans_arrays")
EV = 1.602176634e-19 / 1.054571817e-34
C0 = 2.99792458e8


@pytest.fixture(scope="module")
def out(tmp_path_factory, gpu):
    d = tmp_path_factory.mktemp("trans_arrays")
    subprocess.run([DRIVER, str(d)], check=True, timeout=300)

    def rd(name, dtype=np.complex128):
        return np.fromfile(d / f"{name}.bin", dtype=dtype)
    return rd


def decode(ilist):
    """uvswfi = type + 4 y_sc (indexing.h:60-63): (type, y) with y = y_sc - 1."""
    ilist = np.asarray(ilist, np.uint64)
    return (ilist & 3).astype(int), (ilist >> 2).astype(int) - 1


def remap(A, B, dspec, sspec):
    """dst x src matrix of the bases from A, B (element (y_dst, y_src))."""
    dt, dy = decode(dspec)
    st, sy = decode(sspec)
    T = np.empty((len(dt), len(st)), np.complex128)
    for i in range(len(dt)):
        for j in range(len(st)):
            T[i, j] = A[dy[i], sy[j]] if dt[i] == st[j] else B[dy[i], sy[j]]
    return T


def relerr(a, b):
    return np.max(np.abs(a - b)) / np.max(np.abs(b))


def sph(d):
    r = np.linalg.norm(d)
    return r, np.arccos(d[2] / r), np.arctan2(d[1], d[0])


KD = [(2.0, 0.5, 0.5), (7.0 + 0.3j, 1.2, -2.0), (15.0, np.pi, 1.0), (0.8, 2.0, 0.3)]


def test_get_trans_array(out):
    dst, s1, s2 = (out(n, np.uint64) for n in ("spec_dst", "spec_src1", "spec_src2"))
    assert len(dst) == 30 and len(s1) == 30 and len(s2) == 8
    T1 = out("trans1").reshape(4, len(dst), len(s1))
    T2 = out("trans2").reshape(4, len(s2), len(dst))   # column-major dst x src = row-major src x dst
    o = of.OracleCalc(3)
    for i, (kr, th, ph) in enumerate(KD):
        A, B = o.get_AB_arrays(kr, th, ph, J=3)
        assert relerr(T1[i], remap(A, B, dst, s1)) < 1e-10, i
        A, B = o.get_AB_arrays(kr, th, ph, J=1)
        assert relerr(T2[i].T, remap(A, B, dst, s2)) < 1e-10, i


def test_get_trans_array_lc3p(out):
    dst, s1 = out("spec_dst", np.uint64), out("spec_src1", np.uint64)
    x = out("lc3p_in", np.float64)
    pd, ps, k = x[:3], x[3:6], complex(x[6], x[7])
    r, th, ph = sph(pd - ps)   # vectors.h:359-365: kd = k (dest - src)
    A, B = of.OracleCalc(3).get_AB_arrays(k * r, th, ph, J=3)
    T = out("trans_lc3p").reshape(len(dst), len(s1))
    assert relerr(T, remap(A, B, dst, s1)) < 1e-10


def test_get_trans_array_e32(out):
    A, B = out("e32_A").reshape(8, 8), out("e32_B").reshape(8, 8)
    d2 = np.array([2 + 4 * (n * (n + 1) + m) for n in (1, 2) for m in range(-n, n + 1)] +
                  [1 + 4 * (n * (n + 1) + m) for n in (1, 2) for m in range(-n, n + 1)], np.uint64)
    s2r = d2[::-1]
    T = out("e32_trans").reshape(16, 16)
    assert np.array_equal(T, remap(A, B, d2, s2r))
    assert relerr(out("e32_trans_parts").reshape(16, 16), T) < 1e-12


def test_mixed_basis_translation_matrix(out):
    x = out("mixed_in", np.float64)
    np_ = 6
    pos = x[:24].reshape(np_, 4)[:, :3]
    lmax = x[:24].reshape(np_, 4)[:, 3].astype(int)
    k = complex(x[24], x[25])
    nb = 2 * lmax * (lmax + 2)
    start = np.concatenate([[0], np.cumsum(nb)])
    L = int(start[-1])
    W = out("mixed_W").reshape(L, L)
    o = of.OracleCalc(3)
    default = {l: np.array([2 + 4 * (n * (n + 1) + m) for n in range(1, l + 1) for m in range(-n, n + 1)] +
                           [1 + 4 * (n * (n + 1) + m) for n in range(1, l + 1) for m in range(-n, n + 1)], np.uint64)
               for l in (2, 3)}
    for R in range(np_):
        for Cc in range(np_):
            blk = W[start[R]:start[R + 1], start[Cc]:start[Cc + 1]]
            if R == Cc:
                assert not blk.any()
                continue
            r, th, ph = sph(pos[R] - pos[Cc])
            A, B = o.get_AB_arrays(k * r, th, ph, J=3)
            assert relerr(blk, remap(A, B, default[lmax[R]], default[lmax[Cc]])) < 1e-10, (R, Cc)


def test_finite_find_eigenmodes(out):
    n, m = out("eig_n", np.float64).astype(int)
    assert m == 64 and n >= 1
    ev = out("eig_val")
    pos = out("eig_pos", np.float64).reshape(4, 3)
    lMax = 2

    oc = of.OracleCalc(lMax)

    def Mfun(om):
        td, k = of.mie_tdiag(lMax, 70e-9, -12 + 1j, 1.52 ** 2, om)
        return oc.build_M(pos, np.zeros(4, np.int32), td, k)
    centre, rre, rim = (1.12 - 0.25j) * EV, 0.1 * EV, 0.05 * EV
    z, dz = ob.contour_ellipse(centre, rre, rim, 48)
    V = ob.probe(m, m, 3)
    rank_tol = 1e-5 * 2 * np.pi * rre * 64
    ev_o, _, _, _, _ = ob.beyn_solve(Mfun, m, V, z, dz, centre, rank_tol=rank_tol, rank_sel_min=1)
    assert len(ev_o) == n, (ev / EV, ev_o / EV)
    for lam in ev_o:
        j = np.argmin(np.abs(ev - lam))
        assert abs(ev[j] - lam) <= 1e-9 * abs(lam), (ev[j] / EV, lam / EV)
        s = np.linalg.svd(Mfun(ev[j]), compute_uv=False)
        assert s[-1] / s[0] < 1e-8, s[-1] / s[0]


def test_periodic_find_eigenmodes_is_beyn_on_the_mode_problem(out):
    """qpms_scatsys_periodic_find_eigenmodes on the hexagonal two-particle
    lattice (lMax 2, m = 32, l = m probe columns, 40-node ellipse around
    2.0 - 0.08i eV at an in-zone Bloch vector) returns exactly what
    beyn_solve returns for the same contour and probe seed with the periodic
    builder I - T(omega) W(omega, k) as its callback (scatsystem.c:2327-2370):
    the wrapper adds only the contour, the probe width and the callback."""
    nw, nd, m, tol = out("per_n", np.float64)
    assert m == 32 and nw == nd
    assert np.array_equal(out("per_wrap_sv", np.float64), out("per_direct_sv", np.float64))
    assert np.array_equal(out("per_wrap_val"), out("per_direct_val"))
    sv = out("per_wrap_sv", np.float64)
    assert nw == max(1, (sv > tol).sum())   # rank test of beyn.c:382 with rank_sel_min = 1
