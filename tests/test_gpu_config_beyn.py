"""GPU: the Beyn configuration at full size -- a 24 x 24 square array (N = 576,
pitch 400 nm, spheres r = 40 nm, eps = -12 + 1i, background n = 1.52), lMax 2,
so the mode problem has order 9216; 96 trapezoid nodes, 64 probe columns,
singular-value cut 1e-5 of the contour integral's scale (beyn.c:382 rank
test without a forced minimum rank).

* The contour as the configuration specifies it (ellipse centred at 2.05 eV,
  half-axes 0.05 eV (Re) x 0.005 eV (Im)) encloses no eigenfrequency of this
  array -- its modes lie near Im = -0.03 eV, below the ellipse -- and both
  the mixed-precision path (FP32 LU on MFMA + FP64 refinement) and the FP64
  path report the empty result: no singular value above the cut.
* An ellipse around 2.045 - 0.0335i eV with half-axes 0.004 eV encloses
  several modes: the mixed and FP64 eigenfrequencies agree to 1e-9
  relative (same nodes and probe, so the quadrature error is common), and at
  each the oracle's M(lambda) (oracle pair loop, reference algorithm) is
  numerically singular -- the smallest singular value estimated by inverse
  iteration is below 1e-6 of ||M||_F, and at least 100 times smaller than
  at a point 0.001 eV away (scatsystem.c:2281-2325, beyn.c:543-606)."""
import os

import numpy as np
import pytest

import oracle_ffi as of

pytestmark = pytest.mark.gpu
EV = 1.602176634e-19 / 1.054571817e-34


def served(req, mixed):
    from qpms_amd import serve as sv
    srv = sv.Server(arena_gib=100, nstreams=2, mixed=mixed)
    srv.submit(req)
    got = None
    while srv.pending():
        srv.progress(wait=True)
        for r in srv.poll():
            got = r
    srv.close()
    assert got is not None and got.status == 0, got.message if got else None
    return got


def request(centre_ev, rre_ev, rim_ev):
    from qpms_amd import serve as sv
    from qpms_amd.workloads import square_array
    pos = square_array(24, 24, 400e-9)
    return sv.Req(sv.KIND_BEYN, pos, 2, 40e-9, -12 + 1j, 1.52 ** 2, centre_ev * EV, beyn_rRe=rre_ev * EV,
                  beyn_rIm=rim_ev * EV, beyn_nodes=96, beyn_l=64, beyn_seed=1,
                  rank_tol=sv.beyn_rank_tol(9216, 64, rre_ev * EV))


def test_config_contour_is_empty(gpu):
    req = request(2.05, 0.05, 0.005)
    for mixed in (True, False):
        print(f"config contour, mixed={mixed}", flush=True)
        r = served(req, mixed)
        sv_ = np.array(r.sv[:64])
        assert r.neig == 0, (mixed, r.neig)
        assert np.all(sv_ < req.c.rank_tol), (mixed, sv_[:4], req.c.rank_tol)


def smallest_sv_ratio(M):
    """sigma_min(M) / ||M||_F by inverse iteration with the library's FP64 LU
    on the device (the oracle builds M; the GPU only factorises and solves)."""
    import qpms_amd
    L = qpms_amd.lib()
    n = M.shape[0]
    dA = qpms_amd.DeviceBuffer.from_array(np.ascontiguousarray(M))
    dp = qpms_amd.DeviceBuffer(4 * n)
    assert L.qpms_hip_zgetrf_dev(n, dA.ptr, dp.ptr, None) == 0
    rng = np.random.default_rng(0)
    x = rng.normal(size=n) + 1j * rng.normal(size=n)
    dx = qpms_amd.DeviceBuffer(16 * n)
    for _ in range(4):
        L.qpms_hip_memcpy_h2d(dx.ptr, x.ctypes.data, 16 * n)
        L.qpms_hip_zgetrs_dev(n, 1, dA.ptr, dp.ptr, dx.ptr, None)
        y = dx.to_array(np.complex128, (n,))
        x = y / np.linalg.norm(y)
    for b in (dA, dp, dx):
        b.free()
    return np.linalg.norm(M @ x) / np.linalg.norm(M)


def test_enclosing_contour_mixed_vs_fp64_and_oracle_residual(gpu):
    from qpms_amd.workloads import square_array
    req = request(2.045 - 0.0335j, 0.004, 0.004)
    print("enclosing contour, mixed", flush=True)
    rm = served(req, True)
    print("enclosing contour, FP64", flush=True)
    rf = served(req, False)
    em, ef = rm.eigenvalues(), rf.eigenvalues()
    assert rm.neig == rf.neig >= 3, (rm.neig, rf.neig)
    for z in ef:
        j = int(np.argmin(np.abs(em - z)))
        assert abs(em[j] - z) <= 1e-9 * abs(z), (em[j] / EV, z / EV)
    o = of.oracle()
    o.or_set_threads(min(16, os.cpu_count() or 1))
    oc = of.OracleCalc(2)
    pos = square_array(24, 24, 400e-9)

    def M_at(om):
        td, k = of.mie_tdiag(2, 40e-9, -12 + 1j, 1.52 ** 2, om)
        return oc.build_M(pos, np.zeros(len(pos), np.int32), td, k)
    off = smallest_sv_ratio(M_at(ef[0] + 0.001 * EV))
    centre = complex(req.c.omega, req.c.beyn_centre_im)
    for z in sorted(ef, key=lambda w: abs(w - centre))[:4]:   # the four nearest the centre
        print(f"oracle residual at {z / EV:.6f} eV", flush=True)
        at = smallest_sv_ratio(M_at(z))
        assert at < 1e-6 and at * 100 < off, (z / EV, at, off)
