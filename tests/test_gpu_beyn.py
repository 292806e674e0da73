"""GPU parity of the Beyn contour eigensolver (beyn.c:345-646) against the
numpy oracle run with the same probe matrix V̂, on the reference's own Beyn
test problems: the loaded string of tests/tbeyn.c (m = 400, ellipse centre
150+2i, half-axes 148, 50 nodes) and the quadratic matrix polynomial of
tests/tbeyn2.c style (dim 60, |z| < 0.3). Eigenvalues agree with the oracle
to 1e-9 relative and with the exact eigenvalues (companion linearisation) to
1e-7; eigenvectors agree up to phase."""
import numpy as np
import pytest

import oracle_beyn as ob

pytestmark = pytest.mark.gpu


def inside_ellipse(z, c, rre, rim):
    return ((z.real - c.real) / rre) ** 2 + ((z.imag - c.imag) / rim) ** 2 < 1


def check_against_oracle(res, ev, vecs, S, tol_ev=1e-9, inside=None):
    """Eigenvalues inside the contour agree to tol_ev; the ones the filters let
    through outside it (poorly resolved by the quadrature, so sensitive to
    rounding) to 1e-6."""
    assert res.neig == len(ev)
    g_in = res.eigval[inside(res.eigval)] if inside else res.eigval
    o_in = ev[inside(ev)] if inside else ev
    assert len(g_in) == len(o_in)
    assert ob.match_sets(g_in, o_in) < tol_ev and ob.match_sets(o_in, g_in) < tol_ev
    assert ob.match_sets(res.eigval, ev) < 1e-6 and ob.match_sets(ev, res.eigval) < 1e-6
    l = len(S)
    assert np.max(np.abs(res.ranktest_SV - S)) <= 1e-10 * S[0] * l
    for k in range(res.neig):
        j = int(np.argmin(np.abs(ev - res.eigval[k])))
        ov = abs(np.vdot(vecs[j], res.eigvec[k])) / (np.linalg.norm(vecs[j]) * np.linalg.norm(res.eigvec[k]))
        assert ov > 1 - 1e-8, (k, ov)


@pytest.mark.parametrize("lu_mode", ["fp64", "mixed"])
def test_loaded_string(gpu, monkeypatch, lu_mode):
    import qpms_amd
    monkeypatch.setenv("QPMS_HIP_LU_MODE", lu_mode)
    m, l, c, r, n = 400, 10, 150 + 2j, 148.0, 50
    Mf = ob.loaded_string(m)
    V = ob.probe(m, l, 1)
    res = qpms_amd.beyn_solve(Mf, m, l, c, r, r, n, 1e-4, 1, 1e-4, Vhat=V)
    z, dz = ob.contour_ellipse(c, r, r, n)
    ev, vecs, _, S, _ = ob.beyn_solve(Mf, m, V, z, dz, c, 1e-4, 1, 1e-4)
    check_against_oracle(res, ev, vecs, S, inside=lambda z: inside_ellipse(z, c, r, r))
    assert np.all(res.residuals <= 1e-4)
    truth = ob.loaded_string_eigs(m)
    t_in = truth[inside_ellipse(truth, c, r, r)]
    e_in = res.eigval[inside_ellipse(res.eigval, c, r, r)]
    assert len(e_in) == len(t_in) == 5
    assert ob.match_sets(e_in, t_in) < 1e-7 and ob.match_sets(t_in, e_in) < 1e-7


def test_quadratic_polynomial(gpu):
    import qpms_amd
    dim, l, n = 60, 30, 150
    Mf, w = ob.polynomial(dim, 3)
    V = ob.probe(dim, l, 2)
    res = qpms_amd.beyn_solve(Mf, dim, l, 0, 0.3, 0.3, n, 1e-4, 1, 1e-4, Vhat=V)
    z, dz = ob.contour_ellipse(0, 0.3, 0.3, n)
    ev, vecs, _, S, _ = ob.beyn_solve(Mf, dim, V, z, dz, 0, 1e-4, 1, 1e-4)
    check_against_oracle(res, ev, vecs, S, tol_ev=1e-8, inside=lambda z: np.abs(z) < 0.29)
    t_in = w[np.abs(w) < 0.29]
    e_in = res.eigval[np.abs(res.eigval) < 0.29]
    assert len(e_in) == len(t_in)
    assert ob.match_sets(e_in, t_in) < 1e-5


def test_no_residual_filter_keeps_rank(gpu):
    """res_tol = 0 keeps every rank-test survivor (beyn.c:480-497)."""
    import qpms_amd
    m, l = 400, 10
    Mf = ob.loaded_string(m)
    V = ob.probe(m, l, 4)
    res = qpms_amd.beyn_solve(Mf, m, l, 150 + 2j, 148, 148, 50, 1e-4, 1, 0.0, Vhat=V)
    z, dz = ob.contour_ellipse(150 + 2j, 148, 148, 50)
    ev, vecs, _, S, _ = ob.beyn_solve(Mf, m, V, z, dz, 150 + 2j, 1e-4, 1, 0.0)
    assert res.neig == len(ev) == int(np.sum(S > 1e-4))
    check_against_oracle(res, ev, vecs, S, inside=lambda z: inside_ellipse(z, 150 + 2j, 148, 148))


def test_seeded_probe_is_deterministic(gpu):
    import qpms_amd
    m, l = 400, 10
    Mf = ob.loaded_string(m)
    a = qpms_amd.beyn_solve(Mf, m, l, 150 + 2j, 148, 148, 50, 1e-4, 1, 1e-4)
    b = qpms_amd.beyn_solve(Mf, m, l, 150 + 2j, 148, 148, 50, 1e-4, 1, 1e-4)
    assert np.array_equal(a.eigval, b.eigval)
    truth = ob.loaded_string_eigs(m)
    t_in = truth[inside_ellipse(truth, 150 + 2j, 148, 148)]
    e_in = a.eigval[inside_ellipse(a.eigval, 150 + 2j, 148, 148)]
    assert ob.match_sets(t_in, e_in) < 1e-7
