"""Pin of the numpy Beyn oracle: the eigenvalues it finds inside the contour
agree with the exact eigenvalues of the reference's own Beyn test problems
(tests/tbeyn.c loaded string, tests/tbeyn2.c quadratic polynomial), obtained
independently by companion linearisation and scipy.linalg.eig."""
import numpy as np

import oracle_beyn as ob


def inside(z, c, rre, rim):
    return ((z.real - c.real) / rre) ** 2 + ((z.imag - c.imag) / rim) ** 2 < 1


def test_loaded_string():
    m = 400
    z, dz = ob.contour_ellipse(150 + 2j, 148, 148, 50)
    ev, _, _, _, _ = ob.beyn_solve(ob.loaded_string(m), m, ob.probe(m, 10, 1), z, dz, 150 + 2j, 1e-4, 1, 1e-4)
    truth = ob.loaded_string_eigs(m)
    t_in = truth[inside(truth, 150 + 2j, 148, 148)]
    e_in = ev[inside(ev, 150 + 2j, 148, 148)]
    assert len(e_in) == len(t_in) == 5
    assert ob.match_sets(e_in, t_in) < 1e-7 and ob.match_sets(t_in, e_in) < 1e-7


def test_quadratic_polynomial():
    M, w = ob.polynomial(60, 3)
    z, dz = ob.contour_ellipse(0, 0.3, 0.3, 150)
    ev, _, _, _, _ = ob.beyn_solve(M, 60, ob.probe(60, 30, 2), z, dz, 0, 1e-4, 1, 1e-4)
    t_in = w[np.abs(w) < 0.29]
    e_in = ev[np.abs(ev) < 0.29]
    assert len(e_in) == len(t_in)
    assert ob.match_sets(e_in, t_in) < 1e-5 and ob.match_sets(t_in, e_in) < 1e-5
