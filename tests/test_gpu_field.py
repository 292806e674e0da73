"""GPU parity of the scattered-field evaluation (k_scattered_e through
qpms_hip_scattered_E_dev) against the oracle's field sum, on seeded random
coefficient vectors: lMax 1-4, every Bessel kind, real and complex k, the
thread-per-point and wave-per-point launch shapes, points on a particle's z
axis (both poles) and at a particle's centre (regular waves), the parity
option of the translation-based variant, and full-size launches checked on
a sample of points. Tolerance: max |gpu - oracle| <= 1e-10 max |oracle| over
the compared points."""
import numpy as np
import pytest

import oracle_ffi as of

pytestmark = pytest.mark.gpu
TOL = 1e-10


def rel(a, b):
    return np.max(np.abs(a - b)) / np.max(np.abs(b))


def _case(rng, lMax, npart, npts):
    ne = lMax * (lMax + 2)
    pos = rng.uniform(-1, 1, size=(npart, 3)) * [2e-6, 2e-6, 0.3e-6]
    cv = (rng.normal(size=(npart, 2 * ne)) + 1j * rng.normal(size=(npart, 2 * ne))) * 1e-2
    pts = rng.uniform(-1, 1, size=(npts, 3)) * [2.5e-6, 2.5e-6, 1e-6]
    pts[:3] = pos[0] + [[0, 0, 80e-9], [0, 0, -60e-9], [40e-9, -30e-9, 0]]   # both poles of particle 0
    return pos, cv, pts


@pytest.mark.parametrize("lMax", [1, 2, 3, 4])
@pytest.mark.parametrize("J", [1, 2, 3, 4])
@pytest.mark.parametrize("npart", [3, 5, 80])  # 3: one thread per point; 5, 80: particles split over 4 waves
def test_scattered_field_vs_oracle(gpu, lMax, J, npart):
    import qpms_amd
    rng = np.random.default_rng(1000 * lMax + 10 * J + npart)
    pos, cv, pts = _case(rng, lMax, npart, 37)
    for k in (1.1e7, 1.3e7 * (1 + 0.02j)):
        got = qpms_amd.scattered_E(lMax, pos, k, cv, pts, J=J)
        ref = of.scattered_E(lMax, pos, k, cv.ravel(), pts, J=J)
        assert rel(got, ref) < TOL, k


@pytest.mark.parametrize("lMax", [2, 4])
def test_regular_waves_at_particle_centre(gpu, lMax):
    """kr = 0 for one particle: only its l = 1 regular electric waves count."""
    import qpms_amd
    rng = np.random.default_rng(7 + lMax)
    pos, cv, _ = _case(rng, lMax, 6, 3)
    pts = np.concatenate([pos[:2], pos[:1] + [[0, 0, 1e-9]]])
    got = qpms_amd.scattered_E(lMax, pos, 1.2e7, cv, pts, J=1)
    ref = of.scattered_E(lMax, pos, 1.2e7, cv.ravel(), pts, J=1)
    assert rel(got, ref) < TOL


@pytest.mark.parametrize("lMax", [1, 3])
@pytest.mark.parametrize("npart", [4, 64])
def test_parity_option_is_the_translation_variant(gpu, lMax, npart):
    import qpms_amd
    rng = np.random.default_rng(50 + lMax + npart)
    pos, cv, pts = _case(rng, lMax, npart, 21)
    k = 1.15e7
    got = qpms_amd.scattered_E(lMax, pos, k, cv, pts, parity=True)
    ref = of.scattered_E_alt(lMax, pos, k, cv.ravel(), pts)
    assert rel(got, ref) < TOL


def test_no_points(gpu):
    import qpms_amd
    rng = np.random.default_rng(3)
    pos, cv, _ = _case(rng, 2, 3, 3)
    assert qpms_amd.scattered_E(2, pos, 1e7, cv, np.zeros((0, 3))).shape == (0, 3)


@pytest.mark.parametrize("lMax,npart,npts", [(4, 1024, 4096), (3, 16, 300000)])
def test_full_size_sampled(gpu, lMax, npart, npts):
    """A 1024-particle cluster at lMax 4 (wave per point) and 3e5 points of a
    16-particle cluster (thread per point), compared on 24 sampled points."""
    import qpms_amd
    rng = np.random.default_rng(npart + npts)
    pos, cv, pts = _case(rng, lMax, npart, npts)
    k = 1.2e7
    got = qpms_amd.scattered_E(lMax, pos, k, cv, pts)
    idx = np.concatenate([[0, 1, 2, npts - 1], rng.choice(npts, 20, replace=False)])
    ref = of.scattered_E(lMax, pos, k, cv.ravel(), pts[idx])
    assert rel(got[idx], ref) < TOL
    assert np.all(np.isfinite(got))
