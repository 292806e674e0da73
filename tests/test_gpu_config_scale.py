"""GPU parity at the full size of the translation-build configuration:
N = 2048 particles in the seeded 4 x 4 x 0.5 um slab, lMax 3, W of order
61440 (3.77e9 complex128 entries, 60.4 GB) built on the device at one
frequency (1.6 eV, n = 1.52), then compared block by block with the oracle's
A/B arrays (translations.c:610-698 assembled as scatsystem.c:1180-1217):
12,000 seeded random blocks plus every block of the last block row and the
last block column (the entries past 2^31 and the last partial source tile),
max |gpu - oracle| <= 1e-10 max |oracle| per block; diagonal blocks exactly
zero."""
import ctypes
import os

import numpy as np
import pytest

import oracle_ffi as of

pytestmark = pytest.mark.gpu
EV = 1.602176634e-19 / 1.054571817e-34
C0 = 2.99792458e8


def _hip():
    h = ctypes.CDLL(os.path.join(os.environ.get("ROCM_PATH", "/opt/rocm"), "lib", "libamdhip64.so"))
    h.hipMemcpy2D.argtypes = [ctypes.c_void_p, ctypes.c_size_t, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_size_t,
                              ctypes.c_size_t, ctypes.c_int]
    return h


def block_from_AB(A, B, ne):
    """nb x nb block in the default basis (electric waves, then magnetic):
    same wave type takes A, the other B (translations_inlines.h:31-32)."""
    return np.block([[A, B], [B, A]])


def test_w_build_2048_particles_lmax3(gpu):
    import qpms_amd
    from qpms_amd.workloads import cluster_positions
    pos = cluster_positions()
    n, lMax = len(pos), 3
    ne = lMax * (lMax + 2)
    nb = 2 * ne
    L = n * nb
    assert L == 61440
    k = 1.52 * 1.6 * EV / C0
    calc = qpms_amd.TransCalculator(lMax)
    dpos = qpms_amd.DeviceBuffer.from_array(pos)
    dW = qpms_amd.DeviceBuffer(16 * L * L)
    print(f"W build at order {L} ({16 * L * L / 1e9:.1f} GB)", flush=True)
    gpu.qpms_hip_build_translation_matrix_dev(calc.c, n, dpos.ptr, qpms_amd.c_cplx.of(k), 3, dW.ptr, None)
    gpu.qpms_hip_sync()
    rng = np.random.default_rng(4242)
    pairs = set(map(tuple, rng.integers(0, n, size=(12000, 2))))
    pairs |= {(n - 1, c) for c in range(n)} | {(r, n - 1) for r in range(n)}
    h = _hip()
    blk = np.empty((nb, nb), np.complex128)
    o = of.OracleCalc(lMax)
    worst = 0.0
    for i, (R, Cc) in enumerate(sorted(pairs)):
        src = int(dW.ptr + 16 * (int(R) * nb * L + int(Cc) * nb))
        assert h.hipMemcpy2D(blk.ctypes.data, 16 * nb, src, 16 * L, 16 * nb, nb, 2) == 0
        if R == Cc:
            assert not blk.any(), R
            continue
        d = pos[R] - pos[Cc]
        r = np.linalg.norm(d)
        A, B = o.get_AB_arrays(k * r, np.arccos(d[2] / r), np.arctan2(d[1], d[0]), J=3)
        ref = block_from_AB(A, B, ne)
        err = np.max(np.abs(blk - ref)) / np.max(np.abs(ref))
        worst = max(worst, err)
        assert err <= 1e-10, (R, Cc, err)
        if i % 4000 == 0:
            print(f"  {i} of {len(pairs)} blocks checked, worst {worst:.2e}", flush=True)
    print(f"{len(pairs)} blocks, worst relative difference {worst:.2e}", flush=True)
    dW.free()
    dpos.free()
