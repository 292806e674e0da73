"""Seeded synthetic inputs of the benchmark configurations (shared by
bench.py and the tests).

cluster_positions: the translation-build configuration -- N particles placed
uniform-random in a 4 x 4 x 0.5 um slab with a 120 nm minimum separation, by
seeded rejection sampling (candidates drawn 4096 at a time from one PCG64
stream; a candidate is kept if it is at least dmin from every kept point).
square_array: a centred nx x ny square array in the z = 0 plane.
"""
import numpy as np


def cluster_positions(n=2048, box=(4e-6, 4e-6, 0.5e-6), dmin=120e-9, seed=20261202):
    rng = np.random.default_rng(seed)
    pts = np.empty((n, 3))
    cnt = 0
    while cnt < n:
        cand = rng.uniform(0.0, 1.0, (4096, 3)) * np.asarray(box)
        for p in cand:
            if cnt == 0 or np.min(np.sum((pts[:cnt] - p) ** 2, axis=1)) >= dmin * dmin:
                pts[cnt] = p
                cnt += 1
                if cnt == n:
                    break
    return pts


def square_array(nx, ny, pitch):
    gx = (np.arange(nx) - (nx - 1) / 2) * pitch
    gy = (np.arange(ny) - (ny - 1) / 2) * pitch
    return np.array([(x, y, 0.0) for x in gx for y in gy])
