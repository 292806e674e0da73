#!/usr/bin/env python3
"""Benchmark of the translation-matrix build at scale on one MI355X.

Workload (the single-GPU configuration the headline metric is quoted on):
N = 2048 spheres placed uniform-random in a 4 x 4 x 0.5 um slab with 120 nm
minimum separation (seeded rejection sampling), lMax = 3 (30 VSWFs per
particle), so W is 61440 x 61440 complex128 (60.4 GB, resident in HBM);
background n = 1.52; 64 frequencies evenly spaced over 1.6-2.4 eV. One step
= W built for all 64 frequencies (one k_w_build launch per frequency, every
element written once, J = HANKEL_PLUS, FP64).

Prints one JSON line. `roofline` uses the kernel's algorithmic bytes (16 L^2
per launch) over the average launch duration measured with HIP events on the
library stream; `cpu_baseline` times the C oracle (a restatement of the
reference's pair loop, pthreads over host cores) on a bounded sample of rows
and extrapolates linearly in rows.
"""
import argparse
import ctypes
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

EV = 1.602176634e-19 / 1.054571817e-34   # rad/s per eV
C0 = 2.99792458e8
METRIC = "mode-problem solves/s + p99 latency at 1e-9 rel err, batched-mixed vs plain FP64"
HBM_PEAK = 8.0e12


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


class HipEvents:
    def __init__(self):
        self.h = ctypes.CDLL(os.path.join(os.environ.get("ROCM_PATH", "/opt/rocm"), "lib", "libamdhip64.so"))
        self.h.hipEventElapsedTime.argtypes = [ctypes.POINTER(ctypes.c_float), ctypes.c_void_p, ctypes.c_void_p]

    def create(self):
        e = ctypes.c_void_p()
        assert self.h.hipEventCreate(ctypes.byref(e)) == 0
        return e

    def record(self, e, stream):
        assert self.h.hipEventRecord(e, ctypes.c_void_p(stream)) == 0

    def elapsed_ms(self, a, b):
        assert self.h.hipEventSynchronize(b) == 0
        ms = ctypes.c_float()
        assert self.h.hipEventElapsedTime(ctypes.byref(ms), a, b) == 0
        return ms.value


def cpu_baseline(pos, lMax, k, nrows, threads, chunk=32):
    """Oracle W rows of the first `nrows` particles, built `chunk` particles at a
    time (bounded host memory); returns (rows of the first chunk, seconds)."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_ffi
    o = oracle_ffi.oracle()
    o.or_set_threads(threads)
    oc = oracle_ffi.OracleCalc(lMax)
    first = None
    t0 = time.perf_counter()
    for r0 in range(0, nrows, chunk):
        rows = oc.build_W_rows(pos, k, r0, min(chunk, nrows - r0))
        if first is None:
            first = rows
    dt = time.perf_counter() - t0
    return first, dt


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--nfreq", type=int, default=64)
    ap.add_argument("--npart", type=int, default=2048)
    ap.add_argument("--lmax", type=int, default=3)
    ap.add_argument("--cpu-rows", type=int, default=256, help="particles whose rows the CPU baseline builds")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--pmc-probe", action="store_true",
                    help="run one plain copy kernel of 2 GiB and one W build, then exit (for rocprofv3 --pmc)")
    ap.add_argument("--traffic-bytes", type=float, default=None,
                    help="HBM bytes per launch from rocprofv3 FETCH_SIZE+WRITE_SIZE (scaled), if measured")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU only (asked for --gpus {args.gpus}); "
              "the translation build would shard by destination-row slabs across GPUs (see DESIGN.md).",
              file=sys.stderr)
        sys.exit(2)

    import qpms_amd
    L = qpms_amd.lib()
    assert L.qpms_hip_init() == 0, "no HIP device"
    L.qpms_hip_stream.restype = ctypes.c_void_p
    stream = L.qpms_hip_stream()

    pos = cluster_positions(args.npart)
    lMax = args.lmax
    calc = qpms_amd.TransCalculator(lMax)
    nb = 2 * lMax * (lMax + 2)
    Ldim = nb * args.npart
    bytes_per_build = 16.0 * Ldim * Ldim
    omegas = np.linspace(1.6, 2.4, args.nfreq) * EV
    ks = 1.52 * omegas / C0
    dpos = qpms_amd.DeviceBuffer.from_array(pos)
    dW = qpms_amd.DeviceBuffer(int(bytes_per_build))
    ev = HipEvents()

    if args.pmc_probe:
        # calibration: qh_copy_c64 reads and writes exactly 2 GiB each
        n = (2 << 30) // 16
        src, dst = qpms_amd.DeviceBuffer(n * 16), qpms_amd.DeviceBuffer(n * 16)
        L.qh_copy_c64.argtypes = [ctypes.c_long, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
        L.qh_copy_c64(n, src.ptr, dst.ptr, None)
        L.qpms_hip_build_translation_matrix_dev(calc.c, args.npart, dpos.ptr, qpms_amd.c_cplx.of(ks[0]), 3,
                                                dW.ptr, None)
        L.qpms_hip_sync()
        print(json.dumps({"pmc_probe": True, "copy_bytes_each_way": n * 16, "w_bytes": bytes_per_build}))
        return

    def one_step(times):
        for k in ks:
            a, b = ev.create(), ev.create()
            ev.record(a, stream)
            L.qpms_hip_build_translation_matrix_dev(calc.c, args.npart, dpos.ptr, qpms_amd.c_cplx.of(k), 3,
                                                    dW.ptr, None)
            ev.record(b, stream)
            times.append((a, b))

    def log(msg):
        print(f"[bench] {msg}", file=sys.stderr, flush=True)

    log(f"setup done: L = {Ldim}, W = {bytes_per_build / 1e9:.1f} GB on the device")
    for _ in range(args.warmup):
        one_step([])
        L.qpms_hip_sync()
        log("warmup step done")
    ev_pairs = []
    t0 = time.perf_counter()
    for i in range(args.steps):
        one_step(ev_pairs)
    L.qpms_hip_sync()
    wall = time.perf_counter() - t0
    log(f"timed {args.steps} steps in {wall:.3f} s")
    kernel_ms = [ev.elapsed_ms(a, b) for a, b in ev_pairs]
    avg_ms = float(np.mean(kernel_ms))
    builds = args.steps * args.nfreq
    value = builds / wall

    # parity spot check: rows of the first two particles at the last frequency vs the oracle
    parity = None
    cpu = None
    if not args.no_cpu:
        threads = min(16, os.cpu_count() or 1)
        log(f"cpu baseline: {args.cpu_rows} particle rows on {threads} threads")
        rows, dt = cpu_baseline(pos, lMax, ks[-1], args.cpu_rows, threads)
        log(f"cpu baseline done in {dt:.1f} s")
        rows_per_s = args.cpu_rows / dt
        cpu = {"value": rows_per_s / args.npart, "unit": "translation-matrix builds/s", "cores": threads,
               "kind": "port",
               "sample": f"oracle W rows of {args.cpu_rows} of {args.npart} particles x all sources at 1 frequency "
                         f"({dt:.1f} s), extrapolated linearly to the full matrix"}
        g = np.empty((2 * nb, Ldim), np.complex128)
        L.qpms_hip_memcpy_d2h(g.ctypes.data, dW.ptr, g.nbytes)
        ref = rows[:2 * nb]
        parity = float(np.max(np.abs(g - ref)) / np.max(np.abs(ref)))

    achieved = bytes_per_build / (avg_ms * 1e-3)
    line = {
        "metric": METRIC,
        "value": value,
        "unit": "translation-matrix builds/s (61440^2 complex128, one frequency each)",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": wall / args.steps * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "c128 (fp64)", "data": "synthetic (seeded random cluster)",
        "config": {"workload": "translation-matrix build at scale (configs[1])", "npart": args.npart,
                   "lMax": lMax, "matrix_dim": Ldim, "frequencies_per_step": args.nfreq,
                   "background_n": 1.52, "band_eV": [1.6, 2.4]},
        "roofline": {"bound": "hbm", "achieved": achieved / 1e9, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK,
                     "traffic": args.traffic_bytes,
                     "kernel": "k_w_build<3,false>", "avg_launch_ms": avg_ms,
                     "algorithmic_bytes_per_launch": bytes_per_build},
        "cpu_baseline": cpu,
        "parity_max_rel_err": parity,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
