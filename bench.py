#!/usr/bin/env python3
"""Serving benchmark on one MI355X: mode-problem solves/s and p50/p99 latency
of the batched mixed-precision serving loop against the plain FP64 drop-in
path on the same requests.

Workload: the seeded 10,000-request stream of qpms_amd.serve.serving_stream
(Poisson arrivals, clusters of N particles log-uniform in 16..1024 on a
jittered square lattice, lMax in {2,3,4}, random spheres and frequencies;
70 % plane-wave solves, 20 % 2D-periodic determinants, 10 % 32-node Beyn
contours). A step is one slice of `--slice` consecutive requests; the timed
region is the first steps x slice requests of the stream, replayed at their
arrival times into one serving loop and run until every request is answered
(nothing is cancelled or skipped). `value` = requests answered / wall time of
that replay; latency percentiles are over every request of it (arrival to
result, queueing included). The warm-up steps replay the requests that follow
the timed ones.

The plain FP64 path (qpms_hip_request_plain: the public scattering-system API
one call after another, host buffers, FP64 LU) answers the timed requests one
at a time in stream order within a time budget, never starting a request its
size says cannot finish inside the budget; the serving loop then replays
exactly the requests the plain path covered, so `vs_plain_fp64` compares the
two paths on identical work. Every plain result is compared with the served
one (extinction, log-determinant, Beyn singular values and eigenvalues).

`roofline`: the FP32 LU of a mode-problem matrix of order 44880 -- the order
of the largest request of the default timed segment, a 32-node contour whose
node factorisations are most of the segment's GPU time: tournament-pivoted
panels plus trailing complex GEMMs on MFMA -- timed live with HIP events on
the library stream and priced at 8/3 n^3 real flops against the FP32 matrix
peak; `traffic` comes from the committed rocprofv3 counter summary of the
same probe. `roofline_lu16380` is the same at order 16380. `w_build`: the plain
translation-matrix build (61440^2 complex128, HBM roofline). `cpu_baseline`:
the reference's CPU path of the plane-wave solves of the timed requests (the
oracle's pair loop for M = I - T W, host LAPACK zgetrf/zgetrs as
scatsystem.c:2214/2270 call it), all host threads and one thread.
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
FP32_MFMA_PEAK = 157.3e12
PROFILES = os.path.join(ROOT, "profiles")


def profile_traffic(name):
    """HBM bytes per launch recorded in profiles/<name>.md by its counter
    passes (the line `traffic_bytes_per_launch: <bytes>`), or None."""
    try:
        for ln in open(os.path.join(PROFILES, name + ".md")):
            if ln.startswith("traffic_bytes_per_launch:"):
                return float(ln.split(":", 1)[1])
    except OSError:
        pass
    return None


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


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


def measure_w_build(L, ev, stream, nfreq=8):
    """Translation-matrix build, N = 2048, lMax 3, FP64: HBM roofline."""
    import qpms_amd
    from qpms_amd.workloads import cluster_positions
    pos = cluster_positions()
    calc = qpms_amd.TransCalculator(3)
    Ldim = 30 * len(pos)
    nbytes = 16.0 * Ldim * Ldim
    dpos = qpms_amd.DeviceBuffer.from_array(pos)
    dW = qpms_amd.DeviceBuffer(int(nbytes))
    ks = 1.52 * np.linspace(1.6, 2.4, nfreq) * EV / C0
    L.qpms_hip_build_translation_matrix_dev(calc.c, len(pos), dpos.ptr, qpms_amd.c_cplx.of(ks[0]), 3, dW.ptr, None)
    L.qpms_hip_sync()
    pairs = []
    for k in ks:
        a, b = ev.create(), ev.create()
        ev.record(a, stream)
        L.qpms_hip_build_translation_matrix_dev(calc.c, len(pos), dpos.ptr, qpms_amd.c_cplx.of(k), 3, dW.ptr, None)
        ev.record(b, stream)
        pairs.append((a, b))
    L.qpms_hip_sync()
    avg = float(np.mean([ev.elapsed_ms(a, b) for a, b in pairs]))
    dW.free()
    ach = nbytes / (avg * 1e-3)
    return {"kernel": "k_w_build<3,false>", "workload": "W of 2048 particles, lMax 3 (61440^2 complex128), FP64",
            "bound": "hbm", "achieved": ach / 1e9, "peak": HBM_PEAK / 1e9, "unit": "GB/s", "frac": ach / HBM_PEAK,
            "traffic": profile_traffic("k_w_build_translation_2048p_lmax3"), "avg_launch_ms": avg,
            "algorithmic_bytes_per_launch": nbytes}


def w_probe(L):
    """Counter-pass probe: one plain copy kernel over 2 GiB (k_copy2d reads
    and writes 2 GiB: the FETCH_SIZE / WRITE_SIZE calibration), then one
    translation-matrix build of the N = 2048, lMax 3 workload."""
    import qpms_amd
    from qpms_amd.workloads import cluster_positions
    copy_calibration(L)
    pos = cluster_positions()
    calc = qpms_amd.TransCalculator(3)
    Ld = 30 * len(pos)
    dpos, dW = qpms_amd.DeviceBuffer.from_array(pos), qpms_amd.DeviceBuffer(16 * Ld * Ld)
    L.qpms_hip_build_translation_matrix_dev(calc.c, len(pos), dpos.ptr, qpms_amd.c_cplx.of(1.52 * 2.0 * EV / C0), 3,
                                            dW.ptr, None)
    L.qpms_hip_sync()
    print(json.dumps({"w_probe": "done", "copy_bytes_each_way": 16 << 27, "w_bytes": 16 * Ld * Ld}), flush=True)


def measure_peaks(L, ev, stream):
    """Measured counterparts of the datasheet peaks: STREAM triad over three
    2 GiB arrays and a write-only fill of one (hand-written kernels), FP64 FMA
    chains (hand-written kernel),
    and rocBLAS cgemm / zgemm at order 8192 (the library GEMM the LU's
    trailing update calls)."""
    import qpms_amd
    L.qpms_hip_probe_triad.argtypes = [ctypes.c_size_t, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                       ctypes.c_double, ctypes.c_void_p]
    L.qpms_hip_probe_fma64.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p]

    def timed(fn, reps):
        fn()
        L.qpms_hip_sync()
        a, b = ev.create(), ev.create()
        ev.record(a, stream)
        for _ in range(reps):
            fn()
        ev.record(b, stream)
        return ev.elapsed_ms(a, b) / reps

    out = {}
    n = 1 << 28
    bufs = [qpms_amd.DeviceBuffer(8 * n) for _ in range(3)]
    ms = timed(lambda: L.qpms_hip_probe_triad(n, bufs[0].ptr, bufs[1].ptr, bufs[2].ptr, 0.5, None), 5)
    out["hbm_triad_GBps"] = 24.0 * n / (ms * 1e-3) / 1e9
    L.qpms_hip_probe_fill.argtypes = [ctypes.c_size_t, ctypes.c_void_p, ctypes.c_double, ctypes.c_void_p]
    ms = timed(lambda: L.qpms_hip_probe_fill(n, bufs[0].ptr, 0.5, None), 5)
    out["hbm_write_GBps"] = 8.0 * n / (ms * 1e-3) / 1e9
    for b in bufs:
        b.free()
    blocks, iters = 256 * 8, 1 << 15
    o = qpms_amd.DeviceBuffer(blocks * 256 * 8)
    ms = timed(lambda: L.qpms_hip_probe_fma64(blocks, iters, o.ptr, None), 3)
    out["fp64_fma_TFps"] = 16.0 * iters * 256 * blocks / (ms * 1e-3) / 1e12
    o.free()
    try:
        rb = ctypes.CDLL(os.path.join(os.environ.get("ROCM_PATH", "/opt/rocm"), "lib", "librocblas.so"))
        h = ctypes.c_void_p()
        assert rb.rocblas_create_handle(ctypes.byref(h)) == 0
        assert rb.rocblas_set_stream(h, ctypes.c_void_p(stream)) == 0
        nn = 8192
        for name, fn, esz in (("cgemm", rb.rocblas_cgemm, 8), ("zgemm", rb.rocblas_zgemm, 16)):
            A, B, C = (qpms_amd.DeviceBuffer(nn * nn * esz) for _ in range(3))
            one = (ctypes.c_double * 2)(1, 0) if esz == 16 else (ctypes.c_float * 2)(1, 0)
            zero = (ctypes.c_double * 2)(0, 0) if esz == 16 else (ctypes.c_float * 2)(0, 0)

            def g():
                assert fn(h, 111, 111, nn, nn, nn, one, ctypes.c_void_p(A.ptr), nn, ctypes.c_void_p(B.ptr), nn, zero,
                          ctypes.c_void_p(C.ptr), nn) == 0
            ms = timed(g, 3)
            out[name + "_TFps"] = 8.0 * nn ** 3 / (ms * 1e-3) / 1e12
            for b in (A, B, C):
                b.free()
        rb.rocblas_destroy_handle(h)
    except (OSError, AttributeError, AssertionError) as e:
        log(f"rocBLAS GEMM probe skipped: {e}")
    out["datasheet"] = {"hbm_GBps": HBM_PEAK / 1e9, "fp64_vector_TFps": 78.6, "fp32_matrix_TFps": 157.3,
                        "fp64_matrix_TFps": 78.6}
    return out


def autotune(L):
    """Runs the library's GEMM autotuner (trailing-update GEMM per size class:
    rocBLAS or one of the hand-written 3M / 4M MFMA kernels, the fastest on
    this device) and returns its table."""
    L.qpms_hip_autotune.argtypes = [ctypes.c_void_p]
    t0 = time.perf_counter()
    L.qpms_hip_autotune(None)
    buf = ctypes.create_string_buffer(4096)
    L.qpms_hip_autotune_table(buf, ctypes.c_size_t(4096))
    table = {}
    for ln in buf.value.decode().splitlines():
        b, d, v = ln.split()
        table[f"{'batched' if b == '1' else 'single'}_{d}"] = v
    log(f"GEMM autotune ({time.perf_counter() - t0:.1f} s): " + ", ".join(f"{k} {v}" for k, v in table.items()))
    return table


def copy_calibration(L):
    """One plain copy kernel over 2 GiB (k_copy2d reads and writes 2 GiB):
    the FETCH_SIZE / WRITE_SIZE calibration of a counter pass."""
    import qpms_amd
    n = 1 << 27
    a, b = qpms_amd.DeviceBuffer(16 * n), qpms_amd.DeviceBuffer(16 * n)
    L.qh_copy_c64.argtypes = [ctypes.c_long, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
    L.qh_copy_c64(n, a.ptr, b.ptr, None)
    L.qpms_hip_sync()
    a.free()
    b.free()


def measure_lu(L, ev, stream, n_part=1496, reps=3, calib=False):
    """FP32 LU of a mode-problem matrix of order 30 * n_part (44880 by
    default: the order of the timed segment's largest request), HIP events on
    the library stream. calib: a copy kernel first (counter passes)."""
    import qpms_amd
    if calib:
        copy_calibration(L)
    rng = np.random.default_rng(2)
    n = 30 * n_part
    side = (n_part ** (1 / 3)) * 300e-9
    pos = rng.uniform(0, side, size=(n_part, 3))
    calc = qpms_amd.TransCalculator(3)
    k = 1.52 * 2.0 * EV / C0
    i32, vp, lg = ctypes.c_int, ctypes.c_void_p, ctypes.c_long
    L.qh_getrf_batched_f32.argtypes = [i32, i32, vp, lg, vp, vp, vp, vp]
    L.qh_d2f.argtypes = [lg, vp, vp, vp]
    dpos = qpms_amd.DeviceBuffer.from_array(pos)
    dty = qpms_amd.DeviceBuffer.from_array(np.zeros(n_part, np.int32))
    dtd = qpms_amd.DeviceBuffer.from_array(np.full(30, -0.05 + 0.02j))
    dM = qpms_amd.DeviceBuffer(n * n * 16)
    dF = qpms_amd.DeviceBuffer(n * n * 8)
    dp, dpm = qpms_amd.DeviceBuffer(4 * n), qpms_amd.DeviceBuffer(4 * n)
    dinfo = qpms_amd.DeviceBuffer.from_array(np.zeros(1, np.int32))
    L.qpms_hip_build_modeproblem_matrix_dev(calc.c, n_part, dpos.ptr, dty.ptr, dtd.ptr, qpms_amd.c_cplx.of(k),
                                            dM.ptr, None)
    times = []
    for r in range(reps + 1):
        L.qh_d2f(n * n, dM.ptr, dF.ptr, None)
        a, b = ev.create(), ev.create()
        ev.record(a, stream)
        L.qh_getrf_batched_f32(n, 1, dF.ptr, 0, dp.ptr, dpm.ptr, dinfo.ptr, None)
        ev.record(b, stream)
        ms = ev.elapsed_ms(a, b)
        log(f"LU of order {n}: factorisation {r + 1} of {reps + 1} ({'timed' if r else 'warm-up'}) {ms:.1f} ms")
        if r:
            times.append(ms)
    for x in (dM, dF, dp, dpm, dinfo, dpos, dty, dtd):
        x.free()
    avg = float(np.mean(times))
    flops = 8.0 / 3.0 * n ** 3
    ach = flops / (avg * 1e-3)
    return {"bound": "mfma", "achieved": ach / 1e12, "peak": FP32_MFMA_PEAK / 1e12, "unit": "TFLOP/s",
            "frac": ach / FP32_MFMA_PEAK,
            # HBM bytes per factorisation from the two rocprofv3 --pmc passes of
            # this probe (FETCH_SIZE, WRITE_SIZE, copy-kernel calibration)
            "traffic": profile_traffic(f"lu_fp32_order{n}"),
            "kernel": f"FP32 LU, order {n} (tournament panels + trailing GEMM)",
            "avg_launch_ms": avg, "algorithmic_flops_per_launch": flops}


def meminfo_gib():
    """(free, total) device memory in GiB (hipMemGetInfo)."""
    h = ctypes.CDLL(os.path.join(os.environ.get("ROCM_PATH", "/opt/rocm"), "lib", "libamdhip64.so"))
    f, t = ctypes.c_size_t(), ctypes.c_size_t()
    assert h.hipMemGetInfo(ctypes.byref(f), ctypes.byref(t)) == 0
    return f.value / 2 ** 30, t.value / 2 ** 30


def replay(srv, reqs, label, burst=False):
    """Submit reqs at their arrival times (relative to the first one; all at
    once with burst) and run the loop until every request is answered.
    Returns ({index: (response, seconds from the first submission)}, wall
    seconds)."""
    t_first = reqs[0].arrival
    id2i, got = {}, {}
    nxt = 0
    t0 = time.perf_counter()
    last_log = t0
    while len(got) < len(reqs):
        now = time.perf_counter() - t0
        while nxt < len(reqs) and (burst or reqs[nxt].arrival - t_first <= now):
            id2i[srv.submit(reqs[nxt])] = nxt
            nxt += 1
        srv.progress(wait=False)
        for r in srv.poll():
            got[id2i[r.id]] = (r, time.perf_counter() - t0)
        if time.perf_counter() - last_log > 15:
            log(f"{label}: t={now:.0f}s submitted {nxt}/{len(reqs)} answered {len(got)}")
            last_log = time.perf_counter()
        time.sleep(0.0005)
    return got, time.perf_counter() - t0


def rel_diff(a, b, req):
    from qpms_amd import serve as sv
    if req.kind == sv.KIND_SCATTER:
        return abs(a.sigma_ext - b.sigma_ext) / abs(b.sigma_ext)
    if req.kind == sv.KIND_PERIODIC:
        return abs(a.logabsdet - b.logabsdet) / max(1.0, abs(b.logabsdet))
    return sv.beyn_rel_diff(a, b, req.c.rank_tol)


def plain_estimate_s(req):
    """Seconds of one plain-path request from its order (used only to decide
    whether it can still finish inside the plain path's budget): FP64 LUs at
    30 TFLOP/s and the host round trips of the matrix through pageable
    buffers at 5.5 GB/s (a plane-wave solve moves it four times: build,
    factorise in and out, solve; a contour node twice), plus fixed costs.
    Calibrated on the MI355X box: a plane-wave solve of order 14,790 takes
    2.9 s."""
    from qpms_amd import serve as sv
    L = float(req.dim)
    lu = 8.0 / 3.0 * L ** 3 / 30e12
    if req.kind == sv.KIND_BEYN:
        return req.c.beyn_nodes * (lu + 32.0 * L * L / 5.5e9 + 0.02)
    if req.kind == sv.KIND_PERIODIC:
        return lu + 32.0 * L * L / 5.5e9 + 0.5
    return lu + 64.0 * L * L / 5.5e9 + 0.05


def cpu_model():
    try:
        for ln in open("/proc/cpuinfo"):
            if ln.startswith("model name"):
                return ln.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline_scatter(reqs, served, budget_s):
    """The reference's CPU path (oracle pair loop for M = I - T W, host LAPACK
    zgetrf/zgetrs, extinction) on the plane-wave requests of `reqs` in stream
    order, once on every host thread and once on one thread, each until
    budget_s of wall time is used (whole requests only). Also the largest
    relative difference of the served extinctions from it."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_ffi
    from qpms_amd import serve as sv
    lap = oracle_ffi.host_lapack()
    o = oracle_ffi.oracle()
    cores = min(16, os.cpu_count() or 1)
    pw = [i for i, r in enumerate(reqs) if r.kind == sv.KIND_SCATTER]
    out = {"unit": "plane-wave solves/s", "kind": "port", "cpu_model": cpu_model(), "lapack": lap.name}
    diffs = []
    for tag, threads in (("all", cores), ("one", 1)):
        lap.set_threads(threads)
        o.or_set_threads(threads)
        done, t_used, dims = 0, 0.0, []
        for i in pw:
            r = reqs[i]
            # host seconds: the oracle's pair loop (~25 us per particle pair) and the LU (~12 GFLOP/s a thread)
            est = (float(r.c.npart) ** 2 * 25e-6 + 8.0 / 3.0 * float(r.dim) ** 3 / 12e9) / threads
            if t_used + est > budget_s:
                continue   # this order alone would overrun what is left of the budget
            t0 = time.perf_counter()
            se = oracle_ffi.reference_cpu_solve(r.pos, r.c.lMax, r.c.radius, complex(r.c.eps_in.re, r.c.eps_in.im),
                                                r.c.eps_bg, r.c.omega, (0, 0, 1), np.array([1, 0, 0], complex), lap)
            t_used += time.perf_counter() - t0
            done += 1
            dims.append(r.dim)
            if tag == "all" and i in served:
                diffs.append(abs(served[i][0].sigma_ext - se) / abs(se))
            if t_used > budget_s:
                break
        out[tag] = {"threads": threads, "requests": done, "seconds": t_used,
                    "value": done / t_used if t_used > 0 else None, "orders": dims}
        log(f"cpu baseline ({threads} threads): {done} plane-wave solves in {t_used:.1f} s")
    lap.set_threads(cores)
    out["value"] = out["all"]["value"]
    out["cores"] = cores
    out["sample"] = (f"the reference's CPU path on the plane-wave requests of the timed segment in stream order, "
                     f"skipping those too large for what is left of the {budget_s:.0f} s budget "
                     f"({out['all']['requests']} on {cores} threads, {out['one']['requests']} on 1 thread); the GPU "
                     f"value covers every request kind and order")
    out["served_vs_cpu_max_rel_diff"] = float(max(diffs)) if diffs else None
    return out


def bench_ewald(L, ev, stream):
    """2D-periodic Ewald sums: hexagonal lattice a = 375 nm, shifts {0, +s, -s}
    (s = a/sqrt3), lMax 3 (order 7), eta = 2.5/a, R: 12 rings, K: |K| < 30 2pi/a;
    8192 (beta, omega) points (64 x 64 over the first zone x 2 frequencies in
    1.5-2.3 eV, 5 % moved to within 1e-4 of the light cone) x 3 shifts, FP64."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import ewald_common as ec
    import qpms_amd
    rng = np.random.default_rng(20261202)
    a = ec.A_LAT
    s = a / np.sqrt(3)
    b1 = np.array(ec.RB1)
    b2 = np.array(ec.RB2)
    tri = []
    for eV in (1.5, 2.3):
        k = ec.wavenumber(eV)
        for i in range(64):
            for j in range(64):
                beta = ((i + 0.5) / 64 - 0.5) * b1 + ((j + 0.5) / 64 - 0.5) * b2
                if rng.uniform() < 0.05:   # onto the light cone of the nearest order, within 1e-4 k
                    d = rng.normal(size=2)
                    beta = d / np.linalg.norm(d) * k * (1 + rng.uniform(-1e-4, 1e-4))
                for sh in ((0, 0), (s, 0), (-s, 0)):
                    tri.append([beta[0], beta[1], 0, k, 0, sh[0], sh[1], 0])
    tri = np.array(tri)
    Le = ec.lib_ewald()
    # the point sets from the library's own xyWeb generator (PGen_xyWeb_*)
    R = lattice_points(Le, ec.B1, ec.B2, 12.5 * ec.H_R)
    R3 = np.concatenate([R, np.zeros((len(R), 1))], axis=1)
    K = lattice_points(Le, ec.RB1, ec.RB2, 30 * 2 * np.pi / ec.A_LAT)
    e3c = Le.qpms_ewald3_constants_init(7, -1)
    n = len(tri)
    dtr, dR, dK = (qpms_amd.DeviceBuffer.from_array(x) for x in (tri, R3, K))
    nsc = 64
    dss, dsl = qpms_amd.DeviceBuffer(n * nsc * 16), qpms_amd.DeviceBuffer(n * nsc * 16)
    times = {}
    for part, (ps, pl) in (("short", (dss.ptr, None)), ("long", (None, dsl.ptr))):
        Le.qpms_hip_ewald3_sigma_batch_dev(e3c, ec.ETA, ec.AREA, n, dtr.ptr, len(R3), dR.ptr, len(K), dK.ptr, ps, pl,
                                           None)
        L.qpms_hip_sync()
        a0, a1 = ev.create(), ev.create()
        ev.record(a0, stream)
        Le.qpms_hip_ewald3_sigma_batch_dev(e3c, ec.ETA, ec.AREA, n, dtr.ptr, len(R3), dR.ptr, len(K), dK.ptr, ps, pl,
                                           None)
        ev.record(a1, stream)
        times[part] = ev.elapsed_ms(a0, a1)
    ss = dss.to_array(np.complex128, (n, nsc))
    sl = dsl.to_array(np.complex128, (n, nsc))
    Le.qpms_ewald3_constants_free(e3c)
    cpu = ewald_cpu_baseline(ec, tri, ss, sl, R, K)
    tot = times["short"] + times["long"]
    return {"metric": "Ewald sigma_short + sigma_long triples/s (FP64)", "value": n / (tot * 1e-3),
            "unit": "(beta, omega, shift) triples/s", "triples": n, "R_points": len(R), "K_points": len(K),
            "short_ms": times["short"], "long_ms": times["long"],
            "short_ns_per_point_pair": times["short"] * 1e6 / (n * len(R)),
            "long_ns_per_point_pair": times["long"] * 1e6 / (n * len(K)),
            "cpu_baseline": cpu}


def lattice_points(Le, b1, b2, maxR):
    """xyWeb points of the lattice (b1, b2) inside maxR (origin included,
    strict cut), from the library's PGen_xyWeb generator."""
    import qpms_amd
    if not hasattr(Le, "_pgen_bound"):
        import ewald_common as ec
        Le.PGen_next_cart2.restype = type("PGenCart2", (ctypes.Structure,), {
            "_fields_": [("flags", ctypes.c_int), ("x", ctypes.c_double), ("y", ctypes.c_double)]})
        Le.PGen_next_cart2.argtypes = [ctypes.POINTER(ec.PGen)]
        Le._pgen_bound = True
    g = Le.PGen_xyWeb_new(qpms_amd.cart2_t(*b1), qpms_amd.cart2_t(*b2), 1e-13, qpms_amd.cart2_t(0, 0), 0, True, maxR,
                          False)
    pts = []
    while True:
        r = Le.PGen_next_cart2(ctypes.byref(g))
        if not (r.flags & 2):   # PGEN_NOTDONE
            break
        pts.append((r.x, r.y))
    return np.array(pts)


def ewald_cpu_baseline(ec, tri, ss, sl, R, K):
    """The oracle's sigma_short + sigma_long on 12 triples spread over the
    workload, one host thread: its rate, and the largest relative difference
    of the GPU's sums on those triples."""
    n = len(tri)
    err, cnt = 0.0, 0
    t0 = time.perf_counter()
    for t in range(0, n, n // 12):
        k, beta, sh = tri[t, 3], tri[t, :2], tri[t, 5:7]
        rr = R[1:] if (sh == 0).all() else R   # the origin carries no term for s = 0
        o_s = ec.oracle_sigma_short(7, k, beta, sh, rr)
        o_l = ec.oracle_sigma_long(7, k, beta, sh, K)
        err = max(err, np.max(np.abs(ss[t] - o_s)) / np.max(np.abs(o_s)),
                  np.max(np.abs(sl[t] - o_l)) / np.max(np.abs(o_l)))
        cnt += 1
    t_cpu = time.perf_counter() - t0
    return {"value": cnt / t_cpu, "unit": "(beta, omega, shift) triples/s", "cores": 1, "kind": "port",
            "sample": f"oracle sigma_short + sigma_long on {cnt} triples spread over the workload ({t_cpu:.1f} s)",
            "max_rel_diff_gpu_vs_oracle": err}


def bench_field(L, ev, stream):
    """Scattered-field map: the field of a 32 x 32 square array (pitch 400 nm,
    lMax 3, seeded random outgoing-wave coefficients) on a 256 x 256 grid of
    points 200 nm above it, FP64; one k_scattered_e launch per map."""
    import qpms_amd
    lMax, ne = 3, 15
    g = (np.arange(32) - 15.5) * 400e-9
    pos = np.array([(x, y, 0.0) for x in g for y in g])
    rng = np.random.default_rng(5)
    cv = (rng.normal(size=(len(pos), 2 * ne)) + 1j * rng.normal(size=(len(pos), 2 * ne))) * 1e-2
    u = np.linspace(-7e-6, 7e-6, 256)
    pts = np.array([(x, y, 200e-9) for x in u for y in u])
    k = 1.52 * 2.0 * EV / C0
    dp, dc, dw = (qpms_amd.DeviceBuffer.from_array(a) for a in (pos, cv, pts))
    do = qpms_amd.DeviceBuffer(len(pts) * 48)
    L.qpms_hip_scattered_E_dev.argtypes = [ctypes.c_int, ctypes.c_size_t, ctypes.c_void_p, qpms_amd.c_cplx,
                                           ctypes.c_int, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_void_p,
                                           ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p]

    def run():
        L.qpms_hip_scattered_E_dev(lMax, len(pos), dp.ptr, qpms_amd.c_cplx.of(k), 3, dc.ptr, len(pts), dw.ptr,
                                   do.ptr, 0, None)
    run()
    L.qpms_hip_sync()
    a, b = ev.create(), ev.create()
    ev.record(a, stream)
    reps = 5
    for _ in range(reps):
        run()
    ev.record(b, stream)
    ms = ev.elapsed_ms(a, b) / reps
    got = do.to_array(np.complex128, (len(pts), 3))
    for x in (dp, dc, dw, do):
        x.free()
    pairs = len(pts) * len(pos)
    return {"metric": "scattered-field (point, particle) pairs/s (FP64)", "value": pairs / (ms * 1e-3),
            "unit": "pairs/s", "points": len(pts), "particles": len(pos), "lMax": lMax, "ms_per_map": ms,
            "cpu_baseline": field_cpu_baseline(lMax, pos, k, cv, pts, got, rng.choice(len(pts), 48, replace=False))}


def field_cpu_baseline(lMax, pos, k, cv, pts, got, idx):
    """The oracle's field sum (long double, one host thread) at the sampled
    points: its rate, and the largest relative difference of the GPU's map there."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_ffi as of
    t0 = time.perf_counter()
    ref = of.scattered_E(lMax, pos, k, cv.ravel(), pts[idx])
    t_cpu = time.perf_counter() - t0
    return {"value": len(idx) * len(pos) / t_cpu, "unit": "pairs/s", "cores": 1, "kind": "port",
            "sample": f"oracle field sum at {len(idx)} of the map's points ({t_cpu:.1f} s)",
            "max_rel_diff_gpu_vs_oracle": float(np.max(np.abs(got[idx] - ref)) / np.max(np.abs(ref)))}


def bench_beyn(L, window=None):
    """Beyn contour on a 24 x 24 cluster (pitch 400 nm, r 40 nm, eps -12+1i,
    n 1.52, lMax 2: m = 9216), ellipse 2.05 eV +- 0.05 eV (Re) x 0.005 eV
    (Im), 96 nodes, 64 probe columns: batched FP32 node LUs + FP64
    refinement in the serving loop, against the FP64 path (beyn_solve_v with
    M from the public builder at every node)."""
    from qpms_amd import serve as sv
    g = (np.arange(24) - 11.5) * 400e-9
    pos = np.array([(x, y, 0.0) for x in g for y in g])
    req = sv.Req(sv.KIND_BEYN, pos, 2, 40e-9, -12 + 1j, 1.52 ** 2, 2.05 * EV, beyn_rRe=0.05 * EV,
                 beyn_rIm=0.005 * EV, beyn_nodes=96, beyn_l=64, beyn_seed=1,
                 rank_tol=sv.beyn_rank_tol(9216, 64, 0.05 * EV))
    out = {}
    for mixed in (True, False):
        srv = sv.Server(arena_gib=100, mixed=mixed)
        t0 = time.perf_counter()
        rid = srv.submit(req)
        got = None
        while srv.pending():
            srv.progress(wait=True)
            for r in srv.poll():
                got = r
        dt = time.perf_counter() - t0
        srv.close()
        out["mixed" if mixed else "fp64"] = (got, dt)
        log(f"beyn {'mixed' if mixed else 'fp64'}: {dt:.2f} s, neig {got.neig}")
    t0 = time.perf_counter()
    plain = sv.request_plain(req)
    tp = time.perf_counter() - t0
    log(f"beyn plain FP64 path: {tp:.2f} s")
    gm, gf = out["mixed"][0], out["fp64"][0]
    svm, svp = np.array(gm.sv[:64]), np.array(plain.sv[:64])
    return {"metric": "Beyn contour (m = 9216, 96 nodes, l = 64) wall time", "unit": "s",
            "mixed_s": out["mixed"][1], "fp64_loop_s": out["fp64"][1], "plain_fp64_s": tp,
            "speedup_vs_plain": tp / out["mixed"][1],
            "node_lu_flops": 96 * 8 / 3 * 9216.0 ** 3,
            "mixed_node_lu_TFLOPs": 96 * 8 / 3 * 9216.0 ** 3 / out["mixed"][1] / 1e12,
            "neig": [gm.neig, gf.neig, plain.neig],
            "rank_tol": req.c.rank_tol, "sv_top": [float(x) for x in svp[:8]],
            "max_rel_diff_mixed_vs_plain": sv.beyn_rel_diff(gm, plain, req.c.rank_tol),
            "max_rel_diff_fp64_loop_vs_plain": sv.beyn_rel_diff(gf, plain, req.c.rank_tol),
            "eigval_mixed": [str(z) for z in gm.eigenvalues()], "eigval_plain": [str(z) for z in
                                                                               plain.eigenvalues()]}


def per_kind(n, plain=True):
    """Each of the first n stream requests (order < 25000, contours < 6000)
    served alone, then (plain=True) through the plain path: times and
    differences."""
    from qpms_amd import serve as sv
    reqs = [r.clone(0.0) for r in sv.serving_stream(n=10000, rate=10.0)
            if r.dim <= 25000 and not (r.kind == sv.KIND_BEYN and r.dim > 6000)][:n]
    srv = sv.Server(arena_gib=150, nstreams=8, mixed=True)
    for r in [x for x in reqs if x.dim < 2000][:6]:   # warm-up
        srv.submit(r)
    while srv.pending():
        srv.progress(wait=True)
        srv.poll()
    tot = {k: [0.0, 0.0, 0.0, 0] for k in (0, 1, 2)}
    for i, r in enumerate(reqs):
        t0 = time.perf_counter()
        srv.submit(r)
        g = None
        while srv.pending():
            srv.progress(wait=True)
            for x in srv.poll():
                g = x
        ts = time.perf_counter() - t0
        tp, err = 0.0, 0.0
        if plain:
            t0 = time.perf_counter()
            p = sv.request_plain(r)
            tp = time.perf_counter() - t0
            err = rel_diff(g, p, r)
        a = tot[r.kind]
        a[0] += ts; a[1] += tp; a[2] = max(a[2], err); a[3] += 1
        log(f"{i:3d} kind {r.kind} order {r.dim:6d}: served {ts:.3f} s, plain {tp:.3f} s, rel diff {err:.1e}")
    srv.close()
    return {"metric": "per-kind served-alone vs plain FP64 time", "requests": n,
            "by_kind": {str(k): {"requests": a[3], "served_s": a[0], "plain_s": a[1],
                                 "ratio": a[1] / a[0] if a[0] > 0 else None, "max_rel_diff": a[2]}
                        for k, a in tot.items()}}


def burst(n):
    """The requests of per_kind(n) submitted at once: the loop's wall time to
    drain them, against their served-alone times (per_kind's served_s)."""
    from qpms_amd import serve as sv
    reqs = [r.clone(0.0) for r in sv.serving_stream(n=10000, rate=10.0)
            if r.dim <= 25000 and not (r.kind == sv.KIND_BEYN and r.dim > 6000)][:n]
    srv = sv.Server(arena_gib=150, nstreams=8, mixed=True)
    for r in [x for x in reqs if x.dim < 2000][:6]:   # warm-up
        srv.submit(r)
    while srv.pending():
        srv.progress(wait=True)
        srv.poll()
    t0 = time.perf_counter()
    for r in reqs:
        srv.submit(r)
    got = []
    while srv.pending():
        srv.progress(wait=True)
        got += srv.poll()
    wall = time.perf_counter() - t0
    srv.close()
    bsz = [r.batch for r in got if r.kind == sv.KIND_SCATTER]
    return {"metric": "burst drain wall time", "requests": n, "wall_s": wall, "answered": len(got),
            "batch_sizes": sorted(bsz)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--workload", choices=["serving", "ewald", "beyn", "peaks", "field"], default="serving")
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--slice", type=int, default=2, help="requests of the stream per step")
    ap.add_argument("--rate", type=float, default=2.0,
                    help="request arrival rate (1/s); above the loop's capacity on this stream, so the replay "
                         "measures capacity")
    ap.add_argument("--arena-gib", type=float, default=150.0,
                    help="serving arena (the rest of the device is left to the plain path's own allocations)")
    ap.add_argument("--plain-budget", type=float, default=150.0, help="seconds for the plain FP64 comparison")
    ap.add_argument("--cpu-budget", type=float, default=12.0, help="seconds per thread setting for the CPU leg")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-wbuild", action="store_true")
    ap.add_argument("--lu-probe", action="store_true", help="only the FP32 LU measurement (for rocprofv3)")
    ap.add_argument("--lu-particles", type=int, default=1496,
                    help="particles (lMax 3) of the LU probe's matrix: order 30 x this (1496: 44880, 546: 16380); "
                         "the probe runs a 2 GiB copy kernel first (counter calibration)")
    ap.add_argument("--lu-reps", type=int, default=2, help="timed factorisations of the LU probe (after one warm-up)")
    ap.add_argument("--no-autotune", action="store_true",
                    help="LU probe without the GEMM autotuner (counter passes: the tuner's own factorisations "
                         "crash rocprofv3 --pmc; QPMS_HIP_GEMM then names the variant)")
    ap.add_argument("--w-probe", action="store_true",
                    help="one 2 GiB copy kernel (counter calibration) and one N = 2048 W build (for rocprofv3 --pmc)")
    ap.add_argument("--per-kind", type=int, default=0, help="serve the first N requests one at a time vs plain")
    ap.add_argument("--burst", type=int, default=0, help="submit the per-kind requests at once and drain them")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU only (asked for --gpus {args.gpus}); requests would be "
              "sharded across GPUs by the serving front end, one loop per GPU.", file=sys.stderr)
        sys.exit(2)

    import qpms_amd
    from qpms_amd import serve as sv
    L = qpms_amd.lib()
    assert L.qpms_hip_init() == 0, "no HIP device"
    L.qpms_hip_stream.restype = ctypes.c_void_p
    stream = L.qpms_hip_stream()
    ev = HipEvents()

    if args.w_probe:
        w_probe(L)
        return
    if args.lu_probe:
        if not args.no_autotune:
            autotune(L)
        print(json.dumps(measure_lu(L, ev, stream, n_part=args.lu_particles, reps=args.lu_reps, calib=True)),
              flush=True)
        return
    if args.burst:
        print(json.dumps(burst(args.burst)), flush=True)
        return
    if args.per_kind:
        print(json.dumps(per_kind(args.per_kind, plain=not args.no_plain)), flush=True)
        return
    if args.workload == "ewald":
        print(json.dumps(bench_ewald(L, ev, stream)), flush=True)
        return
    if args.workload == "peaks":
        print(json.dumps(measure_peaks(L, ev, stream)), flush=True)
        return
    if args.workload == "field":
        print(json.dumps(bench_field(L, ev, stream)), flush=True)
        return
    if args.workload == "beyn":
        print(json.dumps(bench_beyn(L)), flush=True)
        return

    peaks = measure_peaks(L, ev, stream)
    log("measured peaks: " + ", ".join(f"{k} {v:.1f}" for k, v in peaks.items() if k != "datasheet"))
    wb = None if args.no_wbuild else measure_w_build(L, ev, stream)
    if wb:
        log(f"W build: {wb['achieved']:.0f} GB/s ({100 * wb['frac']:.1f} % of HBM peak)")
    tune = autotune(L)
    roof = measure_lu(L, ev, stream, n_part=1496, reps=2)
    log(f"FP32 LU n=44880: {roof['avg_launch_ms']:.1f} ms, {roof['achieved']:.1f} TF/s")
    roof16 = measure_lu(L, ev, stream, n_part=546, reps=3)
    log(f"FP32 LU n=16380: {roof16['avg_launch_ms']:.1f} ms, {roof16['achieved']:.1f} TF/s")

    S, K, W = args.slice, args.steps, args.warmup
    reqs = sv.serving_stream(n=10000, rate=args.rate)
    timed, warm = reqs[:K * S], reqs[K * S:(K + W) * S]
    f0, tot = meminfo_gib()
    srv = sv.Server(arena_gib=args.arena_gib, nstreams=8, mixed=True)
    log(f"device memory: {f0:.1f} of {tot:.1f} GiB free before the server; arena "
        f"{srv.arena_bytes() / 2 ** 30:.0f} GiB; {meminfo_gib()[0]:.1f} GiB free after")
    if warm:
        got_w, wall_w = replay(srv, warm, "warm-up")
        log(f"warm-up: {len(warm)} requests ({W} steps) in {wall_w:.1f} s")

    L.qpms_hip_sync()
    got, wall = replay(srv, timed, "timed")
    L.qpms_hip_sync()
    t_first = timed[0].arrival
    lat = np.array([t - (timed[i].arrival - t_first) for i, (r, t) in got.items()])
    failed = sorted(i for i, (r, _) in got.items() if r.status != 0)
    value = len(timed) / wall
    kinds = {str(k): sum(1 for r in timed if r.kind == k) for k in (0, 1, 2)}
    step_done = [max(got[i][1] for i in range(k * S, (k + 1) * S)) for k in range(K)]
    bsz = np.array([r.batch for r, _ in got.values() if r.kind == sv.KIND_SCATTER] or [0])
    log("plane-wave solves by batch size: " + ", ".join(f"{b}: {int(np.sum(bsz == b))}" for b in np.unique(bsz)))
    log(f"timed: {len(timed)} requests ({kinds}) answered in {wall:.1f} s = {value:.3f}/s; {len(failed)} failed; "
        f"p50 {np.percentile(lat, 50):.2f} s p99 {np.percentile(lat, 99):.2f} s")

    plain = None
    if not args.no_plain:
        times, errs, kinds_p, covered, excluded = [], [], [], [], []
        t_start = time.perf_counter()
        for i, r in enumerate(timed):
            left = args.plain_budget - (time.perf_counter() - t_start)
            if left <= 0:
                excluded += list(range(i, len(timed)))
                break
            if plain_estimate_s(r) > left:   # cannot finish inside what is left of the budget
                excluded.append(i)
                continue
            log(f"plain FP64: request {i} kind {r.kind} order {r.dim}")
            t0 = time.perf_counter()
            pr = sv.request_plain(r)
            times.append(time.perf_counter() - t0)
            errs.append(rel_diff(got[i][0], pr, r))
            kinds_p.append(r.kind)
            covered.append(i)
        P = len(times)
        if P:
            if P == len(timed):
                served_wall, served_src = wall, "the timed replay"
            else:   # the serving loop on exactly the plain path's requests, submitted at once as the
                # plain path takes them back to back
                L.qpms_hip_sync()
                got_p, served_wall = replay(srv, [timed[i] for i in covered], "served subset", burst=True)
                L.qpms_hip_sync()
                served_src = f"the {P} compared requests submitted at once"
            plain = {"requests": P, "of_timed": len(timed), "plain_busy_s": float(sum(times)),
                     "solves_per_s": P / sum(times), "served_wall_s": served_wall, "served_solves_per_s": P / served_wall,
                     "served_measured_by": served_src,
                     "excluded": [{"request": i, "kind": timed[i].kind, "order": timed[i].dim,
                                   "estimated_plain_s": round(plain_estimate_s(timed[i]), 1)} for i in excluded],
                     "busy_s_by_kind": {str(k): float(sum(t for t, kk in zip(times, kinds_p) if kk == k))
                                        for k in (0, 1, 2)},
                     "p50_s": float(np.percentile(times, 50)), "p99_s": float(np.percentile(times, 99)),
                     "max_rel_diff_served_vs_fp64": float(np.max(errs)),
                     "max_rel_diff_by_kind": {str(k): float(max([e for e, kk in zip(errs, kinds_p) if kk == k] or [0]))
                                              for k in (0, 1, 2)}}
            log(f"plain FP64: {P} requests in {sum(times):.1f} s busy; served {served_wall:.1f} s on the same; "
                f"excluded {excluded}; max rel diff {plain['max_rel_diff_served_vs_fp64']:.2e}")
    srv.close()
    log(f"device memory after the server: {meminfo_gib()[0]:.1f} GiB free")

    cpu = None if args.no_cpu else cpu_baseline_scatter(timed, got, args.cpu_budget)

    line = {
        "metric": METRIC,
        "value": value,
        "unit": "requests/s (plane-wave solves, periodic determinants, Beyn contours; every request answered)",
        "n_gpus": 1, "steps": K, "warmup": W,
        "ms_per_step": wall / K * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "c64 LU + c128 refinement (solves, contour nodes), c128 (builds, Ewald sums, determinants, Beyn SVD)",
        "data": "synthetic (seeded request stream)",
        "config": {"workload": f"serving stream: the first {K * S} requests (steps of {S}) of the seeded 10,000-request "
                               "stream (Poisson arrivals, N log-uniform 16..1024, lMax 2-4, 70/20/10 plane-wave / "
                               "periodic-determinant / 32-node Beyn), run to completion",
                   "rate_per_s": args.rate, "requests": len(timed), "by_kind": kinds,
                   "max_order": int(max(r.dim for r in timed)), "failed": len(failed),
                   "warmup_requests": len(warm), "step_done_s": [round(x, 3) for x in step_done]},
        "latency_s": {"p50": float(np.percentile(lat, 50)), "p99": float(np.percentile(lat, 99)),
                      "max": float(np.max(lat))},
        "plain_fp64": plain,
        "vs_plain_fp64": (plain["plain_busy_s"] / plain["served_wall_s"]) if plain else None,
        "roofline": roof,
        "roofline_lu16380": roof16,
        "gemm_autotune": tune,
        "w_build": wb,
        "measured_peaks": peaks,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
