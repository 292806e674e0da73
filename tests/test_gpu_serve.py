"""GPU: the serving loop (include/qpms_hip_serve.h) against the plain FP64
drop-in path (the public scattering-system API, one request at a time) on a
small mixed stream: plane-wave solves of several sizes (batched and
unbatched), 2D-periodic determinants and Beyn contours around resonances
of the clusters. Each result must agree with the plain path (extinction and
log-determinant to 1e-9 relative; Beyn eigenfrequencies to 1e-9 relative
and the singular values above the rank cut to 1e-9 of the largest) and be
the same when the request runs alone as when it runs in the mix (1e-12)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

EV = 1.602176634e-19 / 1.054571817e-34


def grid(N, pitch, rng):
    side = int(np.ceil(np.sqrt(N)))
    g = np.array([(i, j) for i in range(side) for j in range(side)][:N], float)
    g -= g.mean(axis=0)
    xy = (g + rng.uniform(-0.1, 0.1, size=g.shape)) * pitch
    return np.column_stack([xy, np.zeros(N)]), side * pitch


def requests():
    from qpms_amd import serve as sv
    rng = np.random.default_rng(11)
    out = []
    for N, lMax in [(4, 2), (9, 2), (16, 3), (12, 2), (25, 3), (6, 4), (50, 4), (20, 2), (3, 1), (30, 2)]:
        pos, _ = grid(N, 400e-9, rng)
        out.append(sv.Req(sv.KIND_SCATTER, pos, lMax, rng.uniform(30e-9, 80e-9),
                          complex(rng.uniform(-20, -5), rng.uniform(0.5, 2)), 1.52 ** 2,
                          rng.uniform(1.5, 2.5) * EV, kdir=(0, 0, 1), E=(1, 0, 0)))
    for N, lMax in [(2, 2), (4, 2), (3, 3)]:
        pos, a = grid(N, 400e-9, rng)
        out.append(sv.Req(sv.KIND_PERIODIC, pos, lMax, 50e-9, -12 + 1j, 1.52 ** 2, 2.0 * EV,
                          lattice=((a, 0), (0, a)), beta=tuple(rng.uniform(-np.pi / a, np.pi / a, 2))))
    # contours around resonances of these clusters (found with the oracle:
    # 3 and 2 eigenfrequencies inside, the rest at least 0.04 eV outside)
    for (N, lMax), (c, rre, rim) in zip([(4, 2), (9, 2)], [(1.12 - 0.25j, 0.1, 0.05), (1.204 - 0.143j, 0.05, 0.04)]):
        pos, _ = grid(N, 400e-9, rng)
        out.append(sv.Req(sv.KIND_BEYN, pos, lMax, 70e-9, -12 + 1j, 1.52 ** 2, c * EV, beyn_rRe=rre * EV,
                          beyn_rIm=rim * EV, beyn_nodes=16, beyn_l=16, beyn_seed=7 + N,
                          rank_tol=sv.beyn_rank_tol(N * 2 * lMax * (lMax + 2), 16, rre * EV)))
    # larger unit cells (the served lattice sums use exponent-bounded cutoffs,
    # the plain path the reference's)
    rng2 = np.random.default_rng(12)
    for N, lMax in [(16, 2), (30, 3)]:
        pos, a = grid(N, 400e-9, rng2)
        out.append(sv.Req(sv.KIND_PERIODIC, pos, lMax, 60e-9, -15 + 1.5j, 1.52 ** 2, 1.9 * EV,
                          lattice=((a, 0), (0, a)), beta=tuple(rng2.uniform(-np.pi / a, np.pi / a, 2))))
    return out


def drain(srv):
    got = {}
    while srv.pending():
        srv.progress(wait=True)
        for r in srv.poll():
            got[r.id] = r
    return got


def close(a, b, kind, tol, req=None):
    from qpms_amd import serve as sv
    assert a.status == 0 and b.status == 0, (a.message, b.message)
    assert a.dim == b.dim
    if kind == sv.KIND_SCATTER:
        assert abs(a.sigma_ext - b.sigma_ext) <= tol * abs(b.sigma_ext), (a.sigma_ext, b.sigma_ext)
    elif kind == sv.KIND_PERIODIC:
        assert abs(a.logabsdet - b.logabsdet) <= tol * max(1.0, abs(b.logabsdet)), (a.logabsdet, b.logabsdet)
        d = (a.argdet - b.argdet + np.pi) % (2 * np.pi) - np.pi
        assert abs(d) <= tol * max(1.0, abs(b.logabsdet)) + 1e-12
    else:
        assert sv.beyn_rel_diff(a, b, req.c.rank_tol) <= tol, (a.sv[:4], b.sv[:4], a.eigenvalues(), b.eigenvalues())


def test_serving_matches_plain_alone_and_mix(gpu):
    from qpms_amd import serve as sv
    reqs = requests()
    plain = [sv.request_plain(r) for r in reqs]
    srv = sv.Server(arena_gib=16, nstreams=4, mixed=True)
    ids = [srv.submit(r) for r in reqs]
    mix = drain(srv)
    assert sorted(mix) == sorted(ids)
    alone = []
    for r in reqs:
        rid = srv.submit(r)
        alone.append(drain(srv)[rid])
    srv.close()
    for r, p, i, a in zip(reqs, plain, ids, alone):
        close(mix[i], p, r.kind, 1e-9, r)
        close(mix[i], a, r.kind, 1e-12, r)
        if r.kind == sv.KIND_SCATTER:
            assert mix[i].rel_residual < 1e-14


def test_served_beyn_matches_oracle(gpu):
    """The served contours against the oracle's Beyn (numpy restatement of
    beyn.c with the oracle's M = I - T W at every node), same probe."""
    import ctypes
    import oracle_beyn as ob
    import oracle_ffi as of
    from qpms_amd import serve as sv
    reqs = [r for r in requests() if r.kind == sv.KIND_BEYN]
    srv = sv.Server(arena_gib=4, nstreams=2, mixed=True)
    ids = [srv.submit(r) for r in reqs]
    got = drain(srv)
    srv.close()
    lib = sv.lib()
    lib.qh_probe_host.argtypes = [ctypes.c_long, ctypes.c_ulonglong, ctypes.c_void_p]
    for r, i in zip(reqs, ids):
        c = r.c
        N, lMax, L, l = c.npart, c.lMax, r.dim, c.beyn_l
        oc = of.OracleCalc(lMax)
        eps_in = complex(c.eps_in.re, c.eps_in.im)

        def Mf(z):
            td, k = of.mie_tdiag(lMax, c.radius, eps_in, c.eps_bg, z)
            return oc.build_M(r.pos, np.zeros(N, np.int32), td, k)
        V = np.empty((L, l), np.complex128)
        lib.qh_probe_host(L * l, c.beyn_seed, V.ctypes.data)
        centre = complex(c.omega, c.beyn_centre_im)
        z, dz = ob.contour_ellipse(centre, c.beyn_rRe, c.beyn_rIm, c.beyn_nodes)
        ev, _, _, S, _ = ob.beyn_solve(Mf, L, V, z, dz, centre, rank_tol=c.rank_tol, rank_sel_min=0)
        g = got[i]
        assert g.status == 0 and g.neig == len(ev) and g.neig >= 2
        keep = S > c.rank_tol
        assert np.max(np.abs(np.array(g.sv[:l])[keep] - S[keep])) <= 1e-9 * S[0]
        assert ob.match_sets(g.eigenvalues(), ev) <= 1e-9


def test_stream_large_requests_match_plain(gpu):
    """Requests of the seeded serving stream at full size (orders up to
    30,000: FP32 LU with lookahead and FP64 refinement, periodic
    determinants) against the plain FP64 API on the same requests."""
    from qpms_amd import serve as sv
    stream = sv.serving_stream(n=400)
    scat = sorted((r for r in stream if r.kind == sv.KIND_SCATTER and r.dim <= 30000), key=lambda r: -r.dim)[:2]
    per = sorted((r for r in stream if r.kind == sv.KIND_PERIODIC and r.dim <= 12000), key=lambda r: -r.dim)[:1]
    reqs = [r.clone(0.0) for r in scat + per]
    assert reqs[0].dim > 20000 and per[0].dim > 6000, [r.dim for r in reqs]
    print("orders", [r.dim for r in reqs], flush=True)
    srv = sv.Server(arena_gib=64, nstreams=2, mixed=True)
    ids = [srv.submit(r) for r in reqs]
    got = drain(srv)
    srv.close()
    for r, i in zip(reqs, ids):
        close(got[i], sv.request_plain(r), r.kind, 1e-9, r)
        if r.kind == sv.KIND_SCATTER:
            assert got[i].rel_residual < 1e-14


def test_serving_fp64_mode_matches_plain(gpu):
    from qpms_amd import serve as sv
    reqs = [r for r in requests() if r.kind == sv.KIND_SCATTER][:6]
    plain = [sv.request_plain(r) for r in reqs]
    srv = sv.Server(arena_gib=8, mixed=False)
    ids = [srv.submit(r) for r in reqs]
    got = drain(srv)
    srv.close()
    for r, p, i in zip(reqs, plain, ids):
        close(got[i], p, r.kind, 1e-10)


def test_admission_by_footprint(gpu):
    """A small arena: requests that never fit are refused at submit; the
    others run, at most as many at a time as fit, and all finish."""
    from qpms_amd import serve as sv
    rng = np.random.default_rng(3)
    srv = sv.Server(arena_gib=0.25, nstreams=4, mixed=True, batch_max_dim=512)
    big, _ = grid(400, 400e-9, rng)   # L = 400 * 48 at lMax 4: 5.9 GB, never fits
    with pytest.raises(MemoryError):
        srv.submit(sv.Req(sv.KIND_SCATTER, big, 4, 50e-9, -12 + 1j, 2.31, 2 * EV))
    reqs = []
    for N in [40, 44, 48, 52, 20, 8]:
        pos, _ = grid(N, 400e-9, rng)
        reqs.append(sv.Req(sv.KIND_SCATTER, pos, 3, 50e-9, -12 + 1j, 2.31, 2 * EV))
    ids = [srv.submit(r) for r in reqs]
    got = drain(srv)
    srv.close()
    assert sorted(got) == sorted(ids)
    for r, i in zip(reqs, ids):
        close(got[i], sv.request_plain(r), r.kind, 1e-9)


def test_deadline_keeps_long_requests_queued(gpu):
    """With a deadline, a request whose estimate does not fit stays queued
    (cancelled when the caller gives up) while short ones run."""
    from qpms_amd import serve as sv
    rng = np.random.default_rng(5)
    srv = sv.Server(arena_gib=48, nstreams=4, mixed=True)
    srv.set_deadline(srv.clock() + 0.5)
    big, _ = grid(600, 400e-9, rng)   # order 600 * 48 = 28800 at lMax 4: seconds of work
    small, _ = grid(20, 400e-9, rng)
    rb = srv.submit(sv.Req(sv.KIND_SCATTER, big, 4, 50e-9, -12 + 1j, 2.31, 2 * EV))
    rs = srv.submit(sv.Req(sv.KIND_SCATTER, small, 2, 50e-9, -12 + 1j, 2.31, 2 * EV))
    got = {}
    while rs not in got:
        srv.progress(wait=True)
        for r in srv.poll():
            got[r.id] = r
    assert got[rs].status == 0 and rb not in got and srv.pending() == 1
    srv.cancel()   # queued requests are answered at once (status -2)
    while True:
        for r in srv.poll():
            got[r.id] = r
        if not srv.pending():
            break
        srv.progress(wait=True)
    srv.close()
    assert got[rb].status == -2


def test_beyn_chunk_graph_replay_bit_identical(gpu, monkeypatch):
    """Beyn contours of 16 nodes in chunks of 4: the first chunk's launches are
    captured into a graph and replayed for the other three, with plane-wave
    solves of other requests running between the replays. The eigenvalues,
    singular values and extinctions equal, bit for bit, those of a server that
    launches every chunk eagerly (the trailing updates are hand-written
    kernels at these orders in both runs)."""
    from qpms_amd import serve as sv
    reqs = [r for r in requests() if r.kind != sv.KIND_PERIODIC][:14]
    assert sum(r.kind == sv.KIND_BEYN for r in reqs) == 2
    monkeypatch.setenv("QPMS_HIP_BEYN_CHUNK", "4")
    out = {}
    for graphs in ("0", "1"):
        monkeypatch.setenv("QPMS_HIP_GRAPHS", graphs)
        srv = sv.Server(arena_gib=8, nstreams=4, mixed=True)
        ids = [srv.submit(r) for r in reqs]
        got = drain(srv)
        srv.close()
        out[graphs] = [got[i] for i in ids]
    for r, a, b in zip(reqs, out["0"], out["1"]):
        assert a.status == 0 and b.status == 0, (a.message, b.message)
        if r.kind == sv.KIND_SCATTER:
            assert a.sigma_ext == b.sigma_ext
        else:
            assert a.neig == b.neig and a.neig > 0
            assert np.array_equal(np.array(a.sv[:r.c.beyn_l]), np.array(b.sv[:r.c.beyn_l]))
            assert np.array_equal(a.eigenvalues(), b.eigenvalues())
