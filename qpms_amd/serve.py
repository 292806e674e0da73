"""ctypes bindings of the serving loop (include/qpms_hip_serve.h) and the
seeded request stream of the serving workload.

The stream: 10,000 requests with Poisson arrivals at a fixed rate; each is a
finite cluster of N particles, N log-uniform in [16, 1024], on a square
lattice (pitch 400 nm) with 10 % positional jitter; lMax drawn from {2, 3, 4};
sphere radius uniform in [30, 80] nm; eps_in with real part uniform in
[-20, -5] and imaginary part in [0.5, 2]; hbar omega uniform in [1.5, 2.5] eV;
background n = 1.52. 70 % plane-wave scattering solves (x-polarised, normal
incidence), 20 % periodic mode-problem determinants (the cluster as the unit
cell of a square lattice of its own extent, Bloch vector uniform in the first
zone), 10 % Beyn contours (32 nodes, ellipse centred at omega with
half-axes 0.05 eV and 0.005 eV, 32 probe columns, singular-value cut
beyn_rank_tol).
"""
import ctypes

import numpy as np

from . import lib, c_cplx

EV = 1.602176634e-19 / 1.054571817e-34   # rad/s per eV
KIND_SCATTER, KIND_PERIODIC, KIND_BEYN = 0, 1, 2
MAX_EIG = 64


class Request(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_int), ("lMax", ctypes.c_int), ("npart", ctypes.c_size_t),
                ("pos", ctypes.POINTER(ctypes.c_double)), ("radius", ctypes.c_double), ("eps_in", c_cplx),
                ("eps_bg", ctypes.c_double), ("omega", ctypes.c_double), ("kdir", ctypes.c_double * 3),
                ("E", c_cplx * 3), ("lattice", (ctypes.c_double * 2) * 2), ("beta", ctypes.c_double * 2),
                ("beyn_rRe", ctypes.c_double), ("beyn_rIm", ctypes.c_double), ("beyn_nodes", ctypes.c_int),
                ("beyn_l", ctypes.c_int), ("beyn_seed", ctypes.c_ulonglong), ("rank_tol", ctypes.c_double),
                ("beyn_centre_im", ctypes.c_double)]


class Response(ctypes.Structure):
    _fields_ = [("id", ctypes.c_ulonglong), ("status", ctypes.c_int), ("kind", ctypes.c_int),
                ("dim", ctypes.c_size_t), ("sigma_ext", ctypes.c_double), ("rel_residual", ctypes.c_double),
                ("refine_steps", ctypes.c_int), ("logabsdet", ctypes.c_double), ("argdet", ctypes.c_double),
                ("neig", ctypes.c_int), ("eigval", c_cplx * MAX_EIG), ("sv", ctypes.c_double * MAX_EIG),
                ("t_submit", ctypes.c_double),
                ("t_start", ctypes.c_double), ("t_done", ctypes.c_double), ("message", ctypes.c_char * 96),
                ("batch", ctypes.c_int), ("padded_dim", ctypes.c_int)]

    def eigenvalues(self):
        return np.array([complex(self.eigval[i].re, self.eigval[i].im) for i in range(self.neig)])


class Opts(ctypes.Structure):
    _fields_ = [("arena_bytes", ctypes.c_size_t), ("nstreams", ctypes.c_int), ("mixed", ctypes.c_int),
                ("batch_max_dim", ctypes.c_size_t), ("batch_max", ctypes.c_int), ("refine_steps", ctypes.c_int)]


def _bind():
    L = lib()
    if getattr(L, "_serve_bound", False):
        return L
    vp = ctypes.c_void_p
    L.qpms_hip_server_create.restype = vp
    L.qpms_hip_server_create.argtypes = [ctypes.POINTER(Opts)]
    L.qpms_hip_server_destroy.argtypes = [vp]
    L.qpms_hip_server_arena_bytes.restype = ctypes.c_size_t
    L.qpms_hip_server_arena_bytes.argtypes = [vp]
    L.qpms_hip_server_clock.restype = ctypes.c_double
    L.qpms_hip_server_clock.argtypes = [vp]
    L.qpms_hip_server_submit.restype = ctypes.c_ulonglong
    L.qpms_hip_server_submit.argtypes = [vp, ctypes.POINTER(Request)]
    L.qpms_hip_server_progress.restype = ctypes.c_int
    L.qpms_hip_server_progress.argtypes = [vp, ctypes.c_int]
    L.qpms_hip_server_poll.restype = ctypes.c_int
    L.qpms_hip_server_poll.argtypes = [vp, ctypes.POINTER(Response), ctypes.c_int]
    L.qpms_hip_server_pending.restype = ctypes.c_size_t
    L.qpms_hip_server_pending.argtypes = [vp]
    L.qpms_hip_server_cancel.argtypes = [vp]
    L.qpms_hip_server_set_deadline.argtypes = [vp, ctypes.c_double]
    L.qpms_hip_request_plain.restype = ctypes.c_int
    L.qpms_hip_request_plain.argtypes = [ctypes.POINTER(Request), ctypes.POINTER(Response)]
    L._serve_bound = True
    return L


class Server:
    """arena_gib=None sizes the arena from the device's free memory."""

    def __init__(self, arena_gib=None, nstreams=8, mixed=True, batch_max_dim=8192, batch_max=32, refine_steps=3):
        self.L = _bind()
        nbytes = 0 if arena_gib is None else int(arena_gib * (1 << 30))
        o = Opts(nbytes, nstreams, int(mixed), batch_max_dim, batch_max, refine_steps)
        self.h = self.L.qpms_hip_server_create(ctypes.byref(o))
        if not self.h:
            raise MemoryError("qpms_hip_server_create failed (see stderr)")
        self._keep = {}

    def arena_bytes(self):
        return self.L.qpms_hip_server_arena_bytes(self.h)

    def submit(self, req):
        rid = self.L.qpms_hip_server_submit(self.h, ctypes.byref(req.c))
        if rid == 0:
            raise MemoryError("request can never fit in the server's arena")
        return rid

    def progress(self, wait=False):
        return self.L.qpms_hip_server_progress(self.h, int(wait))

    def poll(self, max_n=256):
        buf = (Response * max_n)()
        n = self.L.qpms_hip_server_poll(self.h, buf, max_n)
        return [buf[i] for i in range(n)]

    def pending(self):
        return self.L.qpms_hip_server_pending(self.h)

    def clock(self):
        return self.L.qpms_hip_server_clock(self.h)

    def cancel(self):
        self.L.qpms_hip_server_cancel(self.h)

    def set_deadline(self, t):
        """Start nothing estimated to finish after t (server clock, s); 0 clears."""
        self.L.qpms_hip_server_set_deadline(self.h, ctypes.c_double(t))

    def close(self):
        if self.h:
            self.L.qpms_hip_server_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Req:
    """A request plus the host arrays its C struct points to."""

    def __init__(self, kind, pos, lMax, radius, eps_in, eps_bg, omega, kdir=(0, 0, 1), E=(1, 0, 0),
                 lattice=((0, 0), (0, 0)), beta=(0, 0), beyn_rRe=0.0, beyn_rIm=0.0, beyn_nodes=32, beyn_l=32,
                 beyn_seed=1, rank_tol=1e-4, arrival=0.0):
        self.pos = np.ascontiguousarray(pos, np.float64).reshape(-1, 3)
        self.arrival = arrival
        c = Request()
        c.kind, c.lMax, c.npart = kind, lMax, len(self.pos)
        c.pos = self.pos.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
        # omega may be complex for a Beyn request (the contour centre)
        c.radius, c.eps_in, c.eps_bg = radius, c_cplx.of(eps_in), eps_bg
        c.omega, c.beyn_centre_im = complex(omega).real, complex(omega).imag
        for i in range(3):
            c.kdir[i] = kdir[i]
            c.E[i] = c_cplx.of(E[i])
        for i in range(2):
            c.beta[i] = beta[i]
            for j in range(2):
                c.lattice[i][j] = lattice[i][j]
        c.beyn_rRe, c.beyn_rIm, c.beyn_nodes, c.beyn_l = beyn_rRe, beyn_rIm, beyn_nodes, beyn_l
        c.beyn_seed, c.rank_tol = beyn_seed, rank_tol
        self.c = c
        self.dim = len(self.pos) * 2 * lMax * (lMax + 2)

    @property
    def kind(self):
        return self.c.kind

    def clone(self, arrival=0.0):
        c = self.c
        return Req(c.kind, self.pos, c.lMax, c.radius, complex(c.eps_in.re, c.eps_in.im), c.eps_bg,
                   complex(c.omega, c.beyn_centre_im),
                   kdir=tuple(c.kdir), E=tuple(complex(e.re, e.im) for e in c.E),
                   lattice=((c.lattice[0][0], c.lattice[0][1]), (c.lattice[1][0], c.lattice[1][1])),
                   beta=(c.beta[0], c.beta[1]), beyn_rRe=c.beyn_rRe, beyn_rIm=c.beyn_rIm, beyn_nodes=c.beyn_nodes,
                   beyn_l=c.beyn_l, beyn_seed=c.beyn_seed, rank_tol=c.rank_tol, arrival=arrival)


def beyn_rank_tol(dim, l, rRe, rel=1e-5):
    """Absolute singular-value cut for a Beyn request, scaled to the size of
    the contour integral: A0 = sum_n dz_n M(z_n)^-1 V has norm about
    2 pi rRe sqrt(dim l) when an eigenvalue is enclosed, and rounding noise
    about 1e-16 of that times cond(M). `rel` of the scale separates the two."""
    return rel * 2 * np.pi * rRe * np.sqrt(dim * l)


def beyn_rel_diff(a, b, rank_tol):
    """Largest relative difference between two Beyn answers: the singular
    values above the rank cut (relative to the largest) and the eigenvalues
    (matched nearest-first, relative to |z|). inf if the counts differ."""
    sa, sb = np.array(a.sv[:MAX_EIG]), np.array(b.sv[:MAX_EIG])
    keep = sb > rank_tol
    if keep.sum() != (sa > rank_tol).sum() or a.neig != b.neig:
        return float("inf")
    # below the cut both are rounding noise of an (almost) empty contour
    err = float(np.max(np.abs(sa[keep] - sb[keep])) / sb[0]) if keep.any() else 0.0
    ea, eb = list(a.eigenvalues()), list(b.eigenvalues())
    for z in eb:
        j = int(np.argmin([abs(w - z) for w in ea]))
        err = max(err, abs(ea.pop(j) - z) / abs(z))
    return err


def request_plain(req):
    """The same request through the public scattering-system API (FP64)."""
    L = _bind()
    out = Response()
    L.qpms_hip_request_plain(ctypes.byref(req.c), ctypes.byref(out))
    return out


def serving_stream(n=10000, rate=20.0, seed=20261202, pitch=400e-9, n_bg=1.52):
    """The seeded request stream (list of Req in arrival order)."""
    rng = np.random.default_rng(seed)
    t = 0.0
    out = []
    for i in range(n):
        t += rng.exponential(1.0 / rate)
        N = int(np.clip(np.round(np.exp(rng.uniform(np.log(16), np.log(1024)))), 16, 1024))
        side = int(np.ceil(np.sqrt(N)))
        g = np.array([(ix, iy) for ix in range(side) for iy in range(side)][:N], float)
        g -= g.mean(axis=0)
        xy = (g + rng.uniform(-0.1, 0.1, size=g.shape)) * pitch
        pos = np.column_stack([xy, np.zeros(N)])
        lMax = int(rng.choice([2, 3, 4]))
        radius = rng.uniform(30e-9, 80e-9)
        eps_in = complex(rng.uniform(-20, -5), rng.uniform(0.5, 2))
        eV = rng.uniform(1.5, 2.5)
        u = rng.uniform()
        common = dict(pos=pos, lMax=lMax, radius=radius, eps_in=eps_in, eps_bg=n_bg ** 2, omega=eV * EV, arrival=t)
        if u < 0.7:
            out.append(Req(KIND_SCATTER, kdir=(0, 0, 1), E=(1, 0, 0), **common))
        elif u < 0.9:
            a = side * pitch
            beta = rng.uniform(-np.pi / a, np.pi / a, size=2)
            out.append(Req(KIND_PERIODIC, lattice=((a, 0), (0, a)), beta=tuple(beta), **common))
        else:
            dim = N * 2 * lMax * (lMax + 2)
            out.append(Req(KIND_BEYN, beyn_rRe=0.05 * EV, beyn_rIm=0.005 * EV, beyn_nodes=32, beyn_l=32,
                           beyn_seed=1000 + i, rank_tol=beyn_rank_tol(dim, 32, 0.05 * EV), **common))
    return out
