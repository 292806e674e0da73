"""Oracle (test infrastructure): numpy restatement of the reference's Beyn
solver, beyn.c:513-606 (quadrature loop) and 345-510 (SVD, rank test,
reduced matrix, eigendecomposition, residual filter). numpy.linalg calls the
same LAPACK drivers the reference calls through LAPACKE (zgesv = zgetrf +
zgetrs, zgesdd, zgeev)."""
import numpy as np


def contour_ellipse(centre, rre, rim, n):
    """beyn.c:69-86"""
    t = np.arange(n) * 2 * np.pi / n
    z = centre + np.cos(t) * rre + 1j * np.sin(t) * rim
    dz = (-rre * np.sin(t) + 1j * rim * np.cos(t)) * (2 * np.pi / n)
    return z, dz


def process(Mfun, A0, A1, z0, rank_tol, rank_sel_min, res_tol):
    m, l = A0.shape
    U, S, Wh = np.linalg.svd(A0, full_matrices=False)
    K = sum(1 for k in range(l) if k < rank_sel_min or S[k] > rank_tol)
    if K == 0:
        return np.zeros(0, complex), np.zeros((0, m), complex), np.zeros(0), S
    V0 = U[:, :K]
    W0 = Wh[:K, :].conj().T
    B = V0.conj().T @ A1 @ W0 / S[:K][None, :]
    lam, Sv = np.linalg.eig(B)
    V0S = V0 @ Sv
    ev, vecs, res = [], [], []
    for k in range(K):
        z = z0 + lam[k]
        r = 0.0
        if res_tol > 0:
            r = np.linalg.norm(Mfun(z) @ V0S[:, k])
            if r > res_tol:
                continue
        ev.append(z)
        vecs.append(V0S[:, k])
        res.append(r)
    return np.array(ev), np.array(vecs).reshape(len(ev), m), np.array(res), S


def beyn_solve(Mfun, m, Vhat, z_nodes, dz_nodes, z0, rank_tol=1e-4, rank_sel_min=1, res_tol=0.0):
    l = Vhat.shape[1]
    A0 = np.zeros((m, l), complex)
    A1 = np.zeros((m, l), complex)
    A0c = np.zeros((m, l), complex)
    A1c = np.zeros((m, l), complex)
    for n, (z, dz) in enumerate(zip(z_nodes, dz_nodes)):
        X = np.linalg.solve(Mfun(z), Vhat) * dz
        A0 += X
        if n % 2 == 0:
            A0c += 2 * X
        X = X * (z - z0)
        A1 += X
        if n % 2 == 0:
            A1c += 2 * X
    ev_c, _, _, _ = process(Mfun, A0c, A1c, z0, rank_tol, rank_sel_min, res_tol)
    ev, vecs, res, S = process(Mfun, A0, A1, z0, rank_tol, rank_sel_min, res_tol)
    return ev, vecs, res, S, ev_c


def loaded_string(m):
    """Beyn (2012) section 4.11, as tests/tbeyn.c builds it."""
    def M(z):
        d = 2 * m - 4 * z / (6 * m)
        od = -float(m) - z / (6 * m)
        A = np.zeros((m, m), complex)
        idx = np.arange(m)
        A[idx, idx] = d
        A[idx[1:], idx[:-1]] = od
        A[idx[:-1], idx[1:]] = od
        A[m - 1, m - 1] = d / 2 + z / (z - 1)
        return A
    return M


def loaded_string_eigs(m):
    """Exact eigenvalues by linearisation: (z-1) M(z) = z^2 Q2 + z Q1 + Q0."""
    import scipy.linalg as sl
    Kd = np.zeros((m, m))
    Md = np.zeros((m, m))
    idx = np.arange(m)
    Kd[idx, idx] = 2 * m
    Kd[idx[1:], idx[:-1]] = -m
    Kd[idx[:-1], idx[1:]] = -m
    Kd[m - 1, m - 1] = m
    Md[idx, idx] = 4 / (6 * m)
    Md[idx[1:], idx[:-1]] = 1 / (6 * m)
    Md[idx[:-1], idx[1:]] = 1 / (6 * m)
    Md[m - 1, m - 1] = 2 / (6 * m)
    E = np.zeros((m, m))
    E[m - 1, m - 1] = 1
    # M(z) = Kd - z Md + z/(z-1) E
    Q2, Q1, Q0 = -Md, Kd + Md + E, -Kd
    Z = np.zeros((m, m))
    Aa = np.block([[Z, np.eye(m)], [-Q0, -Q1]])
    Bb = np.block([[np.eye(m), Z], [Z, Q2]])
    w = sl.eig(Aa, Bb, right=False)
    w = w[np.isfinite(w)]
    return w[np.abs(w - 1) > 1e-6]


def polynomial(dim, seed):
    rng = np.random.default_rng(seed)
    T0, T1, T2 = (rng.normal(size=(dim, dim)) for _ in range(3))

    def M(z):
        return T0 + z * T1 + z * z * T2

    import scipy.linalg as sl
    Z = np.zeros((dim, dim))
    Aa = np.block([[Z, np.eye(dim)], [-T0, -T1]])
    Bb = np.block([[np.eye(dim), Z], [Z, T2]])
    w = sl.eig(Aa, Bb, right=False)
    return M, w[np.isfinite(w)]


def probe(m, l, seed):
    rng = np.random.default_rng(seed)
    return rng.normal(size=(m, l)) + 1j * rng.normal(size=(m, l))


def match_sets(a, b):
    """Max over a of the distance to the nearest element of b (relative)."""
    if len(a) == 0:
        return 0.0
    return max(np.min(np.abs(b - x)) / max(abs(x), 1e-300) for x in a)


def _orient(o, rre, rim):
    """beyn.c:127-147"""
    return {3: (-1j, rim, rre), 1: (1j, rim, rre), 0: (1, rre, rim), 2: (-1, rre, rim)}[o]


def contour_halfellipse(centre, rre, rim, n, o):
    """beyn.c:120-163"""
    f, l, h = _orient(o, rre, rim)
    nline = n // 2
    narc = n - nline
    t = (np.arange(narc) + 0.5) * np.pi / narc
    z = [centre + f * (np.cos(t) * l + 1j * np.sin(t) * h)]
    dz = [f * (-l * np.sin(t) + 1j * h * np.cos(t)) * (np.pi / narc)]
    t = 0.5 * (1 - nline) + np.arange(nline)
    z.append(centre + f * t * 2 * l / nline)
    dz.append(np.full(nline, f * 2 * l / nline, complex))
    return np.concatenate(z), np.concatenate(dz)


def contour_kidney(centre, rre, rim, rounding, n, o):
    """beyn.c:165-248: the sampling of a rounded half-ellipse by arc length."""
    f, l, h = _orient(o, rre, rim)
    y = rounding * h
    x = np.sqrt((h - y) ** 2 - y ** 2)
    al = np.arcsin(y / (h - y))
    ar = l / h
    tmax = 2 * (x + (np.pi / 2 + al) * y + (np.pi / 2 - al) * h)
    dt = tmax / n
    z, dz = [], []
    b1 = x
    b2 = b1 + (np.pi / 2 + al) * y
    b3 = b2 + (np.pi - 2 * al) * h
    b4 = b3 + (np.pi / 2 + al) * y
    i = 0
    while i * dt <= b1:
        t = i * dt
        z.append(centre + t * ar * f); dz.append(dt * ar * f); i += 1
    while i * dt < b2:
        phi = (i * dt - b1) / y - np.pi / 2
        z.append(centre + (ar * (x + y * np.cos(phi)) + y * (1 + np.sin(phi)) * 1j) * f)
        dz.append(dt * (-ar * np.sin(phi) + np.cos(phi) * 1j) * f); i += 1
    while i * dt < b3:
        phi = (i * dt - b2) / h + al
        z.append(centre + (ar * h * np.cos(phi) + h * np.sin(phi) * 1j) * f)
        dz.append(dt * (-ar * np.sin(phi) + np.cos(phi) * 1j) * f); i += 1
    while i * dt < b4:
        phi = (i * dt - b3) / y + np.pi - al
        z.append(centre + (ar * (-x + y * np.cos(phi)) + y * (1 + np.sin(phi)) * 1j) * f)
        dz.append(dt * (-ar * np.sin(phi) + np.cos(phi) * 1j) * f); i += 1
    while i < n:
        z.append(centre + (i * dt - b4 - x) * ar * f); dz.append(dt * ar * f); i += 1
    return np.array(z), np.array(dz)
