"""CPU: the host side of the serving loop -- the ctypes mirrors of the
request/response/options structs against the C header (offsets measured by
a C compiler), the seeded request stream, and the Beyn comparison helpers."""
import ctypes
import os
import subprocess
import tempfile

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def c_layout():
    """sizeof / offsetof of the serving structs as gcc lays them out."""
    fields = {
        "qpms_hip_request_t": ["kind", "lMax", "npart", "pos", "radius", "eps_in", "eps_bg", "omega", "kdir", "E",
                               "lattice", "beta", "beyn_rRe", "beyn_rIm", "beyn_nodes", "beyn_l", "beyn_seed",
                               "rank_tol", "beyn_centre_im"],
        "qpms_hip_response_t": ["id", "status", "kind", "dim", "sigma_ext", "rel_residual", "refine_steps",
                                "logabsdet", "argdet", "neig", "eigval", "sv", "t_submit", "t_start", "t_done",
                                "message", "batch", "padded_dim"],
        "qpms_hip_server_opts_t": ["arena_bytes", "nstreams", "mixed", "batch_max_dim", "batch_max", "refine_steps"],
    }
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "qpms_hip_serve.h"', "int main(void) {"]
    for st, fs in fields.items():
        lines.append(f'printf("{st} size %zu\\n", sizeof({st}));')
        for f in fs:
            lines.append(f'printf("{st} {f} %zu\\n", offsetof({st}, {f}));')
    lines.append("return 0; }")
    with tempfile.TemporaryDirectory() as d:
        src, exe = os.path.join(d, "l.c"), os.path.join(d, "l")
        open(src, "w").write("\n".join(lines))
        subprocess.run(["gcc", "-std=gnu11", "-I", os.path.join(ROOT, "include"), src, "-o", exe], check=True)
        out = subprocess.run([exe], check=True, capture_output=True, text=True).stdout
    lay = {}
    for ln in out.splitlines():
        st, f, v = ln.split()
        lay[(st, f)] = int(v)
    return lay


def test_struct_mirrors_match_header():
    from qpms_amd import serve as sv
    lay = c_layout()
    for st, cls in [("qpms_hip_request_t", sv.Request), ("qpms_hip_response_t", sv.Response),
                    ("qpms_hip_server_opts_t", sv.Opts)]:
        assert ctypes.sizeof(cls) == lay[(st, "size")], st
        for name, _ in cls._fields_:
            assert getattr(cls, name).offset == lay[(st, name)], (st, name)


def test_stream_is_seeded_and_has_the_stated_mix():
    from qpms_amd import serve as sv
    a = sv.serving_stream(n=3000, rate=10.0)
    b = sv.serving_stream(n=3000, rate=10.0)
    assert [r.dim for r in a] == [r.dim for r in b]
    assert np.array_equal(a[17].pos, b[17].pos)
    kinds = np.array([r.kind for r in a])
    frac = [np.mean(kinds == k) for k in (0, 1, 2)]
    assert abs(frac[0] - 0.7) < 0.03 and abs(frac[1] - 0.2) < 0.03 and abs(frac[2] - 0.1) < 0.02
    t = np.array([r.arrival for r in a])
    assert np.all(np.diff(t) > 0) and abs(len(t) / t[-1] - 10.0) < 0.6   # Poisson arrivals at 10 /s
    n = np.array([r.c.npart for r in a])
    assert n.min() >= 16 and n.max() <= 1024
    for r in a:
        if r.kind == sv.KIND_BEYN:
            assert r.c.rank_tol == pytest.approx(sv.beyn_rank_tol(r.dim, 32, 0.05 * sv.EV))


def fake(neig, sv_vals, eig=()):
    from qpms_amd import serve as sv
    r = sv.Response()
    r.neig = neig
    for i, v in enumerate(sv_vals):
        r.sv[i] = v
    for i, z in enumerate(eig):
        r.eigval[i].re, r.eigval[i].im = z.real, z.imag
    return r


def test_beyn_comparison():
    from qpms_amd import serve as sv
    tol = 10.0
    a = fake(2, [100.0, 50.0, 1.0], [2.0 - 0.1j, 2.1 - 0.2j])
    assert sv.beyn_rel_diff(a, fake(2, [100.0, 50.0, 3.0], [2.1 - 0.2j, 2.0 - 0.1j]), tol) == 0.0   # noise below the cut
    d = sv.beyn_rel_diff(a, fake(2, [100.0, 50.0], [2.0 - 0.1j, 2.1 - 0.2j + 1e-6]), tol)
    assert 1e-7 < d < 1e-6
    assert sv.beyn_rel_diff(a, fake(1, [100.0, 5.0], [2.0 - 0.1j]), tol) == float("inf")   # counts differ
    assert sv.beyn_rel_diff(fake(0, [3.0]), fake(0, [2.0]), tol) == 0.0   # empty contour both ways
