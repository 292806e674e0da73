"""C ABI: struct layouts of include/qpms_hip.h against the reference's.

The reference's field lists (names and declared C types, in order) are
transcribed below with the header lines they come from; an x86-64 SysV
layout of each list is computed here in Python and compared with the
offsets gcc gives for the same fields of this library's header. A field that
is missing, reordered or of another width shifts every offset after it.
Runs without a GPU."""
import os
import subprocess
import tempfile

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

# size / alignment of the scalar types that appear (x86-64 SysV)
SCALAR = {
    "int": (4, 4), "int32_t": (4, 4), "bool": (1, 1), "double": (8, 8), "size_t": (8, 8), "ptr": (8, 8),
    "complex double": (16, 8), "enum": (4, 4),
}
# typedefs of the reference (qpms_types.h:27-44, 310-323; scatsystem.h:43-44, 118)
TYPEDEF = {
    "qpms_l_t": "int", "qpms_y_t": "size_t", "qpms_ss_tmi_t": "int32_t", "qpms_ss_tmgi_t": "int32_t",
    "qpms_ss_pi_t": "int32_t", "qpms_ss_oti_t": "int32_t", "qpms_iri_t": "int", "qpms_normalisation_t": "enum",
    "gsl_sf_legendre_t": "enum", "qpms_tmatrix_operation_kind_t": "enum",
}

# Reference structs as (field, type); "ptr" stands for any pointer, nested
# structs by name.
REF = {
    # qpms_types.h:214
    "cart3_t": [("x", "double"), ("y", "double"), ("z", "double")],
    # qpms_types.h:240
    "csph_t": [("r", "complex double"), ("theta", "double"), ("phi", "double")],
    # materials.h:16-25 (function pointer + params)
    "qpms_epsmu_generator_t": [("function", "ptr"), ("params", "ptr")],
    # translations.h:64-84 (LATTICESUMS32 defined, as in the reference build)
    "qpms_trans_calculator": [("normalisation", "qpms_normalisation_t"), ("lMax", "qpms_l_t"),
                              ("nelem", "qpms_y_t"), ("A_multipliers", "ptr"), ("B_multipliers", "ptr"),
                              ("e3c", "ptr"), ("legendre0", "ptr")],
    # ewald.h:56-131
    "qpms_ewald3_constants_t": [("lMax", "qpms_l_t"), ("nelem_sc", "qpms_y_t"), ("s1_jMaxes", "ptr"),
                                ("s1_constfacs", "ptr"), ("s1_constfacs_base", "ptr"), ("S1_constfacs", "ptr"),
                                ("S1_constfacs_base", "ptr"), ("s1_constfacs_1Dz", "ptr"),
                                ("s1_constfacs_1Dz_base", "ptr"), ("legendre0", "ptr"), ("legendre_plus1", "ptr"),
                                ("legendre_minus1", "ptr"), ("legendre_normconv", "gsl_sf_legendre_t"),
                                ("legendre_csphase", "int")],
    # scatsystem.h:133-159
    "qpms_scatsys_periodic_info_t": [("lattice_basis", "cart3_t[3]"), ("reciprocal_basis", "cart3_t[3]"),
                                     ("unitcell_volume", "double"), ("eta", "double")],
    # scatsystem.h:169-227
    "qpms_scatsys_t": [("lattice_dimension", "int"), ("medium", "qpms_epsmu_generator_t"), ("tmg", "ptr"),
                       ("tmg_count", "qpms_ss_tmgi_t"), ("tm", "ptr"), ("tm_count", "qpms_ss_tmi_t"),
                       ("tm_capacity", "qpms_ss_tmi_t"), ("p", "ptr"), ("p_count", "qpms_ss_pi_t"),
                       ("p_capacity", "qpms_ss_pi_t"), ("sym", "ptr"), ("p_sym_map", "ptr"), ("tm_sym_map", "ptr"),
                       ("orbit_type_count", "qpms_ss_oti_t"), ("orbit_types", "ptr"), ("p_orbitinfo", "ptr"),
                       ("fecv_size", "size_t"), ("saecv_sizes", "ptr"), ("fecv_pstarts", "ptr"),
                       ("saecv_ot_offsets", "ptr"), ("max_bspecn", "size_t"), ("p_by_orbit", "ptr"),
                       ("otspace", "ptr"), ("otspace_end", "ptr"), ("lenscale", "double"), ("c", "ptr"),
                       ("per", "qpms_scatsys_periodic_info_t")],
    # scatsystem.h:239-249; qpms_epsmu_t at qpms_types.h:422-425
    "qpms_scatsys_at_omega_t": [("ss", "ptr"), ("tm", "ptr"), ("omega", "complex double"),
                                ("medium", "qpms_epsmu_t"), ("wavenumber", "complex double")],
    "qpms_epsmu_t": [("eps", "complex double"), ("mu", "complex double")],
    # scatsystem.h:432-441
    "qpms_ss_LU": [("ssw", "ptr"), ("sswk", "ptr"), ("full", "bool"), ("iri", "qpms_iri_t"), ("a", "ptr"),
                   ("ipiv", "ptr")],
    # scatsystem.h:488-492
    "qpms_scatsys_at_omega_k_t": [("ssw", "ptr"), ("k", "double[3]"), ("eta", "double")],
    # beyn.h:82-94
    "beyn_result_t": [("neig", "size_t"), ("vlen", "size_t"), ("eigval", "ptr"), ("eigval_err", "ptr"),
                      ("residuals", "ptr"), ("eigvec", "ptr"), ("ranktest_SV", "ptr"), ("gsl", "ptr")],
    # beyn.h:30-43 (flexible array member of complex double pairs)
    "beyn_contour_t": [("n", "size_t"), ("centre", "complex double"), ("inside_test", "ptr"),
                       ("z_dz", "complex double[0]")],
}


def size_align(t):
    """(size, alignment) of a type string under x86-64 SysV."""
    count = 1
    if t.endswith("]"):
        t, n = t[:-1].split("[")
        count = int(n)
    t = TYPEDEF.get(t, t)
    if t in SCALAR:
        s, a = SCALAR[t]
    else:
        s, a = layout(t)[1:]
    return s * count, a


def layout(name):
    """(offsets by field, size, alignment) of a reference struct."""
    off, align, out = 0, 1, {}
    for f, t in REF[name]:
        s, a = size_align(t)
        off = (off + a - 1) // a * a
        out[f] = off
        off += s
        align = max(align, a)
    return out, (off + align - 1) // align * align, align


def gcc_offsets():
    """offsetof/sizeof of the same fields in include/qpms_hip.h, from gcc."""
    lines = ["#include <stdio.h>", "#include <stddef.h>", '#include "qpms_hip.h"', "int main(void) {"]
    for st, fields in REF.items():
        if st in ("qpms_epsmu_t",):
            continue
        lines.append(f'printf("{st} . %zu\\n", sizeof({st}));')
        for f, _ in fields:
            lines.append(f'printf("{st} {f} %zu\\n", offsetof({st}, {f}));')
    lines.append("return 0; }")
    with tempfile.TemporaryDirectory() as d:
        src, exe = os.path.join(d, "l.c"), os.path.join(d, "l")
        open(src, "w").write("\n".join(lines))
        subprocess.run(["gcc", "-std=gnu11", "-I", os.path.join(ROOT, "include"), src, "-o", exe], check=True)
        out = subprocess.run([exe], check=True, capture_output=True, text=True).stdout
    res = {}
    for ln in out.splitlines():
        st, f, v = ln.split()
        res[(st, f)] = int(v)
    return res


def test_calculator_self_check():
    """The layout calculator on structs whose layout is known by heart."""
    assert layout("cart3_t")[1:] == (24, 8)
    assert layout("csph_t")[1:] == (32, 8)
    off, size, _ = layout("qpms_ss_LU")
    assert off["full"] == 16 and off["iri"] == 20 and off["a"] == 24 and size == 40


def test_layouts_match_reference():
    got = gcc_offsets()
    bad = []
    for st in REF:
        if st == "qpms_epsmu_t":
            continue
        off, size, _ = layout(st)
        for f, o in off.items():
            if got[(st, f)] != o:
                bad.append(f"{st}.{f}: library {got[(st, f)]}, reference {o}")
        if got[(st, ".")] != size:
            bad.append(f"sizeof({st}): library {got[(st, '.')]}, reference {size}")
    assert not bad, "\n".join(bad)


def test_cython_read_fields():
    """The fields the reference's Cython layer reads (qpms_cdefs.pxd:385-387,
    600-620) sit where the reference puts them."""
    got = gcc_offsets()
    ss, _, _ = layout("qpms_scatsys_t")
    for f in ("p_count", "fecv_size", "saecv_sizes", "per"):
        assert got[("qpms_scatsys_t", f)] == ss[f]
    tc, _, _ = layout("qpms_trans_calculator")
    assert got[("qpms_trans_calculator", "lMax")] == tc["lMax"] == 4
    assert got[("qpms_trans_calculator", "nelem")] == tc["nelem"] == 8
