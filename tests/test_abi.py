"""C ABI: the shared library loads and exports every function that the
headers under include/ declare (no compute calls: runs without a GPU)."""
import ctypes
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    inc = os.path.join(ROOT, "include")
    src = "\n".join(open(os.path.join(inc, f)).read() for f in sorted(os.listdir(inc)) if f.endswith(".h"))
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    src = re.sub(r"static inline[^{]*\{[^}]*\}", "", src)
    names = set()
    for m in re.finditer(r"\b([A-Za-z_][A-Za-z0-9_]*)\s*\(([^;{]*?)\)\s*;", src, flags=re.S):
        name = m.group(1)
        if m.group(2).lstrip().startswith("*"):  # function-pointer member or typedef
            continue
        if name in {"if", "while", "return", "sizeof"}:
            continue
        pre = src[max(0, m.start() - 200):m.start()]
        if re.search(r"\(\s*\*\s*$", pre):  # function pointer member
            continue
        names.add(name)
    return sorted(n for n in names if not n.startswith("__"))


def test_header_declares_reference_entry_points():
    names = declared_functions()
    for ref in ["qpms_trans_calculator_get_AB_arrays", "qpms_scatsys_build_translation_matrix_full",
                "qpms_scatsysw_build_modeproblem_matrix_full", "qpms_scatsyswk_build_modeproblem_matrix_full",
                "qpms_scatsys_build_modeproblem_matrix_full", "ewald3_sigma_short", "ewald3_sigma_long",
                "ewald3_sigma0", "beyn_solve"]:
        assert ref in names, ref


def test_library_exports_every_declared_symbol():
    import qpms_amd
    L = ctypes.CDLL(qpms_amd.LIBPATH)
    missing = [n for n in declared_functions() if not hasattr(L, n)]
    assert not missing, f"declared but not exported: {missing}"


def test_struct_layouts():
    """Field offsets of structs that reference callers read (x86-64)."""
    import qpms_amd
    tc = qpms_amd.trans_calculator
    assert tc.lMax.offset == 4 and tc.nelem.offset == 8 and tc.A_multipliers.offset == 16
    assert ctypes.sizeof(qpms_amd.csph_t) == 32
