"""Struct layouts of the C ABI as the three bindings see them, no GPU needed:
the header compiled by gcc (sizeof / offsetof), the ctypes structures of
fmt2d_amd/hip.py and the bind(C) types of the Fortran module
(fmt2d_amd/fortran/fmt2d_hip_mod.f90, compiled by amdflang).  A field in the
wrong place would pass arguments to the library silently wrong."""
import ctypes
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

# C struct name -> (hip.py class, field names in order)
STRUCTS = {
    "fmt2d_grid_info": ("GridInfo", ["nx", "nz", "dx", "dz", "x0", "z0", "region_nnode"]),
    "fmt2d_system_params": ("SystemParams", ["crit_chi", "phase_wt_psed", "phase_wt_pmp", "phase_wt_pgss",
                                             "refl_weight", "alt_mode"]),
    "fmt2d_lsqr_params": ("LsqrParams", ["wdv", "wdv_iface", "w_smooth", "w_smooth_iface", "smooth_type",
                                         "max_iter", "corr_x", "corr_z", "corr_x_iface"]),
    "fmt2d_damping_params": ("DampingParams", ["lam_damp", "lam_damp_iface", "target_dv", "target_dd",
                                               "max_rel_pert"]),
    "fmt2d_damping_result": ("DampingResult", ["wd", "wd_iface", "avg_dmv", "avg_dmd", "solves", "iterations"]),
    "fmt2d_line_search_params": ("LineSearchParams", ["crit_chi", "iface_step_ratio", "first_step"]),
    "fmt2d_line_search_result": ("LineSearchResult", ["step", "rms_orig", "rms_new", "next_first_step",
                                                      "forwards"]),
}


def c_layout(tmp):
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "fmt2d_hip.h"', "int main(void) {"]
    for s, (_, fields) in STRUCTS.items():
        lines.append(f'  printf("{s} %zu", sizeof({s}));')
        for f in fields:
            lines.append(f'  printf(" %zu", offsetof({s}, {f}));')
        lines.append('  printf("\\n");')
    lines += ["  return 0;", "}"]
    src, exe = os.path.join(tmp, "layout.c"), os.path.join(tmp, "layout_c")
    open(src, "w").write("\n".join(lines) + "\n")
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), "-o", exe, src], check=True)
    return parse(subprocess.run([exe], capture_output=True, text=True, check=True).stdout)


def fortran_layout(tmp, fc):
    lines = ["program layout", "  use iso_c_binding", "  use fmt2d_hip", "  implicit none"]
    for k, s in enumerate(STRUCTS):
        lines.append(f"  type({s}), target :: v{k}")
    lines.append("  integer(c_intptr_t) :: b")
    for k, (s, (_, fields)) in enumerate(STRUCTS.items()):
        lines.append(f"  b = transfer(c_loc(v{k}), b)")
        lines.append(f"  write(*, '(A,1X,I0)', advance='no') '{s}', c_sizeof(v{k})")
        for f in fields:
            lines.append(f"  write(*, '(1X,I0)', advance='no') transfer(c_loc(v{k}%{f}), b) - b")
        lines.append("  write(*, '(A)') ''")
    lines.append("end program layout")
    src, exe = os.path.join(tmp, "layout.f90"), os.path.join(tmp, "layout_f")
    open(src, "w").write("\n".join(lines) + "\n")
    mod = os.path.join(ROOT, "fmt2d_amd", "fortran", "fmt2d_hip_mod.f90")
    lib = os.path.join(ROOT, "fmt2d_amd")
    subprocess.run([fc, "-module-dir", tmp, "-o", exe, mod, src, "-L", lib, "-lfmt2d_hip", f"-Wl,-rpath,{lib}"],
                   check=True, cwd=tmp)
    return parse(subprocess.run([exe], capture_output=True, text=True, check=True).stdout)


def parse(text):
    out = {}
    for ln in text.splitlines():
        p = ln.split()
        if p:
            out[p[0]] = [int(x) for x in p[1:]]
    return out


def ctypes_layout():
    from fmt2d_amd import hip
    out = {}
    for s, (cls, fields) in STRUCTS.items():
        c = getattr(hip, cls)
        assert [f[0] for f in c._fields_] == fields, f"{cls} fields differ from {s}"
        out[s] = [ctypes.sizeof(c)] + [getattr(c, f).offset for f in fields]
    return out


def test_ctypes_structs_match_header(tmp_path):
    assert ctypes_layout() == c_layout(str(tmp_path))


def test_fortran_types_match_header(tmp_path):
    from fmt2d_amd.build import fortran_compiler
    from fmt2d_amd import hip
    fc = fortran_compiler()
    if fc is None or shutil.which("gcc") is None:
        pytest.skip("needs amdflang and gcc")
    if not os.path.exists(hip.LIB_PATH):
        pytest.skip("libfmt2d_hip.so not built (python -m fmt2d_amd.build)")
    assert fortran_layout(str(tmp_path), fc) == c_layout(str(tmp_path))
