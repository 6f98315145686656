"""The C-ABI library loads on the host and exports every declared symbol.
(No compute calls here: those need a GPU and live in the gpu-marked tests.)"""
import ctypes as ct
import os
import re

from ufemism_amd import abi, fields

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    text = open(os.path.join(ROOT, "include", "ufe_dyn.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(ufe_[A-Za-z0-9_]+)\s*\(", text)))


def test_header_declarations_are_exported():
    lib = abi.load()
    names = declared_functions()
    assert len(names) >= 40
    for n in names:
        assert hasattr(lib, n), n
    assert set(names) == set(abi.EXPORTS)


def test_field_table_matches_reference_names():
    table = abi.field_table()
    for name, (grid, dtype) in fields.ICE_FIELDS.items():
        if name in ("LHSx_AaAc", "LHSy_AaAc", "resU_AaAc", "resV_AaAc"):
            continue   # SOR intermediates stay in registers on the device
        assert name in table, name
        code = {"Aa": 0, "Aaz": 1, "Aac": 2, "Ac": 3, "Acz": 4, "AaAc": 5}[grid]
        assert table[name][1] == code, name
        assert table[name][2] == (0 if dtype.__name__ == "float64" else 1), name


STRUCTS = {"ufe_mesh_desc": abi.MeshDesc, "ufe_params": abi.Params, "ufe_ssa_stats": abi.SSAStats,
           "ufe_linsolve_opts": abi.LinSolveOpts, "ufe_linsolve_stats": abi.LinSolveStats,
           "ufe_kernel_profile": abi.KernelProfile,
           "ufe_adjoint_opts": abi.AdjointOpts, "ufe_adjoint_stats": abi.AdjointStats,
           "ufe_batch_opts": abi.BatchOpts, "ufe_batch_stats": abi.BatchStats,
           "ufe_sweep_tuning": abi.SweepTuning, "ufe_mesh_geometry_desc": abi.MeshGeometryDesc,
           "ufe_mesh_geometry_stats": abi.MeshGeometryStats, "ufe_remap_stats": abi.RemapStats,
           "ufe_multigrid_info": abi.MultigridInfo}


def test_struct_layouts_match_header(tmp_path):
    """Compile a probe against include/ufe_dyn.h and compare every offset."""
    import subprocess
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "ufe_dyn.h"', "int main(void){"]
    for cname, py in STRUCTS.items():
        lines.append(f'printf("{cname} size %zu\\n", sizeof({cname}));')
        for f in py._fields_:
            lines.append(f'printf("{cname} {f[0]} %zu\\n", offsetof({cname}, {f[0]}));')
    lines.append("return 0;}")
    src = tmp_path / "probe.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "probe"
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split("\n")
    for line in filter(None, out):
        cname, field, val = line.split()
        py = STRUCTS[cname]
        if field == "size":
            assert ct.sizeof(py) == int(val), cname
        else:
            assert getattr(py, field).offset == int(val), (cname, field)


def test_version_and_no_device_context():
    lib = abi.load()
    assert lib.ufe_version() >= 1
    h = ct.c_void_p()
    # creating a context needs a device; without one it reports an error
    r = lib.ufe_ctx_create(99, ct.byref(h))
    assert r != 0 and not h.value


def test_constants_agree_across_the_three_bindings():
    """Status codes, precision modes and preconditioner kinds: the C header's
    enums, the Python binding's constants and the Fortran binding's
    parameters hold the same values."""
    text = open(os.path.join(ROOT, "include", "ufe_dyn.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    header = {}
    for body in re.findall(r"enum\s*\{(.*?)\}", text, flags=re.S):
        for name, val in re.findall(r"\b(UFE_[A-Z0-9_]+)\s*=\s*(\d+)", body):
            header[name] = int(val)
    assert {"UFE_OK", "UFE_HOST_ERROR", "UFE_MIXED", "UFE_PRECOND_MULTIGRID"} <= set(header)
    for name, val in header.items():
        assert hasattr(abi, name), name
        assert getattr(abi, name) == val, (name, getattr(abi, name), val)
    ftext = open(os.path.join(ROOT, "ufemism_amd", "fortran", "ufe_c_binding.f90")).read()
    fortran = {n: int(v) for n, v in re.findall(r"\b(UFE_[A-Z0-9_]+)\s*=\s*(\d+)", ftext)}
    assert {"UFE_OK", "UFE_HOST_ERROR"} <= set(fortran)
    for name, val in fortran.items():
        assert header.get(name) == val, (name, val, header.get(name))
    assert set(abi.STATUS_NAMES) == {v for n, v in header.items()
                                     if n in ("UFE_OK", "UFE_SOR_NOT_CONVERGED", "UFE_SSA_RESET",
                                              "UFE_SSA_UNSTABLE", "UFE_COLOURING_IDENTIFY",
                                              "UFE_COLOURING_FAILED", "UFE_BAD_INPUT", "UFE_HIP_ERROR",
                                              "UFE_NOT_SUPPORTED", "UFE_HOST_ERROR")}


def test_fortran_bind_c_types_match_header(tmp_path):
    """The Fortran host's BIND(C) mirrors (ufe_c_binding.f90) have the C structs'
    size and every component at the C field's offset: a field added to one
    side only shows here, on the CPU."""
    import shutil
    import subprocess
    import pytest
    flang = shutil.which(os.environ.get("FC", "flang"))
    if flang is None:
        pytest.skip("flang not available")
    fsrc = os.path.join(ROOT, "ufemism_amd", "fortran", "ufe_c_binding.f90")
    text = open(fsrc).read()
    types = {}
    for name, body in re.findall(r"TYPE,\s*BIND\(C\)\s*::\s*(\w+)(.*?)END TYPE", text, flags=re.S | re.I):
        comps = []
        for line in body.splitlines():
            if "::" not in line:
                continue
            decl = re.sub(r"\(.*?\)", "", line.split("::", 1)[1])
            comps += [c.strip() for c in decl.split(",") if c.strip()]
        types[name] = comps
    assert {"ufe_params", "ufe_ssa_stats", "ufe_mesh_desc"} <= set(types)
    fl = ["PROGRAM probe", "  USE, INTRINSIC :: ISO_C_BINDING", "  USE ufe_c_binding", "  IMPLICIT NONE"]
    for t in types:
        fl.append(f"  TYPE({t}), TARGET :: v_{t}")
    fl.append("  INTEGER(C_INTPTR_T) :: b")
    for t, comps in types.items():
        fl.append(f"  PRINT '(A,1X,A,1X,I0)', '{t}', 'size', C_SIZEOF(v_{t})")
        fl.append(f"  b = TRANSFER(C_LOC(v_{t}), b)")
        for c in comps:
            fl.append(f"  PRINT '(A,1X,A,1X,I0)', '{t}', '{c}', TRANSFER(C_LOC(v_{t}%{c}), b) - b")
    fl.append("END PROGRAM probe")
    (tmp_path / "probe.f90").write_text("\n".join(fl) + "\n")
    fexe = tmp_path / "fprobe"
    r = subprocess.run([flang, "-module-dir", str(tmp_path), fsrc, str(tmp_path / "probe.f90"), "-o", str(fexe)],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-3000:]
    fout = subprocess.run([str(fexe)], capture_output=True, text=True, check=True).stdout.split("\n")
    cl = ['#include <stdio.h>', '#include <stddef.h>', '#include "ufe_dyn.h"', "int main(void){"]
    for t, comps in types.items():
        cl.append(f'printf("{t} size %zu\\n", sizeof({t}));')
        for c in comps:
            cl.append(f'printf("{t} {c} %zu\\n", offsetof({t}, {c}));')
    cl.append("return 0;}")
    (tmp_path / "cprobe.c").write_text("\n".join(cl))
    cexe = tmp_path / "cprobe"
    r = subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), str(tmp_path / "cprobe.c"), "-o", str(cexe)],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-3000:]
    cout = subprocess.run([str(cexe)], capture_output=True, text=True, check=True).stdout.split("\n")
    f = [ln.split() for ln in fout if ln.strip()]
    c = [ln.split() for ln in cout if ln.strip()]
    assert len(f) == len(c) > 60
    for a, b in zip(f, c):
        assert a[0] == b[0] and a[1].lower() == b[1].lower() and int(a[2]) == int(b[2]), (a, b)
