"""The C ABI boundary without a GPU: every function include/pfem2.h declares is exported by the built
library, the header is plain C (compiles as C99 on its own), and the C++ drivers written against the
facade (include/gpupfem2.hpp) are built and linked against the library."""
import ctypes
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "pfem2.h")
LIB = os.path.join(ROOT, "gpupfem2_amd", "libpfem2_hip.so")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", " ", text, flags=re.S)
    names = re.findall(r"^\s*(?:const\s+)?[A-Za-z_][A-Za-z0-9_]*\s*\**\s*(pfem_[a-z0-9_]+)\s*\(", text, flags=re.M)
    return sorted(set(names))


def test_header_declares_the_boundary():
    names = declared_functions()
    assert len(names) > 60
    for must in ("pfem_ctx_create", "pfem_mesh2d_load_dat", "pfem_csr_create_from_mesh2d", "pfem_dirichlet_apply",
                 "pfem_precond_create", "pfem_precond_create_custom", "pfem_solver_solve", "pfem_particles_advect",
                 "pfem_assemble_pressure", "pfem_pfem2_step", "pfem_export_vtk"):
        assert must in names


def test_library_exports_every_declared_symbol():
    assert os.path.exists(LIB), "build first: python -m gpupfem2_amd.build"
    lib = ctypes.CDLL(LIB)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_header_is_plain_c(tmp_path):
    src = tmp_path / "c_abi.c"
    src.write_text('#include "pfem2.h"\nint main(void) { pfem_ctx c = 0; (void)c; return 0; }\n')
    p = subprocess.run(["gcc", "-std=c99", "-Wall", "-Werror", "-fsyntax-only", "-I" + os.path.join(ROOT, "include"),
                        str(src)], capture_output=True, text=True)
    assert p.returncode == 0, p.stderr


@pytest.mark.parametrize("name", ["poisson2d", "facade_check"])
def test_cpp_drivers_built(name):
    exe = os.path.join(ROOT, "gpupfem2_amd", "examples", name)
    assert os.access(exe, os.X_OK), "build first: python -m gpupfem2_amd.build"
    out = subprocess.run(["ldd", exe], capture_output=True, text=True).stdout
    assert "libpfem2_hip.so" in out and "not found" not in out.split("libpfem2_hip.so")[1].splitlines()[0]
