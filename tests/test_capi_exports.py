"""The C ABI library loads without a GPU and exports every entry point include/fvm_capi.h
declares; the Python and Fortran bindings bind exactly those names."""
import ctypes
import os
import re
import subprocess

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(REPO, "include", "fvm_capi.h")
LIB = os.path.join(REPO, "fvm_amd", "libfvm_amd.so")
FMOD = os.path.join(REPO, "fvm_amd", "fortran", "fvm_capi_mod.f90")


def header_functions():
    txt = open(HEADER).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\bint\s+(fvm_\w+)\s*\(", txt)))


def test_header_declares_the_step_subroutines():
    names = header_functions()
    for n in ("fvm_create", "fvm_destroy", "fvm_step", "fvm_set_field", "fvm_get_field",
              "fvm_contra_vel", "fvm_nl", "fvm_stress_tensor_dy", "fvm_dif", "fvm_p_source",
              "fvm_iter_u", "fvm_correct", "fvm_iter_p", "fvm_pressure_reference",
              "fvm_calccoef"):
        assert n in names, n


@pytest.mark.skipif(not os.path.exists(LIB), reason="libfvm_amd.so not built")
def test_library_exports_every_declared_symbol():
    out = subprocess.run(["nm", "-D", "--defined-only", LIB], check=True, capture_output=True,
                         text=True).stdout
    exported = {ln.split()[-1] for ln in out.splitlines() if ln.strip()}
    missing = [n for n in header_functions() if n not in exported]
    assert not missing, missing
    lib = ctypes.CDLL(LIB)  # loading performs no HIP call
    for n in header_functions():
        assert getattr(lib, n) is not None


def test_python_binding_matches_header():
    from fvm_amd import capi
    assert sorted(capi.EXPORTS) == header_functions()


def test_fortran_binding_names_are_declared():
    names = set(re.findall(r"NAME='(fvm_\w+)'", open(FMOD).read()))
    assert names and names <= set(header_functions()), names - set(header_functions())


def test_struct_layouts_match_header():
    """ctypes mirrors of fvm_params / fvm_step_info have the C layout the header implies."""
    from fvm_amd import capi
    # int ni,nj,nk; (pad) double re,dtime,ref_a; int use_les,nswpp; double resimax;
    # int poisson_mode,check_every,af_mode; (pad) double af_tol
    assert ctypes.sizeof(capi.Params) == 3 * 4 + 4 + 3 * 8 + 2 * 4 + 8 + 3 * 4 + 4 + 8
    assert capi.Params.af_tol.offset == ctypes.sizeof(capi.Params) - 8
    assert capi.StepInfo.af_iters.offset + 3 * 4 <= ctypes.sizeof(capi.StepInfo)
