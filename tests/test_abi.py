"""The C ABI library loads and exports every function include/shadow_channel.h
declares (no device calls: runs without a GPU)."""
import ctypes
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    text = open(os.path.join(ROOT, "include", "shadow_channel.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return re.findall(r"\b(c_\w+|sc_\w+)\s*\(", text)


def test_header_declares_the_swig_surface():
    names = set(declared_functions())
    swig = {"c_init", "c_destroy", "c_primal", "c_tangent", "c_adjoint", "c_ddt_project",
            "c_getsoln", "c_getincresoln", "c_spec2phys", "c_read_solution",
            "c_save_solution", "c_quad", "c_qpts", "c_dimR", "c_dimQ", "c_Nx", "c_Nz",
            "c_nsteps"}
    assert swig <= names


def test_library_exports_all_declared_symbols():
    lib = ctypes.CDLL(os.path.join(ROOT, "shadow-channel_amd", "libshadowchannel.so"))
    for name in declared_functions():
        assert hasattr(lib, name), name


def test_python_module_mirrors_swig_names():
    import c_channel
    for name in ["c_init", "c_destroy", "c_primal", "c_tangent", "c_adjoint",
                 "c_ddt_project", "c_getsoln", "c_getincresoln", "c_spec2phys",
                 "c_read_solution", "c_save_solution", "c_quad", "c_qpts", "c_dimR",
                 "c_dimQ", "c_Nx", "c_Nz", "c_nsteps"]:
        assert callable(getattr(c_channel, name))


def test_inplace_arrays_are_type_checked():
    import numpy as np
    import pytest
    import c_channel
    with pytest.raises(TypeError):
        c_channel.c_primal(0, np.zeros((2, 2, 2, 2)))            # float64, not complex
    with pytest.raises(TypeError):
        c_channel.c_tangent(0, 1, np.zeros((2, 2, 2), complex), 0)
    with pytest.raises(TypeError):
        c_channel.c_adjoint(1, 0, np.zeros((2, 2, 2, 4), complex)[..., ::2], 0, 0.0)
