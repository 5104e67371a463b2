"""The C-ABI library loads and exports every symbol include/emix3d.h declares;
argument checks that fire before any device call behave as documented.
No GPU is needed."""
import ctypes
import os

import pytest

from emix3d_amd import _lib


def test_library_built_in_tree():
    assert os.path.exists(_lib.LIB_PATH), "run python -m emix3d_amd.build"


def test_exports_every_declared_symbol():
    L = _lib.lib()
    declared = _lib.declared_symbols()
    assert len(declared) >= 20
    missing = [s for s in declared if not hasattr(L, s)]
    assert not missing, missing
    # and every declared symbol has a ctypes signature
    assert set(declared) == set(_lib.SIGNATURES), set(declared) ^ set(_lib.SIGNATURES)


def test_reference_symbol_names_exported():
    """The reference Fortran binds these exact names (bind(C, name=...)):
    solver/deriv_vel_cuda.f90:19-39, solver/divergencia_cuda.f90:21-42,
    include/LevelSet/interfaceAdvect.h:3-16."""
    L = _lib.lib()
    for name in ("cudaDeriv", "cuDivergence", "advect_CUDA"):
        assert hasattr(L, name)


def test_version():
    assert _lib.lib().emx_version() == 1


@pytest.mark.parametrize("args,status", [
    ((64, 64, 64, 7), -1),     # unknown scheme
    ((4, 64, 64, 0), -3),      # below the 5 points the closures need
    ((64, 64, 3, 2), -3),
    ((2048, 64, 64, 1), -3),   # xi line longer than one wavefront can own
    ((64, 1024, 64, 1), -3),   # eta line longer than 16 waves can own
])
def test_create_rejects_bad_sizes(args, status):
    L = _lib.lib()
    c = ctypes.c_void_p()
    assert L.emx_create(ctypes.byref(c), *args, None) == status
    assert not c.value
    assert L.emx_last_error()  # non-empty message


def test_null_context_is_an_error():
    L = _lib.lib()
    assert L.emx_destroy(None) == -1
    assert b"null context" in L.emx_last_error()
    assert L.emx_deriv(None, 0, None, None, 1.0) == -1
    assert L.emx_create(None, 8, 8, 8, 0, None) == -1


def test_declared_entry_points_cite_the_reference():
    """Every entry point in the header names the reference routine it replaces
    (file:line) or says it has no counterpart."""
    text = open(_lib.HEADER).read()
    for ref in ("solver/deriv_vel_cuda.f90:19-39", "solver/divergencia_cuda.f90:21-42",
                "include/LevelSet/interfaceAdvect.h:3-16", "solver/flujos_compact.f90",
                "solver/sgdm_incomp.f90", "solver/jacob_malla.f90",
                "src/LevelSet/lsTransportEc.cu:201-418"):
        assert ref in text, ref
