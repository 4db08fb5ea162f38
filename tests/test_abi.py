"""The C ABI: libspectre_amd.so loads on a machine without a GPU and exports
every function include/spectre_amd.h declares, and the ctypes layer binds
exactly those functions.  No call here touches a device."""
import ctypes
import pathlib
import re

from spectre_amd import _lib

HEADER = pathlib.Path(__file__).resolve().parents[1] / "include" / "spectre_amd.h"


def declared_functions():
    text = re.sub(r"/\*.*?\*/", "", HEADER.read_text(), flags=re.S)
    return sorted(set(re.findall(r"\b(sxd_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_the_entry_points():
    names = declared_functions()
    assert len(names) >= 25
    for expected in ("sxd_partial_derivatives", "sxd_gh_rhs", "sxd_scalar_wave_rhs",
                     "sxd_gh_rhs_adjoint", "sxd_rk3_substep", "sxd_domain_create_mixed"):
        assert expected in names


def test_library_exports_every_declared_function():
    lib = ctypes.CDLL(str(_lib.LIB_PATH))
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_ctypes_layer_binds_the_declared_functions():
    assert set(_lib._SIGNATURES) == set(declared_functions())


def test_host_only_calls_work_without_a_device():
    lib = _lib.lib()
    assert lib.sxd_version().startswith(b"spectre_amd")
    x, w = (ctypes.c_double * 5)(), (ctypes.c_double * 5)()
    assert lib.sxd_lgl_points_and_weights(5, x, w) == 0
    assert abs(sum(w) - 2.0) < 1e-14
    assert lib.sxd_lgl_points_and_weights(1, x, w) != 0
    assert lib.sxd_last_error()
