"""CPU-side checks of the drop-in boundary: the built library loads, exports
every entry point include/qew_amd.h declares, and the Python layer declares
them all (no compute calls: no GPU here)."""
import ctypes as C
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "qew_amd.h")
LIB = os.path.join(ROOT, "qew_amd", "libqew_amd.so")


def declared():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:const char \*|qew_status|qew_lbfgs_params|qew_fire_params)\s*(qew_\w+)\s*\(",
                                 text, flags=re.M)))


def test_header_declares_entry_points():
    names = declared()
    assert "qew_eval" in names and "qew_batch_lbfgs" in names and "qew_rk_block" in names
    assert "qew_fire_minimize" in names and "qew_fire_default_params" in names
    assert len(names) >= 30


@pytest.mark.skipif(not os.path.exists(LIB), reason="libqew_amd.so not built")
def test_library_exports_every_declared_symbol():
    lib = C.CDLL(LIB)
    missing = [n for n in declared() if not hasattr(lib, n)]
    assert not missing, missing


@pytest.mark.skipif(not os.path.exists(LIB), reason="libqew_amd.so not built")
def test_python_layer_binds_every_declared_symbol():
    from qew_amd import _lib
    bound = {n for n, _ in _lib.SIGNATURES} | set(_lib._RESTYPE)
    missing = [n for n in declared() if n not in bound]
    assert not missing, missing
    _lib.load()  # argtypes for every symbol resolve


@pytest.mark.skipif(not os.path.exists(LIB), reason="libqew_amd.so not built")
def test_errors_without_device_are_reported_not_raised_across_abi():
    from qew_amd import _lib
    lib = _lib.load()
    # a null context is rejected with QEW_ERR_INVALID_ARGUMENT and a message
    st = lib.qew_ctx_synchronize(None)
    assert st == _lib.QEW_ERR_INVALID_ARGUMENT
    assert b"null context" in lib.qew_last_error()
    p = lib.qew_lbfgs_default_params()
    assert (p.m, p.ftol, p.max_iter, p.max_ls) == (8, 1e-6, 10000, 30)
    f = lib.qew_fire_default_params()
    # src/fire.h:23-33 defaults; dt_max has none
    assert (f.dt_max, f.dt_init, f.ftol, f.max_iter, f.n_min) == (0.0, 0.0, 1e-6, 100000, 5)
    assert (f.f_inc, f.f_dec, f.alpha_start, f.f_alpha) == (1.1, 0.5, 0.1, 0.99)
    st = lib.qew_fire_minimize(None, None, None, None, 4, None, None)
    assert st == _lib.QEW_ERR_INVALID_ARGUMENT and b"null context" in lib.qew_last_error()


def test_dtype_values_match_the_header():
    """The Python layer's table/state types carry the header's enum values
    (QEW_F64, QEW_F32, QEW_BF16, QEW_F32_MID, QEW_F64_MID)."""
    import re
    import qew_amd._lib as L
    src = open(os.path.join(os.path.dirname(LIB), "..", "include", "qew_amd.h")).read()
    m = re.search(r"typedef enum qew_dtype \{([^}]*)\}", src)
    assert m, "qew_dtype enum not found"
    vals = dict((k.strip(), int(v)) for k, v in (e.split("=") for e in m.group(1).split(",")))
    assert vals == {"QEW_F64": L.F64, "QEW_F32": L.F32, "QEW_BF16": L.BF16, "QEW_F32_MID": L.F32_MID,
                    "QEW_F64_MID": L.F64_MID}
