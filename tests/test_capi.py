"""CPU checks of the C-ABI boundary: libhiten_hip.so loads and exports every
entry point include/hiten_hip.h declares, and the Python binding declares the
same set.  No compute entry point is called (there is no GPU here)."""

import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "hiten_hip.h")


def declared_symbols():
    src = open(HEADER).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(hh_\w+)\s*\(", src)))


def test_header_declares_entry_points():
    names = declared_symbols()
    assert "hh_dop853_final" in names and "hh_dop853_event" in names and "hh_rk8_fixed" in names


def test_library_exports_every_declared_symbol():
    from hiten_amd import _lib
    path = _lib.library_path()
    assert os.path.exists(path), "build the library first (python -m hiten_amd.build)"
    lib = ctypes.CDLL(path)
    missing = [n for n in declared_symbols() if not hasattr(lib, n)]
    assert not missing, missing
    lib.hh_abi_version.restype = ctypes.c_int
    assert lib.hh_abi_version() == _lib.ABI_VERSION


def test_binding_covers_header():
    from hiten_amd import _lib
    assert sorted(_lib.EXPORTED) == declared_symbols()


def test_argument_errors_are_reported_without_gpu():
    """Argument validation happens before any HIP call."""
    from hiten_amd import _lib
    lib = _lib.load(require_gpu=False)
    rc = lib.hh_dop853_final(4, 7, None, None, None, 1e-12, 1e-12, 1e4, 1e-15, 10, 1, 0, None,
                             None, None, None, None, 0, 0, 0, None, 0, None, None)
    assert rc == 1
    assert b"dim must be 6 or 42" in lib.hh_last_error()
    # workspace sizes: none for the one-launch forms, queue + spill area for the pair form
    assert lib.hh_dop853_final_workspace(1000, 42, _lib.HH_VARIANT_PLAIN) == 0
    assert lib.hh_dop853_final_workspace(1000, 42, _lib.HH_VARIANT_PERSISTENT) > 0
    assert lib.hh_dop853_final_workspace(1000, 42, _lib.HH_VARIANT_PERSISTENT_PAIR) >= \
        (15 * 1000 + 64) * 4 + 89 * 1000 * 8
    rc = lib.hh_dop853_event(4, None, None, None, 1e-12, 1e-12, 1e4, 1e-15, 10, 9, 0.0, 0,
                             1e-12, 1e-12, 1, 0, None, None, None, None, None, None, None, None,
                             0, 0, None)
    assert rc == 1
    assert b"null pointer" in lib.hh_last_error() or b"ev_index" in lib.hh_last_error()


def test_no_gpu_raises():
    import torch
    from hiten_amd import _lib, batch
    if torch.cuda.is_available():
        pytest.skip("a GPU is visible")
    with pytest.raises(_lib.HipUnavailable):
        batch.eval_math([1.0], [2.0])


def test_rk_fixed_order_checked_without_gpu():
    from hiten_amd import _lib
    lib = _lib.load(require_gpu=False)
    rc = lib.hh_rk_fixed(5, 4, 6, None, None, None, 10, None, 1, 0, None, None, None, None, 0, 0,
                         None)
    assert rc == 1 and b"order must be 4, 6 or 8" in lib.hh_last_error()


def test_integrator_factories_follow_the_reference():
    """rk.py:2871-2979: FixedRK orders 4/6/8 (default 4), RungeKutta 4/6/8 fixed
    and 853 adaptive (default 4); the adaptive order 5 has no kernel."""
    from hiten_amd import integrators as I
    assert type(I.FixedRK()).__name__ == "_RK4"
    assert [I.FixedRK(order=k).order for k in (4, 6, 8)] == [4, 6, 8]
    assert type(I.RungeKutta()).__name__ == "_RK4"
    assert type(I.RungeKutta(order=6)).__name__ == "_RK6"
    assert type(I.RungeKutta(order=853)).__name__ == "_DOP853"
    with pytest.raises(ValueError):
        I.FixedRK(order=5)
    with pytest.raises(ValueError):
        I.RungeKutta(order=7)
    with pytest.raises(NotImplementedError):
        I.AdaptiveRK()
    with pytest.raises(NotImplementedError):
        I.RungeKutta(order=45)


def test_rk_fixed_event_arguments_checked_without_gpu():
    from hiten_amd import _lib
    lib = _lib.load(require_gpu=False)
    rc = lib.hh_rk_fixed_event(7, 4, None, None, None, 10, None, 1, 0, 1, 0.0, 0, 1e-12, 1e-12,
                               None, None, None, None, None, 0, None)
    assert rc == 1 and b"order must be 4, 6 or 8" in lib.hh_last_error()
    rc = lib.hh_rk_fixed_event(4, 4, None, None, None, 10, None, 1, 0, 9, 0.0, 0, 1e-12, 1e-12,
                               None, None, None, None, None, 0, None)
    assert rc == 1 and b"ev_index" in lib.hh_last_error()
