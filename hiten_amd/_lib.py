"""ctypes binding of ``libhiten_hip.so`` (the C-ABI declared in include/hiten_hip.h).

torch is imported before the library is opened so that the process has a
single HIP runtime: torch ships ``libamdhip64.so`` with the same SONAME the
library links against, and the dynamic loader reuses the one already mapped.

There is no fallback: if the library is missing, or no GPU is visible, every
entry point raises :class:`HipUnavailable`.
"""

from __future__ import annotations

import ctypes as C
import os
import threading

import torch  # noqa: F401  (must precede the library load, see module docstring)

from .build import LIB

HH_OK = 0
HH_ERR_ARGS, HH_ERR_HIP = 1, 2
HH_VARIANT_PLAIN = 0
HH_VARIANT_PERSISTENT = 1
HH_VARIANT_WIDE = 2
HH_VARIANT_PAIR = 3
HH_VARIANT_PERSISTENT_PAIR = 4
HH_STATUS_OK, HH_STATUS_MAXSTEPS, HH_STATUS_NONFINITE, HH_STATUS_DROPPED = 0, 1, 2, 4
ABI_VERSION = 3


class HipUnavailable(RuntimeError):
    """The HIP library or a GPU is not available."""


class HipLibraryError(RuntimeError):
    """An entry point of libhiten_hip.so returned an error code."""


class Tuning(C.Structure):
    """hh_tuning_t: scheduling knobs of the persistent 42-d kernels."""
    _fields_ = [("heavy_projected", C.c_int), ("min_age", C.c_int),
                ("pair_projected", C.c_int), ("refill_min", C.c_int),
                ("waves_per_simd", C.c_int), ("exclusive_projected", C.c_int)]

    def as_dict(self) -> dict:
        return {name: getattr(self, name) for name, _ in self._fields_}


_P = C.c_void_p
_I = C.c_int
_D = C.c_double
_LL = C.c_longlong
_ULL = C.c_ulonglong

_SIGNATURES = {
    "hh_last_error": ([], C.c_char_p),
    "hh_abi_version": ([], _I),
    "hh_set_tuning": ([C.POINTER(Tuning)], _I),
    "hh_get_tuning": ([C.POINTER(Tuning)], _I),
    "hh_rk8_fixed": ([_I, _I, _P, _P, _P, _I, _P, _I, _ULL, _P, _P, _P, _I, _I, _P], _I),
    "hh_rk_fixed": ([_I, _I, _I, _P, _P, _P, _I, _P, _I, _ULL, _P, _P, _P, _P, _I, _I, _P], _I),
    "hh_rk_fixed_event": ([_I, _I, _P, _P, _P, _I, _P, _I, _ULL, _I, _D, _I, _D, _D, _P, _P, _P,
                           _P, _P, _I, _P], _I),
    "hh_dop853_final": ([_I, _I, _P, _P, _P, _D, _D, _D, _D, _LL, _I, _ULL, _P, _P, _P, _P, _P,
                         _I, _I, _I, _P, _LL, _P, _P], _I),
    "hh_dop853_final_workspace": ([_I, _I, _I], _LL),
    "hh_dop853_final_static": ([_I, _I, _P, _P, _P, _D, _D, _D, _D, _LL, _I, _ULL, _P, _P, _P, _P,
                                _P, _I, _I, _I, _P, _P], _I),
    "hh_synodic_detect": ([_I, _I, _P, _I, _P, _P, _P, _D, _I, _I, _I, _I, _I, _D, _D, _D, _I,
                           _I, _I, _P, _P, _P, _P, _P], _I),
    "hh_synodic_workspace": ([_I, _I], _LL),
    "hh_synodic_detect_ws": ([_I, _I, _P, _I, _P, _P, _P, _D, _I, _I, _I, _I, _I, _D, _D, _D,
                              _I, _I, _I, _P, _P, _P, _P, _P, _LL, _P], _I),
    "hh_dop853_grid": ([_I, _I, _P, _P, _P, _I, _P, _D, _D, _D, _D, _LL, _I, _ULL, _P, _P, _P, _P,
                        _P, _I, _I, _P], _I),
    "hh_dop853_event": ([_I, _P, _P, _P, _D, _D, _D, _D, _LL, _I, _D, _I, _D, _D, _I, _ULL, _P,
                         _P, _P, _P, _P, _P, _P, _P, _I, _I, _P], _I),
    "hh_eval_rhs": ([_I, _I, _P, _P, _P, _P], _I),
    "hh_eval_math": ([_I, _P, _P, _P, _P], _I),
    "hh_traj_filters": ([_I, _I, _P, _P, _P, _P, _P, _P], _I),
    "hh_memcpy_async": ([_P, _P, _ULL, _I, _P], _I),
    "hh_stream_create_dedicated": ([], _P),
    "hh_stream_destroy": ([_P], _I),
    "hh_host_alloc": ([_ULL, C.POINTER(_P)], _I),
    "hh_host_free": ([_P], _I),
    "hh_event_create": ([], _P),
    "hh_event_destroy": ([_P], None),
    "hh_event_record": ([_P, _P], _I),
    "hh_event_query": ([_P], _I),
    "hh_bank_create": ([_I, _I], _P),
    "hh_bank_destroy": ([_P], None),
    "hh_bank_size": ([_P], _I),
    "hh_bank_add": ([_P, _D, _P, _P, _P, _P, _I, _I, _I, _D, _D, _D, _D, _D], _I),
    "hh_bank_stages": ([_P, _I, _P, _P], _I),
    "hh_bank_event_inputs": ([_P, _I, _P, _P, _P, _I], _I),
    "hh_bank_feed_events": ([_P, _I, _P, _P, _P, _P], _I),
    "hh_bank_stm_inputs": ([_P, _I, _P, _P, _P], _I),
    "hh_bank_jacobians": ([_P, _I, _P, _P, _P, _P], _I),
    "hh_bank_set_steps": ([_P, _I, _P, _P], _I),
    "hh_bank_result": ([_P, _I, _P, _P, _P, _P, _P, _P, _P, _P], _I),
    "hh_bank_release": ([_P, _I], _I),
    "hh_correct": ([_I, _P, _P, _D, _D, _D, _D, _P], _I),
    "hh_correct_sizes": ([_P, _P], _I),
    "hh_plane_crossing": ([_I, _P, _P, _I, _I, _D, _I, _D, _D, _D, _D, _D, _D, _P, _P, _P, _P,
                           _P], _I),
    "hh_eval_solve2": ([_I, _P, _P, _P, _P, _P], _I),
}

EXPORTED = tuple(_SIGNATURES)

_lock = threading.Lock()
_lib = None


def library_path() -> str:
    return LIB


def load(require_gpu: bool = True):
    """Open the library (once).  Raises HipUnavailable if it cannot be used."""
    global _lib
    with _lock:
        if _lib is None:
            if not os.path.exists(LIB):
                raise HipUnavailable(
                    f"{LIB} is missing: build it with `python -m hiten_amd.build`")
            lib = C.CDLL(LIB)
            for name, (argt, rest) in _SIGNATURES.items():
                fn = getattr(lib, name)
                fn.argtypes = argt
                fn.restype = rest
            if lib.hh_abi_version() != ABI_VERSION:
                raise HipUnavailable("libhiten_hip.so ABI version mismatch; rebuild it")
            _lib = lib
    if require_gpu and not torch.cuda.is_available():
        raise HipUnavailable("no HIP device is visible to this process")
    return _lib


def get_tuning() -> dict:
    """Current hh_tuning_t as a dict."""
    t = Tuning()
    check(load(require_gpu=False).hh_get_tuning(C.byref(t)), "hh_get_tuning")
    return t.as_dict()


def set_tuning(**kw) -> dict:
    """Change some hh_tuning_t fields (the others keep their values); returns
    the previous settings."""
    lib = load(require_gpu=False)
    old = get_tuning()
    unknown = set(kw) - set(old)
    if unknown:
        raise ValueError(f"unknown tuning fields {sorted(unknown)}")
    new = dict(old, **{k: int(v) for k, v in kw.items()})
    check(lib.hh_set_tuning(C.byref(Tuning(**new))), "hh_set_tuning")
    return old


def check(rc: int, what: str) -> None:
    if rc != HH_OK:
        msg = _lib.hh_last_error().decode(errors="replace") if _lib is not None else ""
        raise HipLibraryError(f"{what} failed (code {rc}): {msg}")


def stream_handle(stream=None) -> int:
    s = torch.cuda.current_stream() if stream is None else stream
    return int(s.cuda_stream)
