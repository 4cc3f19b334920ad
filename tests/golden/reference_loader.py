"""Load hiten 0.5.4's own pure-Python sources in this container for the
golden-vector scripts beside this file (never imported by a test).

numba is not installed here, so a stand-in ``numba`` module makes
``numba.njit`` the identity (with or without a signature), ``numba.typed.List``
the built-in ``list`` and ``numba.types`` a do-nothing signature builder; the
decorated functions then run as plain Python and NumPy.  Their arithmetic
differs from numba's compiled code only where Python lowers an operation
differently: ``x ** 2`` and ``x ** 3`` on NumPy scalars call the C library's
``pow`` (numba multiplies), and ``np.dot`` / ``np.linalg.norm`` go through
NumPy's BLAS.  Everything that computes is the reference's file; modules that
the loaded code only names in annotations, base classes or unused branches
(logging set-up, the Hamiltonian and symplectic integrators, the return-map
base class) are replaced by empty stand-ins listed in ``_STAND_INS``.

    from tests.golden.reference_loader import load
    mods = load()          # dict: short name -> module
"""
from __future__ import annotations

import importlib.util
import logging
import os
import sys
import types

REF_SRC = "/root/reference/src/hiten"
REF = REF_SRC + "/algorithms"


def _package(name):
    if name in sys.modules:
        return sys.modules[name]
    mod = types.ModuleType(name)
    mod.__path__ = []
    sys.modules[name] = mod
    return mod


def _load(name, path):
    if name in sys.modules and getattr(sys.modules[name], "__file__", None) == path:
        return sys.modules[name]
    spec = importlib.util.spec_from_file_location(name, path)
    mod = importlib.util.module_from_spec(spec)
    sys.modules[name] = mod
    spec.loader.exec_module(mod)
    return mod


class _Sig:
    """numba.types stand-in: any attribute, call or subscript is itself."""

    def __getattr__(self, _name):
        return self

    def __call__(self, *a, **k):
        return self

    def __getitem__(self, _k):
        return self


def _install_numba():
    numba = types.ModuleType("numba")

    def njit(*args, **kwargs):
        if args and callable(args[0]) and not isinstance(args[0], _Sig) and not kwargs:
            return args[0]
        return lambda f: f

    numba.njit = njit
    numba.jit = njit
    numba.types = _Sig()
    typed = types.ModuleType("numba.typed")
    typed.List = list
    numba.typed = typed
    core = types.ModuleType("numba.core")
    registry = types.ModuleType("numba.core.registry")

    class CPUDispatcher:  # nothing is ever a compiled dispatcher here
        pass

    registry.CPUDispatcher = CPUDispatcher
    core.registry = registry
    numba.core = core
    sys.modules.update({"numba": numba, "numba.typed": typed, "numba.core": core,
                        "numba.core.registry": registry, "numba.types": numba.types})


def _stand_ins():
    log = types.ModuleType("hiten.utils.log_config")
    log.logger = logging.getLogger("hiten-reference")
    log.logger.addHandler(logging.NullHandler())
    log.logger.propagate = False
    sys.modules[log.__name__] = log

    ham = types.ModuleType("hiten.algorithms.dynamics.hamiltonian")

    def _hamiltonian_rhs(*a, **k):
        raise NotImplementedError("Hamiltonian systems are not loaded")
    ham._hamiltonian_rhs = _hamiltonian_rhs
    sys.modules[ham.__name__] = ham

    sym = types.ModuleType("hiten.algorithms.integrators.symplectic")

    class _ExtendedSymplectic:
        def __init__(self, *a, **k):
            raise NotImplementedError("symplectic integrators are not loaded")
    sym._ExtendedSymplectic = _ExtendedSymplectic
    sys.modules[sym.__name__] = sym

    rmb = types.ModuleType("hiten.algorithms.poincare.core.backend")

    class _ReturnMapBackend:
        def __init__(self, *a, **k):
            pass
    rmb._ReturnMapBackend = _ReturnMapBackend
    sys.modules[rmb.__name__] = rmb


_STAND_INS = ("hiten.utils.log_config", "hiten.algorithms.dynamics.hamiltonian",
              "hiten.algorithms.integrators.symplectic",
              "hiten.algorithms.poincare.core.backend",
              "hiten.algorithms.types.core (corrector only: the framework base classes "
              "_HitenBaseBackend, whose lifecycle hooks return None as the reference's do, "
              "_DomainPayload, _BackendCall and _HitenBaseInterface)")


def _core_stand_in():
    import dataclasses
    from types import MappingProxyType
    core = types.ModuleType("hiten.algorithms.types.core")

    class _HitenBaseBackend:
        def __init__(self, *a, **k):
            pass

        def on_iteration(self, k, x, r_norm):
            return None

        def on_accept(self, x, *, iterations, residual_norm):
            return None

        def on_failure(self, x, *, iterations, residual_norm):
            return None

        def on_success(self, x, *, iterations, residual_norm):
            return None

    @dataclasses.dataclass(frozen=True)
    class _DomainPayload:
        data: dict = dataclasses.field(default_factory=dict)

        def __post_init__(self):
            object.__setattr__(self, "data", MappingProxyType(dict(self.data)))

    class _BackendCall:
        def __init__(self, *a, **k):
            self.args, self.kwargs = a, k

    class _HitenBaseInterface:
        def __init__(self, *a, **k):
            pass

    core._HitenBaseBackend = _HitenBaseBackend
    core._DomainPayload = _DomainPayload
    core._BackendCall = _BackendCall
    core._HitenBaseInterface = _HitenBaseInterface
    sys.modules[core.__name__] = core


def load_corrector():
    """load(singlehit=True) plus the Newton/Armijo corrector
    (corrector/types.py, protocols.py, stepping/*.py, backends/base.py,
    backends/newton.py, operators.py) and types/exceptions.py."""
    m = load(singlehit=True)
    for p in ("hiten.algorithms.corrector", "hiten.algorithms.corrector.stepping",
              "hiten.algorithms.corrector.backends"):
        _package(p)
    _core_stand_in()
    C = f"{REF}/corrector"
    m["exceptions"] = _load("hiten.algorithms.types.exceptions", f"{REF}/types/exceptions.py")
    m["ctypes"] = _load("hiten.algorithms.corrector.types", f"{C}/types.py")
    m["cprotocols"] = _load("hiten.algorithms.corrector.protocols", f"{C}/protocols.py")
    m["norm"] = _load("hiten.algorithms.corrector.stepping.norm", f"{C}/stepping/norm.py")
    m["sbase"] = _load("hiten.algorithms.corrector.stepping.base", f"{C}/stepping/base.py")
    m["plain"] = _load("hiten.algorithms.corrector.stepping.plain", f"{C}/stepping/plain.py")
    m["armijo"] = _load("hiten.algorithms.corrector.stepping.armijo", f"{C}/stepping/armijo.py")
    stepping = _load("hiten.algorithms.corrector.stepping", f"{C}/stepping/__init__.py")
    m["stepping"] = stepping
    m["cbase"] = _load("hiten.algorithms.corrector.backends.base", f"{C}/backends/base.py")
    m["newton"] = _load("hiten.algorithms.corrector.backends.newton", f"{C}/backends/newton.py")
    m["operators"] = _load("hiten.algorithms.corrector.operators", f"{C}/operators.py")
    return m


def reference_function(path, qualname, namespace):
    """Compile one function of a reference file that cannot be imported
    whole (its module pulls in the whole package): the function's own source
    text, taken from the file at run time, executed in ``namespace``."""
    import ast
    src = open(path).read()
    tree = ast.parse(src)
    parts = qualname.split(".")
    body = tree.body
    node = None
    for name in parts:
        node = next(n for n in body if isinstance(n, (ast.ClassDef, ast.FunctionDef))
                    and n.name == name)
        body = node.body
    fn_src = ast.get_source_segment(src, node)
    lines = fn_src.splitlines()
    indent = len(lines[0]) - len(lines[0].lstrip())
    code = "\n".join(l[indent:] if l.strip() else "" for l in lines)
    ns = dict(namespace)
    exec(compile(code, path, "exec"), ns)
    return ns[parts[-1]]


def load(singlehit: bool = False):
    """Load the propagation path (and, with ``singlehit``, the section-crossing
    backend) of the reference; returns {short name: module}."""
    if not os.path.isdir(REF):
        raise RuntimeError(f"{REF} not found: the golden scripts run where the reference is")
    _install_numba()
    for p in ("hiten", "hiten.utils", "hiten.algorithms", "hiten.algorithms.utils",
              "hiten.algorithms.types", "hiten.algorithms.dynamics",
              "hiten.algorithms.integrators", "hiten.algorithms.integrators.coefficients",
              "hiten.algorithms.poincare", "hiten.algorithms.poincare.core",
              "hiten.algorithms.poincare.singlehit"):
        _package(p)
    _stand_ins()
    m = {}
    m["config"] = _load("hiten.algorithms.utils.config", f"{REF}/utils/config.py")
    m["serialization"] = _load("hiten.algorithms.types.serialization",
                               f"{REF}/types/serialization.py")
    m["configs"] = _load("hiten.algorithms.types.configs", f"{REF}/types/configs.py")
    m["options"] = _load("hiten.algorithms.types.options", f"{REF}/types/options.py")
    m["protocols"] = _load("hiten.algorithms.dynamics.protocols", f"{REF}/dynamics/protocols.py")
    m["itypes"] = _load("hiten.algorithms.integrators.types", f"{REF}/integrators/types.py")
    m["ibase"] = _load("hiten.algorithms.integrators.base", f"{REF}/integrators/base.py")
    for c in ("dop853", "rk4", "rk6", "rk8", "rk45"):
        m[c] = _load(f"hiten.algorithms.integrators.coefficients.{c}",
                     f"{REF}/integrators/coefficients/{c}.py")
    m["iutils"] = _load("hiten.algorithms.integrators.utils", f"{REF}/integrators/utils.py")
    m["dbase"] = _load("hiten.algorithms.dynamics.base", f"{REF}/dynamics/base.py")
    m["rk"] = _load("hiten.algorithms.integrators.rk", f"{REF}/integrators/rk.py")
    m["rtbp"] = _load("hiten.algorithms.dynamics.rtbp", f"{REF}/dynamics/rtbp.py")
    if singlehit:
        m["events"] = _load("hiten.algorithms.poincare.core.events",
                            f"{REF}/poincare/core/events.py")
        m["ptypes"] = _load("hiten.algorithms.poincare.core.types",
                            f"{REF}/poincare/core/types.py")
        m["singlehit"] = _load("hiten.algorithms.poincare.singlehit.backend",
                               f"{REF}/poincare/singlehit/backend.py")
    return m


def load_continuation():
    """load_corrector() plus the predictor-corrector continuation
    (continuation/types.py, stepping/*.py, backends/base.py, backends/pc.py)."""
    m = load_corrector()
    for p in ("hiten.algorithms.continuation", "hiten.algorithms.continuation.stepping",
              "hiten.algorithms.continuation.stepping.np",
              "hiten.algorithms.continuation.stepping.sc",
              "hiten.algorithms.continuation.backends"):
        _package(p)
    K = f"{REF}/continuation"
    m["ktypes"] = _load("hiten.algorithms.continuation.types", f"{K}/types.py")
    m["support"] = _load("hiten.algorithms.continuation.stepping.support",
                         f"{K}/stepping/support.py")
    m["kbase"] = _load("hiten.algorithms.continuation.stepping.base", f"{K}/stepping/base.py")
    m["kplain"] = _load("hiten.algorithms.continuation.stepping.plain", f"{K}/stepping/plain.py")
    m["np_step"] = _load("hiten.algorithms.continuation.stepping.np.base",
                         f"{K}/stepping/np/base.py")
    m["sc_step"] = _load("hiten.algorithms.continuation.stepping.sc.base",
                         f"{K}/stepping/sc/base.py")
    m["kstepping"] = _load("hiten.algorithms.continuation.stepping",
                           f"{K}/stepping/__init__.py")
    m["kbackend"] = _load("hiten.algorithms.continuation.backends.base", f"{K}/backends/base.py")
    m["pc"] = _load("hiten.algorithms.continuation.backends.pc", f"{K}/backends/pc.py")
    return m

