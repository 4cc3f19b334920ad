"""Golden vectors for section detection on stored trajectories.

Runs the reference's own detection code -- hiten 0.5.4,
src/hiten/algorithms/poincare/synodic/backend.py (detect_on_trajectory and
the functions it calls) with poincare/utils.py (_hermite_scalar,
_hermite_der), synodic/events.py (_AffinePlaneEvent), poincare/core/events.py
and poincare/core/types.py -- in this container, on synthetic trajectories,
and stores inputs and hits in synodic_golden.npz next to this script.

Why: the cubic refinement's Newton iteration uses the reference's
_hermite_der, whose h00 and h01 terms are not the derivative of the Hermite
basis, so the refined hits are fixed by the reference's code, not by the
mathematics; these vectors pin them.

How the code is loaded: numba is not installed here, so `numba.njit` is the
identity (the Hermite helpers then run as Python, whose x ** 2 is the C
library's pow where numba multiplies: an ulp apart in ~0.1% of squares).
Two modules that the detection only names in annotations and a base class
(poincare/core/backend.py and synodic/types.py, which pull in the whole
package) are replaced by empty stand-ins.  Everything that computes is the
reference's file.  Trajectories come from the oracle's DOP853 (the C
restatement) on CR3BP states and from closed-form plane values.

    python tests/golden/make_synodic_golden.py
"""
from __future__ import annotations

import importlib.util
import os
import sys
import types

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
REF = "/root/reference/src/hiten/algorithms"


def _package(name):
    mod = types.ModuleType(name)
    mod.__path__ = []
    sys.modules[name] = mod
    return mod


def _load(name, path):
    spec = importlib.util.spec_from_file_location(name, path)
    mod = importlib.util.module_from_spec(spec)
    sys.modules[name] = mod
    spec.loader.exec_module(mod)
    return mod


def load_reference_backend():
    numba = types.ModuleType("numba")

    def njit(*args, **kwargs):
        if args and callable(args[0]) and not kwargs:
            return args[0]
        return lambda f: f
    numba.njit = njit
    sys.modules["numba"] = numba
    for p in ("hiten", "hiten.algorithms", "hiten.algorithms.utils", "hiten.algorithms.poincare",
              "hiten.algorithms.poincare.core", "hiten.algorithms.poincare.synodic"):
        _package(p)
    _load("hiten.algorithms.utils.config", f"{REF}/utils/config.py")
    _load("hiten.algorithms.poincare.utils", f"{REF}/poincare/utils.py")
    _load("hiten.algorithms.poincare.core.events", f"{REF}/poincare/core/events.py")
    _load("hiten.algorithms.poincare.core.types", f"{REF}/poincare/core/types.py")
    stand_in = types.ModuleType("hiten.algorithms.poincare.core.backend")

    class _ReturnMapBackend:
        def __init__(self, *a, **k):
            pass
    stand_in._ReturnMapBackend = _ReturnMapBackend
    sys.modules[stand_in.__name__] = stand_in
    tmod = types.ModuleType("hiten.algorithms.poincare.synodic.types")
    tmod.SynodicBackendRequest = object
    tmod.SynodicBackendResponse = object
    sys.modules[tmod.__name__] = tmod
    _load("hiten.algorithms.poincare.synodic.events", f"{REF}/poincare/synodic/events.py")
    return _load("hiten.algorithms.poincare.synodic.backend", f"{REF}/poincare/synodic/backend.py")


def trajectories():
    """(name, times, states) cases."""
    sys.path.insert(0, ROOT)
    from oracle import oracle as O
    from hiten_amd import workloads as W
    out = []
    mu = 0.012150585609624
    rng = np.random.default_rng(20261203)
    seeds = W.planar_seeds(rng, 4)
    t = np.linspace(0.0, 6.0, 301)
    for i, y0 in enumerate(seeds):
        st, _ = O.dop853(O.System("cr3bp6", mu=mu), y0, t, rtol=1e-12, atol=1e-12,
                         max_step=1e4)
        out.append((f"cr3bp{i}", t, np.asarray(st)))
    # closed-form plane values (y), x = t
    def mk(tt, y, vy=None):
        X = np.zeros((len(tt), 6))
        X[:, 0] = tt
        X[:, 1] = y
        if vy is not None:
            X[:, 4] = vy
        return X
    tq = np.linspace(0.0, 1.0, 21)
    out.append(("quadratic", tq, mk(tq, (tq - 0.37) ** 2 - 0.0121)))
    t6 = np.arange(6.0)
    out.append(("two_roots", t6, mk(t6, [13.0, 13.0, 1.0, 1.0, 13.0, 13.0])))
    out.append(("on_surface", tq, mk(tq, tq - 0.25)))
    ts = np.linspace(0.0, 4.5, 451)
    out.append(("sine", ts, mk(ts, np.sin(2 * np.pi * ts), np.ones_like(ts))))
    return out


sys.path.insert(0, ROOT)
from tests.golden_synodic_options import OPTIONS  # noqa: E402


def main():
    backend = load_reference_backend()._SynodicDetectionBackend()
    arrays = {}
    cases = trajectories()
    for ci, (name, t, X) in enumerate(cases):
        arrays[f"t{ci}"] = np.asarray(t, float)
        arrays[f"x{ci}"] = np.asarray(X, float)
    rows = []
    for oi, opt in enumerate(OPTIONS):
        kw = dict(opt)
        kw.setdefault("plane_coords", ("y", "vy"))
        for ci, (name, t, X) in enumerate(cases):
            hits = backend.detect_on_trajectory(np.asarray(t, float), np.asarray(X, float), **kw)
            for h in hits:
                rows.append([oi, ci, float(h.time), *np.asarray(h.state, float),
                             *np.asarray(h.point2d, float)])
    arrays["hits"] = np.asarray(rows, float).reshape(-1, 11)
    arrays["names"] = np.array([c[0] for c in cases])
    path = os.path.join(HERE, "synodic_golden.npz")
    np.savez_compressed(path, **arrays)
    print(f"{path}: {len(cases)} trajectories x {len(OPTIONS)} option sets, {len(rows)} hits")


if __name__ == "__main__":
    main()
