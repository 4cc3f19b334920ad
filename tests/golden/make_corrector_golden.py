"""Golden vectors for single-shooting orbit correction, written by the
reference's own code.

Runs hiten 0.5.4's Newton backend with its Armijo line search
(corrector/backends/newton.py:39-143, backends/base.py:142-245,
stepping/armijo.py:58-164, stepping/__init__.py make_armijo_stepper), its
single-shooting operators (corrector/operators.py:260-452) on its own
_y_plane_crossing / _z_plane_crossing and _compute_stm, and the Halo
problem's extra Jacobian term (types/services/orbits.py:917-947, compiled
from that file's own text: the module itself pulls in the whole package),
wired as the orbit-correction interface wires them
(corrector/interfaces.py:44-46 for the infinity norm, 100-177 for the problem
and backend call, 48-73 for the half period), on Halo, Lyapunov and vertical
problems, and stores every outcome in ``corrector_golden.npz``.

Each case records: converged (1/0), the exception's class and text when it
failed (ConvergenceError after max_attempts or when the line search found no
productive step; TypeError when an iterate has no section crossing, which
the reference's residual function then subscripts), the corrected state, the
final residual norm and iteration count, the half period (event search on
the corrected state), and the number of event searches.

    python tests/golden/make_corrector_golden.py
"""
from __future__ import annotations

import os
import sys
import time
import types

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from tests.golden.reference_loader import REF, load_corrector, reference_function  # noqa: E402
from tests.golden_corrector_cases import problems  # noqa: E402

OUT = os.path.join(HERE, "corrector_golden.npz")


def main():
    t_start = time.time()
    m = load_corrector()
    rtbp, sh, ops_mod = m["rtbp"], m["singlehit"], m["operators"]
    SS = types.SimpleNamespace(X=0, Y=1, Z=2, VX=3, VY=4, VZ=5)
    halo = reference_function(f"{REF}/types/services/orbits.py",
                              "_HaloOrbitCorrectionService._halo_quadratic_term",
                              {"np": np, "SynodicState": SS})
    CorrectorInput = m["ctypes"].CorrectorInput
    Newton = m["newton"]._NewtonBackend
    make_armijo = m["stepping"].make_armijo_stepper
    rows = []
    xs = []
    texts = []
    for p in problems():
        mu = float(p.mu)
        dom = types.SimpleNamespace(
            dynamics=types.SimpleNamespace(dynsys=rtbp.rtbp_dynsys(mu),
                                           var_dynsys=rtbp.variational_dynsys(mu)),
            initial_state=np.asarray(p.x0, float).copy(), mu=mu)
        service = types.SimpleNamespace(domain_obj=dom)
        crossing = {1: sh._y_plane_crossing, 2: sh._z_plane_crossing}[int(p.event_index)]
        calls = [0]

        def event_func(*, dynsys, x0, forward=1, t_guess=None, t_window=None, _f=crossing):
            calls[0] += 1
            return _f(dynsys=dynsys, x0=x0, forward=forward, t_guess=t_guess, t_window=t_window)

        extra = ((lambda X, Phi, _s=service: halo(_s, X, Phi))
                 if p.extra_jacobian == "halo" else None)
        ops = ops_mod._SingleShootingOrbitOperators(
            domain_obj=dom, control_indices=tuple(p.control_indices),
            residual_indices=tuple(p.residual_indices), target=tuple(p.target),
            extra_jacobian=extra, event_func=event_func, forward=int(p.forward),
            method="adaptive", order=8, steps=500)
        ctrl = list(p.control_indices)
        request = CorrectorInput(
            initial_guess=np.asarray(p.x0, float)[ctrl].copy(),
            residual_fn=ops.build_residual_fn(), jacobian_fn=ops.build_jacobian_fn(),
            norm_fn=lambda r: float(np.linalg.norm(r, ord=np.inf)),
            max_attempts=int(p.max_attempts), tol=float(p.tol), max_delta=float(p.max_delta),
            fd_step=1e-8)
        backend = Newton(stepper_factory=make_armijo(alpha_reduction=float(p.alpha_reduction),
                                                     min_alpha=float(p.min_alpha),
                                                     armijo_c=float(p.armijo_c)))
        x_full = np.full(6, np.nan)
        conv, it, rn, hp, err = 0, -1, np.nan, np.nan, ""
        try:
            out = backend.run(request=request)
            conv, it, rn = 1, int(out.iterations), float(out.residual_norm)
            x_full = np.asarray(p.x0, float).copy()
            x_full[ctrl] = out.x_corrected
            t_hp, _ = event_func(dynsys=dom.dynamics.dynsys, x0=x_full, forward=int(p.forward))
            hp = float(t_hp)
        except Exception as exc:  # the reference's own failure, recorded as data
            err = f"{type(exc).__name__}: {exc}"
        rows.append([mu, conv, it, rn, hp, calls[0]])
        xs.append(x_full)
        texts.append(err)
        print(f"  case {len(rows) - 1}: converged={conv} it={it} |R|={rn:.2e} events={calls[0]} "
              f"{err[:60]}", flush=True)
    np.savez_compressed(OUT, rows=np.asarray(rows, float), x=np.asarray(xs, float),
                        errors=np.asarray(texts))
    print(f"{OUT}: {len(rows)} corrections ({time.time() - t_start:.1f} s)")


if __name__ == "__main__":
    main()
