"""Golden vectors for predictor-corrector continuation, written by the
reference's own loop.

Runs hiten 0.5.4's _PredictorCorrectorContinuationBackend.run
(continuation/backends/pc.py:42-166) with its natural-parameter and secant
steppers and supports (stepping/__init__.py, stepping/np, stepping/sc,
stepping/support.py), chosen as ContinuationPipeline.with_default_engine
chooses them (continuation/base.py:71-81), on requests built as the orbit
continuation interface builds them (continuation/interfaces.py:63-85; its
predictor, interfaces.py:181-194, and the parameter getter of
continuation/config.py:16-21 compiled from those files' own text).  Two
kinds of families:

* ``mock``: the stand-in corrector of tests/test_continuation_cpu.py (it
  converges, rejects or raises on a fixed pattern of the prediction), every
  case of that test on its three seeds;
* ``halo``: Halo families grown from corrected orbits with the reference's
  own corrector (the Newton/Armijo backend on its operators, as in
  make_corrector_golden.py), natural parameter and secant, wrapped as
  interfaces.py:135-168 wraps orbit.correct.

Stored per family: every member, the parameter values, the accepted and
rejected counts, the iterations, the final step, the last residual and the
aux periods.

    python tests/golden/make_continuation_golden.py
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

from tests.golden.reference_loader import (REF, load_continuation,  # noqa: E402
                                           reference_function)
from tests.golden_continuation_cases import HALO_CASES, MOCK_CASES, MOCK_SEEDS, mock  # noqa: E402

OUT = os.path.join(HERE, "continuation_golden.npz")
MU = 0.012150585609624


def main():
    t_start = time.time()
    m = load_continuation()
    ks = m["kstepping"]
    support = m["support"]
    predictor_of = reference_function(f"{REF}/continuation/interfaces.py",
                                      "_OrbitContinuationInterface._predictor_from_problem",
                                      {"np": np, "Callable": object, "_ContinuationProblem": object})
    getter_of = reference_function(f"{REF}/continuation/config.py",
                                   "_make_orbit_parameter_getter",
                                   {"np": np, "Optional": object, "Tuple": object,
                                    "Callable": object})

    def run(seed, corrector, kw):
        stepper = kw.get("stepper", "natural")
        if stepper == "secant":
            backend = m["pc"]._PredictorCorrectorContinuationBackend(
                stepper_factory=ks.make_secant_stepper(),
                support_factory=support._VectorSpaceSecantSupport)
        else:
            backend = m["pc"]._PredictorCorrectorContinuationBackend(
                stepper_factory=ks.make_natural_stepper(),
                support_factory=support._NullStepSupport)
        idx = kw.get("state_indices")
        idx = None if idx is None else np.asarray(idx, dtype=int)
        predictor = predictor_of(None, types.SimpleNamespace(state_indices=idx))
        representation = lambda v: np.asarray(v, dtype=float)  # noqa: E731
        request = m["ktypes"].ContinuationBackendRequest(
            seed_repr=np.asarray(seed, dtype=float),
            stepper_fn=representation if stepper == "secant" else predictor,
            predictor_fn=predictor,
            parameter_getter=getter_of(None if idx is None else tuple(idx)),
            corrector=corrector,
            step=np.asarray(kw["step"], dtype=float),
            target=np.asarray(kw.get("target", (-np.inf, np.inf)), dtype=float),
            max_members=int(kw["max_members"]),
            max_retries_per_step=int(kw.get("max_retries_per_step", 50)),
            shrink_policy=kw.get("shrink_policy"),
            step_min=float(kw.get("step_min", 1e-10)),
            step_max=float(kw.get("step_max", 1.0)),
            metadata={})
        resp = backend.run(request=request)
        return resp.family_repr, resp.info

    A = {}
    fams = []

    def store(tag, fam, info):
        k = len(fams)
        fams.append(tag)
        A[f"f{k}_family"] = np.asarray(fam, dtype=float)
        A[f"f{k}_params"] = np.asarray(info["parameter_values"], dtype=float)
        A[f"f{k}_counts"] = np.array([info["accepted_count"], info["rejected_count"],
                                      info["iterations"]], np.int64)
        A[f"f{k}_final_step"] = np.asarray(info["final_step"], dtype=float)
        A[f"f{k}_residual"] = np.float64(info["residual_norm"])
        A[f"f{k}_periods"] = np.asarray([a.get("period", np.nan) for a in info["aux"]], float)

    for ci, case in enumerate(MOCK_CASES):
        for si, seed in enumerate(MOCK_SEEDS):
            fam, info = run(seed, mock, dict(case, max_members=12))
            store(f"mock case {ci} seed {si}", fam, info)

    # Halo families with the reference's own corrector
    mc = load_continuation()
    rtbp, sh, ops_mod = mc["rtbp"], mc["singlehit"], mc["operators"]
    SS = types.SimpleNamespace(X=0, Y=1, Z=2, VX=3, VY=4, VZ=5)
    halo = reference_function(f"{REF}/types/services/orbits.py",
                              "_HaloOrbitCorrectionService._halo_quadratic_term",
                              {"np": np, "SynodicState": SS})

    def reference_correct(x0, max_attempts):
        dom = types.SimpleNamespace(
            dynamics=types.SimpleNamespace(dynsys=rtbp.rtbp_dynsys(MU),
                                           var_dynsys=rtbp.variational_dynsys(MU)),
            initial_state=np.asarray(x0, float).copy(), mu=MU)
        svc = types.SimpleNamespace(domain_obj=dom)
        ops = ops_mod._SingleShootingOrbitOperators(
            domain_obj=dom, control_indices=(0, 4), residual_indices=(3, 5), target=(0.0, 0.0),
            extra_jacobian=lambda X, Phi: halo(svc, X, Phi), event_func=sh._y_plane_crossing,
            forward=1, method="adaptive", order=8, steps=500)
        req = mc["ctypes"].CorrectorInput(
            initial_guess=np.asarray(x0, float)[[0, 4]].copy(),
            residual_fn=ops.build_residual_fn(), jacobian_fn=ops.build_jacobian_fn(),
            norm_fn=lambda r: float(np.linalg.norm(r, ord=np.inf)), max_attempts=max_attempts,
            tol=1e-12, max_delta=1e-2, fd_step=1e-8)
        be = mc["newton"]._NewtonBackend(stepper_factory=mc["stepping"].make_armijo_stepper())
        out = be.run(request=req)          # raises when the correction fails
        x = np.asarray(x0, float).copy()
        x[[0, 4]] = out.x_corrected
        t_half, _ = sh._y_plane_crossing(dynsys=dom.dynamics.dynsys, x0=x, forward=1)
        return x, float(t_half)

    def halo_corrector(prediction):
        x, t_half = reference_correct(prediction, 20)
        return x, float(np.linalg.norm(x - prediction)), True, {"period": 2.0 * t_half}

    seeds = [reference_correct(g, 20)[0] for g in ([0.8234, 0, 0.0224, 0, 0.1343, 0],
                                                     [0.84, 0, 0.1, 0, 0.22, 0])]
    A["halo_seeds"] = np.asarray(seeds)
    for ci, case in enumerate(HALO_CASES):
        for si, seed in enumerate(seeds):
            fam, info = run(seed, halo_corrector, dict(case))
            store(f"halo case {ci} seed {si}", fam, info)
            print(f"  halo case {ci} seed {si}: {info['accepted_count']} members, "
                  f"{info['rejected_count']} rejected", flush=True)
    A["families"] = np.asarray(fams)
    np.savez_compressed(OUT, **A)
    print(f"{OUT}: {len(fams)} families ({time.time() - t_start:.1f} s)")


if __name__ == "__main__":
    main()
