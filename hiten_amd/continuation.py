"""Batched predictor-corrector continuation of periodic-orbit families.

Restates the reference's _PredictorCorrectorContinuationBackend.run
(continuation/backends/pc.py:42-166) with its steppers -- natural parameter
(stepping/np/base.py:13-30) and secant (stepping/sc/base.py:14-48, tangent
kept as stepping/support.py:45-74 keeps it, seeded as
stepping/__init__.py:63-103) -- the predictor and parameter getter of the
orbit interface (continuation/interfaces.py:183-194, config.py:21-27) and
its corrector call (interfaces.py:135-175: correct the predicted orbit,
residual = ||x_corrected - prediction||, aux period = 2 * half period).

What is batched:

* **the retries of a step.**  After a rejected correction the reference
  shrinks the step (on_reject) and corrects the new prediction; the chain
  of predictions depends only on the last member, the step and the stepper's
  state, so ``window`` of them are corrected at once and the first that
  converges, in order, is taken.  Members, step sizes and the counters
  (iterations, rejected_count) are the sequential loop's: the extra
  corrections of a window are never looked at.
* **families.**  ``continue_families`` advances several independent
  families in lock-step: every round corrects the windows of all families
  that are still growing in one launch.

Corrections run on the device-resident corrector (``correct_device``: one
launch per round) or the host-driven bank (``correct_batch``), as chosen.
The reference's target range does not stop its loop (pc.py:122-126 breaks
only the retry loop, as acceptance does), so it is not an input here.
"""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import Callable, List, Optional, Sequence

import numpy as np

from . import corrector as _corr


@dataclass
class ContinuationResult:
    """ContinuationBackendResponse.family_repr and .info of pc.py."""
    family: List[np.ndarray]
    accepted_count: int
    rejected_count: int
    iterations: int
    parameter_values: tuple
    aux: tuple
    final_step: np.ndarray
    residual_norm: float
    corrections: int = 0          # corrections run, the speculative ones included

    @property
    def success_rate(self) -> float:
        """interfaces.py:101-103"""
        return self.accepted_count / max(self.accepted_count + self.rejected_count, 1)


def _clamp(vec, step_min, step_max):
    """_ContinuationStepBase._clamp_step (stepping/base.py:66-70)."""
    mag = np.clip(np.abs(vec), step_min, step_max)
    return np.sign(vec) * mag


class _Family:
    """One family's loop state between rounds."""

    def __init__(self, seed, make_problem, *, step, max_members, max_retries_per_step, stepper,
                 state_indices, shrink_policy, step_min, step_max):
        self.make_problem = make_problem
        self.max_members = int(max_members)
        self.max_retries = int(max_retries_per_step)
        self.stepper = stepper
        self.idx = None if state_indices is None else np.asarray(state_indices, dtype=int)
        self.shrink_policy = shrink_policy
        self.step_min, self.step_max = step_min, step_max
        seed = np.asarray(seed, dtype=float)
        self.family = [seed.copy()]
        self.params = [self.getter(seed).copy()]
        self.aux: list = []
        self.accepted, self.rejected, self.iterations = 1, 0, 0
        self.step_vec = np.asarray(step, dtype=float).copy()
        self.last_residual = float("nan")
        self.failed = False
        self.attempt = 0
        self.corrections = 0
        self.tangent = None
        if stepper == "secant":
            try:
                diff0 = (self.predictor(seed, np.asarray(step, dtype=float)) - seed).ravel()
                n0 = float(np.linalg.norm(diff0))
                self.tangent = None if n0 == 0.0 else diff0 / n0
            except Exception:
                self.tangent = None

    # -- the reference's helpers -------------------------------------------
    def getter(self, vec):
        v = np.asarray(vec, dtype=float)
        return v if self.idx is None else v[self.idx]

    def predictor(self, last, step):
        if self.idx is None:
            return np.asarray(last, dtype=float) + np.asarray(step, dtype=float)
        last = np.asarray(last, dtype=float).copy()
        for i, d in zip(self.idx, np.asarray(step, dtype=float)):
            last[i] += d
        return last

    def predict(self, last, step):
        if self.stepper != "secant":
            return np.asarray(self.predictor(last, step), dtype=float)
        r_last = np.asarray(last, dtype=float)
        ds = float(step) if np.ndim(step) == 0 else float(np.linalg.norm(step))
        if self.tangent is None:
            n = r_last.copy()
            if n.size > 0:
                n[0] = n[0] + ds
            return n
        return r_last + np.asarray(self.tangent, dtype=float).reshape(r_last.shape) * ds

    def on_reject(self, step):
        if self.shrink_policy is not None:
            try:
                new = self.shrink_policy(step)
            except Exception:
                new = step * 0.5
        else:
            new = step * 0.5
        return _clamp(new, self.step_min, self.step_max)

    # -- rounds --------------------------------------------------------------
    @property
    def growing(self) -> bool:
        return self.accepted < self.max_members and not self.failed

    def window(self, w: int):
        """The next w predictions of the retry chain (fewer if the retries
        run out first): [(step used, prediction, step after its rejection)].
        The shrink policy is applied once per entry here, speculatively, and
        consume() takes the steps from this chain, so the steps recorded are
        the ones the corrections used (a policy must therefore depend only on
        the step it is given, as the reference's default halving does)."""
        out, stp = [], self.step_vec
        last = self.family[-1]
        for k in range(w):
            if self.attempt + k > self.max_retries:
                break
            nxt = self.on_reject(stp)
            out.append((np.asarray(stp, dtype=float), self.predict(last, stp), nxt))
            stp = nxt
        return out

    @staticmethod
    def decided(results) -> int:
        """How many leading results of a window fix its outcome, or 0 while
        that is open: up to the first converged one, else all of them
        (``None`` marks a result still outstanding)."""
        for k, r in enumerate(results):
            if r is None:
                return 0
            if r[2]:
                return k + 1
        return len(results)

    def consume(self, window, results) -> None:
        """Replay the sequential loop over a window's corrections."""
        for (stp, prediction, nxt), res in zip(window, results):
            self.iterations += 1
            corrected, res_norm, converged, aux = res
            if converged:
                self.family.append(corrected)
                self.params.append(self.getter(corrected).copy())
                self.accepted += 1
                self.aux.append(dict(aux))
                if self.stepper == "secant":
                    diff = (np.asarray(self.family[-1], dtype=float)
                            - np.asarray(self.family[-2], dtype=float)).ravel()
                    nrm = float(np.linalg.norm(diff))
                    self.tangent = None if nrm == 0.0 else diff / nrm
                self.step_vec = _clamp(stp, self.step_min, self.step_max)
                self.last_residual = float(res_norm)
                self.attempt = 0
                return
            self.rejected += 1
            self.attempt += 1
            self.step_vec = nxt
            if self.attempt > self.max_retries:
                self.failed = True
                return

    def result(self) -> ContinuationResult:
        lr = self.last_residual
        return ContinuationResult(
            family=self.family, accepted_count=self.accepted, rejected_count=self.rejected,
            iterations=self.iterations, parameter_values=tuple(self.params),
            aux=tuple(self.aux), final_step=np.asarray(self.step_vec, dtype=float).copy(),
            residual_norm=lr if np.isfinite(lr) else float("nan"),
            corrections=self.corrections)


def result_tuple(r: "_corr.CorrectionResult", prediction) -> tuple:
    """(corrected, residual, converged, aux) of one correction, as the
    interface's _correct returns them (interfaces.py:135-175)."""
    x = np.asarray(r.x_corrected, dtype=float)
    return (x, float(np.linalg.norm(x - prediction)), bool(r.converged),
            {"period": float(2.0 * r.half_period)})


#: what a prediction whose correction raised counts as (pc.py:85-90)
FAILED = (None, float("nan"), False, {})


def make_problems(make_problem, predictions) -> list:
    """The correction problem of each prediction, or None where building it
    raised (the reference's corrector call raising, pc.py:85-90)."""
    out = []
    for pred in predictions:
        try:
            out.append(make_problem(pred))
        except Exception:
            out.append(None)
    return out


def _correct_all(problems, predictions, corrector):
    """result_tuple per prediction (FAILED where its problem is None).  A
    callable ``corrector`` is called on the list of predictions instead."""
    if callable(corrector):
        return list(corrector(predictions))
    if corrector not in ("device", "bank"):
        raise ValueError("corrector must be 'device' or 'bank'")
    live = [p for p in problems if p is not None]
    res = iter(_corr.correct_device(live) if corrector == "device"
               else _corr.correct_batch(live))
    return [FAILED if p is None else result_tuple(next(res), pred)
            for p, pred in zip(problems, predictions)]


def continue_families(seeds: Sequence, make_problem: Callable[[np.ndarray], "_corr.CorrectionProblem"],
                      *, step, max_members: int = 100, max_retries_per_step: int = 50,
                      stepper: str = "natural", state_indices=None,
                      shrink_policy: Optional[Callable] = None, step_min: float = 1e-10,
                      step_max: float = 1.0, window: int = 4,
                      corrector="device") -> List[ContinuationResult]:
    """Continue every seed's family; each round corrects the next ``window``
    predictions of every growing family in one launch.  ``make_problem``
    turns a prediction (a 6-d state) into the correction problem of the
    family's orbit kind (e.g. ``lambda x: corrector.halo_problem(mu, x)``);
    ``corrector`` is "device", "bank", or a callable mapping a list of
    predictions to (corrected, residual, converged, aux) tuples."""
    if stepper not in ("natural", "secant"):
        raise ValueError("stepper must be 'natural' or 'secant'")
    if window < 1:
        raise ValueError("window must be >= 1")
    fams = [_Family(s, make_problem, step=step, max_members=max_members,
                    max_retries_per_step=max_retries_per_step, stepper=stepper,
                    state_indices=state_indices, shrink_policy=shrink_policy,
                    step_min=step_min, step_max=step_max) for s in seeds]
    while True:
        live = [f for f in fams if f.growing]
        if not live:
            break
        wins = [f.window(window) for f in live]
        preds = [p for w in wins for _, p, _ in w]
        problems = ([] if callable(corrector) else
                    [q for f, w in zip(live, wins) for q in make_problems(f.make_problem,
                                                                         [p for _, p, _ in w])])
        results = _correct_all(problems, preds, corrector)
        o = 0
        for f, w in zip(live, wins):
            f.corrections += len(w)
            f.consume(w, results[o:o + len(w)])
            o += len(w)
    return [f.result() for f in fams]


def continue_family(seed, make_problem, **kw) -> ContinuationResult:
    """One family (pc.py's run)."""
    return continue_families([seed], make_problem, **kw)[0]
