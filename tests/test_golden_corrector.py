"""Orbit correction against golden vectors written by the reference's own
corrector (tests/golden/corrector_golden.npz: _NewtonBackend with its Armijo
line search on its single-shooting operators, _y_plane_crossing /
_z_plane_crossing, _compute_stm and the Halo term; see
tests/golden/make_corrector_golden.py).

* In the reference's Python arithmetic (the oracle's libm_pow=2) the oracle's
  sequential loop and the product's batched Newton/Armijo state machine
  (correct_batch) reproduce every outcome bit for bit: convergence, the
  corrected state, residual norm, half period, iteration count, error text
  and the number of event searches (residuals, Jacobians and the half
  period's search).
* In numba's arithmetic (libm_pow=0, what the HIP kernels compute) the
  outcomes are the same and converged orbits agree to 1e-10 (the iterates
  move by rounding; one case takes 9 Newton iterations instead of 11).
"""
import os
import re

import numpy as np
import pytest

from hiten_amd import corrector as Cr
from tests.golden_corrector_cases import problems

G = dict(np.load(os.path.join(os.path.dirname(__file__), "golden", "corrector_golden.npz")))
PS = problems()
TOL = 1e-10


def _ref(i):
    r = G["rows"][i]
    err = str(G["errors"][i])
    it = int(r[2])
    if not r[1]:
        m = re.search(r"after (\d+) iterations", err) or re.search(r"at iter (\d+)", err)
        it = int(m.group(1))
    return dict(converged=bool(r[1]), iterations=it, residual_norm=float(r[3]),
                half_period=float(r[4]), events=int(r[5]), x=G["x"][i],
                error=err.split(": ", 1)[1] if err else "")


def _kind(err):
    return err.split(" ")[0]


def test_fixture_covers_the_outcomes():
    refs = [_ref(i) for i in range(len(PS))]
    kinds = {(r["converged"], _kind(r["error"])) for r in refs}
    assert {(True, ""), (False, "Newton"), (False, "Step")} <= kinds
    assert sum(r["converged"] for r in refs) >= 8
    assert any(p.event_index == 2 for p in PS) and any(p.extra_jacobian is None for p in PS)


def _check_bitwise(got, ref):
    assert got["converged"] == ref["converged"]
    assert got["iterations"] == ref["iterations"]
    assert got["error"] == ref["error"]
    assert got["events"] == ref["events"]
    if ref["converged"]:
        assert np.array_equal(got["x"], ref["x"])
        assert got["residual_norm"] == ref["residual_norm"]
        assert got["half_period"] == ref["half_period"]


def test_oracle_loop_bitwise_in_reference_arithmetic(oracle):
    eng = oracle.CorrectorEngine(libm_pow=2)
    for i, p in enumerate(PS):
        w = oracle.correct_reference(p, engine=eng)
        got = dict(w, events=w["n_event_searches"] + w["n_stm"] + int(w["converged"]))
        _check_bitwise(got, _ref(i))


def test_batched_state_machine_bitwise_in_reference_arithmetic(oracle):
    res = Cr.correct_batch(PS, engine=oracle.CorrectorEngine(libm_pow=2))
    for i, g in enumerate(res):
        got = dict(converged=g.converged, iterations=g.iterations, error=g.error,
                   events=g.n_event_searches + g.n_stm + int(g.converged), x=g.x_corrected,
                   residual_norm=g.residual_norm, half_period=g.half_period)
        _check_bitwise(got, _ref(i))


def test_numba_arithmetic_same_outcomes(oracle):
    eng = oracle.CorrectorEngine(libm_pow=0)
    for i, p in enumerate(PS):
        w, ref = oracle.correct_reference(p, engine=eng), _ref(i)
        assert w["converged"] == ref["converged"]
        assert _kind(w["error"]) == _kind(ref["error"])
        if ref["converged"]:
            assert np.abs(w["x"] - ref["x"]).max() <= TOL
            assert abs(w["half_period"] - ref["half_period"]) <= TOL
            assert w["residual_norm"] < 1e-12
