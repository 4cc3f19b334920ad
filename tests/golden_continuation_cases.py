"""Continuation cases shared by tests/test_continuation_cpu.py, the script
that writes tests/golden/continuation_golden.npz and the tests that read it:
a deterministic stand-in corrector with the cases run against it, and Halo
families grown with a real corrector."""
import numpy as np


def mock(pred):
    """Converges unless the first coordinate's 4th decimal digit is 7..9;
    raises (a failed correction) when it is 6; the corrected state moves."""
    pred = np.asarray(pred, dtype=float)
    digit = int(abs(pred[0]) * 1e4) % 10
    if digit == 6:
        raise RuntimeError("no crossing")
    corrected = pred + 1e-4 * np.sin(np.arange(6) + pred[0])
    return corrected, float(np.linalg.norm(corrected - pred)), digit < 7, {"period": 3.0 + pred[0]}


def _shrink_03(s):
    return s * 0.3


MOCK_SEEDS = [np.array([0.8234, 0.0, 0.0224, 0.0, 0.1343, 0.0]),
              np.array([0.84, 0.0, 0.1, 0.0, 0.22, 0.0]),
              np.array([0.81, 0.0, -0.03, 0.0, 0.2, 0.0])]
MOCK_CASES = [
    dict(step=np.array([1e-3]), state_indices=[0], stepper="natural", max_retries_per_step=6),
    dict(step=np.array([3.3e-3, 1e-3]), state_indices=[0, 2], stepper="natural",
         shrink_policy=_shrink_03, max_retries_per_step=6),
    dict(step=np.array([2e-3]), state_indices=[2], stepper="secant", max_retries_per_step=6),
    dict(step=np.full(6, 1e-3), state_indices=None, stepper="secant", step_min=5e-4,
         max_retries_per_step=6),
    dict(step=np.array([7e-3]), state_indices=[0], stepper="natural", max_retries_per_step=1),
]
HALO_CASES = [
    dict(step=np.array([2e-3]), state_indices=[2], stepper="natural", max_members=4,
         max_retries_per_step=4),
    dict(step=np.array([4e-3]), state_indices=[2], stepper="secant", max_members=4,
         max_retries_per_step=4),
    dict(step=np.array([0.08]), state_indices=[2], stepper="natural", max_members=3,
         max_retries_per_step=5),
]
