"""CPU: the serving loop's continuation job (hiten_amd.serving._ContRun)
consumes each window of corrections as the results come back, in any order
and with results of superseded rounds arriving late, and grows the same
family with the same counters as the reference's sequential loop (the
oracle's restatement of continuation/backends/pc.py:42-166) given the same
corrector.  The corrector here is a deterministic stand-in on the CPU whose
rejections follow the prediction, so retries, shrinking steps and retry
exhaustion all occur."""
import math

import numpy as np
import pytest

from hiten_amd import continuation as Ct, corrector as Cr, serving as S
from oracle import continuation as OC


def _stand_in(pred):
    """CorrectionResult of a prediction: rejected where frac(37 z) >= 0.6."""
    z = float(pred[2])
    ok = (37.0 * z) % 1.0 < 0.6
    x = np.asarray(pred, float).copy()
    x[0] += 1e-3 * math.sin(z)
    return Cr.CorrectionResult(bool(ok), x, 1e-13 if ok else 1.0, 3, half_period=1.0 + z)


def _as_tuple(pred):
    return Ct.result_tuple(_stand_in(pred), pred)


class _FakeServer:
    def __init__(self):
        self.queue = []
        self.cont_dirty = []

    def now(self):
        return 0.0

    def _enqueue_correction(self, job):
        self.queue.append(job)


def _serve(req, rng):
    """Drive one _ContRun as the loop does, delivering results in a random
    order, a random number at a time."""
    job = S.Job(0, "continuation", req, 0.0)
    run = S._ContRun(job)
    fake = _FakeServer()
    run.start_round(fake)
    done = False
    while not done:
        if fake.cont_dirty:                   # decided without a launch
            fake.cont_dirty = []
            done = run.advance(fake)
            continue
        assert fake.queue, "a continuation waits for results that never come"
        k = int(rng.integers(1, len(fake.queue) + 1))
        pick = sorted(rng.choice(len(fake.queue), size=k, replace=False), reverse=True)
        got = [fake.queue.pop(int(i)) for i in pick]
        dirty = False
        for sub in got:
            r, rnd, i = sub.parent
            dirty |= r.on_result(rnd, i, _stand_in(sub.request))
        if dirty:
            done = run.advance(fake)
    # corrections of the last round still in flight are dropped
    for sub in fake.queue:
        r, rnd, i = sub.parent
        assert not r.on_result(rnd, i, _stand_in(sub.request))
    return job.result


@pytest.mark.parametrize("stepper,step", [("natural", [0.01]), ("secant", [0.02]),
                                          ("natural", [0.05])])
@pytest.mark.parametrize("window", [1, 2, 5])
def test_out_of_order_results_give_the_sequential_family(stepper, step, window):
    seed = np.array([0.83, 0.0, 0.01, 0.0, 0.15, 0.0])
    kw = dict(step=np.array(step), state_indices=[2], stepper=stepper,
              max_retries_per_step=3)
    fam, info = OC.run(seed, _as_tuple, max_members=7, **kw)
    req = S.ContinuationRequest(seed, lambda x: np.asarray(x, float).copy(), np.array(step),
                                max_members=7, max_retries_per_step=3, stepper=stepper,
                                state_indices=[2], window=window)
    for trial in range(4):
        got = _serve(req, np.random.default_rng(100 * window + trial))
        assert got.accepted_count == info["accepted_count"]
        assert got.rejected_count == info["rejected_count"]
        assert got.iterations == info["iterations"]
        assert len(got.family) == len(fam)
        for a, b in zip(got.family, fam):
            assert np.array_equal(a, b)
        assert got.aux == info["aux"]
        assert np.array_equal(got.final_step, info["final_step"])
        # the batched API gives the same family
        batched = Ct.continue_family(seed, req.make_problem, window=window,
                                     corrector=lambda ps: [_as_tuple(p) for p in ps],
                                     max_members=7, **kw)
        assert batched.iterations == got.iterations
        assert batched.corrections == got.corrections
        for a, b in zip(batched.family, got.family):
            assert np.array_equal(a, b)


def test_exhausted_retries_end_the_family():
    # every prediction past the seed is rejected: 1 + max_retries attempts
    seed = np.array([0.83, 0.0, 0.7 / 37.0, 0.0, 0.15, 0.0])
    req = S.ContinuationRequest(seed, lambda x: np.asarray(x, float).copy(), np.array([0.0]),
                                max_members=4, max_retries_per_step=3, state_indices=[2],
                                window=2)
    got = _serve(req, np.random.default_rng(7))
    assert got.accepted_count == 1 and got.rejected_count == 4 and got.iterations == 4


def test_decided_prefix():
    ok, bad = (None, None, True, None), (None, None, False, None)
    assert Ct._Family.decided([]) == 0
    assert Ct._Family.decided([None, ok]) == 0
    assert Ct._Family.decided([bad, ok, None]) == 2
    assert Ct._Family.decided([ok, None]) == 1
    assert Ct._Family.decided([bad, bad]) == 2
    assert Ct._Family.decided([bad, None]) == 0


def test_stream_with_continuations_keeps_the_two_kind_draws():
    from hiten_amd import workloads as W
    base = W.serving_stream(duration=0.05, rate=2000.0, fan_max=64)
    same = W.serving_stream(duration=0.05, rate=2000.0, fan_max=64, continuation_fraction=0.0)
    assert [r["kind"] for r in base] == [r["kind"] for r in same]
    assert all(a["t"] == b["t"] for a, b in zip(base, same))
    seeds = [np.array([0.83, 0.0, 0.01, 0.0, 0.15, 0.0])]
    mix = W.serving_stream(duration=0.05, rate=2000.0, fan_max=64, correction_fraction=0.5,
                           continuation_fraction=0.3, continuation_seeds=seeds)
    kinds = {r["kind"] for r in mix}
    assert kinds == {"correction", "continuation", "fan"}
    kind, req = S.to_request(next(r for r in mix if r["kind"] == "continuation"))
    assert kind == "continuation" and req.window == 2 and req.max_members == 4
    p = req.make_problem(seeds[0])
    assert p.max_attempts == 10 and p.tol == 1e-12
    with pytest.raises(ValueError):
        W.serving_stream(duration=0.01, continuation_fraction=0.2)


def _raising_make(x):
    """a problem builder that raises in a band of z (where the reference's
    corrector call would raise, pc.py:85-90)"""
    if 0.035 < float(x[2]) < 0.05:
        raise ValueError("no problem here")
    return np.asarray(x, float).copy()


def _raising_corrector(pred):
    _raising_make(pred)
    return _as_tuple(pred)


@pytest.mark.parametrize("window", [1, 3])
def test_failing_problems_count_as_rejections(window):
    seed = np.array([0.83, 0.0, 0.01, 0.0, 0.15, 0.0])
    kw = dict(step=np.array([0.01]), state_indices=[2], stepper="natural",
              max_retries_per_step=3)
    fam, info = OC.run(seed, _raising_corrector, max_members=8, **kw)
    req = S.ContinuationRequest(seed, _raising_make, np.array([0.01]), max_members=8,
                                max_retries_per_step=3, state_indices=[2], window=window)
    got = _serve(req, np.random.default_rng(window))
    assert (got.accepted_count, got.rejected_count, got.iterations) == (
        info["accepted_count"], info["rejected_count"], info["iterations"])
    for a, b in zip(got.family, fam):
        assert np.array_equal(a, b)
    assert info["rejected_count"] > 0
    # the batched API treats them the same way (problems built on the host)
    probs = Ct.make_problems(_raising_make, [np.array([0, 0, 0.04, 0, 0, 0.0]), seed])
    assert probs[0] is None and probs[1] is not None
