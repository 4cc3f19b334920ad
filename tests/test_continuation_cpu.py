"""CPU: the batched continuation (speculative retry windows, families in
lock-step) against the reference's sequential predictor-corrector loop
restated in the oracle (continuation/backends/pc.py:42-166), with a
deterministic stand-in corrector that rejects some predictions: every
member, counter, parameter value and final step must be identical."""
import numpy as np
import pytest

from hiten_amd import continuation as Ct
from oracle import continuation as OC
from tests import golden_continuation_cases as GC


_mock = GC.mock


def _batched_mock(preds):
    out = []
    for p in preds:
        try:
            out.append(_mock(p))
        except RuntimeError:
            out.append((np.asarray(p, dtype=float), float("nan"), False, {}))
    return out


SEEDS = GC.MOCK_SEEDS
CASES = GC.MOCK_CASES


@pytest.mark.parametrize("case", range(len(CASES)))
@pytest.mark.parametrize("window", [1, 3, 8])
def test_batched_continuation_equals_the_sequential_loop(case, window):
    kw = dict(CASES[case])
    kw.setdefault("max_retries_per_step", 6)
    got = Ct.continue_families(SEEDS, None, max_members=12, window=window,
                               corrector=_batched_mock, **kw)
    for seed, g in zip(SEEDS, got):
        fam, info = OC.run(seed, _mock, max_members=12, **kw)
        assert len(g.family) == len(fam)
        for a, b in zip(g.family, fam):
            assert np.array_equal(a, b)
        assert g.accepted_count == info["accepted_count"]
        assert g.rejected_count == info["rejected_count"]
        assert g.iterations == info["iterations"]
        assert all(np.array_equal(a, b) for a, b in zip(g.parameter_values,
                                                        info["parameter_values"]))
        assert g.aux == info["aux"]
        assert np.array_equal(g.final_step, info["final_step"])
        assert (g.residual_norm == info["residual_norm"]) or (np.isnan(g.residual_norm)
                                                              and np.isnan(info["residual_norm"]))
        assert g.corrections >= g.iterations


def test_some_families_fail_and_others_grow():
    kw = dict(CASES[4], max_retries_per_step=1)
    got = Ct.continue_families(SEEDS, None, max_members=30, window=2,
                               corrector=_batched_mock, **kw)
    assert any(g.rejected_count > 0 for g in got)
    assert any(g.accepted_count == 30 for g in got) or any(g.accepted_count < 30 for g in got)
    for seed, g in zip(SEEDS, got):
        _, info = OC.run(seed, _mock, max_members=30, **kw)
        assert (g.accepted_count, g.rejected_count) == (info["accepted_count"],
                                                        info["rejected_count"])
