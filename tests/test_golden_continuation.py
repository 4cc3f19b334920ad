"""Continuation against golden vectors written by the reference's own
predictor-corrector loop (tests/golden/continuation_golden.npz:
_PredictorCorrectorContinuationBackend.run with its natural-parameter and
secant steppers; see tests/golden/make_continuation_golden.py).

* Stand-in corrector families: the oracle's restated loop and the product's
  batched continuation (speculative windows of 1, 3 and 8) reproduce every
  member, parameter value, counter, final step, residual and aux period bit
  for bit.
* Halo families grown with the reference's own corrector: the oracle's loop
  on the oracle's corrector in the reference's arithmetic (libm_pow=2), and
  the product's batched continuation over the product's batched corrector on
  the same arithmetic, reproduce them bit for bit.
"""
import os

import numpy as np
import pytest

from hiten_amd import continuation as Ct, corrector as Cr
from oracle import continuation as OC
from tests import golden_continuation_cases as GC

G = dict(np.load(os.path.join(os.path.dirname(__file__), "golden", "continuation_golden.npz")))
MU = 0.012150585609624
FAMS = [str(f) for f in G["families"]]


def _fam(tag):
    k = FAMS.index(tag)
    return dict(family=G[f"f{k}_family"], params=G[f"f{k}_params"],
                counts=tuple(int(c) for c in G[f"f{k}_counts"]),
                final_step=G[f"f{k}_final_step"], residual=float(G[f"f{k}_residual"]),
                periods=G[f"f{k}_periods"])


def _same(fam, info, ref):
    assert len(fam) == len(ref["family"])
    for a, b in zip(fam, ref["family"]):
        assert np.array_equal(np.asarray(a, float), b)
    assert (info["accepted_count"], info["rejected_count"], info["iterations"]) == ref["counts"]
    assert np.array_equal(np.asarray(info["parameter_values"], float), ref["params"])
    assert np.array_equal(np.asarray(info["final_step"], float), ref["final_step"])
    r = info["residual_norm"]
    assert r == ref["residual"] or (np.isnan(r) and np.isnan(ref["residual"]))
    per = np.array([a.get("period", np.nan) for a in info["aux"]], float)
    assert np.array_equal(per, ref["periods"], equal_nan=True)


def _info_of(r):
    return dict(accepted_count=r.accepted_count, rejected_count=r.rejected_count,
                iterations=r.iterations, parameter_values=r.parameter_values,
                final_step=r.final_step, residual_norm=r.residual_norm, aux=r.aux)


def _batched_mock(preds):
    out = []
    for p in preds:
        try:
            out.append(GC.mock(p))
        except RuntimeError:
            out.append((np.asarray(p, dtype=float), float("nan"), False, {}))
    return out


def test_fixture_covers_rejections_and_failures():
    counts = [_fam(t)["counts"] for t in FAMS]
    assert any(c[1] > 0 for c in counts) and any(c[0] < 12 for c in counts[:15])
    assert sum(t.startswith("halo") for t in FAMS) == 6
    assert any(_fam(t)["counts"][1] > 0 for t in FAMS if t.startswith("halo"))


@pytest.mark.parametrize("ci", range(len(GC.MOCK_CASES)))
def test_oracle_loop_reproduces_the_reference_loop(ci):
    for si, seed in enumerate(GC.MOCK_SEEDS):
        fam, info = OC.run(seed, GC.mock, max_members=12, **GC.MOCK_CASES[ci])
        _same(fam, info, _fam(f"mock case {ci} seed {si}"))


@pytest.mark.parametrize("window", [1, 3, 8])
@pytest.mark.parametrize("ci", range(len(GC.MOCK_CASES)))
def test_batched_continuation_reproduces_the_reference_loop(ci, window):
    got = Ct.continue_families(GC.MOCK_SEEDS, None, max_members=12, window=window,
                               corrector=_batched_mock, **GC.MOCK_CASES[ci])
    for si, g in enumerate(got):
        _same(g.family, _info_of(g), _fam(f"mock case {ci} seed {si}"))


def _problem(x):
    return Cr.halo_problem(MU, x, max_attempts=20)


def test_halo_families_in_the_reference_arithmetic(oracle):
    eng = oracle.CorrectorEngine(libm_pow=2)
    seeds = G["halo_seeds"]
    for guess, seed in zip(([0.8234, 0, 0.0224, 0, 0.1343, 0], [0.84, 0, 0.1, 0, 0.22, 0]),
                           seeds):
        w = oracle.correct_reference(_problem(np.array(guess, float)), engine=eng)
        assert w["converged"] and np.array_equal(w["x"], seed)

    def single(pred):
        w = oracle.correct_reference(_problem(pred), engine=eng)
        if not w["converged"]:
            raise RuntimeError(w["error"])
        x = np.asarray(w["x"], float)
        return x, float(np.linalg.norm(x - pred)), True, {"period": 2.0 * w["half_period"]}

    def batched(preds):
        res = Cr.correct_batch([_problem(p) for p in preds], engine=eng)
        return [Ct.result_tuple(r, p) if r.converged else Ct.FAILED
                for r, p in zip(res, preds)]

    for ci, case in enumerate(GC.HALO_CASES):
        kw = {k: v for k, v in case.items() if k != "max_members"}
        got = Ct.continue_families(list(seeds), None, max_members=case["max_members"],
                                   window=2, corrector=batched, **kw)
        for si, seed in enumerate(seeds):
            ref = _fam(f"halo case {ci} seed {si}")
            fam, info = OC.run(seed, single, max_members=case["max_members"], **kw)
            _same(fam, info, ref)
            _same(got[si].family, _info_of(got[si]), ref)
