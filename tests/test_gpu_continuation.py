"""GPU: batched continuation of Halo families (speculative retry windows,
families in lock-step, one corrector launch per round) against the
reference's sequential predictor-corrector loop in the oracle, each
prediction corrected alone by the same corrector: identical members and
counters, for the device-resident corrector and for the host-driven bank."""
import numpy as np
import pytest

from hiten_amd import continuation as Ct, corrector as Cr, workloads as W
from oracle import continuation as OC

pytestmark = pytest.mark.gpu
MU = W.MU_EM


def say(msg):
    print(f"\n[case] {msg}", flush=True)


def _seeds():
    ps = [Cr.halo_problem(MU, [0.8234, 0, 0.0224, 0, 0.1343, 0], max_attempts=20),
          Cr.halo_problem(MU, [0.84, 0, 0.1, 0, 0.22, 0], max_attempts=20)]
    res = Cr.correct_device(ps)
    assert all(r.converged for r in res)
    return [r.x_corrected for r in res]


def _make(x):
    return Cr.halo_problem(MU, x, max_attempts=20)


def _single(kind):
    def corr(pred):
        f = Cr.correct_device if kind == "device" else Cr.correct_batch
        r = f([_make(pred)])[0]
        x = np.asarray(r.x_corrected, dtype=float)
        return x, float(np.linalg.norm(x - pred)), bool(r.converged), {"period": 2.0 * r.half_period}
    return corr


@pytest.mark.parametrize("kind", ["device", "bank"])
@pytest.mark.parametrize("stepper,step,idx", [("natural", [2e-3], [2]),
                                              ("secant", [4e-3], [2])])
def test_batched_families_equal_the_sequential_loop(gpu, kind, stepper, step, idx):
    seeds = _seeds()
    kw = dict(step=np.array(step), state_indices=idx, stepper=stepper, max_retries_per_step=4)
    say(f"continuation ({kind} corrector, {stepper}): 2 Halo families x 6 members, window 3")
    got = Ct.continue_families(seeds, _make, max_members=6, window=3, corrector=kind, **kw)
    for seed, g in zip(seeds, got):
        fam, info = OC.run(seed, _single(kind), max_members=6, **kw)
        assert g.accepted_count == info["accepted_count"] == 6
        assert g.rejected_count == info["rejected_count"]
        assert g.iterations == info["iterations"]
        for a, b in zip(g.family, fam):
            assert np.array_equal(a, b)
        assert g.aux == info["aux"]
        assert np.array_equal(g.final_step, info["final_step"])
        z = [m[2] for m in g.family]
        assert all(abs(b - a) > 0 for a, b in zip(z, z[1:]))


def test_rejections_shrink_the_step(gpu):
    seeds = _seeds()[:1]
    say("continuation with an oversized step: rejections, shrink, same as the loop")
    kw = dict(step=np.array([0.08]), state_indices=[2], stepper="natural",
              max_retries_per_step=5)
    got = Ct.continue_families(seeds, _make, max_members=3, window=2, corrector="device", **kw)[0]
    fam, info = OC.run(seeds[0], _single("device"), max_members=3, **kw)
    assert (got.accepted_count, got.rejected_count, got.iterations) == (
        info["accepted_count"], info["rejected_count"], info["iterations"])
    for a, b in zip(got.family, fam):
        assert np.array_equal(a, b)


@pytest.mark.parametrize("kind", ["device", "bank"])
def test_halo_families_against_the_reference_loop(gpu, kind):
    """The batched continuation on the GPU correctors against the Halo
    families the reference's own loop and corrector grew
    (tests/golden/continuation_golden.npz): the same accepted, rejected and
    iteration counts, members and aux periods to 1e-10 (the kernels compute
    in numba's arithmetic, the vectors in Python's)."""
    from tests import golden_continuation_cases as GC
    from tests.test_golden_continuation import G, _fam
    seeds = list(G["halo_seeds"])
    for ci, case in enumerate(GC.HALO_CASES):
        kw = {k: v for k, v in case.items() if k != "max_members"}
        say(f"continuation ({kind}) vs the reference loop: case {ci}")
        got = Ct.continue_families(seeds, _make, max_members=case["max_members"], window=2,
                                   corrector=kind, **kw)
        for si, g in enumerate(got):
            ref = _fam(f"halo case {ci} seed {si}")
            assert (g.accepted_count, g.rejected_count, g.iterations) == ref["counts"]
            assert len(g.family) == len(ref["family"])
            for a, b in zip(g.family, ref["family"]):
                assert np.abs(np.asarray(a) - b).max() <= 1e-10
            per = np.array([a["period"] for a in g.aux])
            assert np.abs(per - ref["periods"]).max() <= 1e-10
