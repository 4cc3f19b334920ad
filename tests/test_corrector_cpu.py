"""CPU test: the batched corrector's state machines (speculative line search,
reuse of the accepted trial's residual) against the reference's sequential
Newton-Armijo loop restated literally in the oracle
(corrector/backends/newton.py:39-143, stepping/armijo.py:58-164), both on
the oracle's event search and STM.  Every field must be identical."""

import math

import numpy as np

from hiten_amd import corrector as Cr, workloads as W

MU = W.MU_EM


def _problems():
    xl1 = W.l1_position(MU)
    c2 = (1 - MU) / abs(xl1 + MU)**3 + MU / abs(xl1 - 1 + MU)**3
    wp = np.sqrt((2 - c2 + np.sqrt(9 * c2**2 - 8 * c2)) / 2)
    kap = (wp**2 + 1 + 2 * c2) / (2 * wp)
    lyap = lambda ax: [xl1 - ax, 0, 0, 0, kap * ax * wp, 0]
    ps = [
        Cr.halo_problem(MU, [0.8234, 0, 0.0224, 0, 0.1343, 0], max_attempts=20),
        Cr.halo_problem(MU, [0.84, 0, 0.1, 0, 0.22, 0], max_attempts=20),
        Cr.halo_problem(MU, [0.82, 0, 0.05, 0, 0.15, 0], max_attempts=4),
        Cr.lyapunov_problem(MU, lyap(0.005), max_attempts=20),
        Cr.lyapunov_problem(MU, lyap(0.01), max_attempts=20),
        Cr.halo_problem(MU, [0.8234, 0, 0.0224, 0, 0.1343, 0], max_attempts=20,
                        alpha_reduction=0.7, armijo_c=0.3),
    ]
    rng = np.random.default_rng(20261202)
    for _ in range(10):      # seeds drawn as the serving stream draws them
        ps.append(Cr.halo_problem(MU, [rng.uniform(0.80, 0.88), 0, rng.uniform(-0.2, 0.2), 0,
                                       rng.uniform(-0.3, 0.3), 0], max_attempts=10))
    return ps


def test_batched_corrector_equals_sequential_reference(oracle):
    ps = _problems()
    got = Cr.correct_batch(ps, engine=oracle.CorrectorEngine())
    outcomes = set()
    for p, g in zip(ps, got):
        w = oracle.correct_reference(p)
        assert g.converged == w["converged"]
        assert g.iterations == w["iterations"]
        assert np.array_equal(g.x_corrected, w["x"])
        assert g.residual_norm == w["residual_norm"] or (math.isnan(g.residual_norm)
                                                          and math.isnan(w["residual_norm"]))
        assert g.error == w["error"]
        assert g.n_event_searches == w["n_event_searches"]
        assert g.n_stm == w["n_stm"]
        assert (g.half_period == w["half_period"]) or (math.isnan(g.half_period)
                                                       and math.isnan(w["half_period"]))
        outcomes.add((g.converged, g.error.split(" ")[0]))
    # converged, max-iterations and line-search failures are all exercised
    assert (True, "") in outcomes and (False, "Newton") in outcomes and (False, "Step") in outcomes
