"""GPU test of the Newton-Armijo corrector: the GPU operators (event search,
STM) against the oracle's restatement of the same calls, through the same
host-side Newton/Armijo restatement.  Results must be bit-identical; the
converged orbits must also close on themselves (periodicity)."""

import sys

import numpy as np
import pytest

from hiten_amd import batch, corrector as Cr, workloads as W

pytestmark = pytest.mark.gpu
MU = W.MU_EM


def problems():
    xl1 = W.l1_position(MU)
    c2 = (1 - MU) / abs(xl1 + MU)**3 + MU / abs(xl1 - 1 + MU)**3
    wp = np.sqrt((2 - c2 + np.sqrt(9 * c2**2 - 8 * c2)) / 2)
    kap = (wp**2 + 1 + 2 * c2) / (2 * wp)
    lyap = lambda ax: [xl1 - ax, 0, 0, 0, kap * ax * wp, 0]
    return [
        Cr.halo_problem(MU, [0.8234, 0, 0.0224, 0, 0.1343, 0], max_attempts=20),
        Cr.halo_problem(MU, [0.84, 0, 0.1, 0, 0.22, 0], max_attempts=20),
        Cr.halo_problem(MU, [0.82, 0, 0.05, 0, 0.15, 0], max_attempts=4),     # no convergence
        Cr.lyapunov_problem(MU, lyap(0.005), max_attempts=20),
        Cr.lyapunov_problem(MU, lyap(0.01), max_attempts=20),                # Armijo failure
    ]


def test_corrector_matches_oracle_operators(gpu, oracle):
    print("[case] correct_batch: 3 Halo + 2 Lyapunov problems, GPU vs oracle operators", flush=True)
    gpu_res = Cr.correct_batch(problems())
    ref_res = Cr.correct_batch(problems(), engine=oracle.CorrectorEngine())
    for g, r in zip(gpu_res, ref_res):
        assert g.converged == r.converged
        assert g.iterations == r.iterations
        assert np.array_equal(g.x_corrected, r.x_corrected)
        assert g.residual_norm == r.residual_norm or (np.isnan(g.residual_norm) and np.isnan(r.residual_norm))
        assert (g.half_period == r.half_period) or (np.isnan(g.half_period) and np.isnan(r.half_period))
        assert g.error == r.error
    assert [g.converged for g in gpu_res] == [True, True, False, True, False]
    assert "Armijo" in gpu_res[4].error
    # and the reference's sequential loop on the oracle operators
    for p, g in zip(problems(), gpu_res):
        w = oracle.correct_reference(p)
        assert np.array_equal(g.x_corrected, w["x"]) and g.iterations == w["iterations"]
        assert g.n_event_searches == w["n_event_searches"] and g.error == w["error"]


def test_corrected_orbits_are_periodic(gpu):
    print("[case] corrected orbits close after one period", flush=True)
    res = [r for r in Cr.correct_batch(problems()) if r.converged]
    x0 = np.stack([r.x_corrected for r in res])
    T = np.array([r.period for r in res])
    out = batch.dop853_final(MU, x0, 0.0, T, dim=6).y.cpu().numpy()
    assert np.max(np.abs(out - x0)) < 1e-8
