"""The HIP kernels against golden vectors written by the reference's own code
(tests/golden/integrator_golden.npz; see tests/test_golden_integrators.py for
what the vectors hold and how the oracle reproduces them).

The kernels compute in numba's arithmetic (the oracle's libm_pow=0, which
they equal bit for bit -- asserted here too), so against the reference's
vectors: accepted and rejected step counts, hit flags and crossing outcomes
identical; states, hit times and crossing states within SPREAD_FACTOR times
each case's rounding sensitivity (golden_integrators.bound).
"""
import numpy as np
import pytest
import torch

from tests import golden_integrators as GI

pytestmark = pytest.mark.gpu

STM = list(GI.stm_cases())
ARC = list(GI.arc_cases())
EVENT = list(GI.event_cases())
FIXED = list(GI.fixed_cases())
CROSS = list(GI.crossing_cases())


def _np(t):
    return t.detach().cpu().numpy() if isinstance(t, torch.Tensor) else np.asarray(t)


def _stack(cases, key):
    return np.stack([np.asarray(c[key], float) for c in cases])


@pytest.mark.parametrize("variant", ["auto", "wide", "pair", "persistent", "persistent_pair"])
def test_stm_kernels_vs_reference(gpu, oracle, variant):
    from hiten_amd import batch
    for fwd in (1, -1):
        for steps in (2, 9):
            cs = [c for c in STM if c["forward"] == fwd and c["steps"] == steps]
            if not cs or (steps == 9 and variant != "auto"):
                continue
            y0 = np.zeros((len(cs), 42))
            y0[:, :36] = np.eye(6).ravel()
            y0[:, 36:] = _stack(cs, "x0")
            mu = np.array([c["mu"] for c in cs])
            tf = np.array([c["tf"] for c in cs])
            print(f"stm fwd={fwd} steps={steps} variant={variant} n={len(cs)}", flush=True)
            if steps == 2:
                r = batch.dop853_final(mu, y0, 0.0, tf, dim=42, forward=fwd,
                                       flip_indices=slice(36, 42), variant=variant)
                got = _np(r.y)[:, None, :]
            else:
                r = batch.dop853_grid(mu, y0, 0.0, tf, steps, dim=42, forward=fwd,
                                      flip_indices=slice(36, 42))
                got = _np(r.states)
            na, nr, st = _np(r.n_accepted), _np(r.n_rejected), _np(r.status)
            for i, c in enumerate(cs):
                ref = c["out"][-1:] if steps == 2 else c["out"]
                assert st[i] == 0
                assert (na[i], nr[i]) == GI.counts(c["trace"]), f"stm case {c['k']}"
                assert np.abs(got[i] - ref).max() <= GI.bound(c["spread"], na[i], ref), c["k"]
                want, _, _ = GI.oracle_stm(oracle, c, 0)
                assert np.array_equal(got[i], want[-1:] if steps == 2 else want), c["k"]


def test_compute_stm_dropin_vs_reference(gpu):
    from hiten_amd import dynamics
    for c in STM[:6]:
        print(f"_compute_stm case {c['k']}", flush=True)
        x, times, phi_T, PHI = dynamics._compute_stm(dynamics.variational_dynsys(c["mu"]),
                                                     c["x0"], c["tf"], steps=c["steps"],
                                                     forward=c["forward"])
        assert np.array_equal(times, c["times"])
        assert np.abs(PHI - c["out"]).max() <= GI.bound(c["spread"], 10_000, c["out"])


def test_arcs_vs_reference(gpu, oracle):
    from hiten_amd import batch, dynamics
    for c in ARC:
        print(f"arc case {c['k']}", flush=True)
        sol = dynamics._propagate_dynsys(dynamics.rtbp_dynsys(c["mu"]), c["x0"], 0.0, c["tf"],
                                         forward=c["forward"], steps=c["steps"])
        assert np.array_equal(sol.times, c["times"])
        if c["steps"] == 2:
            r = batch.dop853_final(c["mu"], c["x0"][None], 0.0, c["tf"], dim=6,
                                   forward=c["forward"])
        else:
            r = batch.dop853_grid(c["mu"], c["x0"][None], 0.0, c["tf"], c["steps"], dim=6,
                                  forward=c["forward"])
        na, nr = int(_np(r.n_accepted)[0]), int(_np(r.n_rejected)[0])
        assert (na, nr) == GI.counts(c["trace"])
        assert np.abs(sol.states - c["out"]).max() <= GI.bound(c["spread"], na, c["out"])
        want, _, _ = GI.oracle_arc(oracle, c, 0)
        assert np.array_equal(sol.states, want)


@pytest.mark.parametrize("variant", ["plain", "wide"])
def test_events_vs_reference(gpu, oracle, variant):
    from hiten_amd import batch
    mu = np.array([c["mu"] for c in EVENT])
    tmax = np.array([c["tmax"] for c in EVENT])
    print(f"events variant={variant} n={len(EVENT)}", flush=True)
    r = batch.dop853_event(mu, _stack(EVENT, "x0"), 0.0, tmax, ev_index=1, direction=-1,
                           xtol=1e-13, gtol=1e-13, max_step=1e4, variant=variant)
    hit, th, yh = _np(r.hit), _np(r.t_hit), _np(r.y_hit)
    na, nr = _np(r.n_accepted), _np(r.n_rejected)
    for i, c in enumerate(EVENT):
        assert bool(hit[i]) == c["hit"], c["k"]
        assert (na[i], nr[i]) == GI.counts(c["trace"]), c["k"]
        b = GI.bound(c["spread"], na[i], c["state"])
        if c["hit"]:
            assert abs(th[i] - c["t"]) <= b and np.abs(yh[i] - c["state"]).max() <= b
        h0, t0, y0, _, _ = GI.oracle_event(oracle, c, 0)
        assert bool(hit[i]) == h0 and th[i] == t0 and np.array_equal(yh[i], y0)


def test_fixed_rk8_vs_reference(gpu, oracle):
    from hiten_amd import batch
    x0 = _stack(FIXED, "x0")
    steps = FIXED[0]["steps"]
    print(f"fixed rk8 n={len(FIXED)} steps={steps}", flush=True)
    _, traj = batch.rk_fixed(GI.MU_EM, x0, 0.0, 2.0 * np.pi, steps, order=8, dim=6,
                             return_trajectory=True)
    got = _np(traj)
    assert got.shape == (len(FIXED), steps, 6)
    for i, c in enumerate(FIXED):
        assert np.abs(got[i] - c["out"]).max() <= GI.bound(c["spread"], steps, c["out"])
        assert np.array_equal(got[i], GI.oracle_fixed(oracle, c, 0))


def test_y_plane_crossing_vs_reference(gpu, oracle):
    from hiten_amd import dynamics, singlehit
    for c in CROSS:
        print(f"crossing case {c['k']} mu={c['mu']} forward={c['forward']}", flush=True)
        t, y = singlehit._y_plane_crossing(dynsys=dynamics.rtbp_dynsys(c["mu"]), x0=c["x0"],
                                           forward=c["forward"])
        if np.isnan(c["t"]):
            assert t is None and y is None
            continue
        b = GI.bound(c["spread"], 200, c["state"])
        assert abs(t - c["t"]) <= b and np.abs(np.asarray(y) - c["state"]).max() <= b
        t0, y0 = GI.oracle_crossing(oracle, c, 0)
        assert t == t0 and np.array_equal(np.asarray(y), y0)
