"""GPU tests of the continuous-batching serving loop.

* Every job of a mixed stream (corrections + manifold fans arriving as a
  Poisson process) gives the same bits as the same request run alone, both
  through the loop and through the direct calls (correct_device for the
  device-resident corrector, correct_batch for the host-driven bank).
* Continuation jobs in the mix (Halo families grown from corrected seeds,
  their rounds' corrections run among everyone else's) give the same
  families as continue_family run alone with the same corrector.
* A sample of results against the oracle (the reference's sequential
  Newton-Armijo loop, bit for bit for the bank and to the corrector's
  tolerance for the device corrector, whose Halo-term powers are correctly
  rounded; its event integration per fan trajectory).
* HBM-budget admission: with a budget that holds only a few jobs the loop
  queues the rest, never exceeds the budget, and still completes every job.
"""

import math
import sys

import numpy as np
import pytest

from hiten_amd import batch, continuation as Ct, corrector as Cr, serving as S, workloads as W

pytestmark = pytest.mark.gpu


def say(msg):
    print(f"[case] {msg}", flush=True)
    sys.stdout.flush()


def _stream():
    # 60 ms of a 2k req/s stream, fans capped at 256 trajectories
    return W.serving_stream(duration=0.06, rate=2000.0, fan_max=256)


def _same_correction(a, b):
    assert a.converged == b.converged and a.iterations == b.iterations
    assert np.array_equal(a.x_corrected, b.x_corrected)
    assert (a.residual_norm == b.residual_norm) or (math.isnan(a.residual_norm)
                                                    and math.isnan(b.residual_norm))
    assert a.error == b.error and a.n_event_searches == b.n_event_searches
    assert (a.half_period == b.half_period) or (math.isnan(a.half_period)
                                                and math.isnan(b.half_period))


def _same_fan(a, b):
    assert np.array_equal(a.hit, b.hit)
    assert np.array_equal(a.t_hit, b.t_hit) and np.array_equal(a.y_hit, b.y_hit)


@pytest.fixture(scope="module", params=["device", "bank"])
def mixed_run(gpu, request):
    reqs = _stream()
    say(f"serving loop ({request.param} corrector): {len(reqs)} requests over 60 ms")
    server = S.Server(corrector=request.param)
    jobs = server.run(reqs, time_limit=60.0)
    server.close()
    assert all(not math.isnan(j.finished) for j in jobs)
    return reqs, jobs, request.param


def test_mixed_results_equal_direct_calls(mixed_run, oracle):
    reqs, jobs, kind = mixed_run
    say(f"serving results ({kind} corrector) vs direct calls and the oracle")
    corr = [j for j in jobs if j.kind == "correction"]
    fans = [j for j in jobs if j.kind == "fan"]
    assert corr and fans
    probs = [j.request for j in corr]
    direct = Cr.correct_device(probs) if kind == "device" else Cr.correct_batch(probs)
    for j, d in zip(corr, direct):
        _same_correction(j.result, d)
    if kind == "device":
        for j, b in zip(corr, Cr.correct_batch(probs)):
            assert j.result.converged == b.converged
            if b.converged:
                assert np.max(np.abs(j.result.x_corrected - b.x_corrected)) <= 1e-10
    for j in fans:
        r = j.request
        ev = batch.dop853_event(r.mu, r.y0, 0.0, r.tmax, ev_index=1, direction=-1,
                                xtol=r.xtol, gtol=r.gtol)
        _same_fan(j.result, S.FanResult(ev.hit.cpu().numpy(), ev.t_hit.cpu().numpy(),
                                        ev.y_hit.cpu().numpy(), ev.n_accepted.cpu().numpy()))
    for j in corr[:4]:
        w = oracle.correct_reference(j.request)
        assert j.result.converged == w["converged"]
        if kind == "bank":
            assert np.array_equal(j.result.x_corrected, w["x"])
            assert j.result.iterations == w["iterations"] and j.result.error == w["error"]
        elif w["converged"]:
            assert np.max(np.abs(j.result.x_corrected - w["x"])) <= 1e-10
    for j in fans[:2]:
        r = j.request
        hit, th, yh, _, _ = oracle.batch_dop853_event(r.mu, r.y0, r.tmax, ev_index=1,
                                                      direction=-1, xtol=r.xtol, gtol=r.gtol)
        assert np.array_equal(j.result.hit.astype(bool), hit.astype(bool))
        assert np.array_equal(j.result.t_hit, th) and np.array_equal(j.result.y_hit, yh)


def test_alone_equals_mixed(mixed_run):
    reqs, jobs, kind = mixed_run
    pick = [i for i, r in enumerate(reqs) if r["kind"] == "correction"][:3] + \
           [i for i, r in enumerate(reqs) if r["kind"] == "fan"][:2]
    say(f"alone vs mixed ({kind} corrector): requests {pick}")
    server = S.Server(corrector=kind)
    for i in pick:
        alone = server.run([dict(reqs[i], t=0.0)], time_limit=30.0)[0]
        if alone.kind == "correction":
            _same_correction(alone.result, jobs[i].result)
        else:
            _same_fan(alone.result, jobs[i].result)
    server.close()


def test_alone_equals_mixed_full_config(gpu):
    """0.5 s of the configuration's stream (2k req/s, 70% corrections, fans of
    the full 32-2048 range) through the loop with the device corrector: every
    correction equals correct_device on its problem, every fan equals its
    event search alone, and a sample (the largest fans among them) equals the
    same request run alone through the loop."""
    reqs = W.serving_stream(duration=0.5, rate=2000.0)
    sizes = [len(r["y0"]) for r in reqs if r["kind"] == "fan"]
    say(f"serving loop (device): {len(reqs)} requests over 0.5 s, fans {min(sizes)}-{max(sizes)}")
    assert max(sizes) > 1024 and min(sizes) < 64
    server = S.Server(corrector="device")
    jobs = server.run(reqs, time_limit=120.0)
    assert all(not math.isnan(j.finished) for j in jobs)
    corr = [j for j in jobs if j.kind == "correction"]
    fans = [j for j in jobs if j.kind == "fan"]
    for j, d in zip(corr, Cr.correct_device([j.request for j in corr])):
        _same_correction(j.result, d)
    for j in fans:
        r = j.request
        ev = batch.dop853_event(r.mu, r.y0, 0.0, r.tmax, ev_index=1, direction=-1,
                                xtol=r.xtol, gtol=r.gtol)
        _same_fan(j.result, S.FanResult(ev.hit.cpu().numpy(), ev.t_hit.cpu().numpy(),
                                        ev.y_hit.cpu().numpy(), ev.n_accepted.cpu().numpy()))
    big = sorted(range(len(reqs)), key=lambda i: -len(reqs[i].get("y0", ())))[:3]
    pick = big + [i for i, r in enumerate(reqs) if r["kind"] == "correction"][::97]
    say(f"alone vs mixed: requests {pick}")
    for i in pick:
        alone = server.run([dict(reqs[i], t=0.0)], time_limit=30.0)[0]
        if alone.kind == "correction":
            _same_correction(alone.result, jobs[i].result)
        else:
            _same_fan(alone.result, jobs[i].result)
    server.close()
    summ = S.summarize(jobs, 1.0, server.steps)
    assert sum(v["n"] for v in summ["correction"]["outcomes"].values()) == len(corr)


def test_converging_stream_converges(gpu):
    """Guesses near corrected Halo orbits (the bench's converging stream)
    converge to the tolerance, and the loop says so per outcome."""
    guesses = [Cr.halo_problem(W.MU_EM, [0.8234, 0, 0.0224, 0, 0.1343, 0], max_attempts=20),
               Cr.halo_problem(W.MU_EM, [0.84, 0, 0.1, 0, 0.22, 0], max_attempts=20)]
    seeds = [r.x_corrected for r in Cr.correct_device(guesses)]
    reqs = W.serving_stream(duration=0.05, rate=2000.0, fan_max=256, correction_seeds=seeds)
    say(f"converging stream: {len(reqs)} requests")
    server = S.Server(corrector="device")
    jobs = server.run(reqs, time_limit=60.0)
    server.close()
    corr = [j for j in jobs if j.kind == "correction"]
    conv = [j for j in corr if j.result.converged]
    assert len(conv) >= 0.9 * len(corr)
    assert all(j.result.residual_norm < 1e-12 for j in conv)
    summ = S.summarize(jobs, 1.0, server.steps)
    assert summ["correction"]["outcomes"]["converged"]["n"] == len(conv)


def test_hbm_budget_admission(gpu):
    reqs = _stream()
    largest = max(S.Server.fan_bytes(len(r["y0"])) for r in reqs if r["kind"] == "fan")
    budget = largest + 4 * S.Server.CORRECTION_BYTES
    say(f"admission with a {budget} B budget")
    server = S.Server(hbm_budget=budget)
    jobs = server.run(reqs, time_limit=60.0)
    server.close()
    assert server.max_queue > 0
    assert server.peak_used <= budget
    assert all(not math.isnan(j.finished) and j.admitted >= j.arrival for j in jobs)


def test_dedicated_streams_give_the_same_results(mixed_run):
    reqs, jobs, kind = mixed_run
    if kind != "device":
        pytest.skip("one corrector is enough")
    say("serving loop on dedicated streams: same bits as on torch's pool")
    server = S.Server(dedicated_streams=True, n_streams=4)
    again = server.run(reqs, time_limit=60.0)
    server.close()
    for a, b in zip(again, jobs):
        if a.kind == "correction":
            _same_correction(a.result, b.result)
        else:
            _same_fan(a.result, b.result)


@pytest.mark.parametrize("kind", ["device", "bank"])
def test_continuations_in_the_mix_equal_continue_family(gpu, kind):
    guesses = [Cr.halo_problem(W.MU_EM, [0.8234, 0, 0.0224, 0, 0.1343, 0], max_attempts=20),
               Cr.halo_problem(W.MU_EM, [0.84, 0, 0.1, 0, 0.22, 0], max_attempts=20)]
    seeds = [r.x_corrected for r in Cr.correct_device(guesses)]
    reqs = W.serving_stream(duration=0.04, rate=2000.0, fan_max=256, correction_fraction=0.55,
                            continuation_fraction=0.15, continuation_seeds=seeds)
    n_cont = sum(r["kind"] == "continuation" for r in reqs)
    say(f"serving loop with {n_cont} continuations among {len(reqs)} requests ({kind})")
    assert n_cont > 0
    server = S.Server(corrector=kind)
    jobs = server.run(reqs, time_limit=120.0)
    server.close()
    assert all(not math.isnan(j.finished) for j in jobs)
    conts = [j for j in jobs if j.kind == "continuation"]
    for j in conts:
        r = j.request
        alone = Ct.continue_family(r.seed, r.make_problem, window=r.window, corrector=kind,
                                   **r.family_kwargs())
        g = j.result
        assert (g.accepted_count, g.rejected_count, g.iterations) == (
            alone.accepted_count, alone.rejected_count, alone.iterations)
        for a, b in zip(g.family, alone.family):
            assert np.array_equal(a, b)
        assert g.aux == alone.aux and np.array_equal(g.final_step, alone.final_step)
    assert any(j.result.accepted_count == j.request.max_members for j in conts)
    corr = [j for j in jobs if j.kind == "correction"]
    probs = [j.request for j in corr]
    direct = Cr.correct_device(probs) if kind == "device" else Cr.correct_batch(probs)
    for j, d in zip(corr, direct):
        _same_correction(j.result, d)
    summ = S.summarize(jobs, 1.0, server.steps)
    assert summ["continuation"]["n"] == len(conts)


@pytest.mark.parametrize("kind", ["device", "bank"])
def test_sections_of_different_planes_in_one_loop(gpu, kind):
    """Vertical-orbit corrections (z = 0 section) beside Halo ones (y = 0):
    the loop keeps each problem on its own section."""
    ps = [Cr.vertical_problem(W.MU_EM, [0.84, 0, 0, 0, vy, vz], max_attempts=25)
          for vy, vz in ((0.006, 0.1387), (0.0, 0.1187), (0.02, 0.1687))]
    ps += [Cr.halo_problem(W.MU_EM, [0.8234, 0, 0.0224, 0, 0.1343, 0], max_attempts=20),
           Cr.halo_problem(W.MU_EM, [0.84, 0, 0.1, 0, 0.22, 0], max_attempts=20)]
    say(f"serving loop ({kind}): 3 vertical and 2 Halo corrections")
    server = S.Server(corrector=kind)
    jobs = [server.submit("correction", p, arrival=0.0) for p in ps]
    while server.busy():
        server.step()
    server.close()
    direct = Cr.correct_device(ps) if kind == "device" else Cr.correct_batch(ps)
    for j, d in zip(jobs, direct):
        _same_correction(j.result, d)
    assert jobs[0].result.converged and jobs[3].result.converged
