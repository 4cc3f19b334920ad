"""GPU: section detection on stored trajectories (hh_synodic_detect) against
the reference-generated golden vectors and the oracle.

Hit counts and order are exact; hit times are bit-identical to the oracle's
(same Hermite and Newton arithmetic); hit states agree to 1e-14 relative
(the reference's state interpolation squares through the C library's pow).
"""
import sys

import numpy as np
import pytest
import torch

from hiten_amd import batch, synodic, workloads as W
from oracle import synodic as S

pytestmark = pytest.mark.gpu
sys.path.insert(0, __file__.rsplit("/tests/", 1)[0])
from tests.test_synodic_cpu import golden_cases, oracle_kwargs  # noqa: E402


def say(msg):
    print(f"\n[case] {msg}", flush=True)


@pytest.fixture
def gpu():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")


def _gpu_kwargs(opt):
    kw = dict(opt)
    kw.setdefault("plane_coords", ("y", "vy"))
    return kw


def test_matches_reference_golden_vectors(gpu):
    say("hh_synodic_detect: 8 trajectories x 9 option sets against the reference's hits")
    n = 0
    for oi, ci, t, X, opt, want in golden_cases():
        res = synodic.detect_arrays(t, X[None], **_gpu_kwargs(opt))
        got = synodic.hits_from_arrays(res)[0]
        assert len(got) == len(want), (oi, ci)
        for h, w in zip(got, want):
            assert abs(h.time - w[2]) <= 1e-13 * max(1.0, abs(w[2])), (oi, ci)
            assert np.allclose(h.state, w[3:9], rtol=1e-13, atol=1e-14), (oi, ci)
            assert np.allclose(h.point2d, w[9:11], rtol=1e-13, atol=1e-14), (oi, ci)
            n += 1
    assert n > 100


def _cr3bp_grid(n=64, m=1200, tf=6.0, seed=5):
    w = W.stm_uneven(n=n, seed=seed)
    y0 = w["y0"][:, 36:]
    r = batch.dop853_grid(w["mu"], y0, 0.0, tf, m, dim=6, max_step=1e4)
    return np.linspace(0.0, tf, m), r.states


OPTS = [
    dict(interp_kind="linear", segment_refine=0),
    dict(interp_kind="cubic", segment_refine=0, newton_max_iter=4),
    dict(interp_kind="cubic", segment_refine=20, newton_max_iter=25),
    dict(interp_kind="linear", segment_refine=3, direction=1),
    dict(interp_kind="cubic", segment_refine=20, newton_max_iter=25, direction=-1,
         max_hits_per_traj=3),
]


@pytest.mark.parametrize("opt", OPTS, ids=[str(i) for i in range(len(OPTS))])
@pytest.mark.parametrize("plane", ["y", "x"])
def test_matches_oracle_on_cr3bp_grids(gpu, opt, plane):
    say(f"hh_synodic_detect on 64 CR3BP grids of 1200 samples, plane {plane}, {opt}")
    t, states = _cr3bp_grid()
    normal = synodic.axis_normal(plane)
    offset = 0.0 if plane == "y" else 0.84
    res = synodic.detect_arrays(t, states, normal=normal, offset=offset, **opt)
    got = synodic.hits_from_arrays(res)
    host = states.cpu().numpy()
    kw = dict(opt)
    if "max_hits_per_traj" in kw:
        kw["max_hits"] = kw.pop("max_hits_per_traj")
    total = 0
    for i in range(host.shape[0]):
        want = S.detect(t, host[i], normal=normal, offset=offset, plane_idx=(1, 4), **kw)
        assert len(got[i]) == len(want), i
        for h, wv in zip(got[i], want):
            assert h.time == wv[0], (i, h.time, wv[0])
            assert np.allclose(h.state, wv[1], rtol=1e-14, atol=1e-15), i
            total += 1
    assert total > 0


@pytest.mark.parametrize("opt", OPTS, ids=[str(i) for i in range(len(OPTS))])
def test_two_pass_equals_one_wave(gpu, opt):
    """hh_synodic_detect_ws (parts scanned by waves of their own, then merged)
    against hh_synodic_detect (one wave walks a trajectory): the same hits."""
    say(f"synodic two-pass vs one wave per trajectory, 64 grids, {opt}")
    t, states = _cr3bp_grid()
    kw = dict(normal=synodic.axis_normal("y"), **opt)
    a = synodic.hits_from_arrays(synodic.detect_arrays(t, states, two_pass=False, **kw))
    b = synodic.hits_from_arrays(synodic.detect_arrays(t, states, two_pass=True, **kw))
    assert [len(h) for h in a] == [len(h) for h in b]
    for ha, hb in zip(a, b):
        for u, v in zip(ha, hb):
            assert u.time == v.time and np.array_equal(u.state, v.state)
            assert np.array_equal(u.point2d, v.point2d)


def test_two_pass_crowded_parts_fall_back(gpu):
    """Trajectories crossing the plane every 10 samples list more than 32
    segments in a part of 512: those are detected whole; the result is the
    one-wave kernel's and the oracle's."""
    say("synodic two-pass: crowded parts (51 crossings per 512 segments)")
    m = 1500
    t = np.linspace(0.0, 15.0, m)
    n = 4
    states = np.zeros((n, m, 6))
    for i in range(n):
        w = 2 * np.pi / (0.2 * (1.0 + 0.1 * i))          # a crossing every ~10 samples
        states[i, :, 0] = 0.8 + 0.01 * np.cos(0.3 * t)
        states[i, :, 1] = 0.01 * np.sin(w * t + 0.1 * i)
        states[i, :, 4] = 0.01 * w * np.cos(w * t + 0.1 * i)
    states[1, :, 1] = 0.01 + 0.001 * np.sin(t)            # never crosses
    kw = dict(normal=synodic.axis_normal("y"), interp_kind="cubic", segment_refine=3,
              newton_max_iter=10)
    a = synodic.hits_from_arrays(synodic.detect_arrays(t, states, two_pass=False, **kw))
    b = synodic.hits_from_arrays(synodic.detect_arrays(t, states, two_pass=True, **kw))
    assert len(b[0]) > 100 and len(b[1]) == 0
    for i in range(n):
        want = S.detect(t, states[i], normal=synodic.axis_normal("y"), plane_idx=(1, 4),
                        interp_kind="cubic", segment_refine=3, newton_max_iter=10)
        assert len(a[i]) == len(b[i]) == len(want), i
        for u, v, wv in zip(a[i], b[i], want):
            assert u.time == v.time == wv[0]
            assert np.array_equal(u.state, v.state)


def test_per_trajectory_times_lengths_and_small_cap(gpu):
    say("hh_synodic_detect: per-trajectory times, ragged lengths, cap overflow re-run")
    t, states = _cr3bp_grid(n=16, m=800)
    host = states.cpu().numpy()
    T = np.repeat(t[None], 16, axis=0) * (1.0 + 0.01 * np.arange(16))[:, None]
    lengths = np.array([800, 2, 1, 0] + [400 + 20 * i for i in range(12)], dtype=np.int32)
    opt = dict(interp_kind="cubic", segment_refine=5, newton_max_iter=25)
    res = synodic.detect_arrays(T, host, lengths=lengths, normal=synodic.axis_normal("y"),
                                cap=1, **opt)
    got = synodic.hits_from_arrays(res)
    assert res.t.shape[1] >= max(len(h) for h in got)
    for i in range(16):
        L = int(lengths[i])
        want = S.detect(T[i, :L], host[i, :L], normal=synodic.axis_normal("y"),
                        plane_idx=(1, 4), **opt)
        assert len(got[i]) == len(want), i
        for h, wv in zip(got[i], want):
            assert h.time == wv[0]


def test_backend_run_on_a_request(gpu):
    say("_SynodicDetectionBackend.run on ragged trajectories")
    t, states = _cr3bp_grid(n=6, m=600)
    host = states.cpu().numpy()
    trajs = [(t[: 300 + 50 * i], host[i, : 300 + 50 * i]) for i in range(6)]
    req = synodic.SynodicBackendRequest(trajectories=trajs, normal=synodic.axis_normal("y"),
                                        interp_kind="cubic", segment_refine=20,
                                        newton_max_iter=25, trajectory_indices=list(range(10, 16)))
    resp = synodic._SynodicDetectionBackend().run(req)
    assert len(resp.hits) == 6
    k = 0
    for i, (ti, xi) in enumerate(trajs):
        want = S.detect(ti, xi, normal=synodic.axis_normal("y"), interp_kind="cubic",
                        segment_refine=20, newton_max_iter=25)
        assert [h.time for h in resp.hits[i]] == [w[0] for w in want]
        assert all(h.trajectory_index == 10 + i for h in resp.hits[i])
        k += len(want)
    assert resp.points.shape == (k, 2) and resp.states.shape == (k, 6)
    single = synodic._SynodicDetectionBackend().detect_on_trajectory(
        trajs[0][0], trajs[0][1], normal=synodic.axis_normal("y"), interp_kind="cubic",
        segment_refine=20, newton_max_iter=25, trajectory_index=10)
    assert [h.time for h in single] == [h.time for h in resp.hits[0]]
