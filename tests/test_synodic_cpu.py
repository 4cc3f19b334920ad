"""CPU: the oracle's section detection (oracle/synodic.py) against crossings
fixed by hand -- linear and quadratic plane values, whose linear and cubic
Hermite interpolants are known exactly, a Hermite cubic with two roots in
one segment, on-surface samples, direction filters, deduplication and the
hit cap."""
import math

import numpy as np
import pytest

from oracle import synodic as S

EY = [0.0, 1.0, 0.0, 0.0, 0.0, 0.0]


def _states(times, y, vy=None):
    """States whose y is the plane value, x = t, vy as given (default t, so
    that distinct hits have distinct 2-d points (y, vy))."""
    n = len(times)
    X = np.zeros((n, 6))
    X[:, 0] = times
    X[:, 1] = y
    X[:, 4] = times if vy is None else vy
    return X


@pytest.mark.parametrize("interp", ["linear", "cubic"])
@pytest.mark.parametrize("refine", [0, 20])
def test_linear_plane_value_crosses_at_its_root(interp, refine):
    t = np.linspace(0.0, 1.0, 11)
    X = _states(t, 0.33 - t)
    hits = S.detect(t, X, normal=EY, interp_kind=interp, segment_refine=refine)
    assert len(hits) == 1
    th, st, pt = hits[0]
    assert abs(th - 0.33) < 1e-14
    # x = t along the trajectory: the interpolated state is on the line
    assert abs(st[0] - 0.33) < 1e-14 and abs(st[1]) < 1e-14
    assert pt == (st[1], st[4])


def test_quadratic_crossings_linear_chord_and_cubic_in_segment():
    # g = (t - 0.37)^2 - 0.0121: roots 0.26 and 0.48, inside segments 5, 9
    t = np.linspace(0.0, 1.0, 21)
    g = (t - 0.37) ** 2 - 0.0121
    X = _states(t, g)
    # cubic: the reference's Newton step uses its _hermite_der, which is not
    # the Hermite derivative, so the refined time is only fixed by the
    # reference's code (pinned by tests/golden); here: one hit per segment,
    # clamped to it
    cub = S.detect(t, X, normal=EY, interp_kind="cubic", newton_max_iter=25)
    assert len(cub) == 2
    assert t[5] <= cub[0][0] <= t[6] and t[9] <= cub[1][0] <= t[10]
    # linear: the chord root of each crossing segment
    lin = S.detect(t, X, normal=EY, interp_kind="linear")
    want = []
    for k in (5, 9):
        a = g[k] / (g[k] - g[k + 1])
        want.append((1 - a) * t[k] + a * t[k + 1])
    assert [h[0] for h in lin] == pytest.approx(want, abs=1e-15)
    assert abs(lin[0][0] - 0.26) > 1e-5          # the chord is not the root


@pytest.mark.parametrize("interp", ["linear", "cubic"])
def test_direction_filters(interp):
    t = np.linspace(0.0, 1.0, 21)
    X = _states(t, (t - 0.37) ** 2 - 0.0121)
    down = S.detect(t, X, normal=EY, interp_kind=interp, direction=-1, newton_max_iter=25)
    up = S.detect(t, X, normal=EY, interp_kind=interp, direction=1, newton_max_iter=25)
    assert len(down) == 1 and t[5] <= down[0][0] <= t[6]      # falling through 0.26
    assert len(up) == 1 and t[9] <= up[0][0] <= t[10]         # rising through 0.48


def test_two_roots_inside_one_segment_need_segment_refine():
    # samples g = 13, 13, 1, 1, 13, 13 at t = 0..5: on [2, 3] the Hermite
    # cubic with central-difference slopes -6 and +6 is 1 - 6 s (1 - s),
    # zero at s = 1/2 -+ sqrt(3)/6; the samples themselves never change sign
    t = np.arange(6.0)
    X = _states(t, [13.0, 13.0, 1.0, 1.0, 13.0, 13.0])
    assert S.detect(t, X, normal=EY, interp_kind="cubic") == []
    hits = S.detect(t, X, normal=EY, interp_kind="cubic", segment_refine=20,
                    newton_max_iter=25)
    r = math.sqrt(3.0) / 6.0
    # each root is found in its own sub-interval of width 1/21 and the
    # refinement stays inside it
    assert len(hits) == 2
    for h, root in zip(hits, (2.5 - r, 2.5 + r)):
        m = int((root - 2.0) * 21)
        assert 2.0 + m / 21 - 1e-15 <= h[0] <= 2.0 + (m + 1) / 21 + 1e-15
    # one of them in each direction
    assert len(S.detect(t, X, normal=EY, interp_kind="cubic", segment_refine=20,
                        direction=1)) == 1


@pytest.mark.parametrize("refine", [0, 20])
def test_on_surface_sample_is_one_hit(refine):
    # g = t - 0.25 is exactly 0 at the sample t = 0.25: the crossing that
    # ends there and the on-surface sample are the same hit (time dedup)
    t = np.linspace(0.0, 1.0, 21)
    X = _states(t, t - 0.25)
    hits = S.detect(t, X, normal=EY, segment_refine=refine)
    assert len(hits) == 1 and hits[0][0] == 0.25
    # falling only: neither the rising crossing nor the sample qualifies
    assert S.detect(t, X, normal=EY, direction=-1, segment_refine=refine) == []


def test_point_dedup_and_hit_cap():
    # y = sin(2 pi t): rising crossings at t = 0 (on the surface), 1, 2, 3, 4
    t = np.linspace(0.0, 4.5, 451)
    X = _states(t, np.sin(2 * np.pi * t))            # vy = t: distinct points
    rising = S.detect(t, X, normal=EY, direction=1)
    assert [round(h[0], 4) for h in rising] == [0.0, 1.0, 2.0, 3.0, 4.0]
    two = S.detect(t, X, normal=EY, direction=1, max_hits=2)
    assert [round(h[0], 4) for h in two] == [0.0, 1.0]
    # with vy = 1 every crossing is the 2-d point (0, 1) to rounding: points
    # within dedup_point_tol of the last kept hit are dropped
    X1 = _states(t, np.sin(2 * np.pi * t), vy=np.ones_like(t))
    assert len(S.detect(t, X1, normal=EY, direction=1)) == 1
    # a coarse point tolerance collapses the distinct points too
    assert len(S.detect(t, X, normal=EY, direction=1, dedup_point_tol=10.0)) == 1


def test_affine_normal_and_offset():
    # g = x + y - 1 with x = t, y = 0.5 t: root t = 2/3
    t = np.linspace(0.0, 1.0, 11)
    X = _states(t, 0.5 * t)
    hits = S.detect(t, X, normal=[1.0, 1.0, 0, 0, 0, 0], offset=1.0, interp_kind="cubic",
                    newton_max_iter=25)
    assert len(hits) == 1 and abs(hits[0][0] - 2.0 / 3.0) < 1e-14


def _golden():
    import os
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden",
                        "synodic_golden.npz")
    return np.load(path)


def golden_cases():
    """(option index, trajectory index, times, states, options, hits) of the
    reference-generated vectors (tests/golden/make_synodic_golden.py)."""
    from tests.golden_synodic_options import OPTIONS
    g = _golden()
    hits = g["hits"]
    n_traj = len(g["names"])
    for oi, opt in enumerate(OPTIONS):
        for ci in range(n_traj):
            sel = hits[(hits[:, 0] == oi) & (hits[:, 1] == ci)]
            yield oi, ci, g[f"t{ci}"], g[f"x{ci}"], opt, sel


def oracle_kwargs(opt):
    kw = dict(opt)
    pc = kw.pop("plane_coords", ("y", "vy"))
    idx = {"x": 0, "y": 1, "z": 2, "vx": 3, "vy": 4, "vz": 5}
    kw["plane_idx"] = (idx[pc[0]], idx[pc[1]])
    if "max_hits_per_traj" in kw:
        kw["max_hits"] = kw.pop("max_hits_per_traj")
    return kw


def test_oracle_matches_the_reference_golden_vectors():
    """The reference's own detection code (run once in this container, with
    numba's njit as the identity) against the oracle: same hits, same
    order; times and states to 1e-13 (the golden's Hermite helpers squared
    through the C library's pow, numba -- and the oracle -- multiply)."""
    n = 0
    for oi, ci, t, X, opt, want in golden_cases():
        got = S.detect(t, X, **oracle_kwargs(opt))
        assert len(got) == len(want), (oi, ci)
        for h, w in zip(got, want):
            assert abs(h[0] - w[2]) <= 1e-13 * max(1.0, abs(w[2])), (oi, ci)
            assert np.allclose(h[1], w[3:9], rtol=1e-13, atol=1e-14), (oi, ci)
            assert np.allclose(h[2], w[9:11], rtol=1e-13, atol=1e-14), (oi, ci)
            n += 1
    assert n == len(_golden()["hits"]) and n > 100
