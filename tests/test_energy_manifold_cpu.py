"""CPU tests of the energy filters (oracle) and of the manifold host logic.

Pins:
* the oracle's Jacobi constant against a closed form and against the
  reference's own NumPy expression of the manifold distances
  (services/manifold.py:413-420 evaluates them with numpy, so numpy is the
  reference arithmetic there);
* _max_rel_energy_error's indexing and branch (first sample as reference,
  running maximum that skips NaN) on constructed trajectories with known
  drift;
* the eigen-decomposition restatement on known spectra (the reference's
  docstring example, a symplectic monodromy-like matrix);
* the manifold seed construction on an identity STM grid.
"""

import math

import numpy as np
import pytest

from hiten_amd import manifold as M

MU = 0.012150585609624


def test_jacobi_closed_form(oracle):
    # mu = 1/2, point (0, 0, 1/2) at rest: r1 = r2 = sqrt(1/2),
    # C = 2 (1/2 / r + 1/2 / r) = 2 sqrt(2)
    c = oracle.jacobi([0.0, 0.0, 0.5, 0.0, 0.0, 0.0], 0.5)
    assert abs(c - 2.0 * math.sqrt(2.0)) <= 1e-15          # two ulps
    # velocity enters as -|v|^2
    c2 = oracle.jacobi([0.0, 0.0, 0.5, 0.3, 0.0, 0.4], 0.5)
    assert abs((c - c2) - 0.25) < 1e-15


def test_jacobi_invariant_along_orbit(oracle):
    x0 = np.array([0.82, 0.0, 0.03, 0.0, 0.15, 0.0])
    te = np.linspace(0.0, 3.0, 200)
    states, _ = oracle.dop853(oracle.System("cr3bp6", mu=MU), x0, te, rtol=1e-12, atol=1e-12)
    assert oracle.max_rel_energy_error(states, MU) < 1e-9
    # a wrong sign of the Coriolis term or of v^2 would not conserve C:
    # perturbing the velocity by 1e-4 must show
    bad = states.copy()
    bad[:, 4] *= 1.0 + 1e-4
    assert oracle.max_rel_energy_error(bad, MU) > 1e-6


def test_max_rel_energy_error_known_drift(oracle):
    s0 = np.array([0.8, 0.01, 0.02, 0.01, 0.2, 0.0])
    C0 = oracle.jacobi(s0, MU)
    # sample 1: same state (no drift); sample 2: |v|^2 raised by exactly
    # 2^-10 via vz (vz^2 = 2^-10, exact), so C drops by 2^-10
    s2 = s0.copy()
    s2[5] = 2.0 ** -5
    S = np.stack([s0, s0, s2])
    want = abs((oracle.jacobi(s2, MU)) - C0) / abs(C0)
    assert abs(want - 2.0 ** -10 / abs(C0)) < 1e-15
    assert oracle.max_rel_energy_error(S, MU) == want
    # the first sample is the reference, so reversing the order changes it
    assert oracle.max_rel_energy_error(S[::-1].copy(), MU) == abs(C0 - oracle.jacobi(s2, MU)) / abs(oracle.jacobi(s2, MU))
    # NaN samples never raise the running maximum
    S_nan = np.stack([s0, np.full(6, np.nan), s2])
    assert oracle.max_rel_energy_error(S_nan, MU) == want


def test_traj_filter_distances_match_reference_numpy(oracle):
    rng = np.random.default_rng(5)
    S = rng.uniform(-1.2, 1.2, (7, 33, 6))
    mu = rng.uniform(1e-6, 0.05, 7)
    r1, r2, e = oracle.traj_filters(mu, S)
    for i in range(7):
        x, y, z = S[i, :, 0], S[i, :, 1], S[i, :, 2]
        want1 = np.sqrt((x + mu[i]) ** 2 + y ** 2 + z ** 2).min()
        want2 = np.sqrt((x - 1 + mu[i]) ** 2 + y ** 2 + z ** 2).min()
        assert r1[i] == want1 and r2[i] == want2
        assert e[i] == oracle.max_rel_energy_error(S[i], mu[i])
    S[3, 10, 0] = np.nan
    r1, r2, _ = oracle.traj_filters(mu, S)
    assert np.isnan(r1[3]) and np.isnan(r2[3]) and not np.isnan(r1[2])


def test_eigen_decomposition_docstring_example():
    # linalg/backend.py docstring: continuous system diag(-2, 0, 0.5)
    sn, un, cn, Ws, Wu, Wc = M.eigenvalue_decomposition(np.diag([-2.0, 0.0, 0.5]), delta=1e-4,
                                                        discrete=False)
    assert sn.tolist() == [-2.0] and un.tolist() == [0.5] and cn.tolist() == [0.0]
    assert Ws.shape == (3, 1) and Wu.shape == (3, 1) and Wc.shape == (3, 1)


def test_eigen_decomposition_monodromy_like():
    rng = np.random.default_rng(3)
    lam, th = 37.5, 0.7
    D = np.zeros((6, 6))
    D[0, 0], D[1, 1] = lam, 1.0 / lam
    D[2:4, 2:4] = [[math.cos(th), -math.sin(th)], [math.sin(th), math.cos(th)]]
    D[4, 4] = D[5, 5] = 1.0
    P = rng.normal(size=(6, 6))
    A = P @ D @ np.linalg.inv(P)
    sn, un, cn, Ws, Wu, Wc = M.eigenvalue_decomposition(A, delta=1e-6, discrete=True)
    assert len(sn) == 1 and len(un) == 1 and len(cn) == 4
    assert abs(sn[0] - 1 / lam) < 1e-9 and abs(un[0] - lam) < 1e-9
    for vals, vecs in ((sn, Ws), (un, Wu), (cn, Wc)):
        for v, w in zip(vals, vecs.T):
            assert np.allclose(A @ w, v * w, atol=1e-8)
            k = np.argmax(np.abs(w) >= 1e-14)
            assert abs(w[k] - 1.0) < 1e-15     # pivot-normalised (complex division rounds)
    _, real_s = M.get_real_eigenvectors(Ws, sn)
    assert real_s.shape == (6, 1) and np.all(np.isreal(real_s))


def test_manifold_seeds_identity_stm():
    m = 11
    tt = -np.linspace(0.0, 2.0, m)          # backward STM run: times are -t
    xx = np.arange(m * 6, dtype=float).reshape(m, 6)
    PHI = np.zeros((m, 42))
    PHI[:, :36] = np.eye(6).ravel()
    PHI[:, 36:] = xx
    ev = np.array([0.0, 2.0, 0.0, 0.0, 0.0, 0.0], dtype=complex)
    seeds = M.manifold_seeds(xx, tt, PHI, 2.0, (0.0, 0.5, 0.98), ev, -1, 1e-6)
    # fraction 0.5 of period 2 -> |t| = 1 -> sample 5; fraction 0.98 -> 1.96 -> sample 10
    want = np.stack([xx[0], xx[5], xx[10]])
    want[:, 1] += -1e-6                     # d = 1e-6 / |MAN[:3]|, MAN = -(0, 2, 0, ...)
    assert np.array_equal(seeds, want)


def test_manifold_to_df_rows_and_columns(tmp_path):
    """ManifoldResult.to_df against Manifold.to_df's row-by-row construction
    (system/manifold.py:373-387) restated literally."""
    import pandas as pd
    from hiten_amd.manifold import ManifoldResult
    rng = np.random.default_rng(4)
    times = [np.linspace(0.0, 1.0, 5), np.linspace(0.0, -2.0, 3)]
    states = [rng.standard_normal((5, 6)), rng.standard_normal((3, 6))]
    res = ManifoldResult(states_list=states, times_list=times, successes=2, attempts=2)
    rows = []
    for i, traj in enumerate(res.trajectories):
        t, x = traj.times, traj.states
        for j in range(x.shape[0]):
            rows.append([i, t[j], *x[j, :6]])
    want = pd.DataFrame(rows, columns=["trajectory_id", "time", "x", "y", "z", "vx", "vy", "vz"])
    got = res.to_df()
    assert list(got.columns) == list(want.columns)
    assert np.array_equal(got.to_numpy(), want.to_numpy())
    path = tmp_path / "out" / "m.csv"
    res.to_csv(str(path))
    back = pd.read_csv(path)
    assert back.shape == (8, 8) and back["trajectory_id"].tolist() == [0] * 5 + [1] * 3
    assert ManifoldResult().to_df().shape == (0, 8)


def test_trajectory_container_as_the_reference_defines_it():
    """Trajectory (types/states.py:613-830): validation, views, (time, state)
    access, slicing, reversal and from_solution without a copy."""
    from hiten_amd.integrators import _Solution
    from hiten_amd.trajectory import Trajectory
    t = np.linspace(0.0, -1.0, 5)
    x = np.arange(30.0).reshape(5, 6)
    tr = Trajectory(t, x)
    assert (tr.n_samples, tr.dim, len(tr)) == (5, 6, 5)
    assert tr.t0 == 0.0 and tr.tf == -1.0 and tr.duration == -1.0
    t2, x2 = tr[2]
    assert t2 == t[2] and np.array_equal(x2, x[2])
    sub = tr[1:4]
    assert isinstance(sub, Trajectory) and np.array_equal(sub.times, t[1:4])
    assert [p[0] for p in tr] == list(t)
    rv = tr.reversed()
    assert np.array_equal(rv.times, t[::-1]) and np.array_equal(rv.states, x[::-1])
    sol = _Solution(t, x)
    fs = Trajectory.from_solution(sol)
    assert fs.times is not None and np.shares_memory(fs.states, x)
    assert "decreasing" in repr(tr)
    for bad in ((t[:1], x[:1]), (np.array([0.0, 1.0, 1.0]), x[:3]), (t, x[:4]), (t[None], x)):
        with pytest.raises(ValueError):
            Trajectory(*bad)
    with pytest.raises(ValueError):
        tr.vector_at(0)
