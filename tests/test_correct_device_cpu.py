"""CPU: the pieces of the device-resident corrector that can be checked
without a GPU -- the 2x2 dense step restated as NumPy's LAPACK forms it
(oracle/linalg2.py, the formula the device kernel uses) against
numpy.linalg.solve and numpy.linalg.cond, the record layouts against the
library's, and the host-side validation of problems."""
import ctypes as C

import numpy as np
import pytest

from hiten_amd import _lib, corrector as Cr, workloads as W
from oracle import linalg2 as L2


def _systems(n, seed):
    rng = np.random.default_rng(seed)
    A = rng.standard_normal((n, 2, 2)) * np.exp(rng.uniform(-8, 8, (n, 1, 1)))
    b = rng.standard_normal((n, 2)) * np.exp(rng.uniform(-8, 8, (n, 1)))
    return A, b


def test_lu_solve2_is_numpys_solve_bit_for_bit():
    A, b = _systems(20000, 1)
    want = np.linalg.solve(A, b[:, :, None])[:, :, 0]
    for i in range(len(A)):
        got = L2.lu_solve2(A[i].ravel(), b[i])
        assert got is not None
        assert got[0] == want[i, 0] and got[1] == want[i, 1], i


def test_lu_solve2_pivoting_and_singular_cases():
    # equal magnitudes keep the first row as pivot; a zero pivot column and a
    # zero Schur complement are LAPACK's singular cases
    for A, b in [([2.0, 1.0, -2.0, 3.0], [1.0, 5.0]), ([1e-300, 1.0, 1.0, 1.0], [2.0, 3.0]),
                 ([0.0, 1.0, 3.0, 4.0], [1.0, 1.0])]:
        want = np.linalg.solve(np.array(A).reshape(2, 2), np.array(b))
        assert L2.lu_solve2(A, b) == (want[0], want[1])
    assert L2.lu_solve2([0.0, 1.0, 0.0, 2.0], [1.0, 1.0]) is None
    assert L2.lu_solve2([1.0, 2.0, 2.0, 4.0], [1.0, 1.0]) is None
    with pytest.raises(np.linalg.LinAlgError):
        np.linalg.solve(np.array([[1.0, 2.0], [2.0, 4.0]]), np.ones(2))


def test_cond2_matches_numpy_and_its_threshold_decisions():
    A, _ = _systems(5000, 2)
    c = np.linalg.cond(A)
    got = np.array([L2.cond2(a.ravel()) for a in A])
    assert np.allclose(got, c, rtol=1e-12)
    # matrices with condition numbers straddling the 1e8 threshold
    rng = np.random.default_rng(3)
    for k in range(200):
        s = 10.0 ** rng.uniform(7.9, 8.1)
        U, _ = np.linalg.qr(rng.standard_normal((2, 2)))
        V, _ = np.linalg.qr(rng.standard_normal((2, 2)))
        M = U @ np.diag([1.0, 1.0 / s]) @ V.T
        assert (L2.cond2(M.ravel()) > 1e8) == (np.linalg.cond(M) > 1e8)
    assert L2.cond2([0.0, 0.0, 0.0, 0.0]) == np.inf == np.linalg.cond(np.zeros((2, 2)))
    assert L2.cond2([1.0, 2.0, 2.0, 4.0]) == np.inf


# Exactly singular steps: J + 1e-12 I rounds back to J (diagonal entries of
# at least 2^14, whose ulp exceeds 2e-12) and getf2's Schur complement is
# exactly 0 (a power-of-two pivot), so the reference's solve raises and it
# takes numpy.linalg.lstsq(rcond=None), LAPACK dgelsd
SINGULAR_J = [2.0**20 * np.array([[4.0, 2.0], [2.0, 1.0]]),
              2.0**20 * np.array([[1.0, 1.0], [1.0, 1.0]]),
              2.0**30 * np.array([[2.0, -4.0], [-1.0, 2.0]]),
              2.0**16 * np.array([[8.0, -2.0], [-4.0, 1.0]]),
              2.0**24 * np.array([[-1.0, 3.0], [0.5, -1.5]])]


def test_singular_step_is_the_minimum_norm_solution():
    """The lstsq branch of _solve_delta_dense (backends/base.py:227-230) on
    rank-1 systems, pinned by the closed form of the minimum-norm solution:
    for A = s u v^T, A^+ = A^T / |A|_F^2, so delta = -A^T r / |A|_F^2."""
    rng = np.random.default_rng(11)
    for J in SINGULAR_J:
        J_reg = J + np.eye(2) * 1e-12
        assert np.array_equal(J_reg, J)
        assert np.linalg.cond(J) > 1e8
        with pytest.raises(np.linalg.LinAlgError):
            np.linalg.solve(J_reg, np.ones(2))
        for _ in range(200):
            r = rng.standard_normal(2)
            got = Cr.solve_delta_dense(J, r)
            want = -(J.T @ r) / np.sum(J * J)
            # to rounding on the solution's scale |r| / |A|_F = |r| |A^+|
            scale = np.linalg.norm(r) / np.sqrt(np.sum(J * J))
            assert np.max(np.abs(got - want)) <= 4e-15 * scale


def test_record_layouts_match_the_library():
    lib = _lib.load(require_gpu=False)
    a, b = C.c_int(), C.c_int()
    assert lib.hh_correct_sizes(C.byref(a), C.byref(b)) == 0
    assert (a.value, b.value) == (Cr.PROBLEM_DTYPE.itemsize, Cr.RESULT_DTYPE.itemsize)


def test_pack_problems_and_validation():
    mu = W.MU_EM
    ps = [Cr.halo_problem(mu, [0.82, 0, 0.05, 0, 0.15, 0]),
          Cr.lyapunov_problem(mu, [0.83, 0, 0, 0, 0.01, 0], forward=-1, max_delta=None)]
    rec = Cr.pack_problems(ps)
    assert rec["halo"].tolist() == [1, 0] and rec["forward"].tolist() == [1, -1]
    assert rec["residual_indices"][0, :2].tolist() == [3, 5]
    assert rec["control_indices"][1, :2].tolist() == [4, 5]
    assert np.isinf(rec["max_delta"][1])
    assert Cr.line_search_steps(ps[0]) == 14
    bad = Cr.CorrectionProblem(mu, np.zeros(6), (3, 5, 1), (0, 4, 2), (0, 0, 0), None)
    assert not Cr.device_supported(bad)
    assert not Cr.device_supported(Cr.halo_problem(mu, np.zeros(6), min_alpha=1e-30))
    with pytest.raises(ValueError):
        Cr.pack_problems([bad])
    # the section plane: y by default (a zero record field), z for vertical orbits
    v = Cr.vertical_problem(mu, [0.84, 0, 0, 0, 0.006, 0.1387])
    assert (v.residual_indices, v.control_indices, v.event_index) == ((3, 1), (5, 4), 2)
    rec = Cr.pack_problems(ps + [v])
    assert rec["event_plane"].tolist() == [0, 0, 2]
    assert not Cr.device_supported(Cr.halo_problem(mu, np.zeros(6), event_index=4))


def test_vertical_orbit_on_the_oracle():
    """The oracle's sequential loop on the z = 0 section converges to a
    periodic vertical orbit of L1 (the closed orbit pins the section's
    coordinate and the residual / control choice)."""
    from oracle import oracle as O
    mu = W.MU_EM
    w = O.correct_reference(Cr.vertical_problem(mu, [0.84, 0, 0, 0, 0.006, 0.1387],
                                                max_attempts=25))
    assert w["converged"] and w["residual_norm"] < 1e-12
    x = w["x"]
    assert x[1] == x[2] == x[3] == 0.0 and x[5] > 0.1
    r = O.propagate(mu, x, 0.0, 2 * w["half_period"], steps=2)[1][-1]
    assert np.max(np.abs(r - x)) < 1e-9
    # the orbit leaves the xy-plane: at a quarter period z is far from 0
    q = O.propagate(mu, x, 0.0, 0.5 * w["half_period"], steps=2)[1][-1]
    assert abs(q[2]) > 0.02
