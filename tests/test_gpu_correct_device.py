"""GPU: the device-resident corrector (hh_correct) -- its 2x2 step against
NumPy on this host bit for bit, and whole corrections against the
host-driven bank (correct_batch) and the reference's sequential loop on the
oracle's operators, to the corrector's tolerance (the Halo term's powers are
correctly rounded on the device, the C library's pow is within 0.52 ulp:
iterates may move by an ulp)."""
import ctypes as C
import math

import numpy as np
import pytest
import torch

from hiten_amd import _lib, corrector as Cr, workloads as W

pytestmark = pytest.mark.gpu
MU = W.MU_EM
TOL_X = 1e-10          # corrected states and half periods


def say(msg):
    print(f"\n[case] {msg}", flush=True)


def _problems():
    xl1 = W.l1_position(MU)
    c2 = (1 - MU) / abs(xl1 + MU)**3 + MU / abs(xl1 - 1 + MU)**3
    wp = np.sqrt((2 - c2 + np.sqrt(9 * c2**2 - 8 * c2)) / 2)
    kap = (wp**2 + 1 + 2 * c2) / (2 * wp)
    lyap = lambda ax: [xl1 - ax, 0, 0, 0, kap * ax * wp, 0]
    ps = [
        Cr.halo_problem(MU, [0.8234, 0, 0.0224, 0, 0.1343, 0], max_attempts=20),
        Cr.halo_problem(MU, [0.84, 0, 0.1, 0, 0.22, 0], max_attempts=20),
        Cr.halo_problem(MU, [0.82, 0, 0.05, 0, 0.15, 0], max_attempts=4),     # no convergence
        Cr.lyapunov_problem(MU, lyap(0.005), max_attempts=20),
        Cr.lyapunov_problem(MU, lyap(0.01), max_attempts=20),                # Armijo failure
        Cr.halo_problem(MU, [0.8234, 0, 0.0224, 0, 0.1343, 0], max_attempts=20,
                        alpha_reduction=0.7, armijo_c=0.3),                 # 26 trial steps
        Cr.halo_problem(MU, [3.0, 0, 0.0, 0, 0.0, 0.0], max_attempts=5),       # far away
    ]
    rng = np.random.default_rng(20261202)
    for _ in range(24):      # seeds drawn as the serving stream draws them
        ps.append(Cr.halo_problem(MU, [rng.uniform(0.80, 0.88), 0, rng.uniform(-0.2, 0.2), 0,
                                       rng.uniform(-0.3, 0.3), 0], max_attempts=10))
    return ps


def _same_outcome(g, w):
    assert g.converged == w.converged
    assert g.error.split(" ")[0] == w.error.split(" ")[0]
    if g.converged:
        assert np.max(np.abs(g.x_corrected - w.x_corrected)) <= TOL_X
        assert abs(g.half_period - w.half_period) <= TOL_X
        assert g.residual_norm < 1e-12 and w.residual_norm < 1e-12


def test_step_solve_is_numpys_bit_for_bit(gpu):
    say("hh_eval_solve2: 20000 random 2x2 systems and 400 near the cond threshold vs NumPy")
    rng = np.random.default_rng(7)
    n = 20000
    J = rng.standard_normal((n, 2, 2)) * np.exp(rng.uniform(-8, 8, (n, 1, 1)))
    r = rng.standard_normal((n, 2)) * np.exp(rng.uniform(-8, 8, (n, 1)))
    extra = []
    for _ in range(400):        # regularised and not, either side of cond = 1e8
        s = 10.0 ** rng.uniform(7.5, 8.5)
        U, _ = np.linalg.qr(rng.standard_normal((2, 2)))
        V, _ = np.linalg.qr(rng.standard_normal((2, 2)))
        extra.append(U @ np.diag([1.0, 1.0 / s]) @ V.T)
    J = np.concatenate([J, np.stack(extra)])
    r = np.concatenate([r, rng.standard_normal((400, 2))])
    n = len(J)
    dev = torch.device("cuda:0")
    Jd = torch.from_numpy(J.reshape(n, 4).copy()).to(dev)
    rd = torch.from_numpy(r.copy()).to(dev)
    dd = torch.empty((n, 2), dtype=torch.float64, device=dev)
    cd = torch.empty(n, dtype=torch.float64, device=dev)
    lib = _lib.load()
    _lib.check(lib.hh_eval_solve2(n, Jd.data_ptr(), rd.data_ptr(), dd.data_ptr(), cd.data_ptr(),
                                  None), "hh_eval_solve2")
    d, c = dd.cpu().numpy(), cd.cpu().numpy()
    want = Cr.solve_delta_stack(J, r)
    assert np.array_equal(d, want)
    cw = np.linalg.cond(J)
    # NumPy's cond (dgesdd) carries about eps * cond relative error in
    # sigma_min, the device's accurate determinant does not: compare to that
    assert np.allclose(c, cw, rtol=1e-15 * np.maximum(cw, 1.0) + 1e-13)
    assert ((c > 1e8) == (cw > 1e8)).all()


def test_step_solve_singular_is_lstsq(gpu):
    """Exactly singular regularised systems take the least-squares branch:
    the device's closed-form minimum-norm solution (one Jacobi rotation)
    against numpy.linalg.lstsq (LAPACK dgelsd), which forms it differently:
    equal to rounding, not bit for bit."""
    from test_correct_device_cpu import SINGULAR_J
    say(f"hh_eval_solve2: {len(SINGULAR_J)} exactly singular systems x 20 residuals vs lstsq")
    rng = np.random.default_rng(12)
    J = np.stack([j for j in SINGULAR_J for _ in range(20)])
    r = rng.standard_normal((len(J), 2)) * np.exp(rng.uniform(-6, 6, (len(J), 1)))
    n = len(J)
    dev = torch.device("cuda:0")
    Jd = torch.from_numpy(J.reshape(n, 4).copy()).to(dev)
    rd = torch.from_numpy(r.copy()).to(dev)
    dd = torch.empty((n, 2), dtype=torch.float64, device=dev)
    cd = torch.empty(n, dtype=torch.float64, device=dev)
    lib = _lib.load()
    _lib.check(lib.hh_eval_solve2(n, Jd.data_ptr(), rd.data_ptr(), dd.data_ptr(), cd.data_ptr(),
                                  None), "hh_eval_solve2")
    d, c = dd.cpu().numpy(), cd.cpu().numpy()
    assert np.isinf(c).all()
    for i in range(n):
        want = Cr.solve_delta_dense(J[i], r[i])
        # to rounding on the solution's scale |r| / |A|_F = |r| |A^+|
        scale = np.linalg.norm(r[i]) / np.sqrt(np.sum(J[i] * J[i]))
        assert np.max(np.abs(d[i] - want)) <= 2e-14 * scale, i


def test_device_corrector_matches_the_bank_and_the_reference(gpu, oracle):
    ps = _problems()
    say(f"hh_correct: {len(ps)} problems vs correct_batch and the sequential reference loop")
    got = Cr.correct_device(ps)
    bank = Cr.correct_batch(ps)
    same_iters = 0
    for i, (g, b) in enumerate(zip(got, bank)):
        _same_outcome(g, b)
        same_iters += (g.iterations == b.iterations)
    assert same_iters >= len(ps) - 2
    kinds = {(g.converged, g.error.split(" ")[0]) for g in got}
    assert {(True, ""), (False, "Newton"), (False, "Step")} <= kinds
    for p, g in list(zip(ps, got))[:7]:
        w = oracle.correct_reference(p)
        assert g.converged == w["converged"]
        if g.converged:
            assert np.max(np.abs(g.x_corrected - w["x"])) <= TOL_X
            assert abs(g.half_period - w["half_period"]) <= TOL_X


def test_alone_and_in_a_batch_are_identical(gpu):
    ps = _problems()
    say("hh_correct: each problem alone vs inside the batch, bit for bit")
    together = Cr.correct_device(ps)
    for i in (0, 3, 5, 9, len(ps) - 1):
        alone = Cr.correct_device([ps[i]])[0]
        t = together[i]
        assert np.array_equal(alone.x_corrected, t.x_corrected)
        assert alone.iterations == t.iterations and alone.error == t.error
        assert alone.n_event_searches == t.n_event_searches
        assert (alone.half_period == t.half_period) or (math.isnan(alone.half_period)
                                                        and math.isnan(t.half_period))


def test_invalid_record_and_done_flags(gpu):
    say("hh_correct: a record the kernel must refuse, and every done flag set")
    ps = _problems()[:3]
    rec = Cr.pack_problems(ps)
    rec[1]["n_residuals"] = 3
    dev = torch.device("cuda:0")
    prob = torch.from_numpy(rec.view(np.uint8).copy()).to(dev)
    res = torch.zeros(3 * Cr.RESULT_DTYPE.itemsize, dtype=torch.uint8, device=dev)
    Cr.launch_device(prob.data_ptr(), res.data_ptr(), 3)
    out = res.cpu().numpy().view(Cr.RESULT_DTYPE)
    assert out["done"].tolist() == [1, 1, 1]
    assert out["error"][1] == Cr.HH_CORRECT_INVALID and out["converged"][1] == 0
    assert np.array_equal(out["x"][1], rec["x0"][1])
    good = Cr.results_from_records(out[[0, 2]], [ps[0], ps[2]])
    want = Cr.correct_device([ps[0], ps[2]])
    for g, w in zip(good, want):
        assert np.array_equal(g.x_corrected, w.x_corrected)


def _vertical_problems():
    # seeds about a vertical orbit of L1 (x = 0.84): converged, Newton out of
    # iterations and Armijo failures, each on the z = 0 section
    return [Cr.vertical_problem(MU, [x0, 0, 0, 0, vy, vz], max_attempts=25)
            for x0, vy, vz in ((0.84, 0.006, 0.1387), (0.84, 0.0, 0.1387), (0.84, 0.02, 0.1687),
                               (0.835, 0.0, 0.1387), (0.845, 0.0, 0.1387),
                               (0.86, 0.0, 0.1187))]


def test_vertical_orbits_on_the_z_section(gpu, oracle):
    ps = _vertical_problems()
    say(f"vertical orbits (_z_plane_crossing): {len(ps)} problems, device vs bank vs oracle")
    got = Cr.correct_device(ps)
    bank = Cr.correct_batch(ps)
    kinds = set()
    for p, g, b in zip(ps, got, bank):
        w = oracle.correct_reference(p)
        # no Halo term: the device, the bank and the reference loop compute
        # the same bits
        assert np.array_equal(g.x_corrected, b.x_corrected)
        assert g.iterations == b.iterations == w["iterations"]
        assert g.error == b.error and g.converged == b.converged == w["converged"]
        assert np.array_equal(b.x_corrected, w["x"])
        kinds.add((g.converged, g.error.split(" ")[0]))
        if g.converged:
            assert g.half_period == b.half_period == w["half_period"]
            # a periodic orbit: the state returns after two half periods
            r = oracle.propagate(MU, g.x_corrected, 0.0, 2 * g.half_period, steps=2)[1][-1]
            assert np.max(np.abs(r - g.x_corrected)) < 1e-9
    assert {(True, ""), (False, "Newton"), (False, "Step")} <= kinds


def test_unknown_event_plane_is_refused(gpu):
    say("hh_correct: a record with an unknown event plane is refused")
    rec = Cr.pack_problems(_vertical_problems()[:2])
    assert rec["event_plane"].tolist() == [2, 2]
    rec[0]["event_plane"] = 7
    dev = torch.device("cuda:0")
    prob = torch.from_numpy(rec.view(np.uint8).copy()).to(dev)
    res = torch.zeros(2 * Cr.RESULT_DTYPE.itemsize, dtype=torch.uint8, device=dev)
    Cr.launch_device(prob.data_ptr(), res.data_ptr(), 2)
    out = res.cpu().numpy().view(Cr.RESULT_DTYPE)
    assert out["done"].tolist() == [1, 1]
    assert out["error"][0] == Cr.HH_CORRECT_INVALID and out["error"][1] != Cr.HH_CORRECT_INVALID


def test_device_and_bank_against_the_reference_corrector(gpu):
    """Both correctors on the GPU kernels against the reference's own
    corrector (tests/golden/corrector_golden.npz): the same outcome for every
    problem (converged, stopped at max_attempts, line search failed), and
    converged orbits and half periods to 1e-10."""
    from tests.golden_corrector_cases import problems
    from tests.test_golden_corrector import _kind, _ref
    ps = problems()
    say(f"device corrector and bank vs the reference corrector: {len(ps)} problems")
    for res in (Cr.correct_device(ps), Cr.correct_batch(ps)):
        for i, g in enumerate(res):
            ref = _ref(i)
            assert g.converged == ref["converged"], i
            assert _kind(g.error) == _kind(ref["error"]), (i, g.error, ref["error"])
            if ref["converged"]:
                assert np.abs(g.x_corrected - ref["x"]).max() <= TOL_X, i
                assert abs(g.half_period - ref["half_period"]) <= TOL_X, i
                assert g.residual_norm < 1e-12

