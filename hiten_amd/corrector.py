"""Single-shooting differential correction of periodic orbits on the GPU.

Restates the reference's Newton corrector for the CR3BP orbit families:

* residual / Jacobian operators   corrector/operators.py:319-452
  (residual = x_event[res_idx] - target at the first y = 0 crossing;
  Jacobian = Phi(t_event)[res_idx][:, ctrl_idx] - extra term)
* Newton loop                     corrector/backends/newton.py:39-143
* dense solve with regularisation corrector/backends/base.py:187-245
* Armijo backtracking             corrector/stepping/armijo.py:58-164
* infinity norm                   corrector/interfaces.py:44-46
* Halo extra Jacobian term        algorithms/types/services/orbits.py:917-947
* family defaults                 services/orbits.py:851-915 (Halo),
                                  1238-1302 (Lyapunov)
* half period                     corrector/interfaces.py:48-74

``correct_batch`` runs many independent corrections in lock-step: each
Newton phase of every active orbit -- event search, STM propagation,
line-search trials -- is one lane of one kernel launch.  The Jacobian
repeats the residual's event search in the reference
(operators.py:440); that search is a pure function of the same state, so the
result of the residual's search is reused (bit-identical).
"""

from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass, field
from typing import Callable, Optional, Sequence

import numpy as np

from . import batch
from .dynamics import rtbp_dynsys
from .singlehit import _PlaneEvent, section_crossing_batch

X, Y, Z, VX, VY, VZ = range(6)


class ConvergenceError(RuntimeError):
    """types/exceptions.py: Newton did not converge."""


class BackendError(RuntimeError):
    """types/exceptions.py: the line search found no productive step."""


def halo_quadratic_term(mu: float, X_ev: np.ndarray, Phi: np.ndarray) -> np.ndarray:
    """services/orbits.py:917-947."""
    x, y, z, vx, vy, vz = X_ev
    mu2 = 1 - mu
    rho_1 = 1 / (((x + mu)**2 + y**2 + z**2)**1.5)
    rho_2 = 1 / (((x - mu2)**2 + y**2 + z**2)**1.5)
    omega_x = -(mu2 * (x + mu) * rho_1) - (mu * (x - mu2) * rho_2) + x
    DDx = 2 * vy + omega_x
    DDz = -(mu2 * z * rho_1) - (mu * z * rho_2)
    if abs(vy) < 1e-9:
        vy = np.sign(vy) * 1e-9 if vy != 0 else 1e-9
    return np.array([[DDx], [DDz]]) @ Phi[[Y], :][:, (X, VY)] / vy


@dataclass
class CorrectionProblem:
    """One orbit to correct (OrbitCorrectionConfig + OrbitCorrectionOptions)."""
    mu: float
    x0: np.ndarray
    residual_indices: Sequence[int] = (VX, VZ)
    control_indices: Sequence[int] = (X, VY)
    target: Sequence[float] = (0.0, 0.0)
    extra_jacobian: Optional[str] = "halo"     # "halo" or None
    max_attempts: int = 50
    tol: float = 1e-12
    max_delta: float = 1e-2
    alpha_reduction: float = 0.5
    min_alpha: float = 1e-4
    armijo_c: float = 0.1
    forward: int = 1
    event_index: int = Y        # section y = 0 (Halo, Lyapunov); Z for vertical orbits


def halo_problem(mu: float, x0, **kw) -> CorrectionProblem:
    """Halo defaults (services/orbits.py:851-915)."""
    return CorrectionProblem(mu, np.asarray(x0, float), (VX, VZ), (X, VY), (0.0, 0.0), "halo", **kw)


def lyapunov_problem(mu: float, x0, **kw) -> CorrectionProblem:
    """Lyapunov defaults (services/orbits.py:1238-1302)."""
    return CorrectionProblem(mu, np.asarray(x0, float), (VX, Z), (VY, VZ), (0.0, 0.0), None, **kw)


def vertical_problem(mu: float, x0, **kw) -> CorrectionProblem:
    """Vertical-orbit defaults (services/orbits.py:1429-1446): VX = Y = 0 at
    the z = 0 crossing (_z_plane_crossing), controls VZ and VY."""
    kw.setdefault("event_index", Z)
    return CorrectionProblem(mu, np.asarray(x0, float), (VX, Y), (VZ, VY), (0.0, 0.0), None, **kw)


@dataclass
class CorrectionResult:
    converged: bool
    x_corrected: np.ndarray
    residual_norm: float
    iterations: int
    half_period: float = math.nan
    error: str = ""
    n_event_searches: int = 0
    n_stm: int = 0

    @property
    def period(self) -> float:
        return 2.0 * self.half_period


def _inf_norm(r: np.ndarray) -> float:
    return float(np.linalg.norm(r, ord=np.inf))


def solve_delta_dense(J: np.ndarray, r: np.ndarray, cond_threshold: float = 1e8) -> np.ndarray:
    """backends/base.py:187-245."""
    try:
        cond_J = np.linalg.cond(J)
    except np.linalg.LinAlgError:
        cond_J = np.inf
    if J.shape[0] == J.shape[1]:
        if np.isnan(cond_J) or cond_J > cond_threshold:
            J_reg = J + np.eye(J.shape[0]) * 1e-12
        else:
            J_reg = J
        try:
            return np.linalg.solve(J_reg, -r)
        except np.linalg.LinAlgError:
            return np.linalg.lstsq(J_reg, -r, rcond=None)[0]
    lam = 1e-12 if (np.isnan(cond_J) or cond_J > cond_threshold) else 0.0
    JTJ = J.T @ J + lam * np.eye(J.shape[1])
    JTr = J.T @ r
    try:
        return np.linalg.solve(JTJ, -JTr)
    except np.linalg.LinAlgError:
        return np.linalg.lstsq(J, -r, rcond=None)[0]


class _Engine:
    """GPU evaluation of the operators for a set of orbits."""

    def event_batch(self, mu: np.ndarray, states: np.ndarray, forward: int, index: int = Y):
        """_y_plane_crossing (the crossing of plane y[index] = 0) for each
        row: (t (n,), x_event (n, 6), ok (n,))."""
        n = states.shape[0]
        t = np.full(n, np.nan)
        xe = np.full((n, 6), np.nan)
        ok = np.zeros(n, bool)
        for m in np.unique(mu):
            sel = np.nonzero(mu == m)[0]
            hits = section_crossing_batch(rtbp_dynsys(float(m)), states[sel],
                                          surface=_PlaneEvent("xyz"[index], 0.0, None),
                                          forward=forward)
            for j, h in zip(sel, hits):
                if h is not None:
                    t[j], xe[j], ok[j] = h.time, h.state, True
        return t, xe, ok

    def stm_batch(self, mu: np.ndarray, states: np.ndarray, t_event: np.ndarray) -> np.ndarray:
        """Phi(t_event) of _compute_stm(var_dynsys, x, t_event, steps=500)
        (operators.py:298-307); the last grid point is the dense value at t_event."""
        n = states.shape[0]
        PHI0 = np.zeros((n, 42))
        PHI0[:, :36] = np.eye(6).ravel()
        PHI0[:, 36:] = states
        const = np.isclose(0.0, t_event)
        phi = np.repeat(np.eye(6)[None], n, axis=0)
        live = ~const
        if live.any():
            # two waves per trajectory while the batch is small (latency), one
            # wave per trajectory up to 2048, else the work-stealing form; all
            # give the same bits
            k = int(live.sum())
            variant = "pair" if k <= 1024 else ("wide" if k <= 2048 else "persistent_pair")
            r = batch.dop853_final(mu[live], PHI0[live], 0.0, t_event[live], dim=42,
                                   flip_indices=slice(36, 42), max_step=1e4, variant=variant)
            phi[live] = r.y[:, :36].reshape(-1, 6, 6).cpu().numpy()
        return phi


def solve_delta_stack(J: np.ndarray, r: np.ndarray, cond_threshold: float = 1e8) -> np.ndarray:
    """solve_delta_dense for a stack of square systems; the stacked LAPACK
    calls are the per-matrix ones.  A stack that raises falls back to the
    per-matrix function."""
    if J.shape[1] != J.shape[2]:
        return np.stack([solve_delta_dense(a, b, cond_threshold) for a, b in zip(J, r)])
    try:
        cond_J = np.linalg.cond(J)
        reg = np.isnan(cond_J) | (cond_J > cond_threshold)
        J_reg = np.where(reg[:, None, None], J + np.eye(J.shape[1]) * 1e-12, J)
        return np.linalg.solve(J_reg, -r[:, :, None])[:, :, 0]
    except np.linalg.LinAlgError:
        return np.stack([solve_delta_dense(a, b, cond_threshold) for a, b in zip(J, r)])


HH_BANK_MAX_ITERATIONS = 2
HH_BANK_LINE_SEARCH = 3
_ERRORS = {
    1: "no section crossing for the current state",
}


def _line_search_error(k: int) -> str:
    """The reference's ConvergenceError text when the Armijo step finds no
    productive step at Newton iteration k (backends/newton.py:125-130 around
    stepping/armijo.py:160-164)."""
    return (f"Step strategy failed to produce an update at iter {k}: "
            "Armijo line search failed to find a productive step.")


def _i32(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.int32)


def _f64(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.float64)


def _p(a: np.ndarray) -> int:
    return a.ctypes.data


class _Bank:
    """Corrections advanced together, one row each, in the library's native
    state machine (hh_bank_*): _NewtonBackend.run with its Armijo line search.

    Stages: RESID waits for the event search at x (the residual), JAC for
    the STM to the event time, then the dense solve (NumPy, stacked), then
    ARMIJO for the event searches of every trial step alpha = 1, r, r^2, ...
    >= min_alpha.  The line search is evaluated speculatively -- all trials in
    one batch -- and armijo.py:111-164 is replayed over them in order, so the
    accepted step, the fallback to the best point and the failure are the
    sequential algorithm's; the residual of the accepted trial is the
    residual at the next iterate and is reused.  Rows of one bank share the
    numbers of residual and control indices.
    """

    RESID, JAC, STEP, ARMIJO, DONE = 0, 1, 2, 3, 4

    def __init__(self, nr: int, nc: int):
        from . import _lib
        self._l = _lib.load(require_gpu=False)
        self._h = self._l.hh_bank_create(int(nr), int(nc))
        if not self._h:
            raise ValueError(f"a bank holds 1..6 residuals and controls, not ({nr}, {nc})")
        self.nr, self.nc = nr, nc
        self.problems: list = []          # by row (None once released)
        self.mu = np.zeros(0)             # by row
        self.fwd = np.zeros(0, dtype=np.int64)
        self.evi = np.zeros(0, dtype=np.int64)   # section plane coordinate by row

    def __del__(self):
        h, self._h = getattr(self, "_h", None), None
        if h:
            self._l.hh_bank_destroy(h)

    @property
    def n(self) -> int:
        return len(self.problems)

    def _check(self, rc: int, what: str) -> None:
        if rc != 0:
            raise RuntimeError(f"{what} failed (code {rc})")

    def add(self, p: CorrectionProblem) -> int:
        if int(p.event_index) not in (X, Y, Z):
            raise ValueError("the section plane is x, y or z = 0 (event_index 0, 1 or 2)")
        res, ctrl = _i32(list(p.residual_indices)), _i32(list(p.control_indices))
        x0, tgt = _f64(p.x0), _f64(p.target)
        md = math.inf if p.max_delta is None else float(p.max_delta)
        row = self._l.hh_bank_add(self._h, float(p.mu), _p(x0), _p(res), _p(ctrl), _p(tgt),
                                  int(p.extra_jacobian == "halo"), int(p.forward),
                                  int(p.max_attempts), float(p.tol), md, float(p.alpha_reduction),
                                  float(p.min_alpha), float(p.armijo_c))
        if row < 0:
            raise ValueError("invalid correction problem for the native bank")
        if row == len(self.problems):
            self.problems.append(p)
            if row >= len(self.mu):
                cap = max(64, 2 * len(self.mu))
                self.mu = np.concatenate([self.mu, np.zeros(cap - len(self.mu))])
                self.fwd = np.concatenate([self.fwd, np.zeros(cap - len(self.fwd), np.int64)])
                self.evi = np.concatenate([self.evi, np.zeros(cap - len(self.evi), np.int64)])
        else:
            self.problems[row] = p
        self.mu[row] = float(p.mu)
        self.fwd[row] = int(p.forward)
        self.evi[row] = int(p.event_index)
        return row

    def release(self, row: int) -> None:
        """Give a finished row (result already read) back for reuse."""
        self._check(self._l.hh_bank_release(self._h, int(row)), "hh_bank_release")
        self.problems[row] = None

    def stages(self, rows) -> np.ndarray:
        rows = _i32(rows)
        out = np.empty(len(rows), dtype=np.int32)
        self._check(self._l.hh_bank_stages(self._h, len(rows), _p(rows), _p(out)), "hh_bank_stages")
        return out

    def event_inputs(self, rows):
        """States whose event searches the rows wait for (grouped by row) and
        per-row counts."""
        rows = _i32(rows)
        counts = np.empty(len(rows), dtype=np.int32)
        total = self._l.hh_bank_event_inputs(self._h, len(rows), _p(rows), _p(counts), None, 0)
        if total < 0:
            raise RuntimeError("hh_bank_event_inputs failed")
        states = np.empty((total, 6))
        got = self._l.hh_bank_event_inputs(self._h, len(rows), _p(rows), _p(counts),
                                           _p(states), total)
        if got != total:
            raise RuntimeError("hh_bank_event_inputs failed")
        return states, counts

    def feed_events(self, rows, t, xe, ok) -> None:
        rows = _i32(rows)
        # converted arrays are held by names for the call: a temporary would
        # be freed before the library reads it
        t, xe, ok = _f64(t), _f64(xe), _i32(ok)
        self._check(self._l.hh_bank_feed_events(self._h, len(rows), _p(rows), _p(t), _p(xe),
                                                _p(ok)), "hh_bank_feed_events")

    def stm_inputs(self, rows):
        rows = _i32(rows)
        states, tev = np.empty((len(rows), 6)), np.empty(len(rows))
        self._check(self._l.hh_bank_stm_inputs(self._h, len(rows), _p(rows), _p(states), _p(tev)),
                    "hh_bank_stm_inputs")
        return states, tev

    def feed_stm(self, rows, Phi) -> None:
        """Phi(t_event) (n, 6, 6): Jacobians with the Halo term (native),
        dense solves (NumPy, as backends/base.py:187-245), trial steps."""
        rows = _i32(rows)
        n = len(rows)
        J, r = np.empty((n, self.nr, self.nc)), np.empty((n, self.nr))
        Phi = _f64(Phi)
        self._check(self._l.hh_bank_jacobians(self._h, n, _p(rows), _p(Phi), _p(J), _p(r)),
                    "hh_bank_jacobians")
        d = _f64(solve_delta_stack(J, r))
        self._check(self._l.hh_bank_set_steps(self._h, n, _p(rows), _p(d)), "hh_bank_set_steps")

    def result(self, row: int) -> CorrectionResult:
        x = np.empty(6)
        rn, hp = C.c_double(), C.c_double()
        k, ne, ns, ok, err = (C.c_int() for _ in range(5))
        self._check(self._l.hh_bank_result(self._h, int(row), _p(x), C.byref(rn), C.byref(k),
                                           C.byref(ne), C.byref(ns), C.byref(ok), C.byref(err),
                                           C.byref(hp)), "hh_bank_result")
        p = self.problems[row]
        if err.value == HH_BANK_MAX_ITERATIONS:
            msg = (f"Newton did not converge after {p.max_attempts} iterations "
                   f"(|R|={rn.value:.2e}).")
        elif err.value == HH_BANK_LINE_SEARCH:
            msg = _line_search_error(k.value)
        else:
            msg = _ERRORS.get(err.value, "")
        return CorrectionResult(bool(ok.value), x, rn.value, k.value, half_period=hp.value,
                                error=msg, n_event_searches=ne.value, n_stm=ns.value)


def _drive(bank: _Bank, eng: "_Engine") -> None:
    """Run every row of a bank to completion, one batched launch per
    request kind per round."""
    rows_all = np.arange(bank.n)
    mu, fwd, evi = bank.mu, bank.fwd, bank.evi
    while True:
        st = bank.stages(rows_all)
        live = rows_all[st != bank.DONE]
        if live.size == 0:
            return
        sl = st[st != bank.DONE]
        ev = live[sl != bank.JAC]
        jac = live[sl == bank.JAC]
        for f, e in sorted(set(zip(fwd[ev].tolist(), evi[ev].tolist()))):
            rows = ev[(fwd[ev] == f) & (evi[ev] == e)]
            states, counts = bank.event_inputs(rows)
            t, xe, ok = eng.event_batch(np.repeat(mu[rows], counts), states, int(f), index=int(e))
            bank.feed_events(rows, t, xe, ok)
        if jac.size:
            states, tev = bank.stm_inputs(jac)
            bank.feed_stm(jac, eng.stm_batch(mu[jac], states, tev))


def correct_batch(problems: Sequence[CorrectionProblem], engine: _Engine | None = None,
                  compute_half_period: bool = True) -> list[CorrectionResult]:
    """Newton-Armijo for every problem; each round of the corrections'
    event searches (residuals, all line-search trials) and each round of
    their STMs is one batched launch."""
    eng = engine or _Engine()
    banks: dict = {}
    where = []
    for p in problems:
        key = (len(p.residual_indices), len(p.control_indices))
        b = banks.get(key)
        if b is None:
            b = banks[key] = _Bank(*key)
        where.append((b, b.add(p)))
    for b in banks.values():
        _drive(b, eng)
    out = [b.result(i) for b, i in where]
    if not compute_half_period:
        for r in out:
            r.half_period = math.nan
    return out


# ---- device-resident corrector (hh_correct) --------------------------------

#: hh_correction_t / hh_correction_result_t (include/hiten_hip.h), C layout
PROBLEM_DTYPE = np.dtype([
    ("mu", "f8"), ("x0", "f8", 6), ("target", "f8", 6), ("tol", "f8"), ("max_delta", "f8"),
    ("alpha_reduction", "f8"), ("min_alpha", "f8"), ("armijo_c", "f8"),
    ("residual_indices", "i4", 6), ("control_indices", "i4", 6), ("n_residuals", "i4"),
    ("n_controls", "i4"), ("halo", "i4"), ("forward", "i4"), ("max_attempts", "i4"),
    ("event_plane", "i4")], align=True)
RESULT_DTYPE = np.dtype([
    ("x", "f8", 6), ("residual_norm", "f8"), ("half_period", "f8"), ("iterations", "i4"),
    ("n_event_searches", "i4"), ("n_stm", "i4"), ("converged", "i4"), ("error", "i4"),
    ("n_steps", "i4"), ("done", "i4"), ("reserved", "i4")], align=True)
HH_CORRECT_INVALID = 4
_JMAX = 64
#: hh_correction_t.event_plane of a section coordinate (HH_PLANE_Y / _X / _Z)
EVENT_PLANE = {Y: 0, X: 1, Z: 2}


def line_search_steps(p: CorrectionProblem) -> int:
    """Number of trial steps alpha = 1, r, r^2, ... >= min_alpha."""
    a, n = 1.0, 0
    while a >= p.min_alpha and n <= _JMAX:
        n += 1
        a = a * p.alpha_reduction
    return n


def device_supported(p: CorrectionProblem) -> bool:
    """Whether hh_correct takes the problem: a square 2x2 system (Halo,
    Lyapunov, vertical) on an x, y or z section with at most 64 line-search
    steps."""
    return (len(p.residual_indices) == 2 and len(p.control_indices) == 2
            and all(0 <= int(i) <= 5 for i in (*p.residual_indices, *p.control_indices))
            and int(p.event_index) in EVENT_PLANE and line_search_steps(p) <= _JMAX)


def pack_problems(problems: Sequence[CorrectionProblem]) -> np.ndarray:
    """hh_correction_t records of the problems."""
    out = np.zeros(len(problems), dtype=PROBLEM_DTYPE)
    for i, p in enumerate(problems):
        if not device_supported(p):
            raise ValueError(f"problem {i}: the device corrector takes 2x2 problems with at "
                             f"most {_JMAX} line-search steps")
        r = out[i]
        r["mu"] = float(p.mu)
        r["x0"] = np.asarray(p.x0, float)
        r["target"][:2] = np.asarray(p.target, float)
        r["tol"] = float(p.tol)
        r["max_delta"] = math.inf if p.max_delta is None else float(p.max_delta)
        r["alpha_reduction"] = float(p.alpha_reduction)
        r["min_alpha"] = float(p.min_alpha)
        r["armijo_c"] = float(p.armijo_c)
        r["residual_indices"][:2] = list(p.residual_indices)
        r["control_indices"][:2] = list(p.control_indices)
        r["n_residuals"] = r["n_controls"] = 2
        r["halo"] = int(p.extra_jacobian == "halo")
        r["forward"] = int(p.forward)
        r["max_attempts"] = int(p.max_attempts)
        r["event_plane"] = EVENT_PLANE[int(p.event_index)]
    return out


def results_from_records(rec: np.ndarray, problems: Sequence[CorrectionProblem]
                         ) -> list[CorrectionResult]:
    """CorrectionResult of each hh_correction_result_t record."""
    out = []
    for r, p in zip(rec, problems):
        err = int(r["error"])
        if err == HH_BANK_MAX_ITERATIONS:
            msg = (f"Newton did not converge after {p.max_attempts} iterations "
                   f"(|R|={float(r['residual_norm']):.2e}).")
        elif err == HH_CORRECT_INVALID:
            msg = "invalid correction problem"
        elif err == HH_BANK_LINE_SEARCH:
            msg = _line_search_error(int(r["iterations"]))
        else:
            msg = _ERRORS.get(err, "")
        out.append(CorrectionResult(bool(r["converged"]), np.array(r["x"], dtype=float),
                                    float(r["residual_norm"]), int(r["iterations"]),
                                    half_period=float(r["half_period"]), error=msg,
                                    n_event_searches=int(r["n_event_searches"]),
                                    n_stm=int(r["n_stm"])))
    return out


def launch_device(prob_ptr: int, res_ptr: int, n: int, stream=None,
                  stream_handle: Optional[int] = None) -> None:
    """hh_correct on device (or mapped) records: ``n`` problems at
    ``prob_ptr``, results to ``res_ptr`` (their done flags cleared); on a
    torch stream or a raw stream handle."""
    from . import _lib
    from .singlehit import _default_windows
    (ts, span), (ts2, span2) = _default_windows()
    lib = _lib.load()
    sh = _lib.stream_handle(stream) if stream_handle is None else stream_handle
    _lib.check(lib.hh_correct(int(n), prob_ptr, res_ptr, ts, span, ts2, span2, sh),
               "hh_correct")


def correct_device(problems: Sequence[CorrectionProblem], stream=None) -> list[CorrectionResult]:
    """Every problem's whole Newton-Armijo loop in one launch (hh_correct):
    one workgroup per correction, no host round trip between iterations.
    Agrees with correct_batch to the corrector's tolerance (the Halo term's
    powers are correctly rounded on the device)."""
    import torch
    n = len(problems)
    if n == 0:
        return []
    dev = batch._device()
    rec = pack_problems(problems)
    prob = torch.from_numpy(rec.view(np.uint8)).to(dev)
    res = torch.zeros(n * RESULT_DTYPE.itemsize, dtype=torch.uint8, device=dev)
    launch_device(prob.data_ptr(), res.data_ptr(), n, stream)
    host = res.cpu().numpy().view(RESULT_DTYPE)
    return results_from_records(host, problems)


def correct(problem: CorrectionProblem, device: bool = False, **kw) -> CorrectionResult:
    """Correct one orbit; raises ConvergenceError like _NewtonBackend.run.
    ``device=True`` runs the whole loop in one launch (correct_device)."""
    r = correct_device([problem])[0] if device else correct_batch([problem], **kw)[0]
    if not r.converged:
        raise ConvergenceError(r.error)
    return r
