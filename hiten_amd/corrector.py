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


def halo_problem(mu: float, x0, **kw) -> CorrectionProblem:
    """Halo defaults (services/orbits.py:851-915)."""
    return CorrectionProblem(mu, np.asarray(x0, float), (VX, VZ), (X, VY), (0.0, 0.0), "halo", **kw)


def lyapunov_problem(mu: float, x0, **kw) -> CorrectionProblem:
    """Lyapunov defaults (services/orbits.py:1238-1302)."""
    return CorrectionProblem(mu, np.asarray(x0, float), (VX, Z), (VY, VZ), (0.0, 0.0), None, **kw)


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

    def event_batch(self, mu: np.ndarray, states: np.ndarray, forward: int):
        """_y_plane_crossing for each row: (t (n,), x_event (n, 6), ok (n,))."""
        n = states.shape[0]
        t = np.full(n, np.nan)
        xe = np.full((n, 6), np.nan)
        ok = np.zeros(n, bool)
        for m in np.unique(mu):
            sel = np.nonzero(mu == m)[0]
            hits = section_crossing_batch(rtbp_dynsys(float(m)), states[sel],
                                          surface=_PlaneEvent("y", 0.0, None), forward=forward)
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
            # one wave per trajectory while the batch is small (latency), else
            # the work-stealing form; both give the same bits
            variant = "wide" if int(live.sum()) <= 2048 else "persistent"
            r = batch.dop853_final(mu[live], PHI0[live], 0.0, t_event[live], dim=42,
                                   flip_indices=slice(36, 42), max_step=1e4, variant=variant)
            phi[live] = r.y[:, :36].reshape(-1, 6, 6).cpu().numpy()
        return phi


def _spow(a: np.ndarray, e: float) -> np.ndarray:
    """Elementwise a**e with NumPy's scalar power (the C library pow), which
    is what the reference's per-orbit scalar code evaluates; NumPy's array
    power may round differently."""
    return np.array([v ** e for v in np.asarray(a, dtype=np.float64)], dtype=np.float64)


def halo_quadratic_terms(mu: np.ndarray, X_ev: np.ndarray, Phi: np.ndarray) -> np.ndarray:
    """halo_quadratic_term for a stack of orbits: (n, 2, 2), element for
    element the operations of services/orbits.py:917-947."""
    x, y, z, vx, vy, vz = (X_ev[:, i] for i in range(6))
    mu2 = 1 - mu
    rho_1 = 1 / _spow(_spow(x + mu, 2) + _spow(y, 2) + _spow(z, 2), 1.5)
    rho_2 = 1 / _spow(_spow(x - mu2, 2) + _spow(y, 2) + _spow(z, 2), 1.5)
    omega_x = -(mu2 * (x + mu) * rho_1) - (mu * (x - mu2) * rho_2) + x
    DDx = 2 * vy + omega_x
    DDz = -(mu2 * z * rho_1) - (mu * z * rho_2)
    small = np.abs(vy) < 1e-9
    vy = np.where(small, np.where(vy != 0, np.sign(vy) * 1e-9, 1e-9), vy)
    DD = np.stack([DDx, DDz], axis=1)                      # (n, 2)
    P = Phi[:, Y, :][:, [X, VY]]                           # (n, 2)
    return (DD[:, :, None] * P[:, None, :]) / vy[:, None, None]


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


class _Bank:
    """Corrections advanced together, one row each: _NewtonBackend.run as a
    vectorized state machine.

    Stages: RESID waits for the event search at x (the residual), JAC for
    the STM to the event time, ARMIJO for the event searches of every trial
    step alpha = 1, r, r^2, ... >= min_alpha.  The line search is evaluated
    speculatively -- all trials in one batch -- and armijo.py:111-164 is
    replayed over them in order, so the accepted step, the fallback to the
    best point and the failure are the sequential algorithm's.  The residual
    of the accepted trial is the residual at the next iterate (the same state
    through the same deterministic search) and is reused.  Every arithmetic
    step is the reference's, element for element.  Rows of one bank share the
    numbers of residual and control indices.
    """

    RESID, JAC, ARMIJO, DONE = 0, 1, 2, 3

    def __init__(self, nr: int, nc: int):
        self.nr, self.nc = nr, nc
        self.n = 0
        self.cap = 0
        self.jm = 1
        self.problems: list = []
        self.results: list = []
        self._grow(64, 16)

    def _grow(self, cap: int, jm: int) -> None:
        def g(name, shape, dtype, fill):
            old = getattr(self, name, None)
            a = np.full((cap,) + shape, fill, dtype=dtype)
            if old is not None:
                a[tuple(slice(0, k) for k in old.shape)] = old
            setattr(self, name, a)
        g("mu", (), float, np.nan)
        g("base", (6,), float, np.nan)
        g("ctrl", (self.nc,), int, 0)
        g("res", (self.nr,), int, 0)
        g("tgt", (self.nr,), float, np.nan)
        g("halo", (), bool, False)
        g("fwd", (), int, 1)
        g("max_att", (), int, 0)
        g("tol", (), float, np.nan)
        g("max_delta", (), float, np.inf)
        g("c", (), float, np.nan)
        g("x", (self.nc,), float, np.nan)
        g("k", (), int, 0)
        g("n_ev", (), int, 0)
        g("n_stm", (), int, 0)
        g("stage", (), int, self.DONE)
        g("r", (self.nr,), float, np.nan)
        g("rn", (), float, np.nan)
        g("t_ev", (), float, np.nan)
        g("x_ev", (6,), float, np.nan)
        g("nal", (), int, 0)
        g("alpha", (jm,), float, np.nan)
        g("trial", (jm, self.nc), float, np.nan)
        self.cap, self.jm = cap, jm

    def add(self, p: CorrectionProblem) -> int:
        al, a = [], 1.0
        while a >= p.min_alpha:
            al.append(a)
            a = a * p.alpha_reduction
        if self.n == self.cap or len(al) > self.jm:
            self._grow(max(self.cap, 2 * self.cap if self.n == self.cap else self.cap),
                       max(self.jm, len(al)))
        i = self.n
        self.n += 1
        self.problems.append(p)
        self.results.append(None)
        self.mu[i] = p.mu
        self.base[i] = np.asarray(p.x0, float)
        self.ctrl[i] = list(p.control_indices)
        self.res[i] = list(p.residual_indices)
        self.tgt[i] = np.asarray(p.target, float)
        self.halo[i] = p.extra_jacobian == "halo"
        self.fwd[i] = p.forward
        self.max_att[i] = p.max_attempts
        self.tol[i] = p.tol
        self.max_delta[i] = np.inf if p.max_delta is None else p.max_delta
        self.c[i] = p.armijo_c
        self.x[i] = self.base[i][self.ctrl[i]]
        self.stage[i] = self.RESID
        self.nal[i] = len(al)
        self.alpha[i, :len(al)] = al
        return i

    # -- requests ------------------------------------------------------------
    def full(self, rows: np.ndarray, params: np.ndarray) -> np.ndarray:
        F = self.base[rows].copy()
        np.put_along_axis(F, self.ctrl[rows], params, axis=1)
        return F

    def _layout(self, rows: np.ndarray):
        """(R, jm) mask of the event requests of each row (RESID: column 0,
        ARMIJO: its trials) and the (R, jm, nc) parameters."""
        resid = self.stage[rows] == self.RESID
        cols = np.arange(self.jm)[None, :]
        valid = np.where(resid[:, None], cols == 0, cols < self.nal[rows][:, None])
        params = self.trial[rows].copy()
        params[resid, 0] = self.x[rows[resid]]
        return resid, valid, params

    def event_inputs(self, rows: np.ndarray):
        """States whose event searches the rows wait for (grouped by row, in
        row order) and per-row counts."""
        _, valid, params = self._layout(rows)
        rid = np.broadcast_to(rows[:, None], valid.shape)[valid]
        return self.full(rid, params[valid]), valid.sum(axis=1)

    def stm_inputs(self, rows: np.ndarray):
        return self.full(rows, self.x[rows]), self.t_ev[rows]

    # -- transitions ---------------------------------------------------------
    def _finish(self, rows, ok, errs, t=None) -> None:
        F = self.full(rows, self.x[rows])
        for j, i in enumerate(rows):
            self.results[i] = CorrectionResult(
                bool(ok[j]), F[j], float(self.rn[i]), int(self.k[i]),
                half_period=float(t[j]) if t is not None else math.nan, error=errs[j],
                n_event_searches=int(self.n_ev[i]), n_stm=int(self.n_stm[i]))
        self.stage[rows] = self.DONE

    def _residual(self, rows, t, xe, ok) -> None:
        """Top of the Newton loop for rows whose residual at x is known."""
        self.n_ev[rows] += 1
        if (~ok).any():
            bad = rows[~ok]
            self.rn[bad] = np.nan
            self._finish(bad, np.zeros(len(bad), bool),
                         ["no section crossing for the current state"] * len(bad))
        rows, t, xe = rows[ok], t[ok], xe[ok]
        if rows.size == 0:
            return
        r = np.take_along_axis(xe, self.res[rows], axis=1) - self.tgt[rows]
        rn = np.max(np.abs(r), axis=1)
        self.rn[rows] = rn
        conv = rn < self.tol[rows]
        if conv.any():
            self._finish(rows[conv], np.ones(conv.sum(), bool), [""] * int(conv.sum()), t[conv])
        maxed = ~conv & (self.k[rows] >= self.max_att[rows])
        if maxed.any():
            mr = rows[maxed]
            self._finish(mr, np.zeros(len(mr), bool),
                         [f"Newton did not converge after {self.max_att[i]} iterations "
                          f"(|R|={self.rn[i]:.2e})." for i in mr])
        go = ~conv & ~maxed
        g = rows[go]
        self.r[g], self.t_ev[g], self.x_ev[g] = r[go], t[go], xe[go]
        self.stage[g] = self.JAC

    def feed_events(self, rows: np.ndarray, t: np.ndarray, xe: np.ndarray, ok: np.ndarray) -> None:
        """Results of event_inputs(rows), in the same order."""
        resid, valid, _ = self._layout(rows)
        R = len(rows)
        T = np.full((R, self.jm), np.nan)
        XE = np.full((R, self.jm, 6), np.nan)
        OK = np.zeros((R, self.jm), bool)
        T[valid], XE[valid], OK[valid] = t, xe, ok
        if resid.any():
            self._residual(rows[resid], T[resid, 0], XE[resid, 0], OK[resid, 0])
        a = ~resid
        if not a.any():
            return
        ar, OKa, XEa, Ta, va = rows[a], OK[a], XE[a], T[a], valid[a]
        res = np.broadcast_to(self.res[ar][:, None, :], (len(ar), self.jm, self.nr))
        nt = np.max(np.abs(np.take_along_axis(XEa, res, axis=2) - self.tgt[ar][:, None, :]),
                    axis=2)
        rn = self.rn[ar][:, None]
        good = va & OKa & (nt <= (1.0 - self.c[ar][:, None] * self.alpha[ar]) * rn)
        has = good.any(axis=1)
        jg = np.argmax(good, axis=1)
        cols = np.arange(self.jm)[None, :]
        before = np.where(has[:, None], cols < jg[:, None], True)
        cand = va & OKa & (nt < rn) & before
        jb = np.argmin(np.where(cand, nt, np.inf), axis=1)
        chosen = np.where(has, jg, np.where(cand.any(axis=1), jb, -1))
        self.n_ev[ar] += np.where(has, jg + 1, self.nal[ar])
        fail = chosen < 0
        if fail.any():
            self._finish(ar[fail], np.zeros(int(fail.sum()), bool),
                         ["Armijo line search failed to find a productive step."] * int(fail.sum()))
        m = ~fail
        if m.any():
            rr, cc = ar[m], chosen[m]
            self.x[rr] = self.trial[rr, cc]
            self.k[rr] += 1
            self.stage[rr] = self.RESID
            idx = np.nonzero(m)[0]
            self._residual(rr, Ta[idx, cc], XEa[idx, cc], OKa[idx, cc])

    def feed_stm(self, rows: np.ndarray, Phi: np.ndarray) -> None:
        """Phi(t_event) (n, 6, 6) for rows in stage JAC: Jacobian, Newton step,
        trial steps of the line search."""
        self.n_stm[rows] += 1
        Jm = Phi[np.arange(len(rows))[:, None, None], self.res[rows][:, :, None],
                 self.ctrl[rows][:, None, :]]
        h = self.halo[rows]
        if h.any():
            Jm[h] = Jm[h] - halo_quadratic_terms(self.mu[rows][h], self.x_ev[rows][h], Phi[h])
        d = solve_delta_stack(Jm, self.r[rows])
        md = self.max_delta[rows]
        dn = np.max(np.abs(d), axis=1)
        cap = ~np.isinf(md) & (dn > md)
        if cap.any():
            d[cap] = d[cap] * (md[cap] / dn[cap])[:, None]
        self.trial[rows] = self.x[rows][:, None, :] + self.alpha[rows][:, :, None] * d[:, None, :]
        self.stage[rows] = self.ARMIJO


def _drive(bank: _Bank, eng: "_Engine") -> None:
    """Run every row of a bank to completion, one batched launch per
    request kind per round."""
    while True:
        live = np.nonzero(bank.stage[:bank.n] != bank.DONE)[0]
        if live.size == 0:
            return
        ev = live[bank.stage[live] != bank.JAC]
        st = live[bank.stage[live] == bank.JAC]
        for f in np.unique(bank.fwd[ev]):
            rows = ev[bank.fwd[ev] == f]
            states, counts = bank.event_inputs(rows)
            mu = np.repeat(bank.mu[rows], counts)
            t, xe, ok = eng.event_batch(mu, states, int(f))
            bank.feed_events(rows, t, xe, ok)
        if st.size:
            states, tev = bank.stm_inputs(st)
            bank.feed_stm(st, eng.stm_batch(bank.mu[st], states, tev))


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
    out = [b.results[i] for b, i in where]
    if not compute_half_period:
        for r in out:
            r.half_period = math.nan
    return out


def correct(problem: CorrectionProblem, **kw) -> CorrectionResult:
    """Correct one orbit; raises ConvergenceError like _NewtonBackend.run."""
    r = correct_batch([problem], **kw)[0]
    if not r.converged:
        if "Armijo" in r.error:
            raise ConvergenceError(f"Step strategy failed to produce an update: {r.error}")
        raise ConvergenceError(r.error)
    return r
