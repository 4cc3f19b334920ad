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
            r = batch.dop853_final(mu[live], PHI0[live], 0.0, t_event[live], dim=42,
                                   flip_indices=slice(36, 42), max_step=1e4)
            phi[live] = r.y[:, :36].reshape(-1, 6, 6).cpu().numpy()
        return phi


class _Newton:
    """One correction as a resumable state machine.

    The stages follow _NewtonBackend.run: residual at x (an event search) ->
    Jacobian (the STM to the event time) -> Armijo line search -> residual at
    the new x.  The line search is evaluated speculatively: the residuals of
    every trial step alpha = 1, r, r^2, ... >= min_alpha are requested in one
    batch, then armijo.py:111-164 is replayed over them in order, so the
    accepted step, the fallback to the best point and the failure are the
    sequential ones.  The residual of the accepted trial is the residual at
    the next iterate (the same state through the same deterministic search)
    and is reused.  ``n_event_searches`` counts the searches the sequential
    algorithm makes.
    """

    def __init__(self, p: CorrectionProblem):
        self.p = p
        self.base = np.asarray(p.x0, float).copy()
        self.ctrl = list(p.control_indices)
        self.res = list(p.residual_indices)
        self.tgt = np.asarray(p.target, float)
        self.x = self.base[self.ctrl].copy()
        self.k = 0
        self.n_ev = 0
        self.n_stm = 0
        self.result: Optional[CorrectionResult] = None
        self.kind = "event"                 # what the machine waits for: event / stm / None
        self.inputs = [self.full(self.x)]
        self.stage = "residual"
        self.alphas = []
        alpha = 1.0
        while alpha >= p.min_alpha:
            self.alphas.append(alpha)
            alpha = alpha * p.alpha_reduction

    def full(self, params) -> np.ndarray:
        v = self.base.copy()
        v[self.ctrl] = params
        return v

    def _finish(self, ok: bool, rn: float, err: str = "", t: float = math.nan) -> None:
        self.result = CorrectionResult(ok, self.full(self.x), rn, int(self.k), half_period=t,
                                       error=err, n_event_searches=self.n_ev, n_stm=self.n_stm)
        self.kind = None

    def _residual(self, t, xe, ok) -> None:
        """Top of the Newton loop with the residual at self.x."""
        p = self.p
        self.n_ev += 1
        if not ok:
            self._finish(False, math.nan, "no section crossing for the current state")
            return
        r = xe[self.res] - self.tgt
        rn = _inf_norm(r)
        if rn < p.tol:
            self._finish(True, rn, t=float(t))
            return
        if self.k >= p.max_attempts:
            self._finish(False, rn, f"Newton did not converge after {p.max_attempts} "
                                    f"iterations (|R|={rn:.2e}).")
            return
        self.r, self.rn, self.t_ev, self.x_ev = r, rn, float(t), xe
        self.stage, self.kind = "jacobian", "stm"
        self.inputs = [(self.full(self.x), self.t_ev)]

    def feed(self, out) -> None:
        """Results of the requests in self.inputs: a list of (t, x_event, ok)
        for event requests, a list with one 6x6 Phi for an STM request."""
        p = self.p
        if self.stage == "residual":
            self._residual(*out[0])
        elif self.stage == "jacobian":
            Phi = out[0]
            self.n_stm += 1
            J = Phi[np.ix_(self.res, self.ctrl)]
            if p.extra_jacobian == "halo":
                J = J - halo_quadratic_term(p.mu, self.x_ev, Phi)
            d = solve_delta_dense(J, self.r)
            if p.max_delta is not None and not np.isinf(p.max_delta):
                dn = np.linalg.norm(d, ord=np.inf)
                if dn > p.max_delta:
                    d = d * (p.max_delta / dn)
            self.trials = [self.x + a * d for a in self.alphas]
            self.stage, self.kind = "armijo", "event"
            self.inputs = [self.full(q) for q in self.trials]
        elif self.stage == "armijo":
            best, chosen = None, None
            best_norm = self.rn
            for j, a in enumerate(self.alphas):
                self.n_ev += 1
                t, xe, ok = out[j]
                if not ok:
                    continue
                nt = _inf_norm(xe[self.res] - self.tgt)
                if nt <= (1.0 - p.armijo_c * a) * self.rn:
                    chosen = j
                    break
                if nt < best_norm:
                    best, best_norm = j, nt
            if chosen is None:
                chosen = best
            if chosen is None:
                self._finish(False, self.rn, "Armijo line search failed to find a productive step.")
                return
            self.x = self.trials[chosen]
            self.k += 1
            self.stage = "residual"
            self._residual(*out[chosen])


def _drive(machines, eng: "_Engine") -> None:
    """Run state machines to completion, one batched launch per request kind
    per round."""
    active = [m for m in machines if m.kind is not None]
    while active:
        ev = [m for m in active if m.kind == "event"]
        st = [m for m in active if m.kind == "stm"]
        if ev:
            fwd = {m.p.forward for m in ev}
            for f in fwd:
                grp = [m for m in ev if m.p.forward == f]
                mu = np.concatenate([np.full(len(m.inputs), m.p.mu) for m in grp])
                states = np.stack([s for m in grp for s in m.inputs])
                t, xe, ok = eng.event_batch(mu, states, f)
                o = 0
                for m in grp:
                    k = len(m.inputs)
                    m.feed([(t[o + j], xe[o + j], ok[o + j]) for j in range(k)])
                    o += k
        if st:
            mu = np.array([m.p.mu for m in st])
            states = np.stack([m.inputs[0][0] for m in st])
            tev = np.array([m.inputs[0][1] for m in st])
            phis = eng.stm_batch(mu, states, tev)
            for m, phi in zip(st, phis):
                m.feed([phi])
        active = [m for m in active if m.kind is not None]


def correct_batch(problems: Sequence[CorrectionProblem], engine: _Engine | None = None,
                  compute_half_period: bool = True) -> list[CorrectionResult]:
    """Newton-Armijo for every problem; each round of the corrections'
    event searches (residuals, all line-search trials) and each round of
    their STMs is one batched launch."""
    eng = engine or _Engine()
    machines = [_Newton(p) for p in problems]
    _drive(machines, eng)
    out = [m.result for m in machines]
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
