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


def correct_batch(problems: Sequence[CorrectionProblem], engine: _Engine | None = None,
                  compute_half_period: bool = True) -> list[CorrectionResult]:
    """Newton-Armijo for every problem; each phase is one batched launch."""
    eng = engine or _Engine()
    P = list(problems)
    n = len(P)
    mu = np.array([p.mu for p in P])
    fwd = {p.forward for p in P}
    if len(fwd) != 1:
        raise ValueError("all problems of a batch must share `forward`")
    forward = fwd.pop()
    base = [np.asarray(p.x0, float).copy() for p in P]
    ctrl = [list(p.control_indices) for p in P]
    res = [list(p.residual_indices) for p in P]
    tgt = [np.asarray(p.target, float) for p in P]
    x = [base[i][ctrl[i]].copy() for i in range(n)]
    k = np.zeros(n, int)
    results: list[Optional[CorrectionResult]] = [None] * n
    counts = np.zeros((n, 2), int)

    def full(i, params):
        v = base[i].copy()
        v[ctrl[i]] = params
        return v

    def residuals(idx, params_list):
        states = np.stack([full(i, q) for i, q in zip(idx, params_list)])
        t, xe, ok = eng.event_batch(mu[idx], states, forward)
        for i in idx:
            counts[i, 0] += 1
        r = [xe[j][res[i]] - tgt[i] if ok[j] else None for j, i in enumerate(idx)]
        return r, t, xe, ok

    active = list(range(n))
    while active:
        r, t_ev, x_ev, ok = residuals(active, [x[i] for i in active])
        still, need_jac = [], []
        for j, i in enumerate(active):
            p = P[i]
            if not ok[j]:
                results[i] = CorrectionResult(False, full(i, x[i]), math.nan, int(k[i]),
                                              error="no section crossing for the current state")
                continue
            rn = _inf_norm(r[j])
            if rn < p.tol:
                results[i] = CorrectionResult(True, full(i, x[i]), rn, int(k[i]))
                continue
            if k[i] >= p.max_attempts:
                results[i] = CorrectionResult(False, full(i, x[i]), rn, int(k[i]),
                                              error=f"Newton did not converge after {p.max_attempts} "
                                                    f"iterations (|R|={rn:.2e}).")
                continue
            need_jac.append((i, r[j], rn, t_ev[j], x_ev[j]))
        if not need_jac:
            break
        idx = [e[0] for e in need_jac]
        phis = eng.stm_batch(mu[idx], np.stack([full(i, x[i]) for i in idx]),
                             np.array([e[3] for e in need_jac]))
        for i in idx:
            counts[i, 1] += 1
        deltas = {}
        for (i, r_i, rn, te, xe), Phi in zip(need_jac, phis):
            J = Phi[np.ix_(res[i], ctrl[i])]
            if P[i].extra_jacobian == "halo":
                J = J - halo_quadratic_term(P[i].mu, xe, Phi)
            d = solve_delta_dense(J, r_i)
            if P[i].max_delta is not None and not np.isinf(P[i].max_delta):
                dn = np.linalg.norm(d, ord=np.inf)
                if dn > P[i].max_delta:
                    d = d * (P[i].max_delta / dn)
            deltas[i] = (d, rn)
        # Armijo backtracking (armijo.py:111-164), all searching orbits per launch
        alpha = {i: 1.0 for i in idx}
        best = {i: (x[i], deltas[i][1], 0.0) for i in idx}
        searching = list(idx)
        while searching:
            trial = [x[i] + alpha[i] * deltas[i][0] for i in searching]
            rt, _, _, okt = residuals(searching, trial)
            nxt = []
            for j, i in enumerate(searching):
                cur = deltas[i][1]
                a = alpha[i]
                if okt[j]:
                    nt = _inf_norm(rt[j])
                    if nt <= (1.0 - P[i].armijo_c * a) * cur:
                        x[i] = trial[j]
                        k[i] += 1
                        continue
                    if nt < best[i][1]:
                        best[i] = (trial[j], nt, a)
                alpha[i] = a * P[i].alpha_reduction
                if alpha[i] >= P[i].min_alpha:
                    nxt.append(i)
                elif best[i][2] > 0:
                    x[i] = best[i][0]
                    k[i] += 1
                else:
                    results[i] = CorrectionResult(
                        False, full(i, x[i]), cur, int(k[i]),
                        error="Armijo line search failed to find a productive step.")
            searching = nxt
        active = [i for i in idx if results[i] is None]

    out = []
    for i in range(n):
        r_ = results[i]
        r_.n_event_searches, r_.n_stm = int(counts[i, 0]), int(counts[i, 1])
        out.append(r_)
    if compute_half_period:
        conv = [i for i in range(n) if out[i].converged]
        if conv:
            t, _, ok = eng.event_batch(mu[conv], np.stack([out[i].x_corrected for i in conv]),
                                       forward)
            for j, i in enumerate(conv):
                out[i].half_period = float(t[j]) if ok[j] else math.nan
    return out


def correct(problem: CorrectionProblem, **kw) -> CorrectionResult:
    """Correct one orbit; raises ConvergenceError like _NewtonBackend.run."""
    r = correct_batch([problem], **kw)[0]
    if not r.converged:
        if "Armijo" in r.error:
            raise ConvergenceError(f"Step strategy failed to produce an update: {r.error}")
        raise ConvergenceError(r.error)
    return r
