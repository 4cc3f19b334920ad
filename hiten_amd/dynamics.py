"""CR3BP dynamical systems and propagation entry points (drop-in for
``hiten.algorithms.dynamics``), executed by the HIP kernels.

* ``_RTBPRHS`` / ``rtbp_dynsys``          dynamics/rtbp.py:550-641
* ``_VarEqRHS`` / ``variational_dynsys``  dynamics/rtbp.py:483-547, 668-691
* ``_JacobianRHS`` / ``jacobian_dynsys``  dynamics/rtbp.py:425-480, 644-665
* ``_DirectedSystem``                     dynamics/base.py:186-343
* ``_propagate_dynsys``                   dynamics/base.py:346-458
* ``_validate_initial_state``             dynamics/base.py:461-494
* ``_compute_stm`` / ``_compute_monodromy`` dynamics/rtbp.py:258-378

Batched forms (many trajectories per launch) are ``propagate_batch`` and
``compute_stm_batch``.
"""

from __future__ import annotations

from typing import Literal, Sequence

import numpy as np

from . import batch
from .integrators import AdaptiveRK, RungeKutta, _Solution


class _DynamicalSystem:
    """dynamics/base.py:39-113 (dimension bookkeeping and ``rhs``)."""

    def __init__(self, dim: int):
        if dim <= 0:
            raise ValueError(f"Dimension must be positive, got {dim}")
        self._dim = dim

    @property
    def dim(self) -> int:
        return self._dim

    def validate_state(self, y: np.ndarray) -> None:
        if len(y) != self.dim:
            raise ValueError(f"State vector dimension {len(y)} != system dimension {self.dim}")


class _RTBPRHS(_DynamicalSystem):
    """CR3BP equations of motion in the rotating frame (rtbp.py:30-74)."""

    def __init__(self, mu: float, name: str = "RTBP"):
        super().__init__(6)
        self.name = name
        self.mu = float(mu)

    @property
    def rhs(self):
        mu = self.mu

        def _crtbp_rhs(t, state):
            return batch.eval_rhs(mu, np.asarray(state, float)[None, :], dim=6)[0].cpu().numpy()
        return _crtbp_rhs

    def __repr__(self):
        return f"_RTBPRHS(name='{self.name}', mu={self.mu})"


class _VarEqRHS(_DynamicalSystem):
    """42-d variational system [Phi row-major | x] (rtbp.py:167-255)."""

    def __init__(self, mu: float, name: str = "CR3BP Variational Equations"):
        super().__init__(42)
        self.name = name
        self.mu = float(mu)

    @property
    def rhs(self):
        mu = self.mu

        def _var_eq_rhs(t, y):
            return batch.eval_rhs(mu, np.asarray(y, float)[None, :], dim=42)[0].cpu().numpy()
        return _var_eq_rhs

    def __repr__(self):
        return f"_VarEqRHS(name='{self.name}', mu={self.mu})"


class _JacobianRHS(_DynamicalSystem):
    """6x6 Jacobian of the CR3BP field at a position (rtbp.py:76-165).

    Evaluated through the variational field with Phi = I, whose STM block is
    F @ I = F.
    """

    def __init__(self, mu: float, name: str = "CR3BP Jacobian"):
        super().__init__(3)
        self.name = name
        self.mu = float(mu)

    @property
    def rhs(self):
        mu = self.mu

        def _jac_rhs(t, state):
            y = np.zeros(42)
            y[:36] = np.eye(6).ravel()
            y[36:39] = np.asarray(state, float)[:3]
            d = batch.eval_rhs(mu, y[None, :], dim=42)[0].cpu().numpy()
            return d[:36].reshape(6, 6)
        return _jac_rhs


def rtbp_dynsys(mu: float, name: str = "RTBP") -> _RTBPRHS:
    return _RTBPRHS(mu=mu, name=name)


def variational_dynsys(mu: float, name: str = "VarEq") -> _VarEqRHS:
    return _VarEqRHS(mu=mu, name=name)


def jacobian_dynsys(mu: float, name: str = "Jacobian") -> _JacobianRHS:
    return _JacobianRHS(mu=mu, name=name)


class _DirectedSystem(_DynamicalSystem):
    """Backward-time wrapper (dynamics/base.py:186-343): with fwd == -1 the
    derivative of each component in ``flip_indices`` (all when None) is
    negated.  The kernels apply the flip inside the vector field."""

    def __init__(self, base_or_dim, fwd: int = 1, flip_indices=None):
        if isinstance(base_or_dim, _DynamicalSystem):
            self._base = base_or_dim
            dim = base_or_dim.dim
        else:
            self._base = None
            dim = int(base_or_dim)
        super().__init__(dim)
        self._fwd = 1 if fwd >= 0 else -1
        self._flip_idx = flip_indices

    @property
    def rhs(self):
        if self._base is None:
            raise AttributeError("`rhs` not implemented: no base system is wrapped.")
        base_rhs = self._base.rhs
        fwd = self._fwd
        mask = batch.flip_mask(self.dim, self._flip_idx)
        sign = np.array([-1.0 if (fwd == -1 and (mask >> i) & 1) else 1.0 for i in range(self.dim)])

        def _rhs(t, y):
            return base_rhs(t, y) * sign
        return _rhs

    def __getattr__(self, item):
        if item.startswith("_") or self.__dict__.get("_base") is None:
            raise AttributeError(item)
        return getattr(self._base, item)

    def __repr__(self):
        return f"DirectedSystem(dim={self.dim}, fwd={self._fwd}, flip_idx={self._flip_idx})"


def _validate_initial_state(state, expected_dim=6):
    """dynamics/base.py:461-494."""
    state_np = np.asarray(state, dtype=np.float64)
    if state_np.shape != (expected_dim,):
        raise ValueError(f"Initial state vector must have {expected_dim} elements, but got shape "
                         f"{state_np.shape}")
    return state_np


def _propagate_dynsys(dynsys, state0: Sequence[float], t0: float, tf: float, forward: int = 1,
                      steps: int = 1000, method: Literal["fixed", "adaptive", "symplectic"] = "adaptive",
                      order: int = 8, flip_indices: Sequence[int] | None = None,
                      **kwargs) -> _Solution:
    """Propagate one trajectory (dynamics/base.py:346-458).

    Returns ``_Solution(times * forward, states)`` with ``steps`` rows.
    """
    state0_np = _validate_initial_state(state0, dynsys.dim)
    dynsys_dir = _DirectedSystem(dynsys, forward, flip_indices=flip_indices)
    t_eval = np.linspace(t0, tf, steps)
    if steps >= 2 and np.isclose(t_eval[0], t_eval[-1]):
        return _Solution(forward * t_eval, np.repeat(state0_np[None, :], len(t_eval), axis=0))
    event_fn = kwargs.get("event_fn", None)
    event_cfg = kwargs.get("event_cfg", None)
    event_options = kwargs.get("event_options", None)
    if method == "fixed":
        integrator = RungeKutta(order=order)
    elif method == "adaptive":
        integrator = AdaptiveRK(order=order, max_step=kwargs.get("max_step", 1e4),
                                rtol=kwargs.get("rtol", 1e-12), atol=kwargs.get("atol", 1e-12))
    elif method == "symplectic":
        raise NotImplementedError("symplectic integration applies to polynomial Hamiltonian "
                                  "systems, which hiten_amd does not integrate")
    else:
        raise ValueError(f"unknown method {method!r}")
    sol = integrator.integrate(dynsys_dir, state0_np, t_eval, event_fn=event_fn,
                               event_cfg=event_cfg, event_options=event_options)
    return _Solution(forward * sol.times, sol.states)


def _compute_stm(dynsys, x0, tf, steps=2000, forward=1,
                 method: Literal["fixed", "adaptive", "symplectic"] = "adaptive", order=8, **kwargs):
    """State transition matrix along a trajectory (rtbp.py:258-340).

    Returns ``(x (steps, 6), times (steps,), phi_T (6, 6), PHI (steps, 42))``.
    """
    PHI0 = np.zeros(42, dtype=np.float64)
    PHI0[:36] = np.eye(6, dtype=np.float64).ravel()
    PHI0[36:] = x0
    sol = _propagate_dynsys(dynsys=dynsys, state0=PHI0, t0=0.0, tf=tf, forward=forward,
                            steps=steps, method=method, order=order, flip_indices=slice(36, 42),
                            **kwargs)
    PHI = sol.states
    return PHI[:, 36:42], sol.times, PHI[-1, :36].reshape((6, 6)), PHI


def _compute_monodromy(dynsys, x0, period):
    """rtbp.py:343-378."""
    _, _, M, _ = _compute_stm(dynsys, x0, period)
    return M


# ---------------------------------------------------------------------------
# batched forms
# ---------------------------------------------------------------------------

def propagate_batch(mu, states0, t0, tf, *, forward: int = 1, steps: int = 2,
                    method: str = "adaptive", flip_indices=None, rtol: float = 1e-12,
                    atol: float = 1e-12, max_step: float = 1e4, dim: int = 6, variant: str = "auto",
                    max_steps: int = batch.DEFAULT_MAX_STEPS, order: int = 8):
    """_propagate_dynsys for many trajectories in one launch (``order`` 4, 6
    or 8 with method "fixed"; DOP853 with method "adaptive").

    ``steps == 2`` returns the final states (n, dim) (the dense value at tf);
    otherwise the (n, steps, dim) grid.  ``t0``/``tf``/``mu`` may be scalars or
    per-trajectory arrays.  Result arrays stay on the GPU (torch tensors).
    ``variant`` picks the kernel layout ("auto", "plain", "wide", and
    "persistent" for the 42-d final value); all give the same bits.
    """
    fixed_variant = variant if variant in ("plain", "wide") else "auto"
    if method == "fixed":
        if steps == 2:
            return batch.rk_fixed(mu, states0, t0, tf, steps, order=order, dim=dim,
                                  forward=forward, flip_indices=flip_indices,
                                  variant=fixed_variant)
        _, traj = batch.rk_fixed(mu, states0, t0, tf, steps, order=order, dim=dim,
                                 forward=forward, flip_indices=flip_indices,
                                 return_trajectory=True, variant=fixed_variant)
        return traj
    if order not in (8, 853):
        raise NotImplementedError("adaptive propagation runs DOP853 (order 8) on the GPU")
    if steps == 2:
        return batch.dop853_final(mu, states0, t0, tf, dim=dim, rtol=rtol, atol=atol,
                                  max_step=max_step, forward=forward, flip_indices=flip_indices,
                                  variant=variant, max_steps=max_steps)
    return batch.dop853_grid(mu, states0, t0, tf, steps, dim=dim, rtol=rtol, atol=atol,
                             max_step=max_step, forward=forward, flip_indices=flip_indices,
                             max_steps=max_steps, variant=fixed_variant)


def compute_stm_batch(mu, x0s, tf, *, forward: int = 1, steps: int = 2, rtol: float = 1e-12,
                      atol: float = 1e-12, max_step: float = 1e4, variant: str = "auto"):
    """_compute_stm for many initial states: returns (phi_T (n, 6, 6), x_T (n, 6))
    with ``steps == 2``, else the full (n, steps, 42) PHI grid."""
    x0s = np.atleast_2d(np.asarray(x0s, dtype=np.float64))
    n = x0s.shape[0]
    PHI0 = np.zeros((n, 42))
    PHI0[:, :36] = np.eye(6).ravel()
    PHI0[:, 36:] = x0s
    r = propagate_batch(mu, PHI0, 0.0, tf, forward=forward, steps=steps, dim=42,
                        flip_indices=slice(36, 42), rtol=rtol, atol=atol, max_step=max_step,
                        variant=variant)
    if steps == 2:
        return r.y[:, :36].reshape(n, 6, 6), r.y[:, 36:]
    return r.states
