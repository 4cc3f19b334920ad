"""Energy of CR3BP states and trajectories (hiten common/energy.py).

``crtbp_energy`` is the reference's host-side helper (NumPy scalar code, kept
on the host here too).  ``_max_rel_energy_error`` and the batched
``traj_filters`` run on the GPU through ``hh_traj_filters``: one wave per
trajectory reduces the Jacobi-constant drift and the distances to both
primaries over every sample of a dense-output grid.
"""

from __future__ import annotations

from typing import Sequence

import numpy as np
import torch

from . import _lib, batch


def crtbp_energy(state: Sequence[float], mu: float) -> float:
    """Hamiltonian energy E = T + U_eff (common/energy.py:77-137), with the
    reference's potential -(mu1/r1) - (mu2/r2) - 0.5 (x^2+y^2+z^2) - 0.5 mu1 mu2."""
    x, y, z, vx, vy, vz = state
    mu1 = 1.0 - mu
    mu2 = mu
    r1 = np.sqrt((x + mu2) ** 2 + y ** 2 + z ** 2)
    r2 = np.sqrt((x - mu1) ** 2 + y ** 2 + z ** 2)
    kin = 0.5 * (vx * vx + vy * vy + vz * vz)
    pot = -(mu1 / r1) - (mu2 / r2) - 0.5 * (x * x + y * y + z * z) - 0.5 * mu1 * mu2
    return kin + pot


def energy_to_jacobi(energy: float) -> float:
    """C = -2E (common/energy.py)."""
    return -2.0 * energy


def jacobi_to_energy(jacobi: float) -> float:
    """E = -C/2 (common/energy.py)."""
    return -jacobi / 2.0


def _as_grid(states) -> torch.Tensor:
    if isinstance(states, torch.Tensor):
        t = states.to(device=batch._device(), dtype=torch.float64)
    else:
        t = torch.as_tensor(np.asarray(states, dtype=np.float64), device=batch._device())
    if t.dim() == 2:
        t = t.unsqueeze(0)
    if t.dim() != 3 or t.shape[2] != 6:
        raise ValueError(f"states must have shape (m, 6) or (n, m, 6), got {tuple(t.shape)}")
    return t.contiguous()


def traj_filters(states, mu, stream=None):
    """Per trajectory of ``states`` (n, m, 6): ``(min_r1, min_r2, max_err)``.

    min_r1/min_r2 are the smallest distances to the primaries
    (services/manifold.py:416-420, NaN if a sample is NaN) and max_err is
    ``_max_rel_energy_error`` of the trajectory (common/energy.py:26-74).
    Results are device tensors of shape (n,).
    """
    lib = _lib.load()
    S = _as_grid(states)
    n, m = S.shape[0], S.shape[1]
    if m < 1:
        raise ValueError("a trajectory needs at least one sample")
    mu_t = batch._vec(mu, n, "mu")
    out = torch.empty((3, n), dtype=torch.float64, device=S.device)
    rc = lib.hh_traj_filters(n, m, batch._ptr(mu_t), batch._ptr(S), batch._ptr(out[0]),
                             batch._ptr(out[1]), batch._ptr(out[2]), _lib.stream_handle(stream))
    _lib.check(rc, "hh_traj_filters")
    return out[0], out[1], out[2]


def _max_rel_energy_error(states: np.ndarray, mu: float) -> float:
    """Maximum relative drift of the Jacobi constant along one trajectory
    (common/energy.py:26-74); the first sample is the reference value."""
    S = _as_grid(states)
    if S.shape[0] != 1:
        raise ValueError("one trajectory (m, 6) expected")
    return float(traj_filters(S, mu)[2][0].item())
