"""Invariant-manifold arcs of a periodic orbit, batched on the GPU.

Restates hiten's ``ManifoldServices._run_compute``
(algorithms/types/services/manifold.py:293-442) with every arc of the
manifold propagated in ONE launch:

1. ``_compute_stm(var_dynsys, x0, period, steps=2000, forward)`` -- one 42-d
   DOP853 dense-output run (manifold.py:215-244, forward = -stable).
2. Eigen-decomposition of the monodromy matrix, classified as a discrete
   system with delta = 1e-6 (manifold.py:580-621; linalg/backend.py:104-
   400), real eigenvectors selected as linalg/base.py:116-142 does.
3. One seed per fraction in ``numpy.arange(0, 1, step)``
   (_compute_manifold_section, manifold.py:470-526; _totime 528-560).
4. All seeds propagated together on numpy.linspace(0, tf, steps) with
   flip_indices = slice(0, 6) (manifold.py:392-409): one 6-d DOP853 grid
   launch (or one fixed-step RK launch of the given order for method="fixed").
5. The proximity and Jacobi-drift filters (manifold.py:411-432) for all arcs
   in one ``hh_traj_filters`` launch.

The eigen-decomposition and the seed construction are host NumPy, as in the
reference (a 6x6 problem and a few hundred seeds); every integration and the
per-sample filters run in HIP kernels.
"""

from __future__ import annotations

from dataclasses import dataclass, field
from typing import List, Literal, Optional, Tuple

import numpy as np

from .trajectory import Trajectory
from . import batch
from .dynamics import _compute_stm, variational_dynsys
from .energy import traj_filters

# utils/constants.py: body radii and the Earth-Moon distance, in the units
# the reference's System stores them (m)
EARTH_RADIUS = 6378.137e3
MOON_RADIUS = 1737.4e3
EARTH_MOON_DISTANCE = 384400e3


# ---------------------------------------------------------------------------
# eigen-decomposition (linalg/backend.py)
# ---------------------------------------------------------------------------

def _zero_small_imag_part(ev: complex, tol: float = 1e-12) -> complex:
    """linalg/backend.py:560-585."""
    if abs(ev.imag) < tol:
        return complex(ev.real, 0.0)
    return ev


def _remove_infinitesimals(vec: np.ndarray, tol: float) -> np.ndarray:
    """linalg/backend.py:501-558: zero real/imaginary parts below tol."""
    out = vec.copy()
    for i in range(len(out)):
        re, im = out[i].real, out[i].imag
        if abs(re) < tol:
            re = 0.0
        if abs(im) < tol:
            im = 0.0
        out[i] = re + 1j * im
    return out


def _sort_eigenvalues(eigvals: np.ndarray, eigvecs: np.ndarray):
    """linalg/backend.py:353-391: clean values, pivot-normalise vectors (first
    entry above 1e-14 made 1), deterministic sign for real vectors."""
    vals = np.array([_zero_small_imag_part(ev, tol=1e-14) for ev in eigvals])
    vecs = eigvecs.astype(np.complex128)
    for k in range(vecs.shape[1]):
        vec = vecs[:, k]
        p = 0
        while p < vec.shape[0] and abs(vec[p]) < 1e-14:
            p += 1
        if p < vec.shape[0]:
            pivot = vec[p]
            if abs(pivot) > 1e-14:
                vec = vec / pivot
                vecs[:, k] = vec
        vecs[:, k] = _remove_infinitesimals(vecs[:, k], tol=1e-14)
        if np.max(np.abs(vecs[:, k].imag)) < 1e-14 and p < vecs.shape[0]:
            if vecs[p, k].real < 0:
                vecs[:, k] = -vecs[:, k]
    return vals, vecs


def eigenvalue_decomposition(A: np.ndarray, delta: float = 1e-6, discrete: bool = True):
    """Stable / unstable / center split of A's spectrum
    (linalg/backend.py:104-200): returns ``sn, un, cn, Ws, Wu, Wc``.
    Discrete: |lambda| < 1 - delta stable, > 1 + delta unstable."""
    n = A.shape[0]
    try:
        eigvals, eigvecs = np.linalg.eig(A)
    except np.linalg.LinAlgError:
        eigvals = np.array([], dtype=np.complex128)
        eigvecs = np.zeros((n, 0), dtype=np.complex128)
    val, vec = _sort_eigenvalues(eigvals, eigvecs)
    sn, un, cn, Ws, Wu, Wc = [], [], [], [], [], []
    for v, w in zip(val, vec.T):
        if discrete:
            mag = abs(v)
            if mag < 1 - delta:
                sn.append(v), Ws.append(w)
            elif mag > 1 + delta:
                un.append(v), Wu.append(w)
            else:
                cn.append(v), Wc.append(w)
        else:
            if v.real < -delta:
                sn.append(v), Ws.append(w)
            elif v.real > +delta:
                un.append(v), Wu.append(w)
            else:
                cn.append(v), Wc.append(w)

    def stack(ws):
        return np.column_stack(ws) if ws else np.zeros((n, 0), dtype=np.complex128)

    return (np.array(sn, dtype=np.complex128), np.array(un, dtype=np.complex128),
            np.array(cn, dtype=np.complex128), stack(Ws), stack(Wu), stack(Wc))


def get_real_eigenvectors(vectors: np.ndarray, values: np.ndarray):
    """linalg/base.py:116-142."""
    mask = np.isreal(values)
    real_vals = values[mask].astype(np.complex128)
    if np.any(mask):
        real_vecs = vectors[:, mask]
    else:
        real_vecs = np.zeros((vectors.shape[0], 0), dtype=np.complex128)
    return real_vals, real_vecs


# ---------------------------------------------------------------------------
# seeds (services/manifold.py:470-560)
# ---------------------------------------------------------------------------

def _totime(t: np.ndarray, tf) -> np.ndarray:
    """Index of the sample closest to each target in |t| (manifold.py:528-560)."""
    t = np.abs(t)
    tf = np.atleast_1d(tf)
    idx = np.empty(tf.shape, dtype=int)
    for k, target in enumerate(tf):
        idx[k] = np.argmin(np.abs(target - t))
    return idx


def manifold_seeds(xx: np.ndarray, tt: np.ndarray, PHI: np.ndarray, period: float,
                   fractions, eigvec: np.ndarray, direction: int,
                   displacement: float) -> np.ndarray:
    """_compute_manifold_section for every fraction: (n_fractions, 6)."""
    out = np.empty((len(fractions), 6), dtype=np.float64)
    for i, fraction in enumerate(fractions):
        k = _totime(tt, fraction * period)[0]
        phi = PHI[k, :36].reshape((6, 6))
        MAN = direction * (phi @ eigvec)
        mag = np.linalg.norm(MAN[0:3])
        if mag < 1e-14:
            mag = 1.0
        d = displacement / mag
        x0W = (xx[k, :].copy() + d * MAN.real).flatten()
        if abs(x0W[2]) < 1.0e-15:
            x0W[2] = 0.0
        if abs(x0W[5]) < 1.0e-15:
            x0W[5] = 0.0
        out[i] = x0W.astype(np.float64)
    return out


@dataclass
class ManifoldResult:
    """``(ysos, dysos, states_list, times_list, successes, attempts)`` of
    _run_compute, plus the seeds and per-arc filter values."""
    ysos: list = field(default_factory=list)
    dysos: list = field(default_factory=list)
    states_list: List[np.ndarray] = field(default_factory=list)
    times_list: List[np.ndarray] = field(default_factory=list)
    successes: int = 0
    attempts: int = 0
    seeds: Optional[np.ndarray] = None
    kept: Optional[np.ndarray] = None
    min_r1: Optional[np.ndarray] = None
    min_r2: Optional[np.ndarray] = None
    energy_err: Optional[np.ndarray] = None

    @property
    def success_rate(self) -> float:
        return self.successes / max(self.attempts, 1)

    def as_tuple(self) -> Tuple:
        return (self.ysos, self.dysos, self.states_list, self.times_list, self.successes,
                self.attempts)

    @property
    def trajectories(self) -> List[Trajectory]:
        """Each kept arc as a Trajectory (types/services/manifold.py:207-213)."""
        return [Trajectory(t, x) for t, x in zip(self.times_list, self.states_list)]

    def to_df(self):
        """Manifold.to_df (system/manifold.py:373-387): one row per sample,
        columns trajectory_id, time, x, y, z, vx, vy, vz."""
        import pandas as pd
        cols = ["trajectory_id", "time", "x", "y", "z", "vx", "vy", "vz"]
        if not self.states_list:
            return pd.DataFrame([], columns=cols)
        ids = np.concatenate([np.full(len(t), i, dtype=np.int64)
                              for i, t in enumerate(self.times_list)])
        t = np.concatenate([np.asarray(t, float) for t in self.times_list])
        X = np.concatenate([np.asarray(x, float)[:, :6] for x in self.states_list])
        df = pd.DataFrame(X, columns=cols[2:])
        df.insert(0, "time", t)
        df.insert(0, "trajectory_id", ids)
        return df

    def to_csv(self, filepath: str) -> None:
        """Manifold.to_csv: to_df written without the index."""
        import os
        os.makedirs(os.path.dirname(os.path.abspath(filepath)), exist_ok=True)
        self.to_df().to_csv(filepath, index=False)


def compute_manifold(mu: float, orbit_state, period: float, *, stable: bool = True,
                     direction: Literal["positive", "negative"] = "positive", step: float = 0.02,
                     integration_fraction: float = 0.75, NN: int = 1,
                     displacement: float = 1e-6, dt: float = 1e-3,
                     method: Literal["fixed", "adaptive"] = "adaptive", order: int = 8,
                     energy_tol: float = 1e-6, safe_distance: float = 2.0,
                     primary_radius: float = EARTH_RADIUS, secondary_radius: float = MOON_RADIUS,
                     distance: float = EARTH_MOON_DISTANCE, delta: float = 1e-6,
                     stm_steps: int = 2000) -> ManifoldResult:
    """Manifold.compute (system/manifold.py:226-300) for the orbit through
    ``orbit_state`` with ``period``; defaults are the reference's.

    ``primary_radius``, ``secondary_radius`` and ``distance`` are in the units
    the reference's System holds them (m); the safe radii are formed exactly
    as manifold.py:340-345 forms them, ``distance * 1e3`` included.
    """
    if method not in ("fixed", "adaptive"):
        raise NotImplementedError(f"method {method!r}: only 'fixed' (RK4/6/8) and 'adaptive' "
                                  "(DOP853) integrate the CR3BP here")
    if method == "fixed" and order not in batch.FIXED_ORDERS:
        raise ValueError("RK order must be 4, 6, or 8")
    if method == "adaptive" and order not in (8, 853):
        raise NotImplementedError("adaptive integration is DOP853 only")
    stable_i = 1 if stable else -1
    direction_i = 1 if direction == "positive" else -1
    forward = -stable_i

    dist_m = distance * 1e3
    safe_r1 = safe_distance * (primary_radius / dist_m)
    safe_r2 = safe_distance * (secondary_radius / dist_m)

    xx, tt, phi_T, PHI = _compute_stm(variational_dynsys(mu), np.asarray(orbit_state, float),
                                      period, steps=stm_steps, forward=forward)
    sn, un, _, Ws, Wu, _ = eigenvalue_decomposition(phi_T, delta=delta, discrete=True)
    _, s_real = get_real_eigenvectors(Ws, sn)
    _, u_real = get_real_eigenvectors(Wu, un)
    col = NN - 1
    vecs = s_real if stable_i == 1 else u_real
    if vecs.shape[1] <= col or col < 0:
        kind = "stable" if stable_i == 1 else "unstable"
        raise ValueError(f"Requested {kind} eigenvector {NN} not available. "
                         f"Only {vecs.shape[1]} real {kind} eigenvectors found.")
    eigvec = vecs[:, col]

    fractions = tuple(np.arange(0.0, 1.0, step))
    res = ManifoldResult(attempts=len(fractions))
    if not fractions:
        return res
    seeds = manifold_seeds(xx, tt, PHI, period, fractions, eigvec, direction_i, displacement)
    tf = integration_fraction * 2 * np.pi
    steps = max(int(abs(tf) / dt) + 1, 100)
    n = seeds.shape[0]
    if method == "adaptive":
        states = batch.dop853_grid(mu, seeds, 0.0, tf, steps, dim=6, forward=forward,
                                   flip_indices=slice(0, 6)).states
    else:
        _, states = batch.rk_fixed(mu, seeds, 0.0, tf, steps, order=order, dim=6,
                                   forward=forward, flip_indices=slice(0, 6),
                                   return_trajectory=True)
    r1, r2, err = traj_filters(states, mu)
    r1, r2, err = r1.cpu().numpy(), r2.cpu().numpy(), err.cpu().numpy()
    times = forward * np.linspace(0.0, tf, steps)
    # a NaN minimum compares false, as numpy's r1.min() < safe_r1 does
    kept = ~((r1 < safe_r1) | (r2 < safe_r2)) & ~(err > energy_tol)
    host = states.cpu().numpy() if kept.any() else None
    for i in np.nonzero(kept)[0]:
        res.states_list.append(host[i].copy())
        res.times_list.append(times.copy())
    res.successes = int(kept.sum())
    res.seeds, res.kept = seeds, kept
    res.min_r1, res.min_r2, res.energy_err = r1, r2, err
    return res
