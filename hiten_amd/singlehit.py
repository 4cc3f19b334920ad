"""Single-hit Poincare section crossings on the GPU (drop-in for
``hiten.algorithms.poincare.singlehit``).

* ``_SingleHitBackend._cross_event_driven``  poincare/singlehit/backend.py:164-235
* ``_SingleHitBackend._cross``               backend.py:237-280
* ``_SingleHitBackend.run``                  backend.py:101-162
* ``_plane_crossing_factory`` and the predefined
  ``_x_plane_crossing`` / ``_y_plane_crossing`` / ``_z_plane_crossing``
                                             backend.py:315-372
* ``_PlaneEvent`` / ``_SectionHit``          poincare/core/events.py:124-249,
                                             poincare/core/types.py:45-86

``section_crossing_batch`` runs the same window logic for many seeds at
once: every seed's event integration is one lane of one kernel launch.
"""

from __future__ import annotations

import math
from dataclasses import dataclass
from typing import Optional

import numpy as np
import torch

from . import _lib, batch
from .dynamics import _DirectedSystem, _propagate_dynsys
from .integrators import EventConfig, PlaneEventFn, RungeKutta

_COORD = {"x": 0, "y": 1, "z": 2}


@dataclass
class _PlaneEvent:
    """Coordinate plane y[index] = offset with a crossing direction
    (poincare/core/events.py:124-249)."""
    coord: str
    value: float = 0.0
    direction: Optional[int] = None

    @property
    def index(self) -> int:
        return _COORD[self.coord.lower()]

    @property
    def offset(self) -> float:
        return float(self.value)

    @property
    def normal(self) -> np.ndarray:
        n = np.zeros(6)
        n[self.index] = 1.0
        return n


@dataclass
class _SectionHit:
    """poincare/core/types.py:45-86."""
    time: float
    state: np.ndarray
    point2d: np.ndarray
    trajectory_index: int = 0


def _system_mu(dynsys) -> float:
    base = dynsys._base if isinstance(dynsys, _DirectedSystem) else dynsys
    mu = getattr(base, "mu", None)
    if mu is None or base.dim != 6:
        raise NotImplementedError("section crossings run on the 6-d CR3BP field")
    return float(mu)


class _SingleHitBackend:
    """First crossing of a coordinate plane (poincare/singlehit/backend.py:71-280)."""

    def run(self, seeds, *, dynsys, surface: _PlaneEvent, dt: float = 1e-2,
            t_guess: float | None = None, forward: int = 1, method="adaptive", order: int = 8):
        seeds = np.atleast_2d(np.asarray(seeds, dtype=float))
        hits = section_crossing_batch(dynsys, seeds, surface=surface, t_guess=t_guess,
                                      forward=forward)
        pts, states = [], []
        for h in hits:
            if h is not None:
                pts.append(h.point2d)
                states.append(h.state.copy())
        if pts:
            return np.asarray(pts, float), np.asarray(states, float)
        return np.empty((0, 2)), np.empty((0, 6))

    def _cross_event_driven(self, state0, *, dynsys, surface: _PlaneEvent, t0: float, tmax: float,
                            forward: int) -> _SectionHit | None:
        """backend.py:164-235: align to max(t0, 1e-12), then DOP853 (rtol = atol
        = 1e-12, no max step) until the plane; the event run ignores ``forward``."""
        ev_dir = 0 if surface.direction is None else int(surface.direction)
        integrator = RungeKutta(order=853, rtol=1e-12, atol=1e-12)
        t_start = float(t0)
        if t_start <= 0.0:
            t_start = 1e-12
        y_start = np.asarray(state0, dtype=float).copy()
        sol_align = _propagate_dynsys(dynsys, y_start, 0.0, t_start, forward=forward, steps=2,
                                      method="adaptive", order=8)
        y_start = sol_align.states[-1].copy()
        span = float(max(0.0, tmax - t_start))
        times = np.array([0.0, span], dtype=float)
        sol = integrator.integrate(dynsys, y_start, times,
                                   event_fn=PlaneEventFn(surface.index, surface.offset),
                                   event_cfg=EventConfig(direction=ev_dir, terminal=True))
        t_hit_rel = float(sol.times[-1])
        y_hit = sol.states[-1].copy()
        t_hit = t_start + t_hit_rel
        if t_hit_rel < span and t_hit_rel >= 0.0:
            return _SectionHit(time=t_hit, state=y_hit, point2d=y_hit[:2].copy(), trajectory_index=0)
        return None

    def _cross(self, state0, *, dynsys, surface: _PlaneEvent, t_guess=None, t0_offset: float = 0.15,
               t_window=None, forward: int = 1):
        """backend.py:237-280 (window around the guess, widening retries, then
        the half-period fallback window)."""
        if (t_guess is not None) and (t_window is not None) and (t_window > 0.0):
            t0 = float(max(t_guess - 0.5 * t_window, 0.0))
            tmax = float(t0 + t_window)
        else:
            t_start = float(t_guess) if t_guess is not None else float(np.pi / 2.0 - t0_offset)
            if t_start < 0.0:
                t_start = 0.0
            half_span = float(np.pi) * 0.5
            t0 = float(max(t_start - half_span, 0.0))
            tmax = float(t0 + 2.0 * half_span)
        hit = self._cross_event_driven(state0, dynsys=dynsys, surface=surface, t0=t0, tmax=tmax,
                                       forward=forward)
        if hit is None and t_guess is not None:
            span = float(t_window if (t_window is not None and t_window > 0.0) else 0.1)
            while span < float(np.pi):
                span *= 2.0
                t0_try = float(max(t_guess - 0.5 * span, 0.0))
                hit = self._cross_event_driven(state0, dynsys=dynsys, surface=surface, t0=t0_try,
                                               tmax=t0_try + span, forward=forward)
                if hit is not None:
                    break
        if hit is None:
            if t_guess is not None:
                t_start = float(max(t_guess, 0.0))
            else:
                t_start = float(np.pi / 2.0 - t0_offset)
                if t_start < 0.0:
                    t_start = 0.0
            hit = self._cross_event_driven(state0, dynsys=dynsys, surface=surface, t0=t_start,
                                           tmax=t_start + float(np.pi), forward=forward)
        return hit


def find_crossing(dynsys, state0, surface, forward: int = 1, **kwargs):
    """backend.py:284-312."""
    return _SingleHitBackend().run(np.asarray(state0, float), dynsys=dynsys, surface=surface,
                                   forward=forward)


def _plane_crossing_factory(coord: str, value: float = 0.0, direction: int | None = None):
    """backend.py:315-367: returns ``f(*, dynsys, x0, forward=1, t_guess=None,
    t_window=None) -> (t, state) | (None, None)``."""
    if coord.lower() not in _COORD:
        raise ValueError(f"Unsupported plane coord '{coord}'. Must be one of 'x','y','z'.")
    event = _PlaneEvent(coord=coord, value=value, direction=direction)

    def _section_crossing(*, dynsys, x0, forward: int = 1, t_guess: float | None = None,
                          t_window: float | None = None):
        hit = _SingleHitBackend()._cross(np.asarray(x0, float), dynsys=dynsys, surface=event,
                                         t_guess=t_guess, t_window=t_window, forward=forward)
        if hit is None:
            return None, None
        return hit.time, hit.state

    _section_crossing.surface = event
    return _section_crossing


_x_plane_crossing = _plane_crossing_factory("x", 0.0, None)
_y_plane_crossing = _plane_crossing_factory("y", 0.0, None)
_z_plane_crossing = _plane_crossing_factory("z", 0.0, None)


# ---------------------------------------------------------------------------
# batched crossings
# ---------------------------------------------------------------------------

def _event_window(mu, states, t_start, span, surface, forward, align_steps_kw=None):
    """Align each state to t_start (skipped where numpy.isclose(0, t_start)),
    then one event launch over [0, span].  Returns hit (n,), t_rel (n,),
    y (n, 6) as numpy arrays."""
    n = states.shape[0]
    t_start = np.broadcast_to(np.asarray(t_start, float), (n,))
    span = np.broadcast_to(np.asarray(span, float), (n,))
    y = np.array(states, dtype=float, copy=True)
    need = ~np.isclose(0.0, t_start)
    if need.any():
        r = batch.dop853_final(mu[need] if np.ndim(mu) else mu, y[need], 0.0, t_start[need], dim=6,
                               forward=forward, flip_indices=None)
        y[need] = r.y.cpu().numpy()
    ev_dir = 0 if surface.direction is None else int(surface.direction)
    live = ~np.isclose(0.0, span)
    hit = np.zeros(n, bool)
    t_rel = span.copy()
    y_out = y.copy()
    if live.any():
        r = batch.dop853_event(mu[live] if np.ndim(mu) else mu, y[live], 0.0, span[live],
                               ev_index=surface.index, ev_offset=surface.offset, direction=ev_dir,
                               rtol=1e-12, atol=1e-12, max_step=math.inf)
        h = r.hit.cpu().numpy()
        th = r.t_hit.cpu().numpy()
        yh = r.y_hit.cpu().numpy()
        yl = r.y_last.cpu().numpy()
        idx = np.nonzero(live)[0]
        hit[idx] = h
        t_rel[idx] = np.where(h, th, span[idx])
        y_out[idx] = np.where(h[:, None], yh, yl)
    ok = hit & (t_rel < span) & (t_rel >= 0.0)
    return ok, t_start + t_rel, y_out


def plane_crossing_device(mu, states, surface: _PlaneEvent, forward: int, window, fallback,
                          stream=None):
    """_event_window over ``window`` and then ``fallback`` for the states
    without a valid hit, in one hh_plane_crossing launch.  Returns ok (n,),
    t = start + t_rel (n,) and the crossing states (n, 6) as NumPy arrays
    (the same bits as the two-launch form)."""
    import torch
    lib = _lib.load()
    states = np.atleast_2d(np.asarray(states, dtype=float))
    n = states.shape[0]
    if n == 0:
        return np.zeros(0, bool), np.zeros(0), np.zeros((0, 6))
    dev = batch._device()
    mu_d = torch.as_tensor(np.broadcast_to(np.asarray(mu, float), (n,)).copy(), device=dev)
    y0 = torch.as_tensor(np.ascontiguousarray(states.T), device=dev)
    ok = torch.empty(n, dtype=torch.int32, device=dev)
    t = torch.empty(n, dtype=torch.float64, device=dev)
    y = torch.empty((6, n), dtype=torch.float64, device=dev)
    ev_dir = 0 if surface.direction is None else int(surface.direction)
    _lib.check(lib.hh_plane_crossing(n, mu_d.data_ptr(), y0.data_ptr(), int(forward),
                                     surface.index, surface.offset, ev_dir, 1e-12, 1e-12,
                                     float(window[0]), float(window[1]), float(fallback[0]),
                                     float(fallback[1]), ok.data_ptr(), t.data_ptr(),
                                     y.data_ptr(), None, _lib.stream_handle(stream)),
               "hh_plane_crossing")
    return ok.cpu().numpy().astype(bool), t.cpu().numpy(), y.cpu().numpy().T.copy()


def _default_windows(t_guess=None, t0_offset: float = 0.15):
    """(start, span) of the first search window of _SingleHitBackend._cross
    (backend.py:244-256) and of its fallback window (269-279)."""
    t_start = float(t_guess) if t_guess is not None else float(np.pi / 2.0 - t0_offset)
    t_start = max(t_start, 0.0)
    half = float(np.pi) * 0.5
    t0 = float(max(t_start - half, 0.0))
    tmax = t0 + 2.0 * half
    ts = t0 if t0 > 0.0 else 1e-12
    tsf = max(float(np.pi / 2.0 - t0_offset), 0.0)
    ts2 = tsf if tsf > 0.0 else 1e-12
    return (ts, max(0.0, tmax - ts)), (ts2, max(0.0, tsf + float(np.pi) - ts2))


def section_crossing_batch(dynsys, states, *, surface: _PlaneEvent | None = None, t_guess=None,
                           forward: int = 1, t0_offset: float = 0.15):
    """_SingleHitBackend._cross for many seeds without a time window: the first
    window [max(pi/2 - 0.15 - pi/2, 0), +pi] for every seed in one launch, then
    the fallback window [pi/2 - 0.15, +pi] (or [t_guess, +pi]) for the seeds
    without a hit.  Returns a list of _SectionHit or None."""
    surface = surface or _PlaneEvent("y", 0.0, None)
    mu = _system_mu(dynsys)
    states = np.atleast_2d(np.asarray(states, dtype=float))
    n = states.shape[0]
    (ts, span), (ts2, span2) = _default_windows(t_guess, t0_offset)
    if t_guess is None:
        # both windows in one launch, one state per 8-lane group
        ok, t_hit, y_hit = plane_crossing_device(mu, states, surface, forward,
                                                 (ts, span), (ts2, span2))
        hits = [None] * n
        for i in np.nonzero(ok)[0]:
            hits[i] = _SectionHit(float(t_hit[i]), y_hit[i].copy(), y_hit[i, :2].copy(), int(i))
        return hits
    ok, t_hit, y_hit = _event_window(mu, states, ts, span, surface, forward)
    hits: list = [None] * n
    for i in np.nonzero(ok)[0]:
        hits[i] = _SectionHit(float(t_hit[i]), y_hit[i].copy(), y_hit[i, :2].copy(), int(i))
    miss = np.nonzero(~ok)[0]
    if miss.size and t_guess is not None:
        # the widening retries of backend.py:257-267 depend on the guess; run
        # them seed by seed
        be = _SingleHitBackend()
        for i in miss:
            h = be._cross(states[i], dynsys=dynsys, surface=surface, t_guess=t_guess,
                          forward=forward, t0_offset=t0_offset)
            if h is not None:
                h.trajectory_index = int(i)
            hits[i] = h
        return hits
    if miss.size:
        ok2, t2, y2 = _event_window(mu, states[miss], ts2, span2, surface, forward)
        for j, i in enumerate(miss):
            if ok2[j]:
                hits[i] = _SectionHit(float(t2[j]), y2[j].copy(), y2[j, :2].copy(), int(i))
    return hits
