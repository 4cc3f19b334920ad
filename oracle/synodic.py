"""Oracle: section detection on stored trajectories (synodic Poincare maps).

Test infrastructure only (the checker of hiten_amd's HIP kernel).  A
restatement of the reference's detection on precomputed (time, state)
samples, in the reference's operation order:

  _SynodicDetectionBackend.detect_on_trajectory
      src/hiten/algorithms/poincare/synodic/backend.py:644-784
  _on_surface_indices            backend.py:134-180
  _crossing_indices_and_alpha    backend.py:183-227
  _refine_hits_linear            backend.py:230-265
  _refine_hits_cubic             backend.py:268-366
  _order_and_dedup_hits          backend.py:369-440
  _detect_with_segment_refine    backend.py:443-621
  _hermite_scalar / _hermite_der src/hiten/algorithms/poincare/utils.py:54-142

Arithmetic notes (what fixes each rounding):
  * the plane value g = states @ normal - offset is a BLAS dot product in the
    reference (order implementation-defined); here it is summed left to
    right, which every order matches when the normal has one non-zero
    entry (coordinate planes, the pipeline's default);
  * _hermite_scalar / _hermite_der are numba functions: x ** 2 is x * x;
  * the state interpolation is Python-level: (1 - s) ** 2 and s ** 2 are
    CPython float powers, i.e. the C library's pow(x, 2.0).

How it is pinned (tests/test_synodic_cpu.py): closed-form crossings of
linear and quadratic plane values (where the linear and the cubic Hermite
interpolants are exact), direction filters, on-surface samples, several
crossings inside one segment (segment_refine), deduplication and the hit
cap, against values worked out by hand.
"""
from __future__ import annotations

import numpy as np


def plane_values(states: np.ndarray, normal, offset: float) -> np.ndarray:
    n = [float(v) for v in normal]
    out = np.empty(states.shape[0])
    for i in range(states.shape[0]):
        acc = 0.0
        first = True
        for j in range(6):
            term = float(states[i, j]) * n[j]
            acc = term if first else acc + term
            first = False
        out[i] = acc - float(offset)
    return out


def hermite_scalar(s, y0, y1, dy0, dy1, dt):
    oms = 1.0 - s
    h00 = (1.0 + 2.0 * s) * (oms * oms)
    h10 = s * (oms * oms)
    h01 = (s * s) * (3.0 - 2.0 * s)
    h11 = (s * s) * (s - 1.0)
    return h00 * y0 + h10 * dy0 * dt + h01 * y1 + h11 * dy1 * dt


def hermite_der(s, y0, y1, dy0, dy1, dt):
    oms = 1.0 - s
    dh00 = 6.0 * s * (s - 1.0) + (oms * oms) * 2.0 - 2.0 * oms * (1.0 + 2.0 * s)
    dh10 = (oms * oms) + s * (2.0 * (s - 1.0))
    dh01 = 6.0 * s * oms - 2.0 * s * (3.0 - 2.0 * s)
    dh11 = 2.0 * s * (s - 1.0) + (s * s)
    return dh00 * y0 + dh10 * dy0 * dt + dh01 * y1 + dh11 * dy1 * dt


def _slopes(times, g, k, N):
    if k - 1 >= 0:
        d0 = (g[k + 1] - g[k - 1]) / (times[k + 1] - times[k - 1])
    else:
        d0 = (g[k + 1] - g[k]) / (times[k + 1] - times[k])
    if k + 2 < N:
        d1 = (g[k + 2] - g[k]) / (times[k + 2] - times[k])
    else:
        d1 = (g[k + 1] - g[k]) / (times[k + 1] - times[k])
    return float(d0), float(d1)


def _state_at(times, states, k, s, dt, use_cubic, N):
    if use_cubic and dt > 0.0 and k - 1 >= 0 and k + 2 < N:
        dxdt0 = (states[k + 1] - states[k - 1]) / (times[k + 1] - times[k - 1])
        dxdt1 = (states[k + 2] - states[k]) / (times[k + 2] - times[k])
        h00 = (1.0 + 2.0 * s) * (1.0 - s) ** 2
        h10 = s * (1.0 - s) ** 2
        h01 = s ** 2 * (3.0 - 2.0 * s)
        h11 = s ** 2 * (s - 1.0)
        return h00 * states[k] + h10 * dxdt0 * dt + h01 * states[k + 1] + h11 * dxdt1 * dt
    return states[k] + s * (states[k + 1] - states[k])


def _on_surface_ok(g, k, tol, direction):
    if abs(g[k]) >= tol:
        return False
    if direction is None:
        return True
    nxt = g[k + 1]
    if direction == 1:
        return bool(nxt >= 0.0 or (k - 1 >= 0 and g[k - 1] <= 0.0))
    return bool(nxt <= 0.0 or (k - 1 >= 0 and g[k - 1] >= 0.0))


def _crosses(a, b, direction):
    if direction is None:
        return (a * b <= 0.0) and (a != b)
    if direction == 1:
        return (a < 0.0) and (b >= 0.0)
    return (a > 0.0) and (b <= 0.0)


def _dedup(cands, plane_idx, dedup_time_tol, dedup_point_tol, max_hits):
    """cands: (segment, time, state) in emission order; stable by segment."""
    cands = sorted(cands, key=lambda c: c[0])
    i, j = plane_idx
    hits = []
    for _, th, st in cands:
        pt = (float(st[i]), float(st[j]))
        if hits:
            pt_prev = hits[-1][2]
            if abs(th - hits[-1][0]) <= dedup_time_tol:
                continue
            du = pt[0] - pt_prev[0]
            dv = pt[1] - pt_prev[1]
            if (du * du + dv * dv) <= (dedup_point_tol * dedup_point_tol):
                continue
        hits.append((float(th), np.array(st, dtype=float), pt))
        if max_hits is not None and len(hits) >= max_hits:
            break
    return hits


def detect(times, states, *, normal, offset=0.0, plane_idx=(1, 4), interp_kind="linear",
           segment_refine=0, tol_on_surface=1e-12, dedup_time_tol=1e-9,
           dedup_point_tol=1e-12, max_hits=None, newton_max_iter=4, direction=None):
    """Hits (time, state (6,), point2d) of one trajectory, as the reference's
    detect_on_trajectory returns them."""
    times = np.asarray(times, dtype=float)
    states = np.asarray(states, dtype=float)
    N = times.shape[0]
    if N < 2:
        return []
    g = plane_values(states, normal, offset)
    use_cubic = interp_kind == "cubic"
    cands = []
    r = int(segment_refine)
    if r > 0:
        step = 1.0 / (r + 1)
        for k in range(N - 1):
            t0, t1 = float(times[k]), float(times[k + 1])
            dt = t1 - t0
            gk, gk1 = float(g[k]), float(g[k + 1])
            if use_cubic and dt > 0.0:
                d0, d1 = _slopes(times, g, k, N)
            accept_left = _on_surface_ok(g, k, tol_on_surface, direction)
            if accept_left:
                cands.append((k, t0, states[k].astype(float, copy=True)))
            for m in range(r + 1):
                s_lo = m * step
                s_hi = (m + 1) * step
                if s_hi > 1.0 + 1e-15:
                    break
                if accept_left and m == 0:
                    continue
                if use_cubic and dt > 0.0:
                    g_lo = hermite_scalar(s_lo, gk, gk1, d0, d1, dt)
                    g_hi = hermite_scalar(s_hi, gk, gk1, d0, d1, dt)
                else:
                    g_lo = (1.0 - s_lo) * gk + s_lo * gk1
                    g_hi = (1.0 - s_hi) * gk + s_hi * gk1
                if not _crosses(g_lo, g_hi, direction):
                    continue
                if g_lo == g_hi:
                    s_star = 0.5 * (s_lo + s_hi)
                else:
                    a = g_lo / (g_lo - g_hi)
                    a = min(1.0, max(0.0, a))
                    s_star = s_lo + a * (s_hi - s_lo)
                if use_cubic and dt > 0.0:
                    for _ in range(newton_max_iter):
                        f = hermite_scalar(s_star, gk, gk1, d0, d1, dt)
                        df = hermite_der(s_star, gk, gk1, d0, d1, dt)
                        if df == 0.0:
                            break
                        s_star -= f / df
                        if s_star < s_lo:
                            s_star = s_lo
                            break
                        if s_star > s_hi:
                            s_star = s_hi
                            break
                th = (1.0 - s_star) * t0 + s_star * t1
                cands.append((k, th, _state_at(times, states, k, s_star, dt, use_cubic, N)))
        return _dedup(cands, plane_idx, dedup_time_tol, dedup_point_tol, max_hits)

    for k in range(N - 1):
        on = _on_surface_ok(g, k, tol_on_surface, direction)
        if on:
            cands.append((k, float(times[k]), states[k].astype(float, copy=True)))
            continue
        g0, g1 = float(g[k]), float(g[k + 1])
        if not _crosses(g0, g1, direction):
            continue
        # numpy.minimum / numpy.maximum over the crossing array (NaN propagates)
        with np.errstate(all="ignore"):
            alpha = float(np.minimum(1.0, np.maximum(0.0, np.float64(g0) / np.float64(g0 - g1))))
        t0, t1 = float(times[k]), float(times[k + 1])
        if not use_cubic:
            th = (1.0 - alpha) * t0 + alpha * t1
            cands.append((k, th, states[k] + alpha * (states[k + 1] - states[k])))
            continue
        dt = t1 - t0
        s_star = alpha
        if dt > 0.0:
            d0, d1 = _slopes(times, g, k, N)
            for _ in range(newton_max_iter):
                f = hermite_scalar(s_star, float(g[k]), float(g[k + 1]), d0, d1, dt)
                df = hermite_der(s_star, float(g[k]), float(g[k + 1]), d0, d1, dt)
                if df == 0.0:
                    break
                s_star -= f / df
                if s_star < 0.0:
                    s_star = 0.0
                    break
                if s_star > 1.0:
                    s_star = 1.0
                    break
        th = (1.0 - s_star) * t0 + s_star * t1
        cands.append((k, th, _state_at(times, states, k, s_star, dt, True, N)))
    return _dedup(cands, plane_idx, dedup_time_tol, dedup_point_tol, max_hits)
