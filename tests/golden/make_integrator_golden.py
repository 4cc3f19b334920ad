"""Golden vectors for the adaptive and fixed-step propagation path, written by
the reference's own code.

Runs hiten 0.5.4's pure-Python sources (loaded by ``reference_loader``:
dynamics/rtbp.py, dynamics/base.py, integrators/rk.py, integrators/utils.py,
integrators/coefficients/*.py, poincare/singlehit/backend.py) in this
container on seeded CR3BP inputs and stores inputs, outputs and each adaptive
run's attempt sequence in ``integrator_golden.npz`` next to this script.

Why: the DOP853 step-size sequence (initial step, PI controller, clamps,
endpoint adjustment), the event driver with its in-step bisection and the
section-crossing windows of ``_SingleHitBackend._cross`` are fixed only by
the reference's code; no closed form pins them.

Cases (all through the reference's public seams):

* ``stm``      ``_compute_stm`` (dynamics/rtbp.py:258-340): 42-d state + STM,
               DOP853 rtol = atol = 1e-12, final value (steps = 2) or a
               9-point grid, forward and backward (flip of 36:42).
* ``arc``      ``_propagate_dynsys`` (dynamics/base.py:346-458), 6-d, adaptive,
               forward and backward, final value or 9-point grid.
* ``event``    ``_propagate_dynsys`` with an event function (rk.py:2221-2319 ->
               _integrate_dop853_until_event rk.py:2678-2803): first y = 0
               crossing with vy < 0, xtol = gtol = 1e-13, manifold-fan seeds.
* ``fixed``    ``_propagate_dynsys(method="fixed", order=8)`` (rk.py:531-588):
               200 grid points to 2*pi, every state stored.
* ``crossing`` ``_y_plane_crossing`` (poincare/singlehit/backend.py:315-371):
               CR3BP seeds forward and backward, and mu = 0 circular orbits
               whose crossing lies in the first window (r = 0.5), only in the
               fallback window (r = 0.68) or in neither (r = 0.72).
* ``tables``   the coefficient arrays of integrators/coefficients/rk8.py and
               dop853.py as the reference builds them.

Attempt traces: ``dop853_step_jit_kernel`` and the two step-factor functions
are wrapped in the loaded rk module's namespace (the wrapper calls the
reference function and returns its result unchanged), so each attempt's
(t, h) and each decision's error norm are recorded as the reference's
driver produces them.

The arithmetic is the reference's as Python runs it: ``**`` on NumPy scalars
is the C library's pow (numba multiplies for integer powers) and ``np.dot`` is
NumPy's OpenBLAS ddot.  The oracle's ``libm_pow=2`` mode evaluates exactly
that arithmetic and must reproduce every vector here bit for bit; in its
default mode (numba's lowering, the arithmetic the HIP kernels share) the
step counts and accept/reject sequences must be equal and the states agree
to the integration's own tolerance.

    python tests/golden/make_integrator_golden.py
"""
from __future__ import annotations

import math
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from tests.golden.reference_loader import load  # noqa: E402
from hiten_amd import workloads as W  # noqa: E402

MU_EM = W.MU_EM
OUT = os.path.join(HERE, "integrator_golden.npz")


class Tracer:
    """Records (t, h, err, accepted) per attempt of the reference's drivers."""

    def __init__(self, rk):
        self.rows = []
        self.active = False
        self._pending = None
        step, acc, rej = rk.dop853_step_jit_kernel, rk._pi_accept_factor, rk._pi_reject_factor

        def step_w(f, t, y, h, *a):
            if self.active:
                self._flush(np.nan, 1.0)
                self._pending = (float(t), float(h))
            return step(f, t, y, h, *a)

        def acc_w(err, err_prev, order):
            if self.active and self._pending is not None:
                self._flush(float(err), 1.0)
            return acc(err, err_prev, order)

        def rej_w(err, order):
            if self.active and self._pending is not None:
                self._flush(float(err), 0.0)
            return rej(err, order)

        rk.dop853_step_jit_kernel = step_w
        rk._pi_accept_factor = acc_w
        rk._pi_reject_factor = rej_w

    def _flush(self, err, acc):
        if self._pending is not None:
            self.rows.append((*self._pending, err, acc))
            self._pending = None

    def run(self, fn):
        self.rows, self._pending, self.active = [], None, True
        try:
            out = fn()
        finally:
            # an attempt with no factor call after it is the accepted step that
            # ended the run (the event driver returns before the factor)
            self._flush(np.nan, 1.0)
            self.active = False
        return out, np.asarray(self.rows, float).reshape(-1, 4)


def ulp_spread(run, x0, ref):
    """Largest change of ``run``'s output when one of the six state components
    of ``x0`` (the last six entries) moves up by one ulp: the rounding
    sensitivity of the case, against which agreement in another arithmetic
    is measured."""
    x0 = np.asarray(x0, float)
    worst = 0.0
    for c in range(x0.size - 6, x0.size):
        xp = x0.copy()
        xp[c] = np.nextafter(xp[c], np.inf)
        out = np.asarray(run(xp), float)
        d = np.abs(out - ref)
        d = d[np.isfinite(d)]
        worst = max(worst, float(d.max()) if d.size else 0.0)
    return worst


def main():
    t_start = time.time()
    m = load(singlehit=True)
    rtbp, dbase, rk = m["rtbp"], m["dbase"], m["rk"]
    EventConfig, EventOptions = m["configs"].EventConfig, m["options"].EventOptions
    tracer = Tracer(rk)
    A = {}
    traces = []

    def add_trace(tr):
        traces.append(tr)
        return len(traces) - 1

    # --- stm: C1-style (Earth-Moon) and C3-style (other mass parameters) ----
    w1 = W.stm_uneven(n=16)
    w3 = W.persistent_262k(n=4096)
    pick3 = [i for i in range(4096) if w3["mu"][i] != MU_EM and w3["tf"][i] < 6.0][:6]
    stm_mu, stm_x0, stm_tf, stm_fwd, stm_steps = [], [], [], [], []
    for i in range(16):
        stm_mu.append(MU_EM)
        stm_x0.append(w1["y0"][i, 36:])
        stm_tf.append(w1["tf"][i] * (0.5 if i % 4 == 3 else 1.0))
        stm_fwd.append(-1 if i in (2, 7, 10, 13) else 1)
        stm_steps.append(9 if i in (1, 5, 7, 11, 14) else 2)
    for i in pick3:
        stm_mu.append(w3["mu"][i])
        stm_x0.append(w3["y0"][i, 36:])
        stm_tf.append(w3["tf"][i])
        stm_fwd.append(1)
        stm_steps.append(2)
    stm_out, stm_times, stm_tr, stm_spread = [], [], [], []
    for k in range(len(stm_mu)):
        vs = rtbp.variational_dynsys(float(stm_mu[k]))
        (x, times, phiT, PHI), tr = tracer.run(lambda: rtbp._compute_stm(
            vs, np.asarray(stm_x0[k], float), float(stm_tf[k]), steps=int(stm_steps[k]),
            forward=int(stm_fwd[k])))
        out = np.full((9, 42), np.nan)
        out[:PHI.shape[0]] = PHI
        tt = np.full(9, np.nan)
        tt[:times.shape[0]] = times
        stm_out.append(out)
        stm_times.append(tt)
        stm_tr.append(add_trace(tr))
        stm_spread.append(ulp_spread(lambda xp: rtbp._compute_stm(
            vs, xp, float(stm_tf[k]), steps=int(stm_steps[k]), forward=int(stm_fwd[k]))[3],
            stm_x0[k], PHI))
    A.update(stm_mu=np.asarray(stm_mu), stm_x0=np.asarray(stm_x0), stm_tf=np.asarray(stm_tf),
             stm_forward=np.asarray(stm_fwd, np.int64), stm_steps=np.asarray(stm_steps, np.int64),
             stm_out=np.asarray(stm_out), stm_times=np.asarray(stm_times),
             stm_trace=np.asarray(stm_tr, np.int64), stm_spread=np.asarray(stm_spread))

    # --- arc: 6-d adaptive propagation --------------------------------------
    rng = np.random.default_rng(W.SEED + 10)
    arc_x0 = W.planar_seeds(rng, 12)
    arc_tf = rng.uniform(1.0, 4.0, 12)
    arc_fwd = np.array([1, -1, 1, 1, -1, 1, 1, -1, 1, -1, 1, 1], np.int64)
    arc_steps = np.array([2, 2, 9, 2, 9, 2, 2, 2, 9, 2, 2, 9], np.int64)
    arc_out, arc_times, arc_tr, arc_spread = [], [], [], []
    ds = rtbp.rtbp_dynsys(MU_EM)
    for k in range(len(arc_x0)):
        sol, tr = tracer.run(lambda: dbase._propagate_dynsys(
            ds, arc_x0[k], 0.0, float(arc_tf[k]), forward=int(arc_fwd[k]),
            steps=int(arc_steps[k]), method="adaptive", order=8))
        out = np.full((9, 6), np.nan)
        out[:sol.states.shape[0]] = sol.states
        tt = np.full(9, np.nan)
        tt[:sol.times.shape[0]] = sol.times
        arc_out.append(out)
        arc_times.append(tt)
        arc_tr.append(add_trace(tr))
        arc_spread.append(ulp_spread(lambda xp: dbase._propagate_dynsys(
            ds, xp, 0.0, float(arc_tf[k]), forward=int(arc_fwd[k]), steps=int(arc_steps[k]),
            method="adaptive", order=8).states, arc_x0[k], sol.states))
    A.update(arc_x0=arc_x0, arc_tf=arc_tf, arc_forward=arc_fwd, arc_steps=arc_steps,
             arc_out=np.asarray(arc_out), arc_times=np.asarray(arc_times),
             arc_trace=np.asarray(arc_tr, np.int64), arc_spread=np.asarray(arc_spread))

    # --- event: manifold-fan seeds, first y = 0 crossing with vy < 0 --------
    fan = W.manifold_fan()
    rows = [s * 512 + (37 * s) % 512 for s in range(24)]
    ev_x0 = fan["y0"][rows]
    ev_tmax = fan["tmax"][rows]

    def g_y(t, y):  # the plane y = 0 of the fan configuration
        return y[1]

    ev_out, ev_t, ev_tr, ev_spread = [], [], [], []
    for k in range(len(rows)):
        sol, tr = tracer.run(lambda: dbase._propagate_dynsys(
            ds, ev_x0[k], 0.0, float(ev_tmax[k]), steps=2, method="adaptive", order=8,
            event_fn=g_y, event_cfg=EventConfig(direction=-1, terminal=True),
            event_options=EventOptions(xtol=1e-13, gtol=1e-13)))
        ev_out.append(sol.states[-1])
        ev_t.append(sol.times[-1])
        ev_tr.append(add_trace(tr))

        def ev_run(xp, k=k):
            so = dbase._propagate_dynsys(
                ds, xp, 0.0, float(ev_tmax[k]), steps=2, method="adaptive", order=8,
                event_fn=g_y, event_cfg=EventConfig(direction=-1, terminal=True),
                event_options=EventOptions(xtol=1e-13, gtol=1e-13))
            return np.concatenate([[so.times[-1]], so.states[-1]])
        ev_spread.append(ulp_spread(ev_run, ev_x0[k], np.concatenate([[sol.times[-1]],
                                                                       sol.states[-1]])))
    A.update(ev_rows=np.asarray(rows, np.int64), ev_x0=ev_x0, ev_tmax=ev_tmax,
             ev_state=np.asarray(ev_out), ev_time=np.asarray(ev_t),
             ev_trace=np.asarray(ev_tr, np.int64), ev_spread=np.asarray(ev_spread))

    # --- fixed: RK8 over 200 grid points to 2*pi ----------------------------
    c0 = W.rk8_grid(n=8)
    fx_out, fx_spread = [], []
    for k in range(8):
        sol = dbase._propagate_dynsys(ds, c0["y0"][k], 0.0, 2.0 * math.pi, steps=200,
                                      method="fixed", order=8)
        fx_out.append(sol.states)
        fx_spread.append(ulp_spread(lambda xp: dbase._propagate_dynsys(
            ds, xp, 0.0, 2.0 * math.pi, steps=200, method="fixed", order=8).states,
            c0["y0"][k], sol.states))
    A.update(fixed_x0=c0["y0"], fixed_out=np.asarray(fx_out), fixed_steps=np.int64(200),
             fixed_spread=np.asarray(fx_spread))

    # --- crossing: _y_plane_crossing windows --------------------------------
    cr_mu, cr_x0, cr_fwd = [], [], []
    rng = np.random.default_rng(W.SEED + 11)
    for k, s in enumerate(W.planar_seeds(rng, 8)):
        cr_mu.append(MU_EM)
        cr_x0.append(s)
        cr_fwd.append(-1 if k in (1, 4, 6) else 1)
    for r in (0.5, 0.68, 0.72):
        cr_mu.append(0.0)
        cr_x0.append(np.array([r, 0.0, 0.0, 0.0, r * (r ** -1.5 - 1.0), 0.0]))
        cr_fwd.append(1)
    cr_t, cr_state, cr_spread = [], [], []
    for k in range(len(cr_mu)):
        sysk = rtbp.rtbp_dynsys(float(cr_mu[k]))
        t_hit, st = m["singlehit"]._y_plane_crossing(dynsys=sysk, x0=cr_x0[k],
                                                     forward=int(cr_fwd[k]))
        cr_t.append(np.nan if t_hit is None else float(t_hit))
        cr_state.append(np.full(6, np.nan) if st is None else np.asarray(st, float))

        def cr_run(xp, sysk=sysk, k=k):
            th, sx = m["singlehit"]._y_plane_crossing(dynsys=sysk, x0=xp, forward=int(cr_fwd[k]))
            return np.full(7, np.nan) if th is None else np.concatenate([[th], sx])
        cr_spread.append(ulp_spread(cr_run, cr_x0[k], np.concatenate([[cr_t[-1]], cr_state[-1]])))
    A.update(cross_mu=np.asarray(cr_mu), cross_x0=np.asarray(cr_x0),
             cross_forward=np.asarray(cr_fwd, np.int64), cross_time=np.asarray(cr_t),
             cross_state=np.asarray(cr_state), cross_spread=np.asarray(cr_spread))

    # --- tables --------------------------------------------------------------
    rk8, dop = m["rk8"], m["dop853"]
    A.update(tab_rk8_A=np.asarray(rk8.A, float), tab_rk8_B=np.asarray(rk8.B, float),
             tab_rk8_C=np.asarray(rk8.C, float), tab_dop_A=np.asarray(dop.A, float),
             tab_dop_B=np.asarray(dop.B, float), tab_dop_C=np.asarray(dop.C, float),
             tab_dop_E3=np.asarray(dop.E3, float), tab_dop_E5=np.asarray(dop.E5, float),
             tab_dop_D=np.asarray(dop.D, float))

    # traces, concatenated
    off = np.zeros(len(traces) + 1, np.int64)
    off[1:] = np.cumsum([t.shape[0] for t in traces])
    A["trace_offsets"] = off
    A["trace_rows"] = np.concatenate(traces, axis=0)
    np.savez_compressed(OUT, **A)
    print(f"{OUT}: {len(stm_mu)} stm, {len(arc_x0)} arc, {len(rows)} event, 8 fixed, "
          f"{len(cr_mu)} crossing cases, "
          f"{int(off[-1])} traced attempts ({time.time() - t_start:.1f} s)")


if __name__ == "__main__":
    main()
