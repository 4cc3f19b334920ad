"""Cases of tests/golden/integrator_golden.npz and the oracle runs that check
them.

The vectors were written by the reference's own code (hiten 0.5.4's
_compute_stm, _propagate_dynsys with and without an event function, the RK8
fixed driver and _y_plane_crossing; see tests/golden/make_integrator_golden.py),
with every DOP853 attempt's (t, h, error norm, accepted) recorded.

``libm_pow`` selects the oracle's arithmetic: 2 evaluates ``**`` through the C
library's pow and dot products as NumPy's BLAS forms them -- the arithmetic
the reference's source has when it runs as plain Python, as it did for these
vectors -- and 0 is numba's lowering, which the HIP kernels share bit for bit.
"""
from __future__ import annotations

import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
PATH = os.path.join(HERE, "golden", "integrator_golden.npz")
MU_EM = 0.012150585609624
RTOL = ATOL = 1e-12
# Agreement in another arithmetic is measured against the case's own rounding
# sensitivity: the largest change a one-ulp move of one initial state
# component makes in the reference's output (stored per case), or the
# tolerance-proportional global error N_accepted * rtol * max(1, |y|),
# whichever is larger, times this factor.
SPREAD_FACTOR = 10.0
# Step sizes in another arithmetic: the error estimate is a difference of
# stage derivatives of size ~1 that cancels to ~1e-12, so ulp-level changes of
# the field move it by up to ~1e-4 relative and the next step by the 1/9th
# power of that; over a run the controller keeps the sequences within a few
# percent of each other (largest seen: 1.4e-2).
H_REL = 0.05

_G = None


def golden():
    global _G
    if _G is None:
        _G = dict(np.load(PATH))
    return _G


def trace(i):
    g = golden()
    off = g["trace_offsets"]
    return g["trace_rows"][off[i]:off[i + 1]]


def bound(spread, n_acc, ref):
    ref = np.asarray(ref, float)
    return SPREAD_FACTOR * max(float(spread), n_acc * RTOL * max(1.0, float(np.nanmax(np.abs(ref)))))


def stm_cases():
    g = golden()
    for k in range(len(g["stm_mu"])):
        st = int(g["stm_steps"][k])
        yield dict(k=k, mu=float(g["stm_mu"][k]), x0=g["stm_x0"][k], tf=float(g["stm_tf"][k]),
                   forward=int(g["stm_forward"][k]), steps=st, out=g["stm_out"][k][:st],
                   times=g["stm_times"][k][:st], trace=trace(int(g["stm_trace"][k])),
                   spread=float(g["stm_spread"][k]))


def arc_cases():
    g = golden()
    for k in range(len(g["arc_x0"])):
        st = int(g["arc_steps"][k])
        yield dict(k=k, mu=MU_EM, x0=g["arc_x0"][k], tf=float(g["arc_tf"][k]),
                   forward=int(g["arc_forward"][k]), steps=st, out=g["arc_out"][k][:st],
                   times=g["arc_times"][k][:st], trace=trace(int(g["arc_trace"][k])),
                   spread=float(g["arc_spread"][k]))


def event_cases():
    g = golden()
    for k in range(len(g["ev_x0"])):
        tr = trace(int(g["ev_trace"][k]))
        yield dict(k=k, mu=MU_EM, x0=g["ev_x0"][k], tmax=float(g["ev_tmax"][k]),
                   t=float(g["ev_time"][k]), state=g["ev_state"][k], trace=tr,
                   hit=bool(g["ev_time"][k] < g["ev_tmax"][k]), spread=float(g["ev_spread"][k]))


def fixed_cases():
    g = golden()
    for k in range(len(g["fixed_x0"])):
        yield dict(k=k, mu=MU_EM, x0=g["fixed_x0"][k], steps=int(g["fixed_steps"]),
                   out=g["fixed_out"][k], spread=float(g["fixed_spread"][k]))


def crossing_cases():
    g = golden()
    for k in range(len(g["cross_mu"])):
        yield dict(k=k, mu=float(g["cross_mu"][k]), x0=g["cross_x0"][k],
                   forward=int(g["cross_forward"][k]), t=float(g["cross_time"][k]),
                   state=g["cross_state"][k], spread=float(g["cross_spread"][k]))


# --------------------------------------------------------------------------
# oracle runs (test infrastructure)
# --------------------------------------------------------------------------

def oracle_stm(O, c, libm_pow):
    """_compute_stm through the oracle: (PHI at the grid, trace, info)."""
    phi0 = np.zeros(42)
    phi0[:36] = np.eye(6).ravel()
    phi0[36:] = c["x0"]
    s = O.System("cr3bp42", mu=c["mu"], fwd=c["forward"], flip=slice(36, 42),
                 libm_pow=libm_pow, trace_cap=200_000)
    out, info = O.dop853(s, phi0, np.linspace(0.0, c["tf"], c["steps"]), rtol=RTOL, atol=ATOL,
                         max_step=1e4)
    return out, s.trace(), info


def oracle_arc(O, c, libm_pow):
    s = O.System("cr3bp6", mu=c["mu"], fwd=c["forward"], libm_pow=libm_pow, trace_cap=200_000)
    out, info = O.dop853(s, c["x0"], np.linspace(0.0, c["tf"], c["steps"]), rtol=RTOL, atol=ATOL,
                         max_step=1e4)
    return out, s.trace(), info


def oracle_event(O, c, libm_pow):
    s = O.System("cr3bp6", mu=c["mu"], libm_pow=libm_pow, trace_cap=200_000)
    hit, t, y, _, info = O.dop853_event(s, c["x0"], 0.0, c["tmax"], 1, 0.0, -1, rtol=RTOL,
                                        atol=ATOL, max_step=1e4, xtol=1e-13, gtol=1e-13)
    return hit, t, y, s.trace(), info


def oracle_fixed(O, c, libm_pow):
    s = O.System("cr3bp6", mu=c["mu"], libm_pow=libm_pow)
    return O.rk_fixed(s, c["x0"], np.linspace(0.0, 2.0 * np.pi, c["steps"]))


def oracle_crossing(O, c, libm_pow):
    return O.plane_crossing(c["mu"], c["x0"], forward=c["forward"], libm_pow=libm_pow)


def counts(tr):
    """(accepted, rejected) of a trace."""
    acc = int(tr[:, 3].sum())
    return acc, tr.shape[0] - acc
