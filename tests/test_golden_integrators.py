"""The oracle against golden vectors written by the reference's own code.

tests/golden/integrator_golden.npz holds hiten 0.5.4's outputs for
_compute_stm (dynamics/rtbp.py:258-340), _propagate_dynsys
(dynamics/base.py:346-458) adaptive, fixed-step RK8 and with a y = 0 event
(rk.py:2375-2549, 531-588, 2678-2803 with 2005-2102) and _y_plane_crossing
(poincare/singlehit/backend.py:164-371), each DOP853 run with its attempt
sequence.

* In the reference's Python arithmetic (libm_pow=2) the oracle reproduces
  every output, every attempt's t, h and error norm, and every decision bit
  for bit: the driver, the step controller (initial step, PI factors,
  clamps, endpoint), the dense output, the event bisection and the section
  windows are the reference's.
* In numba's arithmetic (libm_pow=0, what the HIP kernels compute) the
  accepted/rejected counts, the accept/reject sequence, hit flags and
  crossing outcomes are identical, step sizes agree to H_REL and outputs to
  SPREAD_FACTOR times the case's rounding sensitivity.
* The coefficient tables compiled into kernels and oracle equal the
  reference's arrays bit for bit.
"""
import os
import re

import numpy as np
import pytest

from tests import golden_integrators as GI

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _ids(cases):
    return [f"{c['k']}" for c in cases]


STM = list(GI.stm_cases())
ARC = list(GI.arc_cases())
EVENT = list(GI.event_cases())
FIXED = list(GI.fixed_cases())
CROSS = list(GI.crossing_cases())


def test_fixture_covers_the_path():
    assert len(STM) >= 20 and len(ARC) >= 12 and len(EVENT) >= 24 and len(FIXED) == 8
    assert {c["forward"] for c in STM} == {-1, 1} and {c["forward"] for c in ARC} == {-1, 1}
    assert {c["steps"] for c in STM} == {2, 9}
    assert len({c["mu"] for c in STM}) == 3
    hits = [c["hit"] for c in EVENT]
    assert any(hits) and not all(hits)
    rej = sum(GI.counts(c["trace"])[1] for c in STM + ARC + EVENT)
    assert rej > 300, "the fixtures must exercise rejected steps"
    # the closed-form mu = 0 circular orbits: first window, fallback window, no hit
    r = [c for c in CROSS if c["mu"] == 0.0]
    w = [2 * np.pi / 2 / (c["x0"][0] ** -1.5 - 1.0) for c in r]   # half synodic period
    assert w[0] < np.pi < w[1] < np.pi / 2 - 0.15 + np.pi < w[2]
    assert np.isfinite(r[0]["t"]) and np.isfinite(r[1]["t"]) and np.isnan(r[2]["t"])
    # closed form: the crossing time is 1e-12 + the half period to the tolerance
    for c, wk in zip(r[:2], w[:2]):
        assert abs(c["t"] - wk) < 1e-9


# ---------------------------------------------------------------- bit for bit

@pytest.mark.parametrize("c", STM, ids=_ids(STM))
def test_stm_bitwise_in_reference_arithmetic(oracle, c):
    x, times, phi_T, PHI = oracle.compute_stm(c["mu"], c["x0"], c["tf"], steps=c["steps"],
                                              forward=c["forward"], libm_pow=2)
    assert np.array_equal(PHI, c["out"])
    assert np.array_equal(times, c["times"])
    _, tr, info = GI.oracle_stm(oracle, c, 2)
    assert np.array_equal(tr, c["trace"], equal_nan=True)


@pytest.mark.parametrize("c", ARC, ids=_ids(ARC))
def test_arc_bitwise_in_reference_arithmetic(oracle, c):
    times, states = oracle.propagate(c["mu"], c["x0"], 0.0, c["tf"], forward=c["forward"],
                                     steps=c["steps"], libm_pow=2)
    assert np.array_equal(states, c["out"])
    assert np.array_equal(times, c["times"])
    _, tr, _ = GI.oracle_arc(oracle, c, 2)
    assert np.array_equal(tr, c["trace"], equal_nan=True)


@pytest.mark.parametrize("c", EVENT, ids=_ids(EVENT))
def test_event_bitwise_in_reference_arithmetic(oracle, c):
    hit, t, y, tr, _ = GI.oracle_event(oracle, c, 2)
    assert hit == c["hit"]
    if hit:
        assert t == c["t"] and np.array_equal(y, c["state"])
    # the reference's last attempt (the step that crossed) calls no step
    # factor, so its error norm is not recorded there
    ref = c["trace"].copy()
    got = tr.copy()
    if hit:
        got[-1, 2] = np.nan
    assert np.array_equal(got, ref, equal_nan=True)


@pytest.mark.parametrize("c", FIXED, ids=_ids(FIXED))
def test_fixed_rk8_bitwise_in_reference_arithmetic(oracle, c):
    assert np.array_equal(GI.oracle_fixed(oracle, c, 2), c["out"])


@pytest.mark.parametrize("c", CROSS, ids=_ids(CROSS))
def test_crossing_bitwise_in_reference_arithmetic(oracle, c):
    t, y = GI.oracle_crossing(oracle, c, 2)
    if np.isnan(c["t"]):
        assert t is None and y is None
    else:
        assert t == c["t"] and np.array_equal(y, c["state"])


# ---------------------------------------------------- numba's (the kernels') arithmetic

def _check_steps(tr, ref):
    assert GI.counts(tr) == GI.counts(ref)
    assert np.array_equal(tr[:, 3], ref[:, 3]), "accept/reject sequence differs"
    assert tr[0, 1] == pytest.approx(ref[0, 1], rel=4 * np.finfo(float).eps)   # initial step
    assert np.all(np.abs(tr[:, 1] / ref[:, 1] - 1.0) <= GI.H_REL)
    assert np.array_equal(tr[:, 0] == 0.0, ref[:, 0] == 0.0)


@pytest.mark.parametrize("c", STM, ids=_ids(STM))
def test_stm_numba_arithmetic(oracle, c):
    out, tr, info = GI.oracle_stm(oracle, c, 0)
    _check_steps(tr, c["trace"])
    assert info["status"] == 0
    assert np.abs(out - c["out"]).max() <= GI.bound(c["spread"], info["n_accepted"], c["out"])


@pytest.mark.parametrize("c", ARC, ids=_ids(ARC))
def test_arc_numba_arithmetic(oracle, c):
    out, tr, info = GI.oracle_arc(oracle, c, 0)
    _check_steps(tr, c["trace"])
    assert np.abs(out - c["out"]).max() <= GI.bound(c["spread"], info["n_accepted"], c["out"])


@pytest.mark.parametrize("c", EVENT, ids=_ids(EVENT))
def test_event_numba_arithmetic(oracle, c):
    hit, t, y, tr, info = GI.oracle_event(oracle, c, 0)
    _check_steps(tr, c["trace"])
    assert hit == c["hit"]
    b = GI.bound(c["spread"], info["n_accepted"], c["state"])
    if hit:
        assert abs(t - c["t"]) <= b and np.abs(y - c["state"]).max() <= b


@pytest.mark.parametrize("c", FIXED, ids=_ids(FIXED))
def test_fixed_rk8_numba_arithmetic(oracle, c):
    out = GI.oracle_fixed(oracle, c, 0)
    assert np.abs(out - c["out"]).max() <= GI.bound(c["spread"], c["steps"], c["out"])


@pytest.mark.parametrize("c", CROSS, ids=_ids(CROSS))
def test_crossing_numba_arithmetic(oracle, c):
    t, y = GI.oracle_crossing(oracle, c, 0)
    if np.isnan(c["t"]):
        assert t is None
        return
    b = GI.bound(c["spread"], 200, c["state"])
    assert abs(t - c["t"]) <= b and np.abs(y - c["state"]).max() <= b


# ----------------------------------------------------------------- tables

def _c_array(text, name):
    m = re.search(r"HH_TABLE double " + name + r"((?:\[\d+\])+) = \{(.*?)\};", text, re.S)
    assert m, name
    shape = tuple(int(d) for d in re.findall(r"\[(\d+)\]", m.group(1)))
    vals = [float.fromhex(v) for v in re.findall(r"-?0x[0-9a-fA-F.]+p[+-]\d+", m.group(2))]
    return np.array(vals).reshape(shape)


def test_tables_equal_the_reference_coefficients():
    g = GI.golden()
    text = open(os.path.join(ROOT, "hiten_amd", "csrc", "hh_tables.h")).read()
    assert np.array_equal(_c_array(text, "HH_RK8_A"), g["tab_rk8_A"])
    assert np.array_equal(_c_array(text, "HH_RK8_B"), g["tab_rk8_B"])
    assert np.array_equal(_c_array(text, "HH_RK8_C"), g["tab_rk8_C"])
    a = _c_array(text, "HH_DOP_A")
    ra = g["tab_dop_A"]
    assert np.array_equal(a[:ra.shape[0], :ra.shape[1]], ra)
    for name, key in (("HH_DOP_B", "tab_dop_B"), ("HH_DOP_E3", "tab_dop_E3"),
                      ("HH_DOP_E5", "tab_dop_E5")):
        assert np.array_equal(_c_array(text, name), g[key]), name
    c = _c_array(text, "HH_DOP_C")
    assert np.array_equal(c[:g["tab_dop_C"].size], g["tab_dop_C"])
    assert np.array_equal(_c_array(text, "HH_DOP_D"), g["tab_dop_D"])
